#include "SelectQueryJob.h"

#include <chrono>
#include <iostream>
#include <limits>

#include "QueryExecutor.h"

SelectQueryJob::SelectQueryJob(IndexRef index, std::string v1, std::string v2, TYPEID id1,
                               VariableBoundMap *variables_bound, EmptyIntervalDict *ei_dict,
                               bool *is_predicates, bool is_second_var_used,
                               VedasContext *context)
    : index(index), variable_num(2), is_second_var_used(is_second_var_used),
      variables_bound(variables_bound), ei_dict(ei_dict), context(context) {
    variables[0] = v1;
    variables[1] = v2;
    dataid[0] = id1;
    this->is_predicates[0] = is_predicates[0];
    this->is_predicates[1] = is_predicates[1];
}

SelectQueryJob::SelectQueryJob(IndexRef index, std::string v1, TYPEID id1, TYPEID id2,
                               VariableBoundMap *variables_bound, EmptyIntervalDict *ei_dict,
                               bool *is_predicates, VedasContext *context)
    : index(index), variable_num(1), variables_bound(variables_bound), ei_dict(ei_dict),
      context(context) {
    variables[0] = v1;
    dataid[0] = id1;
    dataid[1] = id2;
    this->is_predicates[0] = is_predicates[0];
}

SelectQueryJob::~SelectQueryJob() { delete intermediateResult; }

// SelectQueryJob::startJob, SelectQueryJob.cu:86-270.
int SelectQueryJob::startJob(int gpuId) {
    this->gpuId = gpuId;
    ContextGuard guard(context);
    auto t0 = std::chrono::steady_clock::now();
    vd_ctx *c = context->raw();
    vd_index *idx = index.storage->index();
    uint64_t b = 0, e = 0;
    vd_rel *rel = nullptr;
    FullRelationIR *ir = nullptr;
    // The key lookup and the select run as one device step (vd_select_key),
    // or, with the range resolved ahead of time, the select alone
    // (vd_select: no device work for an unfiltered range, which is a view).
    auto select = [&](int cols, int clip, TYPEID lo, TYPEID hi, const vd_interval_t *skip,
                      int nskip, int distinct) {
        if (has_range) {
            b = range_begin;
            e = range_end;
            context->check(vd_select(c, idx, index.perm, b, e, cols, clip, lo, hi, skip, nskip,
                                     distinct, &rel),
                           "vd_select");
        } else {
            context->check(vd_select_key(c, idx, index.perm, dataid[0], variable_num == 1,
                                         variable_num == 1 ? dataid[1] : 0, cols, clip, lo, hi,
                                         skip, nskip, distinct, &rel, &b, &e),
                           "vd_select_key");
        }
    };
    if (variable_num == 1) {
        select(1, 0, 0, 0, nullptr, 0, 0);
        ir = new FullRelationIR(context, rel);
        ir->setHeader(0, variables[0], is_predicates[0]);
        context->exe_log.emplace_back(gpuId, UPLOAD_OP, ir->getHeaders(""), ir->size(), 1);
        // Bound tightening happens only for an empty result: the data column
        // is device resident (the reference's preload case, :115-119).
        if (variables_bound && variables_bound->count(variables[0]) > 0 && ir->size() == 0)
            QueryExecutor::updateBoundDict(variables_bound, variables[0], 0, 0);
    } else {
        int clip = 0;
        TYPEID lo = 0, hi = 0;
        if (variables_bound && variables_bound->count(variables[0]) > 0) {
            auto bound = (*variables_bound)[variables[0]];
            lo = bound.first;
            hi = bound.second;
            // A bound still covering every ID (a join variable before its
            // first join) removes no row: the range stays a view.
            clip = (lo == std::numeric_limits<TYPEID>::min() &&
                    hi == std::numeric_limits<TYPEID>::max()) ? 0 : 1;
        }
        if (clip) {
            clip_lo = lo;
            clip_hi = hi;
        }
        if (!is_second_var_used) {
            select(2, clip, lo, hi, nullptr, 0, 1);
            ir = new FullRelationIR(context, rel);
            ir->setHeader(0, variables[0], is_predicates[0]);
        } else {
            std::vector<vd_interval_t> skip;
            if (ei_dict) skip = ei_dict->skipIntervals(variables[0]);
            select(3, clip, lo, hi, skip.empty() ? nullptr : skip.data(), (int)skip.size(), 0);
            alt_ok = skip.empty();
            ir = new FullRelationIR(context, rel);
            ir->setHeader(0, variables[0], is_predicates[0]);
            ir->setHeader(1, variables[1], is_predicates[1]);
        }
        context->exe_log.emplace_back(gpuId, UPLOAD_OP, ir->getHeaders(""), ir->size(),
                                      ir->getColumnNum());
    }
    sel_begin = b;
    sel_end = e;
    intermediateResult = ir;
    context->select_rows += ir->size();
    const long long ns = std::chrono::duration_cast<std::chrono::nanoseconds>(
        std::chrono::steady_clock::now() - t0).count();
    context->upload_ns += ns;
    if (planTreeNode) {
        planTreeNode->resultSize = ir->size();
        planTreeNode->nanosecTime = ns;
    }
    return 0;
}

vd_select_job_t SelectQueryJob::lookupJob() const {
    vd_select_job_t j{};
    j.perm = index.perm;
    j.k1 = dataid[0];
    j.k2 = variable_num == 1 ? dataid[1] : VD_UNBOUND;
    j.cols = variable_num == 1 ? 1 : 3;
    return j;
}

vd_index *SelectQueryJob::indexHandle() const { return index.storage->index(); }

vd_rel *SelectQueryJob::alternateView(TYPEID *lo, TYPEID *hi) const {
    if (!alt_ok) return nullptr;
    static const int other[6] = {VD_PERM_SOP, VD_PERM_SPO, VD_PERM_POS,
                                 VD_PERM_PSO, VD_PERM_OPS, VD_PERM_OSP};
    const int alt = other[index.perm];
    if (!index.storage->hasPermutation(alt)) return nullptr;
    // Both permutations sort on the same X first: the constant's rows are
    // the same row range in either.
    vd_rel *v = nullptr;
    context->check(vd_select(context->raw(), index.storage->index(), alt, sel_begin, sel_end, 3,
                             0, 0, 0, nullptr, 0, 0, &v),
                   "vd_select");
    *lo = clip_lo;
    *hi = clip_hi;
    return v;
}

void SelectQueryJob::setRange(uint64_t begin, uint64_t end) {
    has_range = true;
    range_begin = begin;
    range_end = end;
}

IR *SelectQueryJob::getIR() { return intermediateResult; }

IR *SelectQueryJob::releaseIR() {
    IR *r = intermediateResult;
    intermediateResult = nullptr;
    return r;
}

std::string SelectQueryJob::getVariable(size_t i) const { return variables[i]; }
size_t SelectQueryJob::getVariableNum() const { return variable_num; }
TYPEID SelectQueryJob::getId(size_t i) const { return dataid[i]; }

void SelectQueryJob::print() const {
    std::cout << "\tSELECT JOB - select ";
    if (variable_num == 2)
        std::cout << "[" << variables[0] << "," << variables[1] << "] index " << dataid[0]
                  << " perm " << index.perm << "\n";
    else
        std::cout << "[" << variables[0] << "] index " << dataid[0] << " and " << dataid[1]
                  << " perm " << index.perm << "\n";
}

std::string SelectQueryJob::jobTypeName() const { return "Upload Job"; }
