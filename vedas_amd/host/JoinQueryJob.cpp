#include "JoinQueryJob.h"

#include <algorithm>
#include <chrono>
#include <iostream>
#include <vector>

#include "QueryExecutor.h"
#include "SelectQueryJob.h"

JoinQueryJob::JoinQueryJob(QueryJob *leftJob, QueryJob *rightJob, std::string joinVariable,
                           VedasContext *context, VariableBoundMap *variables_bound,
                           EmptyIntervalDict *ei_dict, bool lastJoinForVar)
    : variables_bound(variables_bound), ei_dict(ei_dict), joinVariable(joinVariable),
      context(context), leftJob(leftJob), rightJob(rightJob), lastJoinForVar(lastJoinForVar) {}

JoinQueryJob::~JoinQueryJob() { delete intermediateResult; }

IR *JoinQueryJob::getIR() { return intermediateResult; }

IR *JoinQueryJob::releaseIR() {
    IR *r = intermediateResult;
    intermediateResult = nullptr;
    return r;
}

int JoinQueryJob::startJob(int gpuId) {
    this->gpuId = gpuId;
    ContextGuard guard(context);
    auto t0 = std::chrono::steady_clock::now();
    auto *l = dynamic_cast<FullRelationIR *>(leftJob->getIR());
    auto *r = dynamic_cast<FullRelationIR *>(rightJob->getIR());
    if (!l || !r) throw VedasError(VD_EINVAL, "join operand has no relation");
    intermediateResult = join(l, r);
    const long long ns = std::chrono::duration_cast<std::chrono::nanoseconds>(
        std::chrono::steady_clock::now() - t0).count();
    context->join_ns += ns;
    if (planTreeNode) {
        planTreeNode->resultSize = dynamic_cast<FullRelationIR *>(intermediateResult)->size();
        planTreeNode->nanosecTime = ns;
    }
    return 0;
}

static void split(const std::string &str, char delim, std::vector<std::string> &out) {
    size_t start, end = 0;
    while ((start = str.find_first_not_of(delim, end)) != std::string::npos) {
        end = str.find(delim, start);
        out.push_back(str.substr(start, end - start));
    }
}

// JoinQueryJob::join, JoinQueryJob.cu:125-453.
IR *JoinQueryJob::join(FullRelationIR *lir, FullRelationIR *rir) {
    if (lir->size() > rir->size()) std::swap(lir, rir);

    std::vector<std::string> joinVars;
    split(joinVariable, ',', joinVars);
    const bool join1Var = joinVars.size() == 1;
    if (joinVars.empty() || joinVars.size() > 2)
        throw VedasError(VD_EINVAL, "join variable list must hold 1 or 2 variables");

    // Projection push-down (:137-155): drop columns whose variable no later
    // operator needs (counter 0) and that are not join variables.  A column
    // is dropped at most once.
    auto counter = [&](const std::string &h) -> size_t {
        return query_variable_counter ? (*query_variable_counter)[h] : 1;
    };
    auto isJoinVar = [&](const std::string &h) {
        for (auto &v : joinVars) if (v == h) return true;
        return false;
    };
    for (int i = (int)lir->getColumnNum() - 1; i >= 0; --i) {
        const std::string h = lir->getHeader(i);
        if (counter(h) == 0 && !isJoinVar(h)) lir->removeColumn(i, joinVars[0]);
    }
    for (int i = (int)rir->getColumnNum() - 1; i >= 0; --i) {
        const std::string h = rir->getHeader(i);
        if (counter(h) == 0 && !isJoinVar(h)) rir->removeColumn(i, joinVars[0]);
    }
    if (query_variable_counter)
        for (auto &v : joinVars) (*query_variable_counter)[v] -= 1;

    auto swap_start = std::chrono::steady_clock::now();
    if (join1Var) {
        // An input an IndexSwapJob left unsorted (column 0 = the join
        // variable): its rows whose key the other input lacks cannot join,
        // so they are dropped before the sort when that keeps at most half
        // of them.  The stable sort keeps the kept rows' order, so the join
        // output is the one the full sort gives.  Which input is the left
        // one was decided above on the unfiltered sizes, as without it.
        auto settle = [&](FullRelationIR *x, FullRelationIR *y) {
            if (!x->pendingSort()) return;
            vd_rel *f = nullptr;
            uint64_t kept = 0;
            context->check(vd_semi_filter(context->raw(), x->raw(), 0, y->raw(), 0, 0xFFFFFFFFu,
                                          x->size() / 2, &f, &kept),
                           "vd_semi_filter");
            if (f) {
                context->filtered_rows += x->size() - kept;
                x->replaceRelation(f);
            }
            x->sortByFirstColumn();
            context->swap_rows += x->size();
        };
        settle(lir, rir);
        settle(rir, lir);
        // Implicit index swap when the key is not column 0 (:162-180).
        const size_t lIdx = lir->getColumnId(joinVars[0]);
        if (lIdx != 0) {
            lir->swapColumn(0, lIdx);
            lir->sortByFirstColumn();
            context->exe_log.emplace_back(gpuId, SWAP_OP, "", lir->size(), lir->getColumnNum());
            context->swap_rows += lir->size();
        }
        const size_t rIdx = rir->getColumnId(joinVars[0]);
        if (rIdx != 0) {
            rir->swapColumn(0, rIdx);
            rir->sortByFirstColumn();
            context->exe_log.emplace_back(gpuId, SWAP_OP, "", rir->size(), rir->getColumnNum());
            context->swap_rows += rir->size();
        }
    } else {
        // Two join variables: both relations ordered on (v0, v1) (:181-225).
        const bool aligned =
            lir->getColumnId(joinVars[0]) == rir->getColumnId(joinVars[0]) &&
            lir->getColumnId(joinVars[1]) == rir->getColumnId(joinVars[1]) &&
            ((lir->getColumnId(joinVars[0]) == 0 && lir->getColumnId(joinVars[1]) == 1) ||
             (lir->getColumnId(joinVars[0]) == 1 && lir->getColumnId(joinVars[1]) == 0));
        if (!aligned) {
            rir->swapColumn(0, rir->getColumnId(joinVars[0]));
            rir->swapColumn(1, rir->getColumnId(joinVars[1]));
            lir->swapColumn(0, lir->getColumnId(joinVars[0]));
            lir->swapColumn(1, lir->getColumnId(joinVars[1]));
            rir->sort();
            lir->sort();
        }
    }
    context->swap_index_ns += std::chrono::duration_cast<std::chrono::nanoseconds>(
        std::chrono::steady_clock::now() - swap_start).count();

    const size_t nk = joinVars.size();
    if (lir->getColumnNum() == 0 || rir->getColumnNum() == 0)
        throw VedasError(VD_EINVAL, "join operand without columns");
    // The output re-sorted by the next IndexSwap's variable, made directly
    // when one input is a select whose other permutation is resident.
    TYPEID dlo = 0, dhi = 0;
    FullRelationIR *newIr = nullptr;
    vd_rel *out = nullptr;  // a plain join the driven attempt already built
    if (join1Var && QueryExecutor::ENABLE_DRIVEN_SWAP && !swapHint.empty() &&
        swapHint != joinVars[0])
        newIr = drivenJoin(lir, rir, &dlo, &dhi, &out);
    const bool driven = newIr != nullptr;
    if (!driven) {
        if (!out)
            context->check(vd_join(context->raw(), lir->raw(), rir->raw(), (int)nk, &out),
                           "vd_join");
        newIr = new FullRelationIR(context, out);
        size_t h = 0;
        for (size_t i = 0; i < nk; ++i)
            newIr->setHeader(h++, lir->getHeader(i), lir->getIsPredicate(i));
        for (size_t i = nk; i < lir->getColumnNum(); ++i)
            newIr->setHeader(h++, lir->getHeader(i), lir->getIsPredicate(i));
        for (size_t i = nk; i < rir->getColumnNum(); ++i)
            newIr->setHeader(h++, rir->getHeader(i), rir->getIsPredicate(i));
    }
    const size_t result_size = newIr->size();
    context->join_rows += result_size;
    context->exe_log.emplace_back(gpuId, JOIN_OP, lir->getHeaders("") + " x " + rir->getHeaders(""),
                                  lir->size(), rir->size(), result_size);

    // Bound and empty-interval updates after the join (:388-433); pruning
    // only, the result is not affected.
    if (QueryExecutor::ENABLE_UPDATE_BOUND_AFTER_JOIN && variables_bound) {
        if (result_size > 0) {
            TYPEID ends[2] = {dlo, dhi};
            // vd_join returns the ends of column 0 with its size read-back
            // (a driven join the smallest and largest matched key).
            if (!driven && !vd_rel_key_range(newIr->raw(), &ends[0], &ends[1])) {
                context->check(vd_copy_d2h(context->raw(), &ends[0],
                                           newIr->getRelationRawPointer(0), 4), "vd_copy_d2h");
                context->check(vd_copy_d2h(context->raw(), &ends[1],
                                           newIr->getRelationRawPointer(0) + result_size - 1, 4),
                               "vd_copy_d2h");
            }
            QueryExecutor::updateBoundDict(variables_bound, joinVariable, ends[0], ends[1]);
            if (!lastJoinForVar && result_size > 2 && result_size < EIF_ALPHA && ei_dict) {
                auto e0 = std::chrono::steady_clock::now();
                size_t b[4];
                // the join-ordered key column: column 0 of a plain join, the
                // key column sorted on the host after a driven one
                TYPEID_HOST_VEC keys = newIr->download(driven ? newIr->getColumnId(joinVariable) : 0);
                if (driven) std::sort(keys.begin(), keys.end());
                eifFromSortedColumn(keys, b);
                context->eif_count++;
                ei_dict->updateBound(joinVariable, b[0], b[1], b[2], b[3]);
                context->eif_ns += std::chrono::duration_cast<std::chrono::nanoseconds>(
                    std::chrono::steady_clock::now() - e0).count();
            }
        } else {
            QueryExecutor::updateBoundDict(variables_bound, joinVariable, 0, 0);
        }
    }
    return newIr;
}

// The join followed by the stable re-sort of its output on swapHint, in one
// device step (vd_join_driven): rows of the input X holding swapHint, in
// swapHint order, each emit their key's run of the other input Y.
//  * X a select (columns [key, swapHint]): the same pattern's rows in the
//    other permutation (the alternate view D, columns [swapHint, key],
//    sorted on swapHint then key) drive, with no sort at all.
//  * X the left input, any relation: X itself is re-sorted on swapHint and
//    drives, when the join's output has at least twice X's rows, so
//    re-sorting X costs less than re-sorting the output: the join's count
//    phase runs first, and a smaller output is built as the plain join's
//    (returned in *plain_out, with null).
// The output comes in (swapHint, key, Y order) order, which is the join's
// left-major order re-sorted stably on swapHint: with X on the left each D
// row is one left row, in left order within equal swapHint values; with X
// on the right each D row is one right row, distinct from its neighbours
// (D rows that repeat and match several left rows would interleave
// differently: vd_join_driven then declines).  The output columns are the
// join's [key, left rest, right rest] with column 0 and swapHint's
// exchanged, as IndexSwapJob leaves them.  Null when no form applies.
FullRelationIR *JoinQueryJob::drivenJoin(FullRelationIR *lir, FullRelationIR *rir,
                                         TYPEID *match_lo, TYPEID *match_hi,
                                         vd_rel **plain_out) {
    auto selectOf = [&](FullRelationIR *ir) -> SelectQueryJob * {
        for (QueryJob *j : {leftJob, rightJob})
            if (j && j->getIR() == ir) return dynamic_cast<SelectQueryJob *>(j);
        return nullptr;
    };
    const std::string &key = joinVariable, &v = swapHint;
    auto holds = [&](FullRelationIR *ir) {
        return ir->getColumnNum() == 2 && ir->getHeader(0) == key && ir->getHeader(1) == v;
    };
    auto has = [&](FullRelationIR *ir, const std::string &h) {
        for (size_t i = 0; i < ir->getColumnNum(); ++i)
            if (ir->getHeader(i) == h) return true;
        return false;
    };
    if (lir->getHeader(0) != key || rir->getHeader(0) != key) return nullptr;
    // The join's output headers, then IndexSwapJob's exchange of column 0
    // with swapHint's column.
    std::vector<std::pair<std::string, bool>> cols;
    for (size_t i = 0; i < lir->getColumnNum(); ++i)
        cols.emplace_back(lir->getHeader(i), lir->getIsPredicate(i));
    for (size_t i = 1; i < rir->getColumnNum(); ++i)
        cols.emplace_back(rir->getHeader(i), rir->getIsPredicate(i));
    size_t iv = 0;
    while (iv < cols.size() && cols[iv].first != v) ++iv;
    if (iv == cols.size()) return nullptr;
    std::swap(cols[0], cols[iv]);
    // Output column sources: D's columns (-1 - c) or Y's (c).
    auto sources = [&](FullRelationIR *D, FullRelationIR *Y) {
        std::vector<int> src;
        for (auto &c : cols)
            src.push_back(has(D, c.first) ? -1 - (int)D->getColumnId(c.first)
                                          : (int)Y->getColumnId(c.first));
        return src;
    };
    vd_rel *out = nullptr;
    FullRelationIR *X = nullptr;
    if (holds(lir) && selectOf(lir)) X = lir;
    else if (holds(rir) && selectOf(rir)) X = rir;
    vd_rel *D = nullptr;
    TYPEID clo = 0, chi = 0;
    if (X) D = selectOf(X)->alternateView(&clo, &chi);
    if (D) {
        FullRelationIR *Y = X == lir ? rir : lir;
        if (has(Y, v)) {
            vd_rel_destroy(context->raw(), D);
            return nullptr;
        }
        // D's columns are [swapHint, key]
        std::vector<int> src;
        for (auto &c : cols)
            src.push_back(c.first == v ? -1 : c.first == key ? -2 : (int)Y->getColumnId(c.first));
        const int rc = vd_join_driven(context->raw(), D, 1, Y->raw(), clo, chi, (int)src.size(),
                                      src.data(), X == rir ? 1 : 0, &out, match_lo, match_hi);
        vd_rel_destroy(context->raw(), D);
        context->check(rc, "vd_join_driven");
    } else if (has(lir, v) && !has(rir, v)) {
        if (!vd_join_driven_applies(context->raw(), lir->raw(), 0, rir->raw(), 0xFFFFFFFFu))
            return nullptr;
        // The join itself when its output is small (its count phase tells).
        uint64_t total = 0;
        context->check(vd_join_capped(context->raw(), lir->raw(), rir->raw(), 1,
                                      2 * (uint64_t)lir->size() - 1, plain_out, &total),
                       "vd_join_capped");
        if (*plain_out) return nullptr;
        lir->swapColumn(0, lir->getColumnId(v));
        lir->sortByFirstColumn();
        context->swap_rows += lir->size();
        const std::vector<int> src = sources(lir, rir);
        context->check(vd_join_driven(context->raw(), lir->raw(), (int)lir->getColumnId(key),
                                      rir->raw(), 0, 0xFFFFFFFFu, (int)src.size(), src.data(), 0,
                                      &out, match_lo, match_hi),
                       "vd_join_driven");
        // the size rules were checked before the re-sort; nothing else declines
        if (!out) throw VedasError(VD_EINVAL, "driven join declined after its input was re-sorted");
    }
    if (!out) return nullptr;
    auto *ir = new FullRelationIR(context, out);
    for (size_t i = 0; i < cols.size(); ++i) ir->setHeader(i, cols[i].first, cols[i].second);
    ir->markSortedOnColumn0();
    context->driven_joins += 1;
    return ir;
}

void JoinQueryJob::setOperands(QueryJob *l, QueryJob *r) {
    leftJob = l;
    rightJob = r;
}

void JoinQueryJob::print() const {
    std::cout << "\tJOIN JOB - on " << joinVariable << "\n";
}

void JoinQueryJob::setQueryVarCounter(std::map<std::string, size_t> *c) { query_variable_counter = c; }

unsigned JoinQueryJob::getLeftIRSize() const {
    auto *ir = dynamic_cast<FullRelationIR *>(leftJob->getIR());
    return ir ? (unsigned)ir->size() : 0;
}

unsigned JoinQueryJob::getRightIRSize() const {
    auto *ir = dynamic_cast<FullRelationIR *>(rightJob->getIR());
    return ir ? (unsigned)ir->size() : 0;
}

std::string JoinQueryJob::getJoinVariable() const { return joinVariable; }

std::string JoinQueryJob::jobTypeName() const { return "Join Job"; }
