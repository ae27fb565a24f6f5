#include "QueryExecutor.h"

#include <algorithm>
#include <cctype>
#include <limits>
#include <map>

#include "IndexSwapJob.h"
#include "JoinQueryJob.h"

bool QueryExecutor::ENABLE_FULL_INDEX = false;
// vdQuery and vdBench switch this on (VedasQuery.cu:37, VedasBench.cu:32).
bool QueryExecutor::ENABLE_UPDATE_BOUND_AFTER_JOIN = true;
// The pattern pre-scan of createVariableBound (QueryExecutor.cu:463-478);
// off gives join variables the full range (for comparing intermediates).
bool QueryExecutor::ENABLE_BOUND_PRESCAN = true;
// A join followed by an IndexSwap delivers the swapped order itself where
// one input is a select with its other permutation resident (no
// counterpart in the reference; off runs the join and the re-sort).
bool QueryExecutor::ENABLE_DRIVEN_SWAP = true;
uint64_t QueryExecutor::FILTERED_SWAP_MIN_ROWS = 1u << 20;

QueryExecutor::QueryExecutor(VedasStorage *s, VedasContext *c) : vedasStorage(s), context(c) {}

// QueryExecutor::updateBoundDict, QueryExecutor.cu:583-594.
void QueryExecutor::updateBoundDict(VariableBoundMap *bound, const std::string &variable,
                                    TYPEID lowerBound, TYPEID upperBound) {
    auto varBound = (*bound)[variable];
    if ((upperBound < varBound.first) || (lowerBound > varBound.second))
        lowerBound = upperBound = 0;
    TYPEID new_min = (lowerBound == 0) ? 0 : std::max(varBound.first, lowerBound);
    TYPEID new_max = std::min(varBound.second, upperBound);
    (*bound)[variable] = std::make_pair(new_min, new_max);
}

// QueryExecutor::createVariableBound, QueryExecutor.cu:441-479: every
// variable that occurs more than once starts with the full range, then each
// pattern narrows the bounds of its join variables to the first and last
// value the variable takes over the pattern's index range
// (squeezeQueryBound1Var / squeezeQueryBound2Var, :527-581, 680-735, via
// updateDataBound / updateL2Bound, :648-675, 737-765), in pattern order;
// a pattern with no rows sets every bound to [0, 0] (:470-475).  The
// reference reads those ranges and values from host copies of the index;
// here one vd_index_bounds_multi launch resolves every pattern.  For a
// pattern with one variable the range is that of both constants and the
// value column is the third one (POS for ?s p o, PSO for s p ?o, OSP for
// s ?p o; the reference's alternative O-/S-first indexes hold the same
// rows in the same order).  For two variables the range is that of the
// constant, and each variable is read from the permutation that has it
// right after the constant, the first variable of the reference's order
// first (object then subject for ?s p ?o, object then predicate for
// s ?p ?o, predicate then subject for ?s ?p o).  A permutation that was
// not built gives no bound.
void QueryExecutor::createVariableBound(std::vector<TriplePattern> &patterns) {
    std::map<std::string, int> count;
    for (auto &tp : patterns) {
        if (tp.subjectIsVariable()) count[tp.getSubject()]++;
        if (tp.predicateIsVariable()) count[tp.getPredicate()]++;
        if (tp.objectIsVariable()) count[tp.getObject()]++;
    }
    variables_bound.clear();
    for (auto &c : count)
        if (c.second > 1)
            variables_bound[c.first] = std::make_pair(std::numeric_limits<TYPEID>::min(),
                                                      std::numeric_limits<TYPEID>::max());
    if (!ENABLE_BOUND_PRESCAN || variables_bound.empty()) return;

    struct Item {
        size_t pattern;
        std::string var;  // the variable the value column holds
    };
    std::vector<vd_select_job_t> jobs;
    std::vector<Item> items;
    auto add = [&](size_t i, const std::string &var, int perm, TYPEID k1, TYPEID k2) {
        if (!vedasStorage->hasPermutation(perm)) return false;
        vd_select_job_t j{};
        j.perm = perm;
        j.k1 = k1;
        j.k2 = k2;
        j.cols = 0;
        jobs.push_back(j);
        items.push_back(Item{i, var});
        return true;
    };
    for (size_t i = 0; i < patterns.size(); ++i) {
        TriplePattern &tp = patterns[i];
        const TYPEID s = tp.getSubjectId(), p = tp.getPredicateId(), o = tp.getObjectId();
        if (tp.getVariableNum() == 1) {
            if (tp.subjectIsVariable()) {
                add(i, tp.getSubject(), VD_PERM_POS, p, o) ||
                    add(i, tp.getSubject(), VD_PERM_OPS, o, p);
            } else if (tp.predicateIsVariable()) {
                add(i, tp.getPredicate(), VD_PERM_OSP, o, s) ||
                    add(i, tp.getPredicate(), VD_PERM_SOP, s, o);
            } else {
                add(i, tp.getObject(), VD_PERM_PSO, p, s) ||
                    add(i, tp.getObject(), VD_PERM_SPO, s, p);
            }
        } else if (tp.getVariableNum() == 2) {
            switch (tp.getVariableBitmap()) {
                case 5:  // ?s p ?o
                    add(i, tp.getObject(), VD_PERM_POS, p, VD_UNBOUND);
                    add(i, tp.getSubject(), VD_PERM_PSO, p, VD_UNBOUND);
                    break;
                case 3:  // s ?p ?o
                    add(i, tp.getObject(), VD_PERM_SOP, s, VD_UNBOUND);
                    add(i, tp.getPredicate(), VD_PERM_SPO, s, VD_UNBOUND);
                    break;
                case 6:  // ?s ?p o
                    add(i, tp.getPredicate(), VD_PERM_OPS, o, VD_UNBOUND);
                    add(i, tp.getSubject(), VD_PERM_OSP, o, VD_UNBOUND);
                    break;
            }
        }
    }
    if (jobs.empty()) return;
    const size_t nj = jobs.size();
    std::vector<uint64_t> b(nj), e(nj);
    std::vector<uint32_t> first(nj), last(nj);
    context->check(vd_index_bounds_multi(context->raw(), vedasStorage->index(), jobs.data(),
                                         (int)nj, b.data(), e.data(), first.data(), last.data()),
                   "vd_index_bounds_multi");
    bool zeroResult = false;
    for (size_t k = 0; k < nj; ++k) {
        if (e[k] == b[k]) {
            zeroResult = true;
            continue;
        }
        if (variables_bound.count(items[k].var) > 0)
            updateBoundDict(&variables_bound, items[k].var, first[k], last[k]);
    }
    if (zeroResult)
        for (auto &vb : variables_bound) vb.second = std::make_pair(0u, 0u);
}

void QueryExecutor::query(std::vector<TriplePattern> &patterns, const std::set<std::string> &selected,
                          PlanTreeNode *root, SparqlResult &result) {
    ContextGuard guard(context);
    selected_variables = selected;
    empty_interval_dict = EmptyIntervalDict();
    QueryPlan plan(selected_variables);
    jobCount = 0;
    if (patterns.size() == 1 && (!root || root->op == UPLOAD)) {
        plan.pushJob(createSelectQueryJob(&patterns[0]));
        plan.execute(result);
        return;
    }
    createVariableBound(patterns);
    postorderTraversal(plan, root);
    plan.execute(result);
}

// QueryExecutor::postorderTraversal, QueryExecutor.cu:106-135.
int QueryExecutor::postorderTraversal(QueryPlan &plan, PlanTreeNode *root) {
    if (!root) return -1;
    int j1 = -1, j2 = -1;
    if (root->child1) j1 = postorderTraversal(plan, root->child1);
    if (root->child2) j2 = postorderTraversal(plan, root->child2);
    switch (root->op) {
        case UPLOAD:
            if (root->tp->getVariableNum() == 1) {
                plan.pushJob(createSelectQueryJob(root->tp));
            } else if (root->tp->getVariableNum() == 2) {
                std::string indexUse = (root->tp->getSubject() == root->var) ? "PSO" : "POS";
                plan.pushJob(createSelectQueryJob(root->tp, indexUse));
            } else {
                throw VedasError(VD_EINVAL, "pattern without a constant is not supported");
            }
            break;
        case JOIN:
            if (j1 < 0 || j2 < 0) throw VedasError(VD_EINVAL, "join node needs two children");
            plan.pushJob(new JoinQueryJob(plan.getJob(j1), plan.getJob(j2), root->var, context,
                                          &variables_bound, &empty_interval_dict, !root->reuseVar));
            break;
        case INDEXSWAP:
            if (j1 < 0) throw VedasError(VD_EINVAL, "index swap node needs a child");
            plan.pushJob(new IndexSwapJob(plan.getJob(j1), root->var, context, &variables_bound,
                                          &empty_interval_dict));
            break;
    }
    return jobCount++;
}

SelectQueryJob *QueryExecutor::createSelectQueryJob(TriplePattern *pattern, std::string index_used) {
    switch (pattern->getVariableNum()) {
        case 1: return create1VarSelectQueryJob(pattern);
        case 2: return create2VarSelectQueryJob(pattern, index_used);
        default: throw VedasError(VD_EINVAL, "select needs 1 or 2 variables");
    }
}

static void need(VedasStorage *s, int perm) {
    if (!s->hasPermutation(perm))
        throw VedasError(VD_EINVAL, "pattern needs a permutation that was not built");
}

// QueryExecutor::create1VarSelectQueryJob, QueryExecutor.cu:781-860.
SelectQueryJob *QueryExecutor::create1VarSelectQueryJob(TriplePattern *pattern) {
    bool is_pred[1] = {false};
    if (pattern->subjectIsVariable()) {
        // ?s p o: (p, o) prefix of POS -> [?s]
        need(vedasStorage, VD_PERM_POS);
        return new SelectQueryJob(IndexRef{vedasStorage, VD_PERM_POS}, pattern->getSubject(),
                                  pattern->getPredicateId(), pattern->getObjectId(),
                                  &variables_bound, &empty_interval_dict, is_pred, context);
    }
    if (pattern->predicateIsVariable()) {
        // s ?p o: (s, o) prefix of SOP -> [?p]
        need(vedasStorage, VD_PERM_SOP);
        is_pred[0] = true;
        return new SelectQueryJob(IndexRef{vedasStorage, VD_PERM_SOP}, pattern->getPredicate(),
                                  pattern->getSubjectId(), pattern->getObjectId(),
                                  &variables_bound, &empty_interval_dict, is_pred, context);
    }
    // s p ?o: (p, s) prefix of PSO -> [?o]
    need(vedasStorage, VD_PERM_PSO);
    return new SelectQueryJob(IndexRef{vedasStorage, VD_PERM_PSO}, pattern->getObject(),
                              pattern->getPredicateId(), pattern->getSubjectId(),
                              &variables_bound, &empty_interval_dict, is_pred, context);
}

// QueryExecutor::create2VarSelectQueryJob, QueryExecutor.cu:862-971.
SelectQueryJob *QueryExecutor::create2VarSelectQueryJob(TriplePattern *pattern, std::string index_used) {
    std::transform(index_used.begin(), index_used.end(), index_used.begin(), ::toupper);
    auto used = [&](const std::string &v) {
        return selected_variables.count(v) > 0 || variables_bound.count(v) > 0;
    };
    bool is_pred[2] = {false, false};
    switch (pattern->getVariableBitmap()) {
        case 3:  // s ?p ?o
            if (index_used == "SOP") {
                need(vedasStorage, VD_PERM_SOP);
                is_pred[1] = true;
                return new SelectQueryJob(IndexRef{vedasStorage, VD_PERM_SOP}, pattern->getObject(),
                                          pattern->getPredicate(), pattern->getSubjectId(),
                                          &variables_bound, &empty_interval_dict, is_pred,
                                          used(pattern->getPredicate()), context);
            }
            need(vedasStorage, VD_PERM_SPO);
            is_pred[0] = true;
            return new SelectQueryJob(IndexRef{vedasStorage, VD_PERM_SPO}, pattern->getPredicate(),
                                      pattern->getObject(), pattern->getSubjectId(),
                                      &variables_bound, &empty_interval_dict, is_pred,
                                      used(pattern->getObject()), context);
        case 5:  // ?s p ?o
            if (index_used == "POS") {
                need(vedasStorage, VD_PERM_POS);
                return new SelectQueryJob(IndexRef{vedasStorage, VD_PERM_POS}, pattern->getObject(),
                                          pattern->getSubject(), pattern->getPredicateId(),
                                          &variables_bound, &empty_interval_dict, is_pred,
                                          used(pattern->getSubject()), context);
            }
            need(vedasStorage, VD_PERM_PSO);
            return new SelectQueryJob(IndexRef{vedasStorage, VD_PERM_PSO}, pattern->getSubject(),
                                      pattern->getObject(), pattern->getPredicateId(),
                                      &variables_bound, &empty_interval_dict, is_pred,
                                      used(pattern->getObject()), context);
        case 6:  // ?s ?p o
            if (index_used == "OPS") {
                need(vedasStorage, VD_PERM_OPS);
                is_pred[0] = true;
                return new SelectQueryJob(IndexRef{vedasStorage, VD_PERM_OPS}, pattern->getPredicate(),
                                          pattern->getSubject(), pattern->getObjectId(),
                                          &variables_bound, &empty_interval_dict, is_pred,
                                          used(pattern->getSubject()), context);
            }
            need(vedasStorage, VD_PERM_OSP);
            is_pred[1] = true;
            return new SelectQueryJob(IndexRef{vedasStorage, VD_PERM_OSP}, pattern->getSubject(),
                                      pattern->getPredicate(), pattern->getObjectId(),
                                      &variables_bound, &empty_interval_dict, is_pred,
                                      used(pattern->getPredicate()), context);
        default:
            throw VedasError(VD_EINVAL, "unsupported 2-variable pattern shape");
    }
}
