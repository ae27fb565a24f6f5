#include "QueryExecutor.h"

#include <algorithm>
#include <cctype>
#include <limits>

#include "IndexSwapJob.h"
#include "JoinQueryJob.h"

bool QueryExecutor::ENABLE_FULL_INDEX = false;
// vdQuery and vdBench switch this on (VedasQuery.cu:37, VedasBench.cu:32).
bool QueryExecutor::ENABLE_UPDATE_BOUND_AFTER_JOIN = true;

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

// Join variables start with the full range (QueryExecutor.cu:441-461); the
// pattern pre-scan that narrows them is not performed.
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
