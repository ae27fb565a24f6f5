// Job factory and plan driver (reference QueryExecutor.h:38-125, the
// single-GPU path): the bound pre-scan, then the plan tree's jobs (the tree
// comes from the caller or the QueryGraph planner).
#ifndef VEDAS_HOST_QUERYEXECUTOR_H
#define VEDAS_HOST_QUERYEXECUTOR_H

#include <map>
#include <set>
#include <string>
#include <vector>

#include "EmptyIntervalDict.h"
#include "PlanTreeNode.h"
#include "QueryPlan.h"
#include "SelectQueryJob.h"
#include "SparqlResult.h"
#include "VedasStorage.h"

class QueryExecutor {
  public:
    QueryExecutor(VedasStorage *vedasStorage, VedasContext *context);
    // Runs the plan tree `root` over `patterns`; SELECT `selected`
    // (QueryExecutor::query, QueryExecutor.cu:294-431, single GPU).
    void query(std::vector<TriplePattern> &patterns, const std::set<std::string> &selected,
               PlanTreeNode *root, SparqlResult &result);

    static void updateBoundDict(VariableBoundMap *bound, const std::string &variable,
                                TYPEID lowerBound, TYPEID upperBound);
    static bool ENABLE_FULL_INDEX;
    static bool ENABLE_UPDATE_BOUND_AFTER_JOIN;
    static bool ENABLE_BOUND_PRESCAN;
    static bool ENABLE_DRIVEN_SWAP;
    // Smallest relation an IndexSwapJob feeding a join leaves to that join
    // to filter and sort (0: every IndexSwapJob sorts).
    static uint64_t FILTERED_SWAP_MIN_ROWS;
    // Bounds of the join variables after the last query (for inspection).
    const VariableBoundMap &bounds() const { return variables_bound; }

    SelectQueryJob *createSelectQueryJob(TriplePattern *pattern, std::string index_used = "");
    SelectQueryJob *create1VarSelectQueryJob(TriplePattern *pattern);
    SelectQueryJob *create2VarSelectQueryJob(TriplePattern *pattern, std::string index_used);

  private:
    int postorderTraversal(QueryPlan &plan, PlanTreeNode *root);
    void createVariableBound(std::vector<TriplePattern> &patterns);

    VedasStorage *vedasStorage;
    VedasContext *context;
    VariableBoundMap variables_bound;
    EmptyIntervalDict empty_interval_dict;
    std::set<std::string> selected_variables;
    int jobCount = 0;
};

#endif
