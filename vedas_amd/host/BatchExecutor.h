// Batched execution of many queries (no counterpart in the reference, which
// executes one query's jobs one after the other, QueryPlan.cu:160-205).
//
// All plans advance together in rounds: every round collects the operators
// of all queries whose inputs are ready and runs each kind with one batched
// C-ABI call (vd_select_multi, vd_sort_col0_multi, vd_join_multi), so a
// round costs a fixed handful of launches and synchronisations instead of a
// few per operator per query.  Per query, the operators are those the
// single-query QueryExecutor creates (same index choice, same join
// rules: smaller input left, projection push-down, key to column 0, output
// [key, left rest, right rest]); the result of each query equals its
// single-query result row for row after canonical ordering.  Bound and
// empty-interval pruning (result-neutral) are not applied in batches; a
// large relation an index swap feeds to a join on its variable is
// filtered by that join's other input before its sort, as on the
// one-query path.
#ifndef VEDAS_HOST_BATCHEXECUTOR_H
#define VEDAS_HOST_BATCHEXECUTOR_H

#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "PlanTreeNode.h"
#include "SparqlResult.h"
#include "VedasStorage.h"

struct BatchQuery {
    const std::vector<TriplePattern> *patterns = nullptr;  // not owned
    std::set<std::string> selected;
    PlanTreeNode *root = nullptr;  // may be null for one pattern
};

struct BatchStats {
    int rounds = 0;
    int launches_calls = 0;   // batched C-ABI calls issued
    uint64_t join_rows = 0, select_rows = 0, swap_rows = 0;
};

class BatchExecutor {
  public:
    BatchExecutor(VedasStorage *storage, VedasContext *context);
    // results[i] receives query i's final relation.
    void run(std::vector<BatchQuery> &queries, std::vector<SparqlResult> &results,
             BatchStats *stats = nullptr);

  private:
    struct Rel;
    struct Node;
    struct Query;
    void selectRound(std::vector<std::pair<Query *, int>> &ready, BatchStats &st);
    void joinRound(std::vector<std::pair<Query *, int>> &joins,
                   std::vector<std::pair<Query *, int>> &swaps, BatchStats &st);

    VedasStorage *storage;
    VedasContext *context;
};

#endif
