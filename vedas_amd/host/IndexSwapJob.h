// IndexSwap operator (reference IndexSwapJob.h:10-26): moves the swap
// variable to column 0 and re-sorts the child's relation by it, in place.
#ifndef VEDAS_HOST_INDEXSWAPJOB_H
#define VEDAS_HOST_INDEXSWAPJOB_H

#include <string>

#include "EmptyIntervalDict.h"
#include "FullRelationIR.h"
#include "QueryJob.h"

class IndexSwapJob : public QueryJob {
  public:
    IndexSwapJob(QueryJob *beforeJob, std::string swapVar, VedasContext *context,
                 VariableBoundMap *variables_bound, EmptyIntervalDict *ei_dict);
    ~IndexSwapJob() override;
    IR *getIR() override;
    IR *releaseIR() override;
    int startJob(int gpuId) override;
    void print() const override;
    std::string jobTypeName() const override;
    QueryJob *childJob() const { return beforeJob; }
    const std::string &swapVariable() const { return swapVar; }
    // The parent job is a join on swapVar (QueryPlan::execute sets it): a
    // relation of at least QueryExecutor::FILTERED_SWAP_MIN_ROWS rows is
    // left unsorted for that join, which filters it first.
    void setJoinParent(bool v) { joinParent = v; }

  private:
    bool joinParent = false;
    VedasContext *context;
    VariableBoundMap *variables_bound{nullptr};
    std::string swapVar;
    QueryJob *beforeJob{nullptr};
    EmptyIntervalDict *ei_dict{nullptr};
};

#endif
