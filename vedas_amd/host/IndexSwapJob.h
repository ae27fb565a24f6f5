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

  private:
    VedasContext *context;
    VariableBoundMap *variables_bound{nullptr};
    std::string swapVar;
    QueryJob *beforeJob{nullptr};
    EmptyIntervalDict *ei_dict{nullptr};
};

#endif
