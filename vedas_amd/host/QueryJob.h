// Operator base class (reference QueryJob.h:7-21).
#ifndef VEDAS_HOST_QUERYJOB_H
#define VEDAS_HOST_QUERYJOB_H

#include <string>

#include "IR.h"

struct ExecPlanTreeNode {
    size_t resultSize = 0;
    long long nanosecTime = 0;
};

class QueryJob {
  public:
    QueryJob() = default;
    virtual ~QueryJob() = default;
    virtual IR *getIR() = 0;
    // 0 on success; C-ABI failures throw VedasError.
    virtual int startJob(int gpuId) = 0;
    virtual void print() const = 0;
    virtual std::string jobTypeName() const = 0;
    // Hand the result to the caller (the plan's final relation).
    virtual IR *releaseIR() = 0;
    ExecPlanTreeNode *planTreeNode = nullptr;
    int getGpuId() { return gpuId; }

  protected:
    int gpuId = 0;
    IR *intermediateResult{nullptr};
};

#endif
