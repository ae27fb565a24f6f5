// Join operator (reference JoinQueryJob.h:10-48).
#ifndef VEDAS_HOST_JOINQUERYJOB_H
#define VEDAS_HOST_JOINQUERYJOB_H

#include <map>
#include <string>

#include "EmptyIntervalDict.h"
#include "FullRelationIR.h"
#include "QueryJob.h"

class JoinQueryJob : public QueryJob {
  public:
    JoinQueryJob(QueryJob *leftJob, QueryJob *rightJob, std::string joinVariable,
                 VedasContext *context, VariableBoundMap *variables_bound,
                 EmptyIntervalDict *ei_dict, bool lastJoinForVar = true);
    ~JoinQueryJob() override;
    IR *getIR() override;
    IR *releaseIR() override;
    int startJob(int gpuId) override;
    std::string jobTypeName() const override;
    void setOperands(QueryJob *leftJob, QueryJob *rightJob);
    IR *join(FullRelationIR *lir, FullRelationIR *rir);
    void print() const override;
    void setQueryVarCounter(std::map<std::string, size_t> *query_variable_counter);
    unsigned getLeftIRSize() const;
    unsigned getRightIRSize() const;
    std::string getJoinVariable() const;
    // The variable an IndexSwapJob re-sorts this join's output by next
    // (QueryPlan::execute sets it): the join may then deliver its output in
    // that order directly (drivenJoin).
    void setSwapHint(const std::string &v) { swapHint = v; }
    QueryJob *leftOperand() const { return leftJob; }
    QueryJob *rightOperand() const { return rightJob; }
    VariableBoundMap *variables_bound{nullptr};
    EmptyIntervalDict *ei_dict{nullptr};

  private:
    FullRelationIR *drivenJoin(FullRelationIR *lir, FullRelationIR *rir, TYPEID *match_lo,
                               TYPEID *match_hi, vd_rel **plain_out);
    std::string joinVariable;
    std::string swapHint;
    VedasContext *context;
    QueryJob *leftJob{nullptr};
    QueryJob *rightJob{nullptr};
    bool lastJoinForVar = false;
    std::map<std::string, size_t> *query_variable_counter{nullptr};
};

#endif
