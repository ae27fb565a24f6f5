// Select operator (reference SelectQueryJob.h:11-59).
//
// The reference constructors take the host L1/L2 arrays and data columns of
// one permutation; here an IndexRef (storage + permutation) names the
// HBM-resident arrays, and the lookups run on the device.
#ifndef VEDAS_HOST_SELECTQUERYJOB_H
#define VEDAS_HOST_SELECTQUERYJOB_H

#include <string>

#include "EmptyIntervalDict.h"
#include "FullRelationIR.h"
#include "QueryJob.h"
#include "VedasStorage.h"

struct IndexRef {
    VedasStorage *storage;
    int perm;  // VD_PERM_*
};

class SelectQueryJob : public QueryJob {
  public:
    // 2-variable pattern: rows with X == id1; columns [v1 = Y, v2 = Z]
    // (SelectQueryJob.cu:20-44).
    SelectQueryJob(IndexRef index, std::string v1, std::string v2, TYPEID id1,
                   VariableBoundMap *variables_bound, EmptyIntervalDict *ei_dict,
                   bool *is_predicates, bool is_second_var_used, VedasContext *context);
    // 1-variable pattern: rows with X == id1 and Y == id2; column [v1 = Z]
    // (SelectQueryJob.cu:46-66).
    SelectQueryJob(IndexRef index, std::string v1, TYPEID id1, TYPEID id2,
                   VariableBoundMap *variables_bound, EmptyIntervalDict *ei_dict,
                   bool *is_predicates, VedasContext *context);
    ~SelectQueryJob() override;
    int startJob(int gpuId) override;
    IR *getIR() override;
    IR *releaseIR() override;
    std::string getVariable(size_t i) const;
    size_t getVariableNum() const;
    TYPEID getId(size_t i) const;
    void print() const override;
    std::string jobTypeName() const override;

    // The lookup this select starts with (permutation and constants; the
    // cols field is not used) and the store it reads.
    vd_select_job_t lookupJob() const;
    vd_index *indexHandle() const;
    VedasContext *getContext() const { return context; }
    // Row range of the lookup, resolved ahead of time (QueryPlan::execute
    // batches the lookups of a plan): startJob then only filters it.
    void setRange(uint64_t begin, uint64_t end);
    // After startJob, for a 2-variable select whose rows were only clipped
    // on its first variable: the same constant's range in the permutation
    // with Y and Z exchanged (PSO <-> POS, SPO <-> SOP, OSP <-> OPS), as a
    // view with columns [second variable, first variable], unclipped, and
    // the clip of the first variable ([0, 2^32-1] without one).  Null when
    // that permutation is not built or rows were dropped otherwise (EIF
    // skips, a distinct first column).
    vd_rel *alternateView(TYPEID *clip_lo, TYPEID *clip_hi) const;

  private:
    bool alt_ok = false;  // startJob kept every row inside the clip
    uint64_t sel_begin = 0, sel_end = 0;  // the constant's row range
    TYPEID clip_lo = 0, clip_hi = 0xFFFFFFFFu;
    bool has_range = false;
    uint64_t range_begin = 0, range_end = 0;
    IndexRef index;
    std::string variables[3];
    size_t variable_num;
    TYPEID dataid[2] = {0, 0};
    bool is_predicates[2] = {false, false};
    bool is_second_var_used = true;
    VariableBoundMap *variables_bound{nullptr};
    EmptyIntervalDict *ei_dict{nullptr};
    VedasContext *context;
};

#endif
