// Store statistics and plan size estimates.
//
// Statistics: per leading key of each built permutation, its row count and
// its number of distinct (X,Y) pairs (vd_index_stats), kept on the host.
// For the predicate-first permutations this gives, per predicate, the
// triple count and the distinct subjects (PSO) / objects (POS) -- the role
// of the reference's Histogram and DataMetaInfo (Histogram.cu:170-231,
// DataMetaInfo.h:7-20), which its planner and this build's admission use.
//
// PlanEstimator: expected rows of every plan node under independence
// (|L join R| = |L| |R| / max(d_L(v), d_R(v))) and the HBM bytes a plan
// holds at its peak, for memory admission in the serving loop.
#ifndef VEDAS_HOST_STATISTICS_H
#define VEDAS_HOST_STATISTICS_H

#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "PlanTreeNode.h"
#include "VedasStorage.h"

class Statistics {
  public:
    explicit Statistics(VedasStorage *storage);
    uint64_t triples() const { return n_; }
    // Rows with leading key x in permutation perm (0 if absent or not built).
    uint64_t rows(int perm, TYPEID x) const;
    // Distinct (x, Y) pairs in permutation perm.
    uint64_t pairs(int perm, TYPEID x) const;
    bool has(int perm) const { return keys_[perm].size() > 0 || built_[perm]; }

  private:
    uint64_t n_ = 0;
    bool built_[VD_NUM_PERMS] = {};
    std::vector<TYPEID> keys_[VD_NUM_PERMS];
    std::vector<uint64_t> rows_[VD_NUM_PERMS], pairs_[VD_NUM_PERMS];
};

struct PlanEstimate {
    double rows = 0;        // expected rows of the final relation
    double peak_bytes = 0;  // expected peak HBM bytes of the plan
    double work_rows = 0;   // sum of expected rows over all operators
};

class PlanEstimator {
  public:
    explicit PlanEstimator(const Statistics &st) : st_(st) {}
    PlanEstimate estimate(const std::vector<TriplePattern> &patterns, PlanTreeNode *root) const;

  private:
    struct Est {
        double rows = 0;
        std::map<std::string, double> distinct;  // per variable
        int cols = 0;
    };
    Est node(PlanTreeNode *n, const std::vector<TriplePattern> &pats, PlanEstimate &acc,
             double &live) const;
    Est select(const TriplePattern &tp) const;
    const Statistics &st_;
};

#endif
