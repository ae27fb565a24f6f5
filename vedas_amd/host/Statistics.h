// Store statistics and plan size estimates.
//
// Statistics, computed on the device when the store is loaded:
//  * per leading key of PSO / POS, its row count and its number of distinct
//    (X,Y) pairs (vd_index_stats): per predicate, the triples and the
//    distinct subjects / objects;
//  * equal-depth and equal-width histograms of the subject and object
//    columns (vd_index_histogram over SPO / OSP), as the reference's
//    Histogram builds them at index build time (Histogram.cu:170-231,
//    VedasStorage.cu:316-317, 518-536) with its default sizes (depth 1000,
//    width 100; VedasStorage.cu:526-527), here over every key.
//
// PlanEstimator: expected rows of every plan node and the HBM bytes a plan
// holds at its peak, for memory admission in the serving loop.  Join sizes
// come from key-distribution moments: a variable that entered a relation
// through a patterns in which it is the subject and b in which it is the
// object has, under independence, the distribution s(v)^a o(v)^b of its
// value v (s, o: the subject and object key shares), so
//   |A join B on v| = |A| |B| M(a1+a2, b1+b2) / (M(a1,b1) M(a2,b2)),
//   M(a,b) = sum_v s(v)^a o(v)^b,
// with M read from the equal-depth histograms as piecewise-constant shares
// (a key holding more rows than the depth has a bucket of its own, so the
// heavy keys that dominate these sums are exact).  Without the histograms
// (no SPO / OSP) the estimate falls back to |A| |B| / max(d_A, d_B).
#ifndef VEDAS_HOST_STATISTICS_H
#define VEDAS_HOST_STATISTICS_H

#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "PlanTreeNode.h"
#include "VedasStorage.h"

struct KeyHistogram {
    std::vector<TYPEID> lowers, uppers;
    std::vector<uint64_t> freqs;
    uint64_t rows() const;
};

class Statistics {
  public:
    explicit Statistics(VedasStorage *storage);
    uint64_t triples() const { return n_; }
    // Rows with leading key x in permutation perm (0 if absent or not built).
    uint64_t rows(int perm, TYPEID x) const;
    // Distinct (x, Y) pairs in permutation perm.
    uint64_t pairs(int perm, TYPEID x) const;
    bool has(int perm) const { return keys_[perm].size() > 0 || built_[perm]; }

    static constexpr unsigned kDepth = 1000;  // Histogram default depth
    static constexpr unsigned kWidth = 100;   // Histogram default width
    // Histograms of the subject (pos 0) and object (pos 2) columns.
    bool hasHistograms() const { return !depth_[0].lowers.empty() && !depth_[2].lowers.empty(); }
    const KeyHistogram &depthHistogram(int pos) const { return depth_[pos]; }
    const KeyHistogram &widthHistogram(int pos) const { return width_[pos]; }
    // sum over keys v of s(v)^a o(v)^b (key shares from the equal-depth
    // histograms); 1 for a == b == 0 is not a key count but the empty sum's
    // neutral value used by the estimator.
    double moment(int a, int b) const;

  private:
    uint64_t n_ = 0;
    bool built_[VD_NUM_PERMS] = {};
    std::vector<TYPEID> keys_[VD_NUM_PERMS];
    std::vector<uint64_t> rows_[VD_NUM_PERMS], pairs_[VD_NUM_PERMS];
    KeyHistogram depth_[3], width_[3];
    mutable std::mutex mu_;
    mutable std::map<std::pair<int, int>, double> moments_;
};

struct PlanEstimate {
    double rows = 0;        // expected rows of the final relation
    double peak_bytes = 0;  // expected peak HBM bytes of the plan
    double work_rows = 0;   // sum of expected rows over all operators
};

class PlanEstimator {
  public:
    // driven: the plan runs on the one-query path (QueryPlan), whose joins
    // followed by an IndexSwap may be driven joins; the batched executor
    // has none.
    explicit PlanEstimator(const Statistics &st, bool driven = true)
        : st_(st), driven_(driven) {}
    PlanEstimate estimate(const std::vector<TriplePattern> &patterns, PlanTreeNode *root) const;

  private:
    struct Var {
        double distinct = 1;
        int a = -1, b = -1;  // subject / object occurrences (-1: unknown)
    };
    struct Est {
        double rows = 0;
        std::map<std::string, Var> vars;
        int cols = 0;
        std::string key;    // the variable of column 0
        bool view = false;  // rows of an index permutation (no device memory)
    };
    // swapped_next: an IndexSwap re-sorts this node's output next
    Est node(PlanTreeNode *n, const std::vector<TriplePattern> &pats, PlanEstimate &acc,
             double &live, bool swapped_next = false) const;
    Est select(const TriplePattern &tp, const std::string &var) const;
    void rekey(Est &e, const std::string &var, PlanEstimate &acc, double &live) const;
    const Statistics &st_;
    bool driven_;
};

#endif
