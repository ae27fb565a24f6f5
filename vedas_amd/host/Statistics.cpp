#include "Statistics.h"

#include <algorithm>
#include <cmath>

Statistics::Statistics(VedasStorage *storage) {
    n_ = storage->getTripleSize();
    VedasContext *ctx = storage->context();
    vd_ctx *c = ctx->raw();
    // Predicate-first permutations: per-predicate counts are what size
    // estimates need, and there are few predicates.
    for (int perm : {VD_PERM_PSO, VD_PERM_POS}) {
        if (!storage->hasPermutation(perm)) continue;
        built_[perm] = true;
        vd_perm_view_t v = storage->view(perm);
        const uint64_t m = v.n_l1;
        keys_[perm].resize(m);
        rows_[perm].resize(m);
        pairs_[perm].resize(m);
        if (m == 0) continue;
        void *dr = nullptr, *dp = nullptr;
        ctx->check(vd_dev_alloc(c, m * 8, &dr), "vd_dev_alloc");
        ctx->check(vd_dev_alloc(c, m * 8, &dp), "vd_dev_alloc");
        try {
            ctx->check(vd_index_stats(c, storage->index(), perm, (uint64_t *)dr, (uint64_t *)dp),
                       "vd_index_stats");
            ctx->check(vd_copy_d2h(c, keys_[perm].data(), v.l1_values, m * 4), "vd_copy_d2h");
            ctx->check(vd_copy_d2h(c, rows_[perm].data(), dr, m * 8), "vd_copy_d2h");
            ctx->check(vd_copy_d2h(c, pairs_[perm].data(), dp, m * 8), "vd_copy_d2h");
        } catch (...) {
            vd_dev_free(c, dr);
            vd_dev_free(c, dp);
            throw;
        }
        vd_dev_free(c, dr);
        vd_dev_free(c, dp);
    }
}

uint64_t Statistics::rows(int perm, TYPEID x) const {
    const auto &k = keys_[perm];
    auto it = std::lower_bound(k.begin(), k.end(), x);
    return (it != k.end() && *it == x) ? rows_[perm][it - k.begin()] : 0;
}

uint64_t Statistics::pairs(int perm, TYPEID x) const {
    const auto &k = keys_[perm];
    auto it = std::lower_bound(k.begin(), k.end(), x);
    return (it != k.end() && *it == x) ? pairs_[perm][it - k.begin()] : 0;
}

PlanEstimator::Est PlanEstimator::select(const TriplePattern &tp) const {
    Est e;
    const bool sv = tp.subjectIsVariable(), pv = tp.predicateIsVariable(),
               ov = tp.objectIsVariable();
    if (pv) {
        // predicate variables: no per-key statistics kept; a coarse share
        e.rows = std::max(1.0, (double)st_.triples() * 1e-3);
        for (auto t : {tp.getSubject(), tp.getPredicate(), tp.getObject()})
            if (!t.empty() && t[0] == '?') e.distinct[t] = e.rows, e.cols++;
        return e;
    }
    const TYPEID p = tp.getPredicateId();
    const double cnt = (double)st_.rows(VD_PERM_PSO, p);
    const double ds = std::max(1.0, (double)st_.pairs(VD_PERM_PSO, p));
    const double dob = std::max(1.0, (double)st_.pairs(VD_PERM_POS, p));
    if (sv && ov) {
        e.rows = cnt;
        e.distinct[tp.getSubject()] = ds;
        e.distinct[tp.getObject()] = dob;
        e.cols = 2;
    } else if (sv) {
        e.rows = cnt / dob;  // triples per object of p
        e.distinct[tp.getSubject()] = e.rows;
        e.cols = 1;
    } else if (ov) {
        e.rows = cnt / ds;  // triples per subject of p
        e.distinct[tp.getObject()] = e.rows;
        e.cols = 1;
    } else {
        e.rows = 1;
    }
    return e;
}

PlanEstimator::Est PlanEstimator::node(PlanTreeNode *n, const std::vector<TriplePattern> &pats,
                                       PlanEstimate &acc, double &live) const {
    if (n->op == UPLOAD) {
        Est e = select(*n->tp);
        acc.work_rows += e.rows;
        live += e.rows * e.cols * 4.0;
        acc.peak_bytes = std::max(acc.peak_bytes, live);
        return e;
    }
    if (n->op == INDEXSWAP) {
        Est e = node(n->child1, pats, acc, live);
        // sort keys, their buffer and fresh columns while the input lives
        const double tmp = e.rows * (16.0 + 4.0 * e.cols);
        acc.peak_bytes = std::max(acc.peak_bytes, live + tmp);
        acc.work_rows += e.rows;
        return e;
    }
    Est a = node(n->child1, pats, acc, live);
    Est b = node(n->child2, pats, acc, live);
    const std::string &v = n->var;
    const double da = a.distinct.count(v) ? a.distinct.at(v) : std::max(1.0, a.rows);
    const double db = b.distinct.count(v) ? b.distinct.at(v) : std::max(1.0, b.rows);
    Est e;
    e.rows = a.rows * b.rows / std::max(1.0, std::max(da, db));
    e.cols = a.cols + b.cols - 1;
    for (auto *x : {&a, &b})
        for (auto &kv : x->distinct) {
            const double d = std::min(kv.second, std::max(1.0, e.rows));
            auto it = e.distinct.find(kv.first);
            e.distinct[kv.first] = (it == e.distinct.end()) ? d : std::min(it->second, d);
        }
    e.distinct[v] = std::min(da, db);
    const double left = std::min(a.rows, b.rows);
    const double temps = left * 24.0 + (a.rows + b.rows) * 16.0;  // counts/offsets, key re-sorts
    const double out = e.rows * e.cols * 4.0;
    acc.peak_bytes = std::max(acc.peak_bytes, live + temps + out);
    live += out - (a.rows * a.cols + b.rows * b.cols) * 4.0;
    live = std::max(live, 0.0);
    acc.work_rows += e.rows;
    return e;
}

PlanEstimate PlanEstimator::estimate(const std::vector<TriplePattern> &patterns,
                                     PlanTreeNode *root) const {
    PlanEstimate acc;
    double live = 0;
    if (!root) {
        if (patterns.empty()) return acc;
        Est e = select(patterns[0]);
        acc.rows = e.rows;
        acc.work_rows = e.rows;
        acc.peak_bytes = e.rows * e.cols * 4.0;
        return acc;
    }
    Est e = node(root, patterns, acc, live);
    acc.rows = e.rows;
    return acc;
}
