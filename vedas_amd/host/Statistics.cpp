#include "Statistics.h"

#include <algorithm>
#include <cmath>

uint64_t KeyHistogram::rows() const {
    uint64_t r = 0;
    for (uint64_t f : freqs) r += f;
    return r;
}

namespace {

KeyHistogram load_histogram(VedasContext *ctx, vd_index *idx, int perm, int kind, unsigned param) {
    KeyHistogram h;
    uint64_t cap = 1 << 16, nb = 0;
    for (;;) {
        h.lowers.resize(cap);
        h.uppers.resize(cap);
        h.freqs.resize(cap);
        const int rc = vd_index_histogram(ctx->raw(), idx, perm, kind, param, 0xFFFFFFFFu,
                                          h.lowers.data(), h.uppers.data(), h.freqs.data(), cap,
                                          &nb);
        if (rc == VD_ERANGE && nb > cap) {
            cap = nb;
            continue;
        }
        ctx->check(rc, "vd_index_histogram");
        break;
    }
    h.lowers.resize(nb);
    h.uppers.resize(nb);
    h.freqs.resize(nb);
    return h;
}

// Key share per ID inside each bucket: freq / (rows * IDs in the bucket).
struct Piece {
    double lo, hi;  // IDs [lo, hi]
    double share;   // per ID
};

std::vector<Piece> pieces(const KeyHistogram &h) {
    std::vector<Piece> p;
    const double n = (double)std::max<uint64_t>(1, h.rows());
    for (size_t i = 0; i < h.freqs.size(); ++i) {
        const double span = (double)h.uppers[i] - (double)h.lowers[i] + 1.0;
        p.push_back(Piece{(double)h.lowers[i], (double)h.uppers[i], (double)h.freqs[i] / (n * span)});
    }
    return p;
}

}  // namespace

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
    // Subject and object histograms from a subject-first and an
    // object-first permutation.
    const int spos[2][2] = {{VD_PERM_SPO, VD_PERM_SOP}, {VD_PERM_OSP, VD_PERM_OPS}};
    for (int k = 0; k < 2; ++k) {
        int perm = -1;
        for (int q : spos[k])
            if (perm < 0 && storage->hasPermutation(q)) perm = q;
        if (perm < 0 || n_ == 0) continue;
        const int pos = k == 0 ? 0 : 2;
        depth_[pos] = load_histogram(ctx, storage->index(), perm, VD_HIST_EQUAL_DEPTH, kDepth);
        width_[pos] = load_histogram(ctx, storage->index(), perm, VD_HIST_EQUAL_WIDTH, kWidth);
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

double Statistics::moment(int a, int b) const {
    if (a == 0 && b == 0) return 1.0;
    std::lock_guard<std::mutex> lk(mu_);
    auto it = moments_.find({a, b});
    if (it != moments_.end()) return it->second;
    const std::vector<Piece> S = pieces(depth_[0]), O = pieces(depth_[2]);
    double m = 0;
    if (b == 0 || a == 0) {
        const auto &P = b == 0 ? S : O;
        const int e = b == 0 ? a : b;
        for (const Piece &p : P) m += (p.hi - p.lo + 1.0) * std::pow(p.share, e);
    } else {
        // keys where both shares are non-zero: overlaps of the two bucket lists
        size_t i = 0, j = 0;
        while (i < S.size() && j < O.size()) {
            const double lo = std::max(S[i].lo, O[j].lo), hi = std::min(S[i].hi, O[j].hi);
            if (hi >= lo) m += (hi - lo + 1.0) * std::pow(S[i].share, a) * std::pow(O[j].share, b);
            if (S[i].hi < O[j].hi) ++i; else ++j;
        }
    }
    moments_[{a, b}] = m;
    return m;
}

PlanEstimator::Est PlanEstimator::select(const TriplePattern &tp, const std::string &var) const {
    Est e;
    e.view = true;
    const bool sv = tp.subjectIsVariable(), pv = tp.predicateIsVariable(),
               ov = tp.objectIsVariable();
    auto put = [&](const std::string &v, double d, int a, int b) {
        Var x;
        x.distinct = std::max(1.0, d);
        x.a = a;
        x.b = b;
        e.vars[v] = x;
    };
    if (pv) {
        // predicate variables: no per-key statistics kept; a coarse share
        e.rows = tp.estimate_rows ? (double)tp.estimate_rows
                                  : std::max(1.0, (double)st_.triples() * 1e-3);
        if (sv) put(tp.getSubject(), e.rows, 1, 0);
        put(tp.getPredicate(), e.rows, -1, -1);
        if (ov) put(tp.getObject(), e.rows, 0, 1);
        e.cols = (int)e.vars.size();
        e.key = var.empty() ? e.vars.begin()->first : var;
        return e;
    }
    const TYPEID p = tp.getPredicateId();
    const double cnt = (double)st_.rows(VD_PERM_PSO, p);
    const double ds = std::max(1.0, (double)st_.pairs(VD_PERM_PSO, p));
    const double dob = std::max(1.0, (double)st_.pairs(VD_PERM_POS, p));
    if (sv && ov) {
        e.rows = cnt;
        put(tp.getSubject(), ds, 1, 0);
        put(tp.getObject(), dob, 0, 1);
        e.cols = 2;
        e.key = var == tp.getObject() ? tp.getObject() : tp.getSubject();
    } else if (sv) {
        e.rows = tp.estimate_rows ? (double)tp.estimate_rows : cnt / dob;  // per object of p
        put(tp.getSubject(), e.rows, 1, 0);
        e.cols = 1;
        e.key = tp.getSubject();
    } else if (ov) {
        e.rows = tp.estimate_rows ? (double)tp.estimate_rows : cnt / ds;  // per subject of p
        put(tp.getObject(), e.rows, 0, 1);
        e.cols = 1;
        e.key = tp.getObject();
    } else {
        e.rows = 1;
    }
    return e;
}

// Column `var` to the front with a re-sort: a view is first copied, and the
// sort's keys, key buffer and fresh columns live while the input does.
void PlanEstimator::rekey(Est &e, const std::string &var, PlanEstimate &acc,
                          double &live) const {
    if (e.key == var) return;
    const double bytes = e.rows * e.cols * 4.0;
    const double tmp = e.rows * (16.0 + 4.0 * e.cols);
    if (e.view) live += bytes;
    acc.peak_bytes = std::max(acc.peak_bytes, live + tmp);
    e.key = var;
    e.view = false;
    acc.work_rows += e.rows;
}

PlanEstimator::Est PlanEstimator::node(PlanTreeNode *n, const std::vector<TriplePattern> &pats,
                                       PlanEstimate &acc, double &live,
                                       bool swapped_next) const {
    if (n->op == UPLOAD) {
        Est e = select(*n->tp, n->var);
        acc.work_rows += e.rows;  // a view: no device memory
        return e;
    }
    if (n->op == INDEXSWAP) {
        Est e = node(n->child1, pats, acc, live, true);
        rekey(e, n->var, acc, live);
        return e;
    }
    Est a = node(n->child1, pats, acc, live);
    Est b = node(n->child2, pats, acc, live);
    const std::string &v = n->var;
    if (b.rows < a.rows) std::swap(a, b);  // the smaller relation is the left one
    rekey(a, v, acc, live);
    rekey(b, v, acc, live);
    const Var va = a.vars.count(v) ? a.vars.at(v) : Var{std::max(1.0, a.rows)};
    const Var vb = b.vars.count(v) ? b.vars.at(v) : Var{std::max(1.0, b.rows)};
    Est e;
    Var ve;
    if (st_.hasHistograms() && va.a >= 0 && vb.a >= 0) {
        const double den = st_.moment(va.a, va.b) * st_.moment(vb.a, vb.b);
        const double num = st_.moment(va.a + vb.a, va.b + vb.b);
        e.rows = den > 0 ? a.rows * b.rows * num / den : 0.0;
        ve.a = va.a + vb.a;
        ve.b = va.b + vb.b;
    } else {
        e.rows = a.rows * b.rows / std::max(1.0, std::max(va.distinct, vb.distinct));
    }
    e.cols = a.cols + b.cols - 1;
    for (auto *x : {&a, &b})
        for (auto &kv : x->vars) {
            Var d = kv.second;
            d.distinct = std::min(d.distinct, std::max(1.0, e.rows));
            auto it = e.vars.find(kv.first);
            if (it == e.vars.end()) e.vars[kv.first] = d;
            else it->second.distinct = std::min(it->second.distinct, d.distinct);
        }
    ve.distinct = std::min(va.distinct, vb.distinct);
    e.vars[v] = ve;
    e.key = v;
    // first-match and offset arrays of the left rows, the output columns;
    // a join an IndexSwap follows may be driven (the driver's first-match
    // and offset arrays and a run table of at most 16 bytes per input row,
    // for inputs of 2^20 rows or more)
    double temps = a.rows * 8.0;
    if (driven_ && swapped_next && a.rows + b.rows >= (double)(1u << 20))
        temps = std::max(temps, 8.0 * std::max(a.rows, b.rows) + 16.0 * (a.rows + b.rows));
    const double out = e.rows * e.cols * 4.0;
    acc.peak_bytes = std::max(acc.peak_bytes, live + temps + out);
    for (auto *x : {&a, &b})
        if (!x->view) live -= x->rows * x->cols * 4.0;
    live = std::max(live, 0.0) + out;
    acc.work_rows += e.rows;
    return e;
}

PlanEstimate PlanEstimator::estimate(const std::vector<TriplePattern> &patterns,
                                     PlanTreeNode *root) const {
    PlanEstimate acc;
    double live = 0;
    if (!root) {
        if (patterns.empty()) return acc;
        Est e = select(patterns[0], "");
        acc.rows = e.rows;
        acc.work_rows = e.rows;
        acc.peak_bytes = 0;  // a view
        return acc;
    }
    Est e = node(root, patterns, acc, live);
    acc.rows = e.rows;
    // the finalise step's unique writes a copy of the result
    acc.peak_bytes = std::max(acc.peak_bytes, live + e.rows * e.cols * 4.0);
    return acc;
}
