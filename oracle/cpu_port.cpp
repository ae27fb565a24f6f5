// cpu_port.cpp -- multi-threaded host port of the VEDAS query operators,
// the CPU baseline the benchmark times beside the GPU (the reference's CPU
// path is Thrust's host / OpenMP backend; ModernGPU's join has no host
// backend, so the join is a transcription of its algorithm).
//
// TEST AND BASELINE INFRASTRUCTURE ONLY: used by tests/ and bench.py's
// cpu_baseline leg, never by the vedas_amd package.  Results are bit-equal
// to the oracle's (vedas_oracle.c) for every thread count; the tests in
// tests/test_cpu_port.py check that.
//
//   cp_select        SelectQueryJob::startJob range copy with the bound
//                    clip (SelectQueryJob.cu:86-270)
//   cp_join_bounds / cp_join_expand
//                    mgpu::inner_join as mirrored in JoinQueryJob.cu:68-101
//                    (per-left-row lower / upper bound, exclusive scan of
//                    the counts) and the gather of :322-380
//   cp_sort_col0     stable sort by column 0 (FullRelationIR.cu:594-713):
//                    LSD radix sort of (key, row) with per-thread digit
//                    histograms, then a gather of every column
//   cp_unique_adjacent  FullRelationIR::removeDuplicate (:338-522)
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

template <class F>
void par(int nt, F &&f) {
    if (nt < 1) nt = 1;
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(f, t, nt);
    f(0, nt);
    for (auto &x : th) x.join();
}

inline uint64_t lo_of(uint64_t n, int t, int nt) { return n * (uint64_t)t / (uint64_t)nt; }

inline uint64_t lb32(const uint32_t *a, uint64_t lo, uint64_t hi, uint32_t v) {
    while (lo < hi) {
        const uint64_t m = lo + (hi - lo) / 2;
        if (a[m] < v) lo = m + 1; else hi = m;
    }
    return lo;
}

inline uint64_t ub32(const uint32_t *a, uint64_t lo, uint64_t hi, uint32_t v) {
    while (lo < hi) {
        const uint64_t m = lo + (hi - lo) / 2;
        if (v < a[m]) hi = m; else lo = m + 1;
    }
    return lo;
}

void par_copy(uint32_t *dst, const uint32_t *src, uint64_t n, int nt) {
    if (n < (1u << 20)) nt = 1;
    par(nt, [&](int t, int k) {
        const uint64_t a = lo_of(n, t, k), b = lo_of(n, t + 1, k);
        if (b > a) std::memcpy(dst + a, src + a, (b - a) * 4);
    });
}

}  // namespace

extern "C" {

// Rows [b, e) of a permutation's Y / Z columns, clipped to Y in [lo, hi]
// when clip: cols 1 -> Z, 2 -> Y, 3 -> Y and Z.  Returns the row count;
// *first receives the first row copied.  With out0 == NULL only the range
// is computed.
uint64_t cp_select(const uint32_t *y, const uint32_t *z, uint64_t b, uint64_t e, int cols,
                   int clip, uint32_t lo, uint32_t hi, uint32_t *out0, uint32_t *out1,
                   uint64_t *first, int nthreads) {
    if (clip) {
        if (lo > hi) {
            e = b;
        } else {
            const uint64_t nb = lb32(y, b, e, lo);
            e = ub32(y, nb, e, hi);
            b = nb;
        }
    }
    if (first) *first = b;
    const uint64_t m = e - b;
    if (!out0) return m;
    if (cols == 1) {
        par_copy(out0, z + b, m, nthreads);
    } else {
        par_copy(out0, y + b, m, nthreads);
        if (cols == 3) par_copy(out1, z + b, m, nthreads);
    }
    return m;
}

// Join on column 0: lower[i] = first right row with key >= L[i], off[i] =
// output rows before left row i (off[nl] = total).  Returns the total.
uint64_t cp_join_bounds(const uint32_t *lk, uint64_t nl, const uint32_t *rk, uint64_t nr,
                        uint64_t *lower, uint64_t *off, int nthreads) {
    if (nl < 4096) nthreads = 1;
    std::vector<uint64_t> part(nthreads + 1, 0);
    par(nthreads, [&](int t, int k) {
        const uint64_t a = lo_of(nl, t, k), b = lo_of(nl, t + 1, k);
        uint64_t acc = 0;
        uint64_t hint = a < b ? lb32(rk, 0, nr, lk[a]) : 0;
        for (uint64_t i = a; i < b; ++i) {
            // left keys ascend: each search starts where the previous ended
            const uint64_t l = (i > a && lk[i] == lk[i - 1]) ? lower[i - 1] : lb32(rk, hint, nr, lk[i]);
            const uint64_t u = ub32(rk, l, nr, lk[i]);
            lower[i] = l;
            off[i] = u - l;
            acc += u - l;
            hint = l;
        }
        part[t + 1] = acc;
    });
    for (int t = 0; t < nthreads; ++t) part[t + 1] += part[t];
    par(nthreads, [&](int t, int k) {
        const uint64_t a = lo_of(nl, t, k), b = lo_of(nl, t + 1, k);
        uint64_t acc = part[t];
        for (uint64_t i = a; i < b; ++i) {
            const uint64_t c = off[i];
            off[i] = acc;
            acc += c;
        }
    });
    off[nl] = part[nthreads];
    return part[nthreads];
}

// Output rows: for each left row in order its matches in right order,
// columns [key, left 1.., right 1..].
void cp_join_expand(const uint32_t *const *L, uint64_t nl, int cl, const uint32_t *const *R,
                    int cr, const uint64_t *lower, const uint64_t *off, uint32_t *const *out,
                    int nthreads) {
    const uint64_t total = off[nl];
    if (total < (1u << 16)) nthreads = 1;
    par(nthreads, [&](int t, int k) {
        // this thread's output rows [oa, ob); its first left row by search
        const uint64_t oa = lo_of(total, t, k), ob = lo_of(total, t + 1, k);
        if (oa >= ob) return;
        uint64_t i = std::upper_bound(off, off + nl + 1, oa) - off - 1;
        for (uint64_t w = oa; w < ob; ++i) {
            const uint64_t s0 = off[i], s1 = off[i + 1];
            if (s1 <= w) continue;
            const uint64_t r0 = lower[i] + (w - s0);
            const uint64_t cnt = std::min(s1, ob) - w;
            for (int c = 0; c < cl; ++c) {
                const uint32_t v = L[c][i];
                uint32_t *o = out[c] + w;
                for (uint64_t j = 0; j < cnt; ++j) o[j] = v;
            }
            for (int c = 1; c < cr; ++c)
                std::memcpy(out[cl + c - 1] + w, R[c] + r0, cnt * 4);
            w += cnt;
        }
    });
}

// Stable sort of the rows of `cols` by column 0 (in place).
int cp_sort_col0(uint32_t *const *cols, int ncols, uint64_t n, int nthreads) {
    if (n <= 1) return 0;
    if (n < (1u << 16)) nthreads = 1;
    const uint32_t *key = cols[0];
    uint32_t kmax = 0;
    {
        std::vector<uint32_t> mx(nthreads, 0);
        par(nthreads, [&](int t, int k) {
            uint32_t m = 0;
            for (uint64_t i = lo_of(n, t, k); i < lo_of(n, t + 1, k); ++i) m = std::max(m, key[i]);
            mx[t] = m;
        });
        for (uint32_t m : mx) kmax = std::max(kmax, m);
    }
    std::vector<uint64_t> a(n), b(n);  // (key << 32 | row) pairs
    par(nthreads, [&](int t, int k) {
        for (uint64_t i = lo_of(n, t, k); i < lo_of(n, t + 1, k); ++i)
            a[i] = ((uint64_t)key[i] << 32) | (uint32_t)i;
    });
    const int nb = 256;
    std::vector<uint64_t> hist((size_t)nthreads * nb);
    for (int shift = 32; shift < 64 && (kmax >> (shift - 32)) != 0; shift += 8) {
        std::fill(hist.begin(), hist.end(), 0);
        par(nthreads, [&](int t, int k) {
            uint64_t *h = &hist[(size_t)t * nb];
            for (uint64_t i = lo_of(n, t, k); i < lo_of(n, t + 1, k); ++i) ++h[(a[i] >> shift) & 255];
        });
        uint64_t acc = 0;
        for (int d = 0; d < nb; ++d)
            for (int t = 0; t < nthreads; ++t) {
                const uint64_t v = hist[(size_t)t * nb + d];
                hist[(size_t)t * nb + d] = acc;
                acc += v;
            }
        par(nthreads, [&](int t, int k) {
            uint64_t *h = &hist[(size_t)t * nb];
            for (uint64_t i = lo_of(n, t, k); i < lo_of(n, t + 1, k); ++i)
                b[h[(a[i] >> shift) & 255]++] = a[i];
        });
        a.swap(b);
    }
    std::vector<uint32_t> tmp(n);
    for (int c = 0; c < ncols; ++c) {
        uint32_t *col = cols[c];
        par(nthreads, [&](int t, int k) {
            for (uint64_t i = lo_of(n, t, k); i < lo_of(n, t + 1, k); ++i) tmp[i] = col[(uint32_t)a[i]];
        });
        par_copy(col, tmp.data(), n, nthreads);
    }
    return 0;
}

// Drop rows equal in every column to the row before them; returns the new
// row count (columns compacted in place).
uint64_t cp_unique_adjacent(uint32_t *const *cols, int ncols, uint64_t n, int nthreads) {
    if (n <= 1) return n;
    if (n < (1u << 16)) nthreads = 1;
    std::vector<uint8_t> keep(n);
    std::vector<uint64_t> part(nthreads + 1, 0);
    par(nthreads, [&](int t, int k) {
        uint64_t c = 0;
        for (uint64_t i = lo_of(n, t, k); i < lo_of(n, t + 1, k); ++i) {
            bool diff = i == 0;
            for (int j = 0; j < ncols && !diff; ++j) diff = cols[j][i] != cols[j][i - 1];
            keep[i] = diff;
            c += diff;
        }
        part[t + 1] = c;
    });
    for (int t = 0; t < nthreads; ++t) part[t + 1] += part[t];
    const uint64_t m = part[nthreads];
    if (m == n) return n;
    std::vector<uint32_t> tmp(m);
    for (int j = 0; j < ncols; ++j) {
        par(nthreads, [&](int t, int k) {
            uint64_t w = part[t];
            for (uint64_t i = lo_of(n, t, k); i < lo_of(n, t + 1, k); ++i)
                if (keep[i]) tmp[w++] = cols[j][i];
        });
        par_copy(cols[j], tmp.data(), m, nthreads);
    }
    return m;
}

}  // extern "C"
