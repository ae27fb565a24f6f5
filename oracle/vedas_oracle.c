/*
 * vedas_oracle.c -- CPU restatement of the VEDAS query path, used ONLY as
 * the checker: tests, the smoke check and the bench's CPU-baseline leg call
 * it; the product (libvedas_hip.so / libvedas.so) never links or calls it.
 *
 * Plain C, single thread, written for clarity rather than speed.  Each
 * function restates the reference (Remixman/Vedas) operator it names.
 * Pinning (what checks each function, independently of this code):
 *   or_build_perm   - Python sorted() of tuples + itertools.groupby run
 *                     starts (tests/test_oracle.py), golden fixtures.
 *   or_lookup       - brute-force linear scan of all triples.
 *   or_select       - brute-force filter of the linear scan.
 *   or_join         - nested-loop join over all row pairs.
 *   or_sort_col0    - Python sorted() with a key (stable by definition).
 *   or_sort_lex     - Python sorted() of row tuples.
 *   or_unique_adjacent - itertools.groupby.
 *   end to end      - a naive nested-loop BGP evaluator over the raw
 *                     triples (set of bindings) in tests/test_oracle.py.
 * The reference's own tests pin none of this path (its only unit tests
 * cover the empty-interval dictionary), and its Thrust / ModernGPU
 * boundary cannot be built here (no nvcc, ModernGPU not vendored).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Permutation column order, same numbering as VD_PERM_* (0=SPO, 1=SOP,
 * 2=PSO, 3=POS, 4=OSP, 5=OPS); entries index (s, p, o). */
static const int PERM_COLS[6][3] = {
    {0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

typedef struct { uint32_t x, y, z; } trip_t;

static int cmp_trip(const void *a, const void *b) {
    const trip_t *u = (const trip_t *)a, *v = (const trip_t *)b;
    if (u->x != v->x) return u->x < v->x ? -1 : 1;
    if (u->y != v->y) return u->y < v->y ? -1 : 1;
    if (u->z != v->z) return u->z < v->z ? -1 : 1;
    return 0;
}

/* Index build of one permutation.  Restates VedasStorage::createIndex1Level
 * and createIndex (VedasStorage.cu:270-289, 306-358): sort (X,Y,Z) tuples
 * lexicographically (thrust::sort on a zip = tuple operator<), keep
 * duplicates; L1 = reduce_by_key over X + exclusive scan (first row of each
 * distinct X); L2 = reduce_by_key over (X,Y) + exclusive scan, storing Y
 * and the first row; Y and Z columns are xy_data / xyz_data
 * (VedasStorage.cu:40-48, 51-70).  Arrays l1v..z hold n entries each. */
int or_build_perm(const uint32_t *s, const uint32_t *p, const uint32_t *o,
                  uint64_t n, int perm, uint32_t *l1v, uint32_t *l1o,
                  uint64_t *n_l1, uint32_t *l2v, uint32_t *l2o, uint64_t *n_l2,
                  uint32_t *y, uint32_t *z) {
    if (perm < 0 || perm > 5) return -1;
    const uint32_t *cols[3] = {s, p, o};
    const uint32_t *cx = cols[PERM_COLS[perm][0]], *cy = cols[PERM_COLS[perm][1]],
                   *cz = cols[PERM_COLS[perm][2]];
    trip_t *t = (trip_t *)malloc((n ? n : 1) * sizeof(trip_t));
    if (!t) return -2;
    for (uint64_t i = 0; i < n; ++i) {
        t[i].x = cx[i]; t[i].y = cy[i]; t[i].z = cz[i];
    }
    qsort(t, n, sizeof(trip_t), cmp_trip);
    uint64_t a = 0, b = 0;
    for (uint64_t i = 0; i < n; ++i) {
        y[i] = t[i].y;
        z[i] = t[i].z;
        if (i == 0 || t[i].x != t[i - 1].x) {
            l1v[a] = t[i].x; l1o[a] = (uint32_t)i; ++a;
        }
        if (i == 0 || t[i].x != t[i - 1].x || t[i].y != t[i - 1].y) {
            l2v[b] = t[i].y; l2o[b] = (uint32_t)i; ++b;
        }
    }
    *n_l1 = a;
    *n_l2 = b;
    free(t);
    return 0;
}

static uint64_t lb32(const uint32_t *a, uint64_t lo, uint64_t hi, uint32_t v) {
    while (lo < hi) {
        uint64_t m = lo + (hi - lo) / 2;
        if (a[m] < v) lo = m + 1; else hi = m;
    }
    return lo;
}

static uint64_t ub32(const uint32_t *a, uint64_t lo, uint64_t hi, uint32_t v) {
    while (lo < hi) {
        uint64_t m = lo + (hi - lo) / 2;
        if (v < a[m]) hi = m; else lo = m + 1;
    }
    return lo;
}

/* Row range of X == k1 (and Y == k2).  Restates
 * QueryExecutor::findL2OffsetFromL1 (QueryExecutor.cu:596-608) and the
 * L2-index form of findDataOffsetFromL2 (:610-633): lower_bound on the L1
 * values, group end = next L1 offset or n; for k2, the group's L2 entries
 * are found by lower_bound on the L2 offsets, then lower/upper_bound on
 * their Y values.  Two reference edge cases are defined here: an absent k1
 * gives an empty range (the reference takes the next key's group), and an
 * L2 position one past the last entry maps to n (the reference reads
 * offsets[end] out of bounds). */
void or_lookup(const uint32_t *l1v, const uint32_t *l1o, uint64_t n_l1,
               const uint32_t *l2v, const uint32_t *l2o, uint64_t n_l2,
               uint64_t n, uint32_t k1, int has_k2, uint32_t k2,
               uint64_t *begin, uint64_t *end) {
    uint64_t i = lb32(l1v, 0, n_l1, k1);
    if (i >= n_l1 || l1v[i] != k1) {
        *begin = *end = (i < n_l1) ? l1o[i] : n;
        return;
    }
    uint64_t b = l1o[i];
    uint64_t e = (i + 1 < n_l1) ? l1o[i + 1] : n;
    if (has_k2) {
        uint64_t g0 = lb32(l2o, 0, n_l2, (uint32_t)b);
        uint64_t g1 = lb32(l2o, 0, n_l2, (uint32_t)e);
        if (e >= n) g1 = n_l2;
        uint64_t j0 = lb32(l2v, g0, g1, k2);
        uint64_t j1 = ub32(l2v, g0, g1, k2);
        b = (j0 < n_l2) ? l2o[j0] : n;
        e = (j1 < n_l2) ? l2o[j1] : n;
        if (e < b) e = b;
    }
    *begin = b;
    *end = e;
}

/* Select of rows [b, e) of a permutation.  Restates SelectQueryJob::startJob
 * (SelectQueryJob.cu:86-270): cols 1 = the Z column alone (1-variable
 * select, :92-121); cols 3 = (Y, Z) (2-variable select); cols 2 = Y alone
 * (second variable unused, :151-167) with thrust::unique when distinct
 * (:158).  Bound clip (:130-141): lower_bound(lo) / upper_bound(hi) over Y.
 * Skip intervals: EmptyIntervalDict::getUploadRelationData
 * (EmptyIntervalDict.cu:108-170) -- rows with Y in [lo_k, hi_k] removed,
 * found by lower_bound(lo_k) / upper_bound(hi_k).  skip holds nskip
 * (lo, hi) pairs, sorted by lo and disjoint.  Returns the row count. */
uint64_t or_select(const uint32_t *ycol, const uint32_t *zcol, uint64_t b,
                   uint64_t e, int cols, int clip, uint32_t lo, uint32_t hi,
                   const uint32_t *skip, int nskip, int distinct,
                   uint32_t *out0, uint32_t *out1) {
    if (clip) {
        if (lo > hi) {
            e = b;
        } else {
            uint64_t nb = lb32(ycol, b, e, lo);
            e = ub32(ycol, nb, e, hi);
            b = nb;
        }
    }
    uint64_t m = 0;
    uint64_t cur = b;
    for (int k = 0; k <= nskip; ++k) {
        uint64_t stop = e, next = e;
        if (k < nskip) {
            if (skip[2 * k] > skip[2 * k + 1]) continue;
            stop = lb32(ycol, cur, e, skip[2 * k]);
            next = ub32(ycol, stop, e, skip[2 * k + 1]);
        }
        for (uint64_t r = cur; r < stop; ++r) {
            if (cols == 1) {
                out0[m] = zcol[r];
            } else if (cols == 2) {
                out0[m] = ycol[r];
            } else {
                out0[m] = ycol[r];
                out1[m] = zcol[r];
            }
            ++m;
        }
        cur = next;
    }
    if (distinct && cols == 2 && m > 1) {
        uint64_t w = 1;
        for (uint64_t r = 1; r < m; ++r)
            if (out0[r] != out0[w - 1]) out0[w++] = out0[r];
        m = w;
    }
    return m;
}

static inline uint64_t key_at(const uint32_t *const *c, int nk, uint64_t i) {
    return nk == 1 ? (uint64_t)c[0][i] : (((uint64_t)c[0][i] << 32) | c[1][i]);
}

/* Bag inner join on the first nk columns (1 or 2; for 2 the key is
 * c0 << 32 | c1 as in JoinQueryJob::mergeColumns, JoinQueryJob.cu:111-118).
 * Restates mgpu::inner_join as mirrored in JoinQueryJob::innerJoinMulti
 * (JoinQueryJob.cu:68-101): lower and upper bound of each left key in the
 * right keys, output pairs (i, lower_i + r) for r < upper_i - lower_i in
 * left-major order; then the gather of JoinQueryJob.cu:322-380 with output
 * columns [keys, left columns after the keys, right columns after the
 * keys] (:287-292).  With out == NULL only counts.  Returns rows. */
uint64_t or_join(const uint32_t *const *L, uint64_t nl, int cl,
                 const uint32_t *const *R, uint64_t nr, int cr, int nk,
                 uint32_t *const *out) {
    uint64_t w = 0;
    for (uint64_t i = 0; i < nl; ++i) {
        const uint64_t k = key_at(L, nk, i);
        uint64_t lo = 0, hi = nr;
        while (lo < hi) {
            uint64_t m = lo + (hi - lo) / 2;
            if (key_at(R, nk, m) < k) lo = m + 1; else hi = m;
        }
        uint64_t up = lo, top = nr;
        while (up < top) {
            uint64_t m = up + (top - up) / 2;
            if (k < key_at(R, nk, m)) top = m; else up = m + 1;
        }
        for (uint64_t j = lo; j < up; ++j, ++w) {
            if (!out) continue;
            int oc = 0;
            for (int c = 0; c < cl; ++c) out[oc++][w] = L[c][i];
            for (int c = nk; c < cr; ++c) out[oc++][w] = R[c][j];
        }
    }
    return w;
}

/* Stable sort of all rows by column 0 (FullRelationIR::sortByFirstColumn,
 * FullRelationIR.cu:594-713, as used by IndexSwapJob.cu:28-95).  The
 * reference's thrust::sort with a first-column comparator does not promise
 * stability; the stable order is the one fixed here.  Bottom-up merge sort
 * of row ids. */
int or_sort_col0(uint32_t *const *cols, int ncols, uint64_t n) {
    if (n <= 1) return 0;
    uint32_t *id = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t *tmp = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t *buf = (uint32_t *)malloc(n * sizeof(uint32_t));
    if (!id || !tmp || !buf) { free(id); free(tmp); free(buf); return -2; }
    const uint32_t *k = cols[0];
    for (uint64_t i = 0; i < n; ++i) id[i] = (uint32_t)i;
    for (uint64_t width = 1; width < n; width *= 2) {
        for (uint64_t lo = 0; lo < n; lo += 2 * width) {
            uint64_t mid = lo + width < n ? lo + width : n;
            uint64_t hi = lo + 2 * width < n ? lo + 2 * width : n;
            uint64_t a = lo, b = mid, w = lo;
            while (a < mid && b < hi) tmp[w++] = (k[id[b]] < k[id[a]]) ? id[b++] : id[a++];
            while (a < mid) tmp[w++] = id[a++];
            while (b < hi) tmp[w++] = id[b++];
        }
        uint32_t *t = id; id = tmp; tmp = t;
    }
    for (int c = 0; c < ncols; ++c) {
        for (uint64_t i = 0; i < n; ++i) buf[i] = cols[c][id[i]];
        memcpy(cols[c], buf, n * sizeof(uint32_t));
    }
    free(id); free(tmp); free(buf);
    return 0;
}

static const uint32_t *const *g_lex_cols;
static int g_lex_n;
static int cmp_lex(const void *a, const void *b) {
    uint32_t i = *(const uint32_t *)a, j = *(const uint32_t *)b;
    for (int c = 0; c < g_lex_n; ++c) {
        uint32_t u = g_lex_cols[c][i], v = g_lex_cols[c][j];
        if (u != v) return u < v ? -1 : 1;
    }
    return i < j ? -1 : (i > j);
}

/* Lexicographic sort over all columns (FullRelationIR::sort,
 * FullRelationIR.cu:218-336, comparators cmp_first_colK). */
int or_sort_lex(uint32_t *const *cols, int ncols, uint64_t n) {
    if (n <= 1) return 0;
    uint32_t *id = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t *buf = (uint32_t *)malloc(n * sizeof(uint32_t));
    if (!id || !buf) { free(id); free(buf); return -2; }
    for (uint64_t i = 0; i < n; ++i) id[i] = (uint32_t)i;
    g_lex_cols = (const uint32_t *const *)cols;
    g_lex_n = ncols;
    qsort(id, n, sizeof(uint32_t), cmp_lex);
    for (int c = 0; c < ncols; ++c) {
        for (uint64_t i = 0; i < n; ++i) buf[i] = cols[c][id[i]];
        memcpy(cols[c], buf, n * sizeof(uint32_t));
    }
    free(id); free(buf);
    return 0;
}

/* Remove rows equal to the previous row in every column; no sort
 * (FullRelationIR::removeDuplicate, FullRelationIR.cu:338-522:
 * thrust::unique over the zip of all columns).  Returns the new count. */
uint64_t or_unique_adjacent(uint32_t *const *cols, int ncols, uint64_t n) {
    if (n <= 1) return n;
    uint64_t w = 1;
    for (uint64_t r = 1; r < n; ++r) {
        int same = 1;
        for (int c = 0; c < ncols; ++c)
            if (cols[c][r] != cols[c][w - 1]) { same = 0; break; }
        if (same) continue;
        for (int c = 0; c < ncols; ++c) cols[c][w] = cols[c][r];
        ++w;
    }
    return w;
}

/* ---- checkers for results too large for a row-by-row comparison -------- *
 * These do not restate the reference; they test invariants of its outputs
 * (sort order, run heads, multisets) so that full-size GPU results can be
 * checked in O(n) time and memory.  Pinned in tests/test_oracle.py against
 * sorted() / Counter on small inputs and against planted faults. */

static inline uint64_t or_fmix64(uint64_t h) {
    h ^= h >> 33;
    h *= 0xFF51AFD7ED558CCDull;
    h ^= h >> 33;
    h *= 0xC4CEB9FE1A85EC53ull;
    h ^= h >> 33;
    return h;
}

/* Hash of one row (values in column order). */
static inline uint64_t or_row_hash(const uint32_t *v, int ncols) {
    uint64_t h = 0x243F6A8885A308D3ull;
    for (int c = 0; c < ncols; ++c)
        h = or_fmix64(h ^ ((uint64_t)v[c] + 0x9E3779B97F4A7C15ull * (uint64_t)(c + 1)));
    return h;
}

/* Order-independent digest of the bag of rows of a column store: the sums
 * (mod 2^64) of two different 64-bit mixes of each row's hash.  Equal bags
 * give equal digests; a different bag gives a different digest except with
 * probability about 2^-64 per sum.  out[0..1] receive the sums. */
void or_bag_digest(const uint32_t *const *cols, int ncols, uint64_t n, uint64_t *out) {
    uint64_t a = 0, b = 0;
    uint32_t v[32];
    if (ncols > 32) ncols = 32;
    for (uint64_t i = 0; i < n; ++i) {
        for (int c = 0; c < ncols; ++c) v[c] = cols[c][i];
        const uint64_t h = or_row_hash(v, ncols);
        a += h;
        b += or_fmix64(h ^ 0x5851F42D4C957F2Dull);
    }
    out[0] = a;
    out[1] = b;
}

/* Bag digest of a permutation's triples in (s, p, o) column order: rows of
 * (X from L1, Y, Z) mapped back through the permutation's column order, so
 * that it equals or_bag_digest over the input s, p, o columns exactly when
 * the permutation holds the same multiset of triples.  Requires a valid L1
 * (or_check_perm). */
void or_perm_digest(const uint32_t *l1v, const uint32_t *l1o, uint64_t n_l1,
                    const uint32_t *y, const uint32_t *z, uint64_t n, int perm,
                    uint64_t *out) {
    uint64_t a = 0, b = 0;
    uint64_t g = 0;
    for (uint64_t r = 0; r < n; ++r) {
        while (g + 1 < n_l1 && l1o[g + 1] <= r) ++g;
        uint32_t t[3], v[3];
        t[0] = l1v[g]; t[1] = y[r]; t[2] = z[r];
        for (int k = 0; k < 3; ++k) v[PERM_COLS[perm][k]] = t[k];
        const uint64_t h = or_row_hash(v, 3);
        a += h;
        b += or_fmix64(h ^ 0x5851F42D4C957F2Dull);
    }
    out[0] = a;
    out[1] = b;
}

/* Invariants of one permutation's arrays (X implicit through L1):
 *   1  L1 empty for n > 0, or not empty for n == 0
 *   2  l1o[0] != 0, L1 offsets not strictly increasing or >= n
 *   3  L1 values not strictly increasing
 *   4  rows not in lexicographic (X, Y, Z) order
 *   5  an (X, Y) run head with no L2 entry, or an L2 entry off a head
 *   6  an L2 value differs from the row's Y
 *   7  L2 entries left over after the last row
 * Returns 0 when all hold, else the code of the first failure; *where
 * receives the row or entry index. */
int or_check_perm(const uint32_t *l1v, const uint32_t *l1o, uint64_t n_l1,
                  const uint32_t *l2v, const uint32_t *l2o, uint64_t n_l2,
                  const uint32_t *y, const uint32_t *z, uint64_t n, uint64_t *where) {
    *where = 0;
    if ((n == 0) != (n_l1 == 0)) return 1;
    if (n == 0) return n_l2 == 0 ? 0 : 7;
    if (l1o[0] != 0) return 2;
    for (uint64_t i = 1; i < n_l1; ++i) {
        *where = i;
        if (l1o[i] <= l1o[i - 1] || l1o[i] >= n) return 2;
        if (l1v[i] <= l1v[i - 1]) return 3;
    }
    uint64_t g = 0, k = 0;
    for (uint64_t r = 0; r < n; ++r) {
        *where = r;
        int new_x = 0;
        if (g + 1 < n_l1 && l1o[g + 1] == r) { ++g; new_x = 1; }
        const int head = (r == 0) || new_x || y[r] != y[r - 1];
        if (r > 0 && !new_x) {
            if (y[r] < y[r - 1] || (y[r] == y[r - 1] && z[r] < z[r - 1])) return 4;
        }
        if (head) {
            if (k >= n_l2 || l2o[k] != r) return 5;
            if (l2v[k] != y[r]) return 6;
            ++k;
        } else if (k < n_l2 && l2o[k] == r) {
            return 5;
        }
    }
    *where = k;
    return k == n_l2 ? 0 : 7;
}
