// Index build (triple permutations with L1/L2 offset indexes), range
// lookup and range select.
//
// Build of one permutation XYZ:
//   1. k_pack_xyz      : key = X << (bY+bZ) | Y << bZ | Z   (the whole
//                        triple in one 64-bit word; bX+bY+bZ <= 64) and the
//                        digit histograms of every radix pass.
//   2. radix_sort_u64  : LSD onesweep passes over bits [0, bX+bY+bZ).
//   3. k_unpack_count  : Y and Z columns, plus per-block counts of L1 heads
//                        (X changes) and L2 heads ((X,Y) changes).
//   4. scan of the block counts.
//   5. k_compact_heads : L1 = (X, first row), L2 = (Y, first row) entries.
// A permutation XYZ whose rows a sibling permutation already holds in an
// order that, among rows of equal X, is (Y, Z) (Y before Z in the sibling's
// key) is derived instead of sorted in full: its first radix pass reads
// the sibling's sorted keys, moving their fields into the XYZ layout as it
// loads them, and the passes sort on X's bits only (LSD passes keep the
// sibling's order among equal X); X's digit histograms are counted from the
// input column.  Of SPO, PSO, POS and OSP only SPO is
// sorted on all bits; PSO and OSP follow from SPO and POS from OSP.
// The resulting arrays match the reference's createIndex1Level /
// createIndex output (VedasStorage.cu:270-358): L1 values/offsets are the
// distinct leading IDs and their first rows, L2 values/offsets the second
// IDs of every distinct (X,Y) pair and their first rows, and the Y and Z
// columns are the reference's xy_data / xyz_data arrays.

#include <algorithm>
#include <mutex>
#include <type_traits>

#include "vd_radix_dev.hpp"
#include "vd_swap_dev.hpp"

#include "vd_index.hpp"

namespace vd {
namespace {

constexpr int PK_BLOCK = radix::HIST_BLOCK;
constexpr int PK_ITEMS = 16;
constexpr int UN_BLOCK = 256;
constexpr int UN_ITEMS = 8;
constexpr int UN_TILE = UN_BLOCK * UN_ITEMS;

// Which of (s, p, o) is X, Y, Z for each permutation.
constexpr int PERM_COLS[VD_NUM_PERMS][3] = {
    {0, 1, 2},  // SPO
    {0, 2, 1},  // SOP
    {1, 0, 2},  // PSO
    {1, 2, 0},  // POS
    {2, 0, 1},  // OSP
    {2, 1, 0},  // OPS
};

__global__ __launch_bounds__(PK_BLOCK) void k_pack_xyz(
    const uint32_t *__restrict__ x, const uint32_t *__restrict__ y,
    const uint32_t *__restrict__ z, uint64_t n, int sy, int sx,
    uint64_t *__restrict__ keys, int kbits, int npass, uint32_t *ghist) {
    __shared__ radix::LdsHist s;
    radix::hist_zero(&s);
    __syncthreads();
    for (uint64_t base = (uint64_t)blockIdx.x * (PK_BLOCK * PK_ITEMS); base < n;
         base += (uint64_t)gridDim.x * (PK_BLOCK * PK_ITEMS)) {
#pragma unroll 4
        for (int j = 0; j < PK_ITEMS; ++j) {
            const uint64_t i = base + (uint64_t)j * PK_BLOCK + threadIdx.x;
            if (i < n) {
                const uint64_t key = ((uint64_t)x[i] << sx) | ((uint64_t)y[i] << sy) |
                                     (uint64_t)z[i];
                keys[i] = key;
                radix::hist_add(&s, key, 0, kbits, npass);
            }
        }
    }
    __syncthreads();
    radix::hist_flush(&s, ghist, npass);
}

// Field layout of a permutation's packed key: shift and mask of the s, p
// and o columns.
struct PackLayout {
    int sh[3];
    uint32_t mask[3];
};

// Keys of a derived permutation: each sorted key of the sibling permutation
// (layout `from`) with its fields moved to layout `to`, in the sibling's
// order; counts the digit histograms of column xc (the new X) under the
// plan of *gmax into g8 / g9.
__global__ __launch_bounds__(PK_BLOCK) void k_repack(const uint64_t *__restrict__ src, uint64_t n,
                                                     PackLayout from, PackLayout to, int xc,
                                                     uint64_t *__restrict__ keys, uint32_t *g8,
                                                     uint32_t *g9, const uint32_t *gmax,
                                                     int bits) {
    __shared__ uint32_t s_h[radix::HIST_WAVES][swap::HW];
    for (int i = threadIdx.x; i < radix::HIST_WAVES * swap::HW; i += PK_BLOCK)
        (&s_h[0][0])[i] = 0;
    const swap::Plan plan = swap::plan_for(*gmax, bits);
    __syncthreads();
    uint32_t *h = s_h[threadIdx.x / 64];
    for (uint64_t base = (uint64_t)blockIdx.x * (PK_BLOCK * PK_ITEMS); base < n;
         base += (uint64_t)gridDim.x * (PK_BLOCK * PK_ITEMS)) {
#pragma unroll 4
        for (int j = 0; j < PK_ITEMS; ++j) {
            const uint64_t i = base + (uint64_t)j * PK_BLOCK + threadIdx.x;
            if (i < n) {
                const uint64_t v = src[i];
                uint64_t key = 0;
                uint32_t x = 0;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const uint32_t f = (uint32_t)(v >> from.sh[c]) & from.mask[c];
                    key |= (uint64_t)f << to.sh[c];
                    x = c == xc ? f : x;
                }
                keys[i] = key;
                swap::hist_add(h, x, plan);
            }
        }
    }
    __syncthreads();
    swap::hist_flush(&s_h[0][0], radix::HIST_WAVES, plan, g8, g9);
}

// Sorted triples, one row = (X, Y, Z).  PackedRows: the whole triple in one
// 64-bit word.  SplitRows: X in its own column and (Y, Z) packed in a 64-bit
// word, for IDs that need more than 64 bits in total.
struct Row {
    uint32_t x, y, z;
    uint64_t xy;  // identity of the (X, Y) pair
};

struct PackedRows {
    const uint64_t *k;
    int sx, sy;
    uint32_t ymask, zmask;
    __device__ __forceinline__ Row at(uint64_t i) const {
        const uint64_t v = k[i];
        return Row{(uint32_t)(v >> sx), (uint32_t)(v >> sy) & ymask, (uint32_t)v & zmask, v >> sy};
    }
};

struct SplitRows {
    const uint32_t *xs;
    const uint64_t *lo;
    int sy;
    uint32_t ymask, zmask;
    __device__ __forceinline__ Row at(uint64_t i) const {
        const uint64_t v = lo[i];
        const uint32_t x = xs[i], y = (uint32_t)(v >> sy) & ymask;
        return Row{x, y, (uint32_t)v & zmask, ((uint64_t)x << 32) | y};
    }
};

// Columns Y and Z; per-block counts of L1 heads (X changes) and L2 heads
// ((X,Y) changes); rows equal to the row before them added to *dups.
template <typename R>
__global__ __launch_bounds__(UN_BLOCK) void k_unpack_count(
    R rows, uint64_t n, uint32_t *__restrict__ ycol, uint32_t *__restrict__ zcol,
    uint32_t *__restrict__ cnt_l1, uint32_t *__restrict__ cnt_l2, uint64_t *dups) {
    __shared__ uint32_t s_tmp[UN_BLOCK / 64];
    const uint64_t base = (uint64_t)blockIdx.x * UN_TILE;
    uint32_t c1 = 0, c2 = 0, c3 = 0;
#pragma unroll
    for (int j = 0; j < UN_ITEMS; ++j) {
        const uint64_t i = base + (uint64_t)j * UN_BLOCK + threadIdx.x;
        if (i < n) {
            const Row r = rows.at(i);
            ycol[i] = r.y;
            zcol[i] = r.z;
            if (i == 0) {
                c1 += 1; c2 += 1;
            } else {
                const Row q = rows.at(i - 1);
                c1 += r.x != q.x;
                c2 += r.xy != q.xy;
                c3 += r.xy == q.xy && r.z == q.z;
            }
        }
    }
    const uint32_t t1 = dev::block_reduce<UN_BLOCK>(c1, s_tmp);
    const uint32_t t2 = dev::block_reduce<UN_BLOCK>(c2, s_tmp);
    const uint32_t t3 = dev::block_reduce<UN_BLOCK>(c3, s_tmp);
    if (threadIdx.x == 0) {
        cnt_l1[blockIdx.x] = t1;
        cnt_l2[blockIdx.x] = t2;
        if (t3) atomicAdd((unsigned long long *)dups, (unsigned long long)t3);
    }
}

template <typename R>
__global__ __launch_bounds__(UN_BLOCK) void k_compact_heads(
    R rows, uint64_t n, const uint64_t *__restrict__ off_l1,
    const uint64_t *__restrict__ off_l2, uint32_t *__restrict__ l1v,
    uint32_t *__restrict__ l1o, uint32_t *__restrict__ l2v,
    uint32_t *__restrict__ l2o) {
    constexpr int WAVES = UN_BLOCK / 64;
    constexpr int SLOTS = UN_ITEMS * WAVES;
    __shared__ uint32_t s_off[2][SLOTS];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t base = (uint64_t)blockIdx.x * UN_TILE;
    // Rows are read striped (row base + j * UN_BLOCK + threadIdx.x: every
    // load coalesced, as in k_unpack_count).  In row order, slot (j, w) = the
    // 64 rows of band j held by wave w comes before (j, w + 1) and (j + 1,
    // 0): a head's rank in the tile is the heads of the slots before its own
    // plus those of lower lanes in its slot.
    uint64_t b1[UN_ITEMS], b2[UN_ITEMS];
    uint32_t xv[UN_ITEMS], yv[UN_ITEMS];
#pragma unroll
    for (int j = 0; j < UN_ITEMS; ++j) {
        const uint64_t i = base + (uint64_t)j * UN_BLOCK + threadIdx.x;
        bool h1 = false, h2 = false;
        xv[j] = yv[j] = 0;
        if (i < n) {
            const Row r = rows.at(i);
            xv[j] = r.x;
            yv[j] = r.y;
            if (i == 0) {
                h1 = h2 = true;
            } else {
                const Row q = rows.at(i - 1);
                h1 = r.x != q.x;
                h2 = r.xy != q.xy;
            }
        }
        b1[j] = __ballot(h1);
        b2[j] = __ballot(h2);
        if (lane == 0) {
            s_off[0][j * WAVES + w] = (uint32_t)__popcll(b1[j]);
            s_off[1][j * WAVES + w] = (uint32_t)__popcll(b2[j]);
        }
    }
    __syncthreads();
    if (threadIdx.x < 2) {  // exclusive scan of the slot counts, in row order
        uint32_t run = 0;
        for (int k = 0; k < SLOTS; ++k) {
            const uint32_t c = s_off[threadIdx.x][k];
            s_off[threadIdx.x][k] = run;
            run += c;
        }
    }
    __syncthreads();
    const uint64_t t1 = off_l1[blockIdx.x], t2 = off_l2[blockIdx.x];
#pragma unroll
    for (int j = 0; j < UN_ITEMS; ++j) {
        const uint64_t i = base + (uint64_t)j * UN_BLOCK + threadIdx.x;
        if ((b1[j] >> lane) & 1u) {
            const uint64_t o = t1 + s_off[0][j * WAVES + w] +
                               __builtin_amdgcn_mbcnt_hi((uint32_t)(b1[j] >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)b1[j], 0u));
            l1v[o] = xv[j];
            l1o[o] = (uint32_t)i;
        }
        if ((b2[j] >> lane) & 1u) {
            const uint64_t o = t2 + s_off[1][j * WAVES + w] +
                               __builtin_amdgcn_mbcnt_hi((uint32_t)(b2[j] >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)b2[j], 0u));
            l2v[o] = yv[j];
            l2o[o] = (uint32_t)i;
        }
    }
}

// Split path, phase 1: key = Y << bZ | Z, value = X.
__global__ __launch_bounds__(PK_BLOCK) void k_pack_split(
    const uint32_t *__restrict__ x, const uint32_t *__restrict__ y,
    const uint32_t *__restrict__ z, uint64_t n, int sy, uint64_t *__restrict__ keys,
    uint32_t *__restrict__ vals, int kbits, int npass, uint32_t *ghist) {
    __shared__ radix::LdsHist s;
    radix::hist_zero(&s);
    __syncthreads();
    for (uint64_t base = (uint64_t)blockIdx.x * (PK_BLOCK * PK_ITEMS); base < n;
         base += (uint64_t)gridDim.x * (PK_BLOCK * PK_ITEMS)) {
#pragma unroll 4
        for (int j = 0; j < PK_ITEMS; ++j) {
            const uint64_t i = base + (uint64_t)j * PK_BLOCK + threadIdx.x;
            if (i < n) {
                const uint64_t key = ((uint64_t)y[i] << sy) | (uint64_t)z[i];
                keys[i] = key;
                vals[i] = x[i];
                radix::hist_add(&s, key, 0, kbits, npass);
            }
        }
    }
    __syncthreads();
    radix::hist_flush(&s, ghist, npass);
}

// Split path, phase 2: key = X << 32 | position, sorted on X's bits only.
__global__ __launch_bounds__(PK_BLOCK) void k_pack_xpos(
    const uint32_t *__restrict__ xs, uint64_t n, uint64_t *__restrict__ keys, int hi_bit,
    int npass, uint32_t *ghist) {
    __shared__ radix::LdsHist s;
    radix::hist_zero(&s);
    __syncthreads();
    for (uint64_t base = (uint64_t)blockIdx.x * (PK_BLOCK * PK_ITEMS); base < n;
         base += (uint64_t)gridDim.x * (PK_BLOCK * PK_ITEMS)) {
#pragma unroll 4
        for (int j = 0; j < PK_ITEMS; ++j) {
            const uint64_t i = base + (uint64_t)j * PK_BLOCK + threadIdx.x;
            if (i < n) {
                const uint64_t key = ((uint64_t)xs[i] << 32) | i;
                keys[i] = key;
                radix::hist_add(&s, key, 32, hi_bit, npass);
            }
        }
    }
    __syncthreads();
    radix::hist_flush(&s, ghist, npass);
}

// X column and (Y,Z) words in final order.
__global__ __launch_bounds__(256) void k_split_gather(const uint64_t *__restrict__ k2,
                                                      const uint64_t *__restrict__ lo, uint64_t n,
                                                      uint32_t *__restrict__ xs,
                                                      uint64_t *__restrict__ lo_out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t k = k2[i];
        xs[i] = (uint32_t)(k >> 32);
        lo_out[i] = lo[(uint32_t)k];
    }
}

// ---- lookups -------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_lookup_batch(
    PermDev pd, const uint32_t *__restrict__ k1, const uint32_t *__restrict__ k2,
    uint64_t nq, uint64_t *__restrict__ begin, uint64_t *__restrict__ end,
    uint64_t *__restrict__ sizes) {
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const uint32_t c2 = k2 ? k2[q] : VD_UNBOUND;
    uint64_t b, e;
    lookup_one(pd, k1[q], c2, c2 != VD_UNBOUND, &b, &e);
    if (begin) begin[q] = b;
    if (end) end[q] = e;
    if (sizes) sizes[q] = e - b;
}

__global__ __launch_bounds__(256) void k_index_stats(
    const uint32_t *__restrict__ l1o, uint64_t n_l1, const uint32_t *__restrict__ l2o,
    uint64_t n_l2, uint64_t n, uint64_t *__restrict__ rows, uint64_t *__restrict__ pairs) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_l1) return;
    const uint64_t b = l1o[i], e = (i + 1 < n_l1) ? l1o[i + 1] : n;
    rows[i] = e - b;
    pairs[i] = dev::lower_bound(l2o, 0, n_l2, e) - dev::lower_bound(l2o, 0, n_l2, b);
}

// ---- runs (a permutation's L1 / L2 arrays) and rows ---------------------------------
// out[r] = values[k] for the run k holding row r: run k covers rows
// [offsets[k], offsets[k + 1]) (the last run ends at n).  A thread expands
// RUN_ROWS consecutive rows, searching only for its first.
constexpr int RUN_ROWS = 16;
__global__ __launch_bounds__(256) void k_runs_expand(const uint32_t *__restrict__ values,
                                                     const uint32_t *__restrict__ offsets,
                                                     uint64_t nruns, uint64_t n,
                                                     uint32_t *__restrict__ out) {
    const uint64_t r0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * RUN_ROWS;
    if (r0 >= n) return;
    uint64_t k = dev::upper_bound(offsets, 0, nruns, (uint32_t)r0);
    k = k > 0 ? k - 1 : 0;  // offsets[0] is 0 in a valid array
    uint64_t next = k + 1 < nruns ? offsets[k + 1] : n;
    const uint64_t r1 = r0 + RUN_ROWS < n ? r0 + RUN_ROWS : n;
    for (uint64_t r = r0; r < r1; ++r) {
        while (r >= next) {
            ++k;
            next = k + 1 < nruns ? offsets[k + 1] : n;
        }
        out[r] = values[k];
    }
}

// head[r] = 1 where row r starts an (X, Y) run: an L1 offset or a change of Y.
__global__ __launch_bounds__(256) void k_mark_l1(const uint32_t *__restrict__ l1o, uint64_t n_l1,
                                                 uint64_t n, uint32_t *__restrict__ head) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_l1 && l1o[i] < n) head[l1o[i]] = 1;
}

__global__ __launch_bounds__(256) void k_mark_y(const uint32_t *__restrict__ y, uint64_t n,
                                                uint32_t *__restrict__ head) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    if (r == 0 || y[r] != y[r - 1]) head[r] = 1;
}

// L2 entries at the heads: pos = exclusive scan of head.
__global__ __launch_bounds__(256) void k_scatter_heads(const uint32_t *__restrict__ head,
                                                       const uint64_t *__restrict__ pos,
                                                       const uint32_t *__restrict__ y, uint64_t n,
                                                       uint32_t *__restrict__ l2v,
                                                       uint32_t *__restrict__ l2o) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n || !head[r]) return;
    l2v[pos[r]] = y[r];
    l2o[pos[r]] = (uint32_t)r;
}

// ---- histograms of a permutation's leading column -----------------------------
// Only the rows whose leading key is below a limit count; the L1 arrays give
// every key's run of rows, so no row is read.  pos(i) = first row of L1
// entry i, or rows_in (the rows below the limit) for i == m.
__device__ __forceinline__ uint64_t hist_pos(const uint32_t *l1o, uint64_t m, uint64_t rows_in,
                                             uint64_t i) {
    return i < m ? l1o[i] : rows_in;
}

// Equal width: bucket b holds keys [start + b w, min(start + (b + 1) w - 1, end)].
__global__ __launch_bounds__(256) void k_hist_width(
    const uint32_t *__restrict__ l1v, const uint32_t *__restrict__ l1o, uint64_t m,
    uint64_t rows_in, uint32_t start, uint32_t end, uint32_t width, uint64_t nb,
    uint64_t *__restrict__ freq) {
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const uint64_t lo = (uint64_t)start + b * width;
    uint64_t hi = lo + width - 1;
    if (hi > end) hi = end;
    const uint64_t i0 = dev::lower_bound(l1v, 0, m, (uint32_t)lo);
    const uint64_t i1 = dev::upper_bound(l1v, 0, m, (uint32_t)hi);
    freq[b] = hist_pos(l1o, m, rows_in, i1) - hist_pos(l1o, m, rows_in, i0);
}

// Equal depth: a bucket that starts at L1 entry g ends with the entry that
// holds row min(pos(g) + depth, rows_in - 1) (a key's run is never split);
// nxt[g] = the entry after it, where the next bucket starts.
__global__ __launch_bounds__(256) void k_hist_depth_next(
    const uint32_t *__restrict__ l1o, uint64_t m, uint64_t rows_in, uint32_t depth,
    uint32_t *__restrict__ nxt) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= m) return;
    uint64_t e = (uint64_t)l1o[g] + depth;
    if (e > rows_in - 1) e = rows_in - 1;
    nxt[g] = (uint32_t)dev::upper_bound(l1o, 0, m, (uint32_t)e);
}

// Buckets of the chain of starts: lower = first key, upper = last key,
// freq = rows.
__global__ __launch_bounds__(256) void k_hist_depth_gather(
    const uint32_t *__restrict__ l1v, const uint32_t *__restrict__ l1o, uint64_t m,
    uint64_t rows_in, const uint32_t *__restrict__ nxt, const uint32_t *__restrict__ chain,
    uint64_t nb, uint32_t *__restrict__ lowers, uint32_t *__restrict__ uppers,
    uint64_t *__restrict__ freq) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    const uint32_t g = chain[k], g2 = nxt[g];
    lowers[k] = l1v[g];
    uppers[k] = l1v[g2 - 1];
    freq[k] = hist_pos(l1o, m, rows_in, g2) - l1o[g];
}

// be[0] = L1 entries with a key below *limit, be[1] = their rows.
__global__ void k_hist_limit(const uint32_t *__restrict__ l1v, const uint32_t *__restrict__ l1o,
                             uint64_t n_l1, uint64_t n, const uint32_t *__restrict__ limit,
                             uint64_t *__restrict__ be) {
    if (threadIdx.x != 0) return;
    const uint64_t m = dev::lower_bound(l1v, 0, n_l1, *limit);
    be[0] = m;
    be[1] = m < n_l1 ? l1o[m] : n;
}

// Rows of query q are output rows [off[q], off[q+1]) (nq + 1 offsets);
// output row o of query q copies source row begin[q] + o - off[q].  A block
// copies GR_TILE consecutive output rows.  It finds the queries of its
// first and last row once: inside one query (a large range) the tile is a
// plain copy of consecutive source rows; across up to GR_LDS queries their
// offsets and first rows are staged in LDS and each row searches there;
// across more (ranges of a few rows) each row searches the offsets in
// memory between the two queries found.
constexpr int GR_BLOCK = 256;
constexpr int GR_ITEMS = 16;
constexpr uint64_t GR_TILE = GR_BLOCK * GR_ITEMS;
constexpr int GR_LDS = 1024;
__global__ __launch_bounds__(GR_BLOCK) void k_gather_ranges(
    const uint64_t *__restrict__ off, const uint64_t *__restrict__ begin,
    uint64_t nq, uint64_t total, const uint32_t *__restrict__ y,
    const uint32_t *__restrict__ z, uint32_t *__restrict__ oy,
    uint32_t *__restrict__ oz) {
    __shared__ uint64_t s_off[GR_LDS + 1];
    __shared__ uint64_t s_beg[GR_LDS];
    __shared__ uint64_t s_q[2];
    const uint64_t o0 = (uint64_t)blockIdx.x * GR_TILE;
    const uint64_t o1 = o0 + GR_TILE < total ? o0 + GR_TILE : total;
    if (threadIdx.x < 2) {
        const uint64_t x = threadIdx.x == 0 ? o0 : o1 - 1;
        s_q[threadIdx.x] = dev::upper_bound(off, 0, nq + 1, x) - 1;
    }
    __syncthreads();
    const uint64_t qa = s_q[0], qb = s_q[1];
    const uint64_t nqs = qb - qa + 1;
    if (nqs == 1) {
        const uint64_t src0 = begin[qa] + (o0 - off[qa]);
        uint32_t vy[GR_ITEMS], vz[GR_ITEMS];
#pragma unroll
        for (int k = 0; k < GR_ITEMS; ++k) {
            const uint64_t o = o0 + (uint64_t)k * GR_BLOCK + threadIdx.x;
            if (o < o1) {
                if (oy) vy[k] = y[src0 + (o - o0)];
                if (oz) vz[k] = z[src0 + (o - o0)];
            }
        }
#pragma unroll
        for (int k = 0; k < GR_ITEMS; ++k) {
            const uint64_t o = o0 + (uint64_t)k * GR_BLOCK + threadIdx.x;
            if (o < o1) {
                if (oy) oy[o] = vy[k];
                if (oz) oz[o] = vz[k];
            }
        }
        return;
    }
    if (nqs <= (uint64_t)GR_LDS) {
        for (uint64_t t = threadIdx.x; t <= nqs; t += GR_BLOCK) {
            s_off[t] = off[qa + t];  // qb + 1 <= nq: off has that entry
            if (t < nqs) s_beg[t] = begin[qa + t];
        }
        __syncthreads();
        for (uint64_t o = o0 + threadIdx.x; o < o1; o += GR_BLOCK) {
            const uint64_t k = dev::upper_bound(s_off, 0, nqs + 1, o) - 1;
            const uint64_t src = s_beg[k] + (o - s_off[k]);
            if (oy) oy[o] = y[src];
            if (oz) oz[o] = z[src];
        }
        return;
    }
    for (uint64_t o = o0 + threadIdx.x; o < o1; o += GR_BLOCK) {
        const uint64_t q = dev::upper_bound(off, qa, qb + 1, o) - 1;
        const uint64_t src = begin[q] + (o - off[q]);
        if (oy) oy[o] = y[src];
        if (oz) oz[o] = z[src];
    }
}

// ---- select ------------------------------------------------------------------
constexpr int MAX_PIECES = 8;

struct Pieces {
    int np;
    uint64_t src[MAX_PIECES];
    uint64_t dst[MAX_PIECES + 1];  // dst[np] = total rows
};

// Binary searches of one thread, or wave-cooperative ones (every lane of a
// wave calls with the same arguments; see dev::wave_bound).
template <bool WAVE>
struct Search {
    static __device__ __forceinline__ uint64_t lower(const uint32_t *a, uint64_t lo, uint64_t hi,
                                                     uint32_t v) {
        return WAVE ? dev::wave_bound<false>(a, lo, hi, v) : dev::lower_bound(a, lo, hi, v);
    }
    static __device__ __forceinline__ uint64_t upper(const uint32_t *a, uint64_t lo, uint64_t hi,
                                                     uint32_t v) {
        return WAVE ? dev::wave_bound<true>(a, lo, hi, v) : dev::upper_bound(a, lo, hi, v);
    }
};

// [b, e) minus clip and skip intervals as pieces of consecutive rows.
// Intervals must be sorted by lo and disjoint.
template <bool WAVE>
__device__ Pieces make_pieces(const uint32_t *__restrict__ ycol, uint64_t b, uint64_t e,
                              int clip, uint32_t lo, uint32_t hi, const vd_interval_t *sk,
                              int nskip) {
    using S = Search<WAVE>;
    if (clip) {
        if (lo > hi) {
            e = b;
        } else if (WAVE) {
            // both ends at once, half a wave each: the first row >= lo and
            // the first row > hi (independent searches over [b, e))
            const bool up = dev::lane_id() >= 32;
            const uint64_t r = dev::group_bound<32>(b, e, [&](uint64_t q) {
                const uint32_t x = ycol[q];
                return up ? hi < x : !(x < lo);
            });
            const uint64_t nb = __shfl(r, 0), ne = __shfl(r, 32);
            b = nb;
            e = ne;
        } else {
            const uint64_t nb = S::lower(ycol, b, e, lo);
            e = S::upper(ycol, nb, e, hi);
            b = nb;
        }
    }
    Pieces P;
    P.np = 0;
    uint64_t cur = b, acc = 0;
    for (int i = 0; i < nskip; ++i) {
        if (sk[i].lo > sk[i].hi) continue;
        const uint64_t a = S::lower(ycol, cur, e, sk[i].lo);
        const uint64_t c = S::upper(ycol, a, e, sk[i].hi);
        if (a > cur) {
            P.src[P.np] = cur;
            P.dst[P.np] = acc;
            acc += a - cur;
            ++P.np;
        }
        cur = c;
    }
    if (e > cur) {
        P.src[P.np] = cur;
        P.dst[P.np] = acc;
        acc += e - cur;
        ++P.np;
    }
    P.dst[P.np] = acc;
    return P;
}

struct SkipSet {
    vd_interval_t s[4];
};

// Pieces as read-back words: np, src[0..MAX_PIECES), dst[0..MAX_PIECES].
constexpr int PIECE_WORDS = 1 + MAX_PIECES + MAX_PIECES + 1;
__device__ __forceinline__ void pieces_to_words(const Pieces &P, uint64_t *w) {
    w[0] = (uint64_t)P.np;
#pragma unroll
    for (int i = 0; i < MAX_PIECES; ++i) w[1 + i] = P.src[i];
#pragma unroll
    for (int i = 0; i <= MAX_PIECES; ++i) w[1 + MAX_PIECES + i] = P.dst[i];
}
Pieces pieces_from_words(const uint64_t *w) {
    Pieces P{};
    P.np = (int)w[0];
    for (int i = 0; i < MAX_PIECES; ++i) P.src[i] = w[1 + i];
    for (int i = 0; i <= MAX_PIECES; ++i) P.dst[i] = w[1 + MAX_PIECES + i];
    return P;
}

// One wave: pieces of a given row range.
__global__ __launch_bounds__(64) void k_select_pieces(const uint32_t *__restrict__ ycol,
                                                      uint64_t b, uint64_t e, int clip,
                                                      uint32_t lo, uint32_t hi, SkipSet sk,
                                                      int nskip, Pieces *out, Readback rb) {
    const Pieces P = make_pieces<true>(ycol, b, e, clip, lo, hi, sk.s, nskip);
    if (threadIdx.x == 0) {
        *out = P;
        pieces_to_words(P, rb.dst);  // the host reads them back from here
    }
    dev::rb_publish(rb);
}

// Pieces plus the key's row range, so a lookup and a select cost one
// read-back together.
struct KeyPieces {
    Pieces p;
    uint64_t begin, end;
};


__global__ __launch_bounds__(64) void k_lookup_pieces(PermDev pd, uint32_t k1, int has_k2,
                                                      uint32_t k2, int clip, uint32_t lo,
                                                      uint32_t hi, SkipSet sk, int nskip,
                                                      KeyPieces *out, Readback rb) {
    uint64_t b, e;
    wave_lookup(pd, k1, k2, has_k2 != 0, &b, &e);
    const Pieces P = make_pieces<true>(pd.y, b, e, clip, lo, hi, sk.s, nskip);
    if (threadIdx.x == 0) {
        out->begin = b;
        out->end = e;
        out->p = P;
        // read back: begin, end, then the pieces
        rb.dst[0] = b;
        rb.dst[1] = e;
        pieces_to_words(P, rb.dst + 2);
    }
    dev::rb_publish(rb);
}

__global__ __launch_bounds__(256) void k_copy_pieces(
    const Pieces *__restrict__ P, const uint32_t *__restrict__ y,
    const uint32_t *__restrict__ z, uint32_t *__restrict__ o0,
    uint32_t *__restrict__ o1, int cols) {
    const int np = P->np;
    const uint64_t total = P->dst[np];
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; o < total;
         o += stride) {
        int k = 0;
        while (k + 1 < np && P->dst[k + 1] <= o) ++k;
        const uint64_t src = P->src[k] + (o - P->dst[k]);
        if (cols == 1) {
            o0[o] = z[src];
        } else if (cols == 2) {
            o0[o] = y[src];
        } else {
            o0[o] = y[src];
            o1[o] = z[src];
        }
    }
}

unsigned stream_grid(uint64_t n) {
    uint64_t g = ceil_div(n, 256 * 4);
    if (g > 16384) g = 16384;
    return g ? (unsigned)g : 1u;
}

// Y/Z columns and L1/L2 arrays from sorted rows.
template <typename R>
void finish_perm(vd_ctx *ctx, vd_perm_arrays &pa, uint64_t n, const R &rows) {
    pa.n = n;
    pa.y = (uint32_t *)dalloc(ctx, n * 4);
    pa.z = (uint32_t *)dalloc(ctx, n * 4);
    const uint64_t tiles = ceil_div(n, UN_TILE);
    DBuf<uint32_t> c1(ctx, tiles), c2(ctx, tiles);
    DBuf<uint64_t> o1(ctx, tiles + 1), o2(ctx, tiles + 1);
    constexpr bool split = std::is_same<R, SplitRows>::value;
    uint64_t tot[3];
    {
        DBuf<uint64_t> t(ctx, 3);  // L1 heads, L2 heads, repeated rows
        fill_u64(ctx, t.p + 2, 0, 1);
        launch(ctx, "index_unpack", (double)n * (split ? 20.0 : 16.0), dim3((unsigned)tiles),
               dim3(UN_BLOCK), 0, k_unpack_count<R>, rows, n, pa.y, pa.z, c1.p, c2.p, t.p + 2);
        scan_u32_to_u64(ctx, c1.p, o1.p, tiles, o1.p + tiles);
        scan_u32_to_u64(ctx, c2.p, o2.p, tiles, o2.p + tiles);
        copy_d2d(ctx, t.p, o1.p + tiles, 8);
        copy_d2d(ctx, t.p + 1, o2.p + tiles, 8);
        read_words(ctx, tot, t.p, 3);
    }
    pa.n_l1 = tot[0];
    pa.n_l2 = tot[1];
    pa.dups_known = true;
    pa.dups = tot[2];
    pa.l1v = (uint32_t *)dalloc(ctx, pa.n_l1 * 4);
    pa.l1o = (uint32_t *)dalloc(ctx, pa.n_l1 * 4);
    pa.l2v = (uint32_t *)dalloc(ctx, pa.n_l2 * 4);
    pa.l2o = (uint32_t *)dalloc(ctx, pa.n_l2 * 4);
    launch(ctx, "index_heads", (double)n * (split ? 12.0 : 8.0) + (pa.n_l1 + pa.n_l2) * 8.0,
           dim3((unsigned)tiles), dim3(UN_BLOCK), 0, k_compact_heads<R>, rows, n,
           (const uint64_t *)o1.p, (const uint64_t *)o2.p, pa.l1v, pa.l1o, pa.l2v, pa.l2o);
}

PackLayout pack_layout(int perm, const int bits[3]) {
    const int xi = PERM_COLS[perm][0], yi = PERM_COLS[perm][1], zi = PERM_COLS[perm][2];
    PackLayout L{};
    L.sh[zi] = 0;
    L.sh[yi] = bits[zi];
    L.sh[xi] = bits[zi] + bits[yi];
    for (int c = 0; c < 3; ++c) L.mask[c] = bits[c] >= 32 ? 0xFFFFFFFFu : ((1u << bits[c]) - 1u);
    return L;
}

// Whether permutation p can be derived from q: among rows of equal X_p,
// q's order is (Y_p, Z_p), i.e. Y_p comes before Z_p in q's key.
bool derivable(int q, int p) {
    int pos[3];
    for (int k = 0; k < 3; ++k) pos[PERM_COLS[q][k]] = k;
    return pos[PERM_COLS[p][1]] < pos[PERM_COLS[p][2]];
}

// Every permutation of the mask when the whole triple fits one 64-bit key
// (bits[0] + bits[1] + bits[2] <= 64, n > 0).  Build order: a permutation
// with a built sibling it can be derived from is derived (a stable sort on
// its X bits), the others are sorted on all bits; a sibling's sorted keys
// are kept until its last derived permutation is built.
void build_packed(vd_ctx *ctx, vd_index *idx, uint32_t mask, const uint32_t *cols[3],
                  const int bits[3]) {
    const uint64_t n = idx->n;
    int order[VD_NUM_PERMS], from[VD_NUM_PERMS], uses[VD_NUM_PERMS] = {};
    bool planned[VD_NUM_PERMS] = {};
    int nord = 0;
    for (;;) {
        int pick = -1, src = -1;
        for (int p = 0; p < VD_NUM_PERMS && pick < 0; ++p) {
            if (!(mask & VD_PERM_BIT(p)) || planned[p]) continue;
            for (int k = 0; k < nord; ++k)
                if (derivable(order[k], p)) {
                    pick = p;
                    src = order[k];
                    break;
                }
        }
        for (int p = 0; p < VD_NUM_PERMS && pick < 0; ++p)
            if ((mask & VD_PERM_BIT(p)) && !planned[p]) pick = p;
        if (pick < 0) break;
        planned[pick] = true;
        from[pick] = src;
        if (src >= 0) ++uses[src];
        order[nord++] = pick;
    }

    const int kbits = bits[0] + bits[1] + bits[2];
    const uint64_t pk_tiles = ceil_div(n, PK_BLOCK * PK_ITEMS);
    DBuf<uint64_t> held[VD_NUM_PERMS];  // sorted keys kept as a sibling
    DBuf<uint64_t> spare;               // a free buffer of n keys
    DBuf<uint32_t> hist(ctx, swap::HBUF), gofs(ctx, swap::HBUF), dmax(ctx, 1);
    // A column's digit histograms do not depend on the order of its rows:
    // those of every derived permutation's X are counted once from the
    // input columns (an upper bound of the maximum: the column's bit length).
    DBuf<uint32_t> xhist[3], xmax[3];
    for (int k = 0; k < nord; ++k) {
        const int p = order[k], xi = PERM_COLS[p][0];
        if (from[p] < 0 || xhist[xi].p) continue;
        xhist[xi] = DBuf<uint32_t>(ctx, swap::HBUF);
        xmax[xi] = DBuf<uint32_t>(ctx, 1);
        fill_u32(ctx, xhist[xi].p, 0, swap::HBUF);
        fill_u32(ctx, xmax[xi].p, pack_layout(p, bits).mask[xi], 1);
        swap_hist(ctx, cols[xi], n, xhist[xi].p, xhist[xi].p + swap::H8, xmax[xi].p);
    }
    for (int k = 0; k < nord; ++k) {
        const int p = order[k];
        const PackLayout to = pack_layout(p, bits);
        const int xi = PERM_COLS[p][0], yi = PERM_COLS[p][1];
        DBuf<uint64_t> keys = spare.p ? std::move(spare) : DBuf<uint64_t>(ctx, n);
        DBuf<uint64_t> tmp(ctx, n);
        uint64_t *sorted = nullptr;
        if (from[p] < 0) {
            const int npass = radix_passes(0, kbits);
            DBuf<uint32_t> h(ctx, radix_hist_words(0, kbits));
            fill_u32(ctx, h.p, 0, radix_hist_words(0, kbits));
            launch(ctx, "index_pack", (double)n * 20.0, dim3(radix::hist_grid(pk_tiles)),
                   dim3(PK_BLOCK), 0, k_pack_xyz, cols[xi], cols[yi], cols[PERM_COLS[p][2]], n,
                   to.sh[yi], to.sh[xi], keys.p, kbits, npass, h.p);
            sorted = radix_sort_u64(ctx, keys.p, tmp.p, n, 0, kbits, h.p, true);
        } else {
            // The first pass reads the sibling's sorted keys and moves their
            // fields into this permutation's layout as it loads them.
            const int q = from[p];
            const PackLayout fl = pack_layout(q, bits);
            int shifts[4], dbits = 8;
            const uint32_t *g[4];
            const int npass = swap_passes(ctx, xhist[xi].p, xhist[xi].p + swap::H8, xmax[xi].p,
                                          n, gofs.p, to.sh[xi], shifts, g, &dbits);
            if (npass == 0 || n <= 1) {  // one X value: the sibling's order
                fill_u32(ctx, hist.p, 0, swap::HBUF);
                fill_u32(ctx, dmax.p, to.mask[xi], 1);
                launch(ctx, "index_repack", (double)n * 16.0, dim3(radix::hist_grid(pk_tiles)),
                       dim3(PK_BLOCK), 0, k_repack, (const uint64_t *)held[q].p, n, fl, to, xi,
                       keys.p, hist.p, hist.p + swap::H8, (const uint32_t *)dmax.p,
                       ctx->radix_bits);
                sorted = keys.p;
            } else {
                RadixKeyIO kio{};
                kio.in_k = held[q].p;
                for (int c = 0; c < 3; ++c) {
                    kio.rp.from_sh[c] = fl.sh[c];
                    kio.rp.mask[c] = fl.mask[c];
                    kio.rp.to_sh[c] = to.sh[c];
                }
                RadixCols none{};
                sorted = radix_sort_shifts(ctx, keys.p, tmp.p, n, shifts, npass, g, 0, none,
                                           none, none, nullptr, dbits, &kio);
            }
            if (--uses[q] == 0) {
                if (!spare.p) spare = std::move(held[q]);
                held[q].reset();
            }
        }
        DBuf<uint64_t> &res = sorted == keys.p ? keys : tmp;
        DBuf<uint64_t> &other = sorted == keys.p ? tmp : keys;
        vd_perm_arrays &pa = idx->perm[p];
        const uint32_t ymask = to.mask[yi], zmask = to.mask[PERM_COLS[p][2]];
        finish_perm(ctx, pa, n, PackedRows{sorted, to.sh[xi], to.sh[yi], ymask, zmask});
        pa.built = true;
        if (uses[p] > 0) held[p] = std::move(res);
        if (!spare.p) spare = std::move(other);
    }
}

// One permutation when the triple needs more than 64 bits: two LSD phases,
// (Y,Z) with X carried as a value, then a stable sort on X (key X << 32 |
// position) and a gather of the (Y,Z) words.
void build_perm_split(vd_ctx *ctx, vd_index *idx, int perm, const uint32_t *cols[3],
                      const int bits[3], uint64_t *keys, uint64_t *tmp) {
    const uint64_t n = idx->n;
    vd_perm_arrays &pa = idx->perm[perm];
    const int xi = PERM_COLS[perm][0], yi = PERM_COLS[perm][1],
              zi = PERM_COLS[perm][2];
    const int bz = bits[zi], by = bits[yi], bx = bits[xi];
    const int sy = bz;
    const uint32_t ymask = by >= 32 ? 0xFFFFFFFFu : ((1u << by) - 1u);
    const uint32_t zmask = bz >= 32 ? 0xFFFFFFFFu : ((1u << bz) - 1u);
    const uint64_t pk_tiles = ceil_div(n, PK_BLOCK * PK_ITEMS);
    const int lbits = by + bz;
    DBuf<uint32_t> xa(ctx, n), xb(ctx, n);
    DBuf<uint32_t> hist(ctx, radix_hist_words(0, lbits));
    fill_u32(ctx, hist.p, 0, radix_hist_words(0, lbits));
    launch(ctx, "index_pack", (double)n * 24.0, dim3(radix::hist_grid(pk_tiles)), dim3(PK_BLOCK), 0,
           k_pack_split, cols[xi], cols[yi], cols[zi], n, sy, keys, xa.p, lbits,
           radix_passes(0, lbits), hist.p);
    uint64_t *lo = keys, *lo_tmp = tmp;
    uint32_t *xv = xa.p, *xv_tmp = xb.p;
    radix_sort_u64_kv(ctx, lo, lo_tmp, xv, xv_tmp, n, 0, lbits, hist.p, true);
    // lo/xv hold phase-1 order; lo_tmp / xv_tmp are free.
    DBuf<uint64_t> k2(ctx, n), k2t(ctx, n);
    const int hi = 32 + bx;
    const int npass2 = radix_passes(32, hi);
    DBuf<uint32_t> hist2(ctx, radix_hist_words(32, hi));
    fill_u32(ctx, hist2.p, 0, radix_hist_words(32, hi));
    launch(ctx, "index_pack", (double)n * 12.0, dim3(radix::hist_grid(pk_tiles)), dim3(PK_BLOCK), 0,
           k_pack_xpos, (const uint32_t *)xv, n, k2.p, hi, npass2, hist2.p);
    uint64_t *s2 = radix_sort_u64(ctx, k2.p, k2t.p, n, 32, hi, hist2.p, true);
    launch(ctx, "index_gather", (double)n * 24.0, dim3(stream_grid(n)), dim3(256), 0,
           k_split_gather, (const uint64_t *)s2, (const uint64_t *)lo, n, xv_tmp, lo_tmp);
    finish_perm(ctx, pa, n, SplitRows{xv_tmp, lo_tmp, sy, ymask, zmask});
    pa.built = true;
}

void free_perm(vd_ctx *ctx, vd_perm_arrays &pa) {
    dfree(ctx, pa.l1v); dfree(ctx, pa.l1o);
    dfree(ctx, pa.l2v); dfree(ctx, pa.l2o);
    dfree(ctx, pa.y); dfree(ctx, pa.z);
    pa = vd_perm_arrays();
}

}  // namespace

// Frees the arrays of every built permutation when the last owner lets go.
std::shared_ptr<void> make_hold(vd_ctx *owner, const vd_perm_arrays (&perm)[VD_NUM_PERMS]) {
    std::vector<vd_perm_arrays> arrays(perm, perm + VD_NUM_PERMS);
    return std::shared_ptr<void>(nullptr, [owner, arrays](void *) mutable {
        try {
            for (auto &pa : arrays) free_perm(owner, pa);
        } catch (...) {
        }
    });
}

PermDev perm_dev(const vd_perm_arrays &pa) {
    return PermDev{pa.n, pa.n_l1, pa.l1v, pa.l1o, pa.y};
}

const vd_perm_arrays &get_perm(const vd_index *idx, int perm) {
    VD_REQUIRE(idx && perm >= 0 && perm < VD_NUM_PERMS, VD_EINVAL,
               "bad index or permutation");
    VD_REQUIRE(idx->perm[perm].built, VD_EINVAL, "permutation not built");
    return idx->perm[perm];
}

}  // namespace vd

namespace vd {
namespace {

// Argument checks shared by the select entry points; the skip intervals
// sorted by lo.
SkipSet check_select(int cols, int clip, const vd_interval_t *skip, int nskip, int distinct) {
    VD_REQUIRE(cols >= 1 && cols <= 3, VD_EINVAL, "cols must be 1, 2 or 3");
    VD_REQUIRE(nskip >= 0 && nskip <= 4 && (nskip == 0 || skip), VD_EINVAL,
               "at most 4 skip intervals");
    VD_REQUIRE(!(clip || nskip || distinct) || cols != 1, VD_EINVAL,
               "clip, skip and distinct need Y as column 0");
    VD_REQUIRE(!distinct || cols == 2, VD_EINVAL, "distinct needs cols == 2");
    SkipSet sk{{{1, 0}, {1, 0}, {1, 0}, {1, 0}}};
    for (int i = 0; i < nskip; ++i) sk.s[i] = skip[i];
    std::sort(sk.s, sk.s + nskip, [](const vd_interval_t &a, const vd_interval_t &b) {
        return a.lo < b.lo;
    });
    for (int i = 1; i < nskip; ++i)
        VD_REQUIRE(sk.s[i].lo > sk.s[i - 1].hi || sk.s[i - 1].lo > sk.s[i - 1].hi, VD_EINVAL,
                   "skip intervals overlap");
    return sk;
}

// Value bounds of a select's columns (Y and / or Z by `cols`).
void select_bounds(vd_rel *r, const vd_perm_arrays &pa, int cols) {
    if (!pa.bounded) return;
    int c = 0;
    if (cols & 2) rel_set_bound(r, c++, pa.ymax);
    if (cols & 1) rel_set_bound(r, c++, pa.zmax);
}

// Rows [src, src + rows) of the permutation's columns as a view.
vd_rel *range_view(const vd_index *idx, const vd_perm_arrays &pa, int cols, uint64_t src,
                   uint64_t rows) {
    uint32_t *vc[2];
    int nc = 0;
    if (cols & 2) vc[nc++] = pa.y + src;
    if (cols & 1) vc[nc++] = pa.z + src;
    vd_rel *r = rel_view(nc, rows, vc, idx->hold);
    select_bounds(r, pa, cols);
    // rows of distinct triples with X (and, for Z alone, Y) fixed
    r->distinct = cols != 2 && idx->distinct();
    return r;
}

// The relation of computed pieces: one run of rows is a view, several are
// copied; distinct then removes adjacent duplicates of Y.
vd_rel *select_pieces_rel(vd_ctx *ctx, const vd_index *idx, const vd_perm_arrays &pa, int cols,
                          int distinct, const Pieces &hp, const Pieces *dp) {
    const uint64_t rows = hp.dst[hp.np];
    if (hp.np <= 1 && !distinct)
        return range_view(idx, pa, cols, hp.np == 1 ? hp.src[0] : 0, rows);
    const int ocols = cols == 3 ? 2 : 1;
    vd_rel *r = rel_new(ctx, ocols, rows);
    select_bounds(r, pa, cols);
    try {
        launch(ctx, "select_copy", (double)rows * 8.0 * ocols, dim3(stream_grid(rows)),
               dim3(256), 0, k_copy_pieces, dp, (const uint32_t *)pa.y, (const uint32_t *)pa.z,
               r->cols[0], ocols == 2 ? r->cols[1] : (uint32_t *)nullptr, cols);
        r->distinct = cols != 2 && idx->distinct();
        if (distinct && rows > 1) {
            int rc = vd_unique_adjacent(ctx, r);
            if (rc != VD_OK) throw Error(rc, ctx->last_error);
        }
        // Y alone, sorted, after the adjacent unique: every value once
        if (distinct) r->distinct = true;
    } catch (...) {
        rel_free(ctx, r);
        throw;
    }
    return r;
}

}  // namespace
}  // namespace vd

using namespace vd;

// ---- batched range select (with or without allocation) ---------------------
// Scratch of vd_select_batch_noalloc: each query's first row, then its row
// count, then the row-offset scan's tile totals.  Every word a call reads is
// written earlier in the same call, so nothing needs clearing between calls.
// Each GPU holds its own copy for the life of the process; calls that
// overlap on one GPU share it (and race).
constexpr uint64_t NA_QUERIES = VD_NOALLOC_MAX_QUERIES;
constexpr uint64_t NA_SCAN_WORDS = 2048;
static __device__ __align__(256) uint64_t scratch[2 * NA_QUERIES + NA_SCAN_WORDS];

namespace {
constexpr int NA_MAX_DEVICES = 64;
std::mutex g_na_mu;
uint64_t *g_na_ptr[NA_MAX_DEVICES];  // this process's address of scratch on each GPU

uint64_t *scratch_ptr(int device) {
    std::lock_guard<std::mutex> lk(g_na_mu);
    VD_REQUIRE(device >= 0 && device < NA_MAX_DEVICES && g_na_ptr[device], VD_EINVAL,
               "no-allocation scratch not resolved on this device");
    return g_na_ptr[device];
}

// Rows of nq lookups written back to back: begin / sizes hold nq words each,
// ws (may be null: the scan allocates) the scan's tile totals.
void select_batch_impl(vd_ctx *ctx, const vd_index *idx, int perm, const uint32_t *d_k1,
                       const uint32_t *d_k2, uint64_t nq, uint64_t *d_row_offsets,
                       uint32_t *d_out_y, uint32_t *d_out_z, uint64_t max_rows,
                       uint64_t *total_rows, uint64_t *begin, uint64_t *sizes, uint64_t *ws) {
    const vd_perm_arrays &pa = get_perm(idx, perm);
    VD_REQUIRE(d_k1 && d_row_offsets, VD_EINVAL, "null select arrays");
    launch(ctx, "select_lookup", (double)nq * 24.0, dim3((unsigned)ceil_div(nq, 256)), dim3(256),
           0, k_lookup_batch, perm_dev(pa), d_k1, d_k2, nq, begin, (uint64_t *)nullptr, sizes);
    if (ws)
        scan_u64_ws(ctx, sizes, d_row_offsets, nq, d_row_offsets + nq, ws, NA_SCAN_WORDS);
    else
        scan_u64(ctx, sizes, d_row_offsets, nq, d_row_offsets + nq);
    const uint64_t total = read_u64(ctx, d_row_offsets + nq);
    *total_rows = total;
    if (max_rows == 0) return;
    VD_REQUIRE(total <= max_rows, VD_ERANGE, "select batch output exceeds max_rows");
    VD_REQUIRE(total == 0 || d_out_y || d_out_z, VD_EINVAL, "no output columns");
    if (total == 0) return;
    launch(ctx, "select_gather", (double)total * 4.0 * ((d_out_y ? 2 : 0) + (d_out_z ? 2 : 0)),
           dim3((unsigned)ceil_div(total, GR_TILE)), dim3(GR_BLOCK), 0, k_gather_ranges,
           (const uint64_t *)d_row_offsets, (const uint64_t *)begin, nq, total,
           (const uint32_t *)pa.y, (const uint32_t *)pa.z, d_out_y, d_out_z);
}
}  // namespace

namespace vd {
void noalloc_init(int device) {
    std::lock_guard<std::mutex> lk(g_na_mu);
    VD_REQUIRE(device >= 0 && device < NA_MAX_DEVICES, VD_EINVAL, "device number too large");
    if (g_na_ptr[device]) return;
    void *p = nullptr;
    VD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_na_ptr[device] = static_cast<uint64_t *>(p);
}
}  // namespace vd


extern "C" {

int vd_index_build(vd_ctx *ctx, const uint32_t *d_s, const uint32_t *d_p,
                   const uint32_t *d_o, uint64_t n, uint32_t perm_mask,
                   vd_index **out) {
    if (!ctx || !out) return VD_EINVAL;
    *out = nullptr;
    vd_index *idx = new (std::nothrow) vd_index();
    if (!idx) return VD_ENOMEM;
    uint32_t maxv[3] = {0, 0, 0};  // the columns' maxima
    try {
        VD_REQUIRE(perm_mask != 0 && perm_mask < (1u << VD_NUM_PERMS), VD_EINVAL,
                   "empty or invalid permutation mask");
        VD_REQUIRE(n == 0 || (d_s && d_p && d_o), VD_EINVAL, "null triples");
        VD_REQUIRE(n < (1ull << 30), VD_ERANGE,
                   "index build supports fewer than 2^30 triples per call");
        idx->owner = ctx;
        idx->n = n;
        idx->mask = perm_mask;
        const uint32_t *cols[3] = {d_s, d_p, d_o};
        int bits[3] = {1, 1, 1};
        if (n > 0)
            for (int c = 0; c < 3; ++c) {
                maxv[c] = max_u32(ctx, cols[c], n);
                bits[c] = std::max(1, bits_for(maxv[c]));
            }
        if (n > 0 && bits[0] + bits[1] + bits[2] <= 64) {
            build_packed(ctx, idx, perm_mask, cols, bits);
        } else {
            DBuf<uint64_t> keys(ctx, n), tmp(ctx, n);
            for (int p = 0; p < VD_NUM_PERMS; ++p) {
                if (!(perm_mask & VD_PERM_BIT(p))) continue;
                if (n == 0) {
                    vd_perm_arrays &pa = idx->perm[p];
                    pa.built = true;
                    pa.y = (uint32_t *)dalloc(ctx, 4);
                    pa.z = (uint32_t *)dalloc(ctx, 4);
                    pa.l1v = (uint32_t *)dalloc(ctx, 4);
                    pa.l1o = (uint32_t *)dalloc(ctx, 4);
                    pa.l2v = (uint32_t *)dalloc(ctx, 4);
                    pa.l2o = (uint32_t *)dalloc(ctx, 4);
                    continue;
                }
                build_perm_split(ctx, idx, p, cols, bits, keys.p, tmp.p);
            }
        }
    } catch (const Error &e) {
        ctx->last_error = e.what();
        for (auto &pa : idx->perm) free_perm(ctx, pa);
        delete idx;
        return e.code;
    } catch (const std::exception &e) {
        ctx->last_error = e.what();
        for (auto &pa : idx->perm) free_perm(ctx, pa);
        delete idx;
        return VD_EINVAL;
    }
    for (int p = 0; p < VD_NUM_PERMS; ++p) {
        vd_perm_arrays &pa = idx->perm[p];
        if (!pa.built) continue;
        pa.bounded = true;
        pa.ymax = maxv[PERM_COLS[p][1]];
        pa.zmax = maxv[PERM_COLS[p][2]];
    }
    try {
        idx->hold = make_hold(ctx, idx->perm);
    } catch (...) {
        for (auto &pa : idx->perm) free_perm(ctx, pa);
        delete idx;
        return VD_ENOMEM;
    }
    *out = idx;
    return VD_OK;
}

namespace {

// Rows of runs: out[r] = values of the run holding row r.
void runs_expand(vd_ctx *ctx, const uint32_t *values, const uint32_t *offsets, uint64_t nruns,
                 uint64_t n, uint32_t *out) {
    if (n == 0) return;
    VD_REQUIRE(nruns > 0, VD_EINVAL, "rows without runs");
    launch(ctx, "runs_expand", 8.0 * (double)n,
           dim3((unsigned)ceil_div(ceil_div(n, (uint64_t)RUN_ROWS), 256)), dim3(256), 0,
           k_runs_expand, values, offsets, nruns, n, out);
}

// The L2 of rows with the given L1 and Y column: (X, Y) run heads.
void l2_from_rows(vd_ctx *ctx, const uint32_t *l1o, uint64_t n_l1, const uint32_t *y, uint64_t n,
                  uint32_t **l2v, uint32_t **l2o, uint64_t *n_l2) {
    DBuf<uint32_t> head(ctx, n);
    DBuf<uint64_t> pos(ctx, n + 1);
    fill_u32(ctx, head.p, 0, n);
    launch(ctx, "l2_heads", 4.0 * (double)n_l1, dim3((unsigned)ceil_div(n_l1, 256)), dim3(256), 0,
           k_mark_l1, l1o, n_l1, n, head.p);
    launch(ctx, "l2_heads", 8.0 * (double)n, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
           k_mark_y, y, n, head.p);
    scan_u32_to_u64(ctx, head.p, pos.p, n, pos.p + n);
    const uint64_t m = read_u64(ctx, pos.p + n);
    *l2v = (uint32_t *)dalloc(ctx, (m ? m : 1) * 4);
    *l2o = (uint32_t *)dalloc(ctx, (m ? m : 1) * 4);
    launch(ctx, "l2_heads", 16.0 * (double)n, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
           k_scatter_heads, (const uint32_t *)head.p, (const uint64_t *)pos.p, y, n, *l2v, *l2o);
    *n_l2 = m;
}

}  // namespace

int vd_runs_expand(vd_ctx *ctx, const uint32_t *values, const uint32_t *offsets, uint64_t nruns,
                   uint64_t n, uint32_t *out) {
    if (!ctx || (n && (!values || !offsets || !out))) return VD_EINVAL;
    VD_TRY(ctx, runs_expand(ctx, values, offsets, nruns, n, out));
}

int vd_index_from_arrays(vd_ctx *ctx, uint32_t perm_mask, const vd_perm_view_t *views,
                         vd_index **out) {
    if (!ctx || !out || !views) return VD_EINVAL;
    *out = nullptr;
    vd_index *idx = new (std::nothrow) vd_index();
    if (!idx) return VD_ENOMEM;
    try {
        VD_REQUIRE(perm_mask != 0 && perm_mask < (1u << VD_NUM_PERMS), VD_EINVAL,
                   "empty or invalid permutation mask");
        idx->owner = ctx;
        idx->mask = perm_mask;
        bool first = true;
        for (int p = 0; p < VD_NUM_PERMS; ++p) {
            if (!(perm_mask & VD_PERM_BIT(p))) continue;
            const vd_perm_view_t &v = views[p];
            VD_REQUIRE(first || v.n == idx->n, VD_EINVAL, "permutations of different sizes");
            const bool have_l2 = v.l2_values && v.l2_offsets;
            VD_REQUIRE(v.n < (1ull << 32) && v.n_l1 <= v.n &&
                           (!have_l2 || (v.n_l2 <= v.n && v.n_l1 <= v.n_l2)),
                       VD_EINVAL, "inconsistent permutation sizes");
            VD_REQUIRE(v.n == 0 || (v.col_z && v.l1_values && v.l1_offsets &&
                                    (v.col_y || have_l2)),
                       VD_EINVAL, "null permutation array");
            first = false;
            idx->n = v.n;
            vd_perm_arrays &pa = idx->perm[p];
            auto copy = [&](const uint32_t *src, uint64_t cnt) {
                uint32_t *d = (uint32_t *)dalloc(ctx, (cnt ? cnt : 1) * 4);
                if (cnt) copy_d2d(ctx, d, src, cnt * 4);
                return d;
            };
            pa.n = v.n;
            pa.n_l1 = v.n_l1;
            pa.z = copy(v.col_z, v.n);
            pa.l1v = copy(v.l1_values, v.n_l1);
            pa.l1o = copy(v.l1_offsets, v.n_l1);
            if (have_l2) {
                pa.n_l2 = v.n_l2;
                pa.l2v = copy(v.l2_values, v.n_l2);
                pa.l2o = copy(v.l2_offsets, v.n_l2);
            }
            if (v.col_y) {
                pa.y = copy(v.col_y, v.n);
            } else {  // Y of each row from its (X, Y) run
                pa.y = (uint32_t *)dalloc(ctx, (v.n ? v.n : 1) * 4);
                runs_expand(ctx, pa.l2v, pa.l2o, pa.n_l2, v.n, pa.y);
            }
            if (!have_l2) {
                if (v.n) {
                    l2_from_rows(ctx, pa.l1o, pa.n_l1, pa.y, v.n, &pa.l2v, &pa.l2o, &pa.n_l2);
                } else {
                    pa.l2v = copy(nullptr, 0);
                    pa.l2o = copy(nullptr, 0);
                }
            }
            pa.built = true;
        }
        idx->hold = make_hold(ctx, idx->perm);
    } catch (const Error &e) {
        ctx->last_error = e.what();
        for (auto &pa : idx->perm) free_perm(ctx, pa);
        delete idx;
        return e.code;
    } catch (const std::exception &e) {
        ctx->last_error = e.what();
        for (auto &pa : idx->perm) free_perm(ctx, pa);
        delete idx;
        return VD_EINVAL;
    }
    *out = idx;
    return VD_OK;
}

int vd_index_free(vd_ctx *ctx, vd_index *idx) {
    if (!idx) return VD_EINVAL;
    vd_ctx *owner = idx->owner ? idx->owner : ctx;
    VD_TRY(owner, {
        idx->hold.reset();  // arrays go when no select view still reads them
        delete idx;
    });
}

uint32_t vd_index_perms(const vd_index *idx) { return idx ? idx->mask : 0; }

int vd_index_view(const vd_index *idx, int perm, vd_perm_view_t *out) {
    if (!idx || !out || perm < 0 || perm >= VD_NUM_PERMS || !idx->perm[perm].built)
        return VD_EINVAL;
    const vd_perm_arrays &pa = idx->perm[perm];
    out->n = pa.n;
    out->n_l1 = pa.n_l1;
    out->n_l2 = pa.n_l2;
    out->l1_values = pa.l1v;
    out->l1_offsets = pa.l1o;
    out->l2_values = pa.l2v;
    out->l2_offsets = pa.l2o;
    out->col_y = pa.y;
    out->col_z = pa.z;
    return VD_OK;
}

int vd_index_lookup_batch(vd_ctx *ctx, const vd_index *idx, int perm,
                          const uint32_t *d_k1, const uint32_t *d_k2,
                          uint64_t nq, uint64_t *d_begin, uint64_t *d_end) {
    if (!ctx) return VD_EINVAL;
    VD_TRY(ctx, {
        const vd_perm_arrays &pa = get_perm(idx, perm);
        VD_REQUIRE(nq == 0 || (d_k1 && d_begin && d_end), VD_EINVAL,
                   "null lookup arrays");
        launch(ctx, "index_lookup", (double)nq * 24.0,
               dim3((unsigned)ceil_div(nq, 256)), dim3(256), 0, k_lookup_batch,
               perm_dev(pa), d_k1, d_k2, nq, d_begin, d_end, (uint64_t *)nullptr);
    });
}

int vd_index_stats(vd_ctx *ctx, const vd_index *idx, int perm, uint64_t *d_rows,
                   uint64_t *d_pairs) {
    if (!ctx) return VD_EINVAL;
    VD_TRY(ctx, {
        const vd_perm_arrays &pa = get_perm(idx, perm);
        VD_REQUIRE(pa.n_l1 == 0 || (d_rows && d_pairs), VD_EINVAL, "null output arrays");
        launch(ctx, "index_stats", (double)pa.n_l1 * 24.0,
               dim3((unsigned)ceil_div(pa.n_l1, 256)), dim3(256), 0, k_index_stats,
               (const uint32_t *)pa.l1o, pa.n_l1, (const uint32_t *)pa.l2o, pa.n_l2, pa.n,
               d_rows, d_pairs);
    });
}

int vd_index_histogram(vd_ctx *ctx, const vd_index *idx, int perm, int kind, uint32_t param,
                       uint32_t id_limit, uint32_t *lowers, uint32_t *uppers, uint64_t *freqs,
                       uint64_t cap, uint64_t *nbuckets) {
    if (!ctx || !nbuckets || (cap && (!lowers || !uppers || !freqs))) return VD_EINVAL;
    *nbuckets = 0;
    VD_TRY(ctx, {
        VD_REQUIRE(kind == VD_HIST_EQUAL_WIDTH || kind == VD_HIST_EQUAL_DEPTH, VD_EINVAL,
                   "unknown histogram kind");
        VD_REQUIRE(param > 0, VD_EINVAL, "histogram width / depth must be positive");
        const vd_perm_arrays &pa = get_perm(idx, perm);
        // the L1 entries with a key below id_limit, and their rows
        std::vector<uint32_t> hv;
        uint64_t m = 0, rows_in = 0;
        if (pa.n_l1) {
            DBuf<uint32_t> k(ctx, 1);
            DBuf<uint64_t> be(ctx, 2);
            VD_HIP(hipMemcpyAsync(k.p, &id_limit, 4, hipMemcpyHostToDevice, ctx->stream));
            // first L1 entry >= id_limit, as a one-constant lookup's begin
            launch(ctx, "index_lookup", 24.0, dim3(1), dim3(64), 0, k_hist_limit,
                   (const uint32_t *)pa.l1v, (const uint32_t *)pa.l1o, pa.n_l1, pa.n,
                   (const uint32_t *)k.p, be.p);
            uint64_t r[2];
            read_words(ctx, r, be.p, 2);
            m = r[0];
            rows_in = r[1];
        }
        if (m == 0) return VD_OK;
        uint32_t ends[2];
        download(ctx, &ends[0], pa.l1v, 4);
        download(ctx, &ends[1], pa.l1v + m - 1, 4);
        if (kind == VD_HIST_EQUAL_WIDTH) {
            const uint64_t nb = (uint64_t)(ends[1] - ends[0]) / param + 1;
            *nbuckets = nb;
            VD_REQUIRE(nb <= cap, VD_ERANGE, "more histogram buckets than cap");
            DBuf<uint64_t> f(ctx, nb);
            launch(ctx, "histogram", 16.0 * (double)nb, dim3((unsigned)ceil_div(nb, 256)),
                   dim3(256), 0, k_hist_width, (const uint32_t *)pa.l1v,
                   (const uint32_t *)pa.l1o, m, rows_in, ends[0], ends[1], param, nb, f.p);
            download(ctx, freqs, f.p, nb * sizeof(uint64_t));
            for (uint64_t b = 0; b < nb; ++b) {
                const uint64_t lo = (uint64_t)ends[0] + b * param;
                lowers[b] = (uint32_t)lo;
                uppers[b] = (uint32_t)std::min<uint64_t>(lo + param - 1, ends[1]);
            }
        } else {
            // every L1 entry's successor bucket start on the device, the chain
            // of starts from entry 0 on the host (one step per bucket), then
            // the buckets' keys and rows on the device
            DBuf<uint32_t> nxt(ctx, m);
            launch(ctx, "histogram", 8.0 * (double)m, dim3((unsigned)ceil_div(m, 256)), dim3(256),
                   0, k_hist_depth_next, (const uint32_t *)pa.l1o, m, rows_in, param, nxt.p);
            std::vector<uint32_t> hn(m);
            download(ctx, hn.data(), nxt.p, m * sizeof(uint32_t));
            std::vector<uint32_t> chain;
            for (uint64_t g = 0; g < m; g = hn[g]) chain.push_back((uint32_t)g);
            const uint64_t nb = chain.size();
            *nbuckets = nb;
            VD_REQUIRE(nb <= cap, VD_ERANGE, "more histogram buckets than cap");
            DBuf<uint32_t> dc(ctx, nb);
            upload(ctx, dc.p, chain.data(), nb * sizeof(uint32_t));
            DBuf<uint32_t> lo(ctx, nb), hi(ctx, nb);
            DBuf<uint64_t> f(ctx, nb);
            launch(ctx, "histogram", 24.0 * (double)nb, dim3((unsigned)ceil_div(nb, 256)),
                   dim3(256), 0, k_hist_depth_gather, (const uint32_t *)pa.l1v,
                   (const uint32_t *)pa.l1o, m, rows_in, (const uint32_t *)nxt.p,
                   (const uint32_t *)dc.p, nb, lo.p, hi.p, f.p);
            download(ctx, lowers, lo.p, nb * 4);
            download(ctx, uppers, hi.p, nb * 4);
            download(ctx, freqs, f.p, nb * 8);
        }
    });
}

int vd_index_lookup(vd_ctx *ctx, const vd_index *idx, int perm, uint32_t k1,
                    int has_k2, uint32_t k2, uint64_t *begin, uint64_t *end) {
    if (!ctx || !begin || !end) return VD_EINVAL;
    VD_TRY(ctx, {
        const vd_perm_arrays &pa = get_perm(idx, perm);
        DBuf<uint32_t> k(ctx, 2);
        DBuf<uint64_t> be(ctx, 2);
        uint32_t hk[2] = {k1, has_k2 ? k2 : VD_UNBOUND};
        VD_HIP(hipMemcpyAsync(k.p, hk, 8, hipMemcpyHostToDevice, ctx->stream));
        launch(ctx, "index_lookup", 24.0, dim3(1), dim3(64), 0, k_lookup_batch,
               perm_dev(pa), (const uint32_t *)k.p, (const uint32_t *)(k.p + 1), (uint64_t)1,
               be.p, be.p + 1, (uint64_t *)nullptr);
        uint64_t r[2];
        read_words(ctx, r, be.p, 2);
        *begin = r[0];
        *end = r[1];
    });
}

int vd_select(vd_ctx *ctx, const vd_index *idx, int perm, uint64_t begin,
              uint64_t end, int cols, int clip, uint32_t lo, uint32_t hi,
              const vd_interval_t *skip, int nskip, int distinct, vd_rel **out) {
    if (!ctx || !out) return VD_EINVAL;
    *out = nullptr;
    VD_TRY(ctx, {
        const vd_perm_arrays &pa = get_perm(idx, perm);
        VD_REQUIRE(begin <= end && end <= pa.n, VD_EINVAL, "row range out of bounds");
        const SkipSet sk = check_select(cols, clip, skip, nskip, distinct);
        if (!clip && !nskip && !distinct) {
            // An unfiltered range is already contiguous in the permutation's
            // columns: the relation reads it in place.
            *out = range_view(idx, pa, cols, begin, end - begin);
            return VD_OK;
        }
        DBuf<Pieces> dp(ctx, 1);
        const Readback rb = rb_next(ctx);
        launch(ctx, "select_pieces", 0.0, dim3(1), dim3(64), 0, k_select_pieces,
               (const uint32_t *)pa.y, begin, end, clip, lo, hi, sk, nskip, dp.p, rb);
        uint64_t w[PIECE_WORDS];
        rb_wait(ctx, rb, w, PIECE_WORDS);
        const Pieces hp = pieces_from_words(w);
        *out = select_pieces_rel(ctx, idx, pa, cols, distinct, hp, dp.p);
    });
}

int vd_select_key(vd_ctx *ctx, const vd_index *idx, int perm, uint32_t k1, int has_k2,
                  uint32_t k2, int cols, int clip, uint32_t lo, uint32_t hi,
                  const vd_interval_t *skip, int nskip, int distinct, vd_rel **out,
                  uint64_t *begin, uint64_t *end) {
    if (!ctx || !out) return VD_EINVAL;
    *out = nullptr;
    VD_TRY(ctx, {
        const vd_perm_arrays &pa = get_perm(idx, perm);
        const SkipSet sk = check_select(cols, clip, skip, nskip, distinct);
        DBuf<KeyPieces> dk(ctx, 1);
        const Readback rb = rb_next(ctx);
        launch(ctx, "select_pieces", 0.0, dim3(1), dim3(64), 0, k_lookup_pieces, perm_dev(pa),
               k1, has_k2, k2, clip, lo, hi, sk, nskip, dk.p, rb);
        uint64_t w[2 + PIECE_WORDS];
        rb_wait(ctx, rb, w, 2 + PIECE_WORDS);
        if (begin) *begin = w[0];
        if (end) *end = w[1];
        *out = select_pieces_rel(ctx, idx, pa, cols, distinct, pieces_from_words(w + 2),
                                 &dk.p->p);
    });
}

int vd_select_batch(vd_ctx *ctx, const vd_index *idx, int perm,
                    const uint32_t *d_k1, const uint32_t *d_k2, uint64_t nq,
                    uint64_t *d_row_offsets, uint32_t *d_out_y,
                    uint32_t *d_out_z, uint64_t max_rows, uint64_t *total_rows) {
    if (!ctx || !total_rows) return VD_EINVAL;
    *total_rows = 0;
    VD_TRY(ctx, {
        if (nq == 0) {
            get_perm(idx, perm);
        } else {
            DBuf<uint64_t> begin(ctx, nq), sizes(ctx, nq);
            select_batch_impl(ctx, idx, perm, d_k1, d_k2, nq, d_row_offsets, d_out_y, d_out_z,
                              max_rows, total_rows, begin.p, sizes.p, nullptr);
        }
    });
}

int vd_select_batch_noalloc(vd_ctx *ctx, const vd_index *idx, int perm,
                            const uint32_t *d_k1, const uint32_t *d_k2, uint64_t nq,
                            uint64_t *d_row_offsets, uint32_t *d_out_y,
                            uint32_t *d_out_z, uint64_t max_rows, uint64_t *total_rows) {
    if (!ctx || !total_rows) return VD_EINVAL;
    *total_rows = 0;
    VD_TRY(ctx, {
        VD_REQUIRE(nq <= VD_NOALLOC_MAX_QUERIES, VD_ERANGE,
                   "more queries than VD_NOALLOC_MAX_QUERIES");
        uint64_t *sp = scratch_ptr(ctx->device);
        if (nq == 0) {
            get_perm(idx, perm);
        } else {
            select_batch_impl(ctx, idx, perm, d_k1, d_k2, nq, d_row_offsets, d_out_y, d_out_z,
                              max_rows, total_rows, sp, sp + NA_QUERIES, sp + 2 * NA_QUERIES);
        }
    });
}

}  // extern "C"
