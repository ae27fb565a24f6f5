// Batched operators: one operator applied to many independent inputs (the
// queries of a batch) with a fixed number of launches and one stream
// synchronisation.  Inputs are laid end to end in a virtual index space:
//  * per-input tile ranges (tbase) map a block to its input, so a tile never
//    spans two inputs (bounds, max, pack, unique);
//  * per-input row bases (lbase / rbase) map a global row or output index
//    back to its input (join expansion, sort gather, select gather).
// The kernels are the single-input ones of vd_join / vd_relops / vd_index
// with that extra indirection.

#include <algorithm>
#include <vector>

#include "vd_index.hpp"
#include "vd_radix_dev.hpp"

namespace vd {
namespace {

constexpr int MB_BLOCK = 256;
constexpr int MB_ITEMS = 8;
constexpr int MB_TILE = MB_BLOCK * MB_ITEMS;  // rows per tile
constexpr int MJ_BLOCK = 256;
constexpr int MJ_RCAP = 8192;  // staged right keys (32 KB)
// The bounds kernel's blocks: 512 threads x 4 rows per MB_TILE-row tile
// (its 32 KB of staged keys allow 4 blocks per CU: 32 waves).
constexpr int MJB_BLOCK = 512;
constexpr int MJB_ITEMS = 4;
static_assert(MJB_BLOCK * MJB_ITEMS == MB_TILE, "bounds blocks cover a tile");

__device__ __forceinline__ int find_seg(const uint64_t *base, int n, uint64_t x) {
    // last j in [0, n) with base[j] <= x (base[0] == 0, non-decreasing)
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (base[mid] <= x) lo = mid + 1; else hi = mid;
    }
    return lo - 1;
}

// An LDS-broadcast block index made wave-uniform, so the per-input
// descriptor behind it is read with scalar loads.
__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

unsigned grid_cap(uint64_t n) {
    uint64_t g = ceil_div(n, 256 * 4);
    if (g > 16384) g = 16384;
    return g ? (unsigned)g : 1u;
}

// ---- select ------------------------------------------------------------------
struct SelJobDev {
    PermDev pd;
    uint32_t k1, k2;
};

// Words of a one-block kernel read back straight from the block (each
// thread's stores, its system-scope fence, one barrier, then the sequence
// word), so no copy kernel runs after it.  rb.dst == nullptr: nothing.
__device__ __forceinline__ void block_publish(const Readback &rb) {
    if (!rb.dst) return;
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0)
        __hip_atomic_store(rb.dst + RB_WORDS, rb.seq, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_SYSTEM);
}

// One wave per lookup job (wave-cooperative searches: about log64 of the
// range in dependent loads).
constexpr int LK_BLOCK = 512;
constexpr int LK_JOBS = LK_BLOCK / 64;  // jobs per block

// range[2j], range[2j+1] = the rows [begin, end) matching job j (also to
// rb.dst when the launch is one block and rb.dst is set).
__global__ __launch_bounds__(LK_BLOCK) void k_sel_lookup_multi(const SelJobDev *jobs, int nj,
                                                               uint64_t *range, Readback rb) {
    const int j = (int)((blockIdx.x * blockDim.x + threadIdx.x) / 64);
    if (j < nj) {  // the same in every lane of the wave
        const SelJobDev &J = jobs[j];
        uint64_t b, e;
        wave_lookup(J.pd, J.k1, J.k2, J.k2 != VD_UNBOUND, &b, &e);
        if (dev::lane_id() == 0) {
            range[2 * j] = b;
            range[2 * j + 1] = e;
            if (rb.dst) {
                rb.dst[2 * j] = b;
                rb.dst[2 * j + 1] = e;
            }
        }
    }
    block_publish(rb);
}

// The lookup of k_sel_lookup_multi plus the first and last value of the
// column after the constants over the rows found (sorted there, so its
// minimum and maximum): out[3j] = begin, out[3j+1] = end, out[3j+2] =
// first | last << 32 (0 for an empty range).
struct BoundJobDev {
    PermDev pd;
    const uint32_t *col;
    uint32_t k1, k2;
};

__global__ __launch_bounds__(LK_BLOCK) void k_sel_bounds_multi(const BoundJobDev *jobs, int nj,
                                                               uint64_t *out, Readback rb) {
    const int j = (int)((blockIdx.x * blockDim.x + threadIdx.x) / 64);
    if (j < nj) {  // the same in every lane of the wave
        const BoundJobDev &J = jobs[j];
        uint64_t b, e;
        wave_lookup(J.pd, J.k1, J.k2, J.k2 != VD_UNBOUND, &b, &e);
        if (dev::lane_id() == 0) {
            const uint64_t ends =
                e > b ? ((uint64_t)J.col[b] | ((uint64_t)J.col[e - 1] << 32)) : 0;
            out[3 * j] = b;
            out[3 * j + 1] = e;
            out[3 * j + 2] = ends;
            if (rb.dst) {
                rb.dst[3 * j] = b;
                rb.dst[3 * j + 1] = e;
                rb.dst[3 * j + 2] = ends;
            }
        }
    }
    block_publish(rb);
}

// ---- join --------------------------------------------------------------------
struct JPairDev {
    const uint32_t *lc[VD_MAX_COLS];
    const uint32_t *rc[VD_MAX_COLS];
    uint32_t *oc[VD_MAX_COLS];
    uint64_t nl, nr;
    uint64_t lbase;  // first global left row
    uint64_t obase;  // first global output row
    int ncl, ncr;
};

// Right-row range of every bounds tile, all tiles in parallel: a group of
// MR_GROUP lanes per tile searches with MR_GROUP samples per round.
constexpr int MR_GROUP = 16;
__global__ __launch_bounds__(256) void k_join_ranges_multi(const JPairDev *P, int np,
                                                           const uint64_t *tbase, uint64_t ntiles,
                                                           uint64_t *ranges) {
    const uint64_t g = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / MR_GROUP;
    const uint64_t b = g / 2;
    if (b >= ntiles) return;  // whole groups leave together
    const int j = find_seg(tbase, np, b);
    const JPairDev &D = P[j];
    const uint32_t *lk = D.lc[0], *rk = D.rc[0];
    const uint64_t base = (b - tbase[j]) * (uint64_t)MB_TILE;
    // even group: first right row >= the tile's first key; odd group: first
    // right row > its last key (one code path, so both run at once)
    const bool upper = g % 2 != 0;
    const uint32_t key = lk[upper ? min(base + MB_TILE, D.nl) - 1 : base];
    const uint64_t end = dev::group_bound<MR_GROUP>(0, D.nr, [&](uint64_t q) {
        return upper ? rk[q] > key : rk[q] >= key;
    });
    if (threadIdx.x % MR_GROUP == 0) ranges[2 * b + (upper ? 1 : 0)] = end;
}

// As k_join_bounds (vd_join.hip): the right range, or a directory of every
// k-th key of it, in LDS; then the tile's exclusive scan of the counts.
// Block b is tile b of the batched row space, in which every input's left
// rows start on a tile (its rows past nl are padding with no matches):
// lower / loc are written for all OFF_TILE rows of the tile.
static_assert(MB_TILE == OFF_TILE, "bounds tiles are the offset tiles");
__global__ __launch_bounds__(MJB_BLOCK) void k_join_bounds_multi(
    const JPairDev *P, int np, const uint64_t *tbase, const uint64_t *ranges, uint32_t *lower,
    uint32_t *loc, uint64_t *tile_total, uint64_t *wide) {
    static_assert(MJ_RCAP * 4 >= MB_TILE * 12, "counts and offsets reuse the staging buffer");
    __shared__ uint32_t s_r[MJ_RCAP];
    __shared__ uint64_t s_tmp[MJB_BLOCK / 64];
    __shared__ int s_j;
    if (threadIdx.x == 0) s_j = find_seg(tbase, np, blockIdx.x);
    __syncthreads();
    const int j = uniform(s_j);
    const JPairDev &D = P[j];
    const uint32_t *lk = D.lc[0], *rk = D.rc[0];
    const uint64_t base = (blockIdx.x - tbase[j]) * (uint64_t)MB_TILE;
    const uint64_t nl = D.nl;
    const uint64_t rlo = ranges[2 * blockIdx.x], rhi = ranges[2 * blockIdx.x + 1], m = rhi - rlo;
    const uint64_t k = m <= (uint64_t)MJ_RCAP ? 1 : (m + MJ_RCAP - 1) / MJ_RCAP;
    const uint64_t nd = (m + k - 1) / k;
    // Staging and the left keys in batches of loads issued before their use.
    constexpr int SB = 8;
    for (uint64_t t0 = 0; t0 < nd; t0 += (uint64_t)SB * MJB_BLOCK) {
        uint32_t v[SB];
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            const uint64_t t = t0 + (uint64_t)u * MJB_BLOCK + threadIdx.x;
            if (t < nd) v[u] = rk[rlo + t * k];
        }
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            const uint64_t t = t0 + (uint64_t)u * MJB_BLOCK + threadIdx.x;
            if (t < nd) s_r[t] = v[u];
        }
    }
    uint32_t lkey[MJB_ITEMS];
#pragma unroll
    for (int it = 0; it < MJB_ITEMS; ++it) {
        const uint64_t i = base + (uint64_t)it * MJB_BLOCK + threadIdx.x;
        lkey[it] = i < nl ? lk[i] : 0u;
    }
    __syncthreads();
    // first staged entry >= key (t0) and > key (t1) of every row
    uint32_t t0[MJB_ITEMS], t1[MJB_ITEMS];
    dev::lockstep_bounds<MJB_ITEMS>(s_r, (uint32_t)nd, lkey, t0, t1);
    const uint64_t row0 = (uint64_t)blockIdx.x * MB_TILE;  // tile b in the row space
    // bounds relative to rlo (with a directory: exact within its gaps)
    uint32_t lo[MJB_ITEMS], hi[MJB_ITEMS];
    if (k > 1) {
        bool valid[MJB_ITEMS];
#pragma unroll
        for (int it = 0; it < MJB_ITEMS; ++it)
            valid[it] = base + (uint64_t)it * MJB_BLOCK + threadIdx.x < nl;
        dev::lockstep_gap_bounds<MJB_ITEMS>([&](uint32_t q) { return rk[rlo + q]; }, k, m, lkey,
                                           valid, t0, t1, lo, hi);
    } else {
#pragma unroll
        for (int it = 0; it < MJB_ITEMS; ++it) {
            lo[it] = t0[it];
            hi[it] = t1[it];
        }
    }
    uint32_t cnt[MJB_ITEMS];
#pragma unroll
    for (int it = 0; it < MJB_ITEMS; ++it) {
        const uint64_t i = base + (uint64_t)it * MJB_BLOCK + threadIdx.x;
        uint32_t c = 0, low = 0;
        if (i < nl) {
            low = (uint32_t)(rlo + lo[it]);
            c = hi[it] - lo[it];
        }
        lower[row0 + (uint64_t)it * MJB_BLOCK + threadIdx.x] = low;
        cnt[it] = c;
    }
    __syncthreads();  // the staged right keys are no longer read
    uint64_t *s_off = reinterpret_cast<uint64_t *>(s_r);
    uint32_t *s_cnt = reinterpret_cast<uint32_t *>(s_off + MB_TILE);
#pragma unroll
    for (int it = 0; it < MJB_ITEMS; ++it) s_cnt[it * MJB_BLOCK + threadIdx.x] = cnt[it];
    __syncthreads();
    // Thread t scans rows t*ITEMS .. t*ITEMS+ITEMS-1 of the tile.
    uint64_t run = 0;
    uint32_t mine[MJB_ITEMS];
#pragma unroll
    for (int it = 0; it < MJB_ITEMS; ++it) {
        mine[it] = s_cnt[threadIdx.x * MJB_ITEMS + it];
        run += mine[it];
    }
    uint64_t tot;
    uint64_t ex = dev::block_excl_scan<MJB_BLOCK>(run, s_tmp, &tot);
#pragma unroll
    for (int it = 0; it < MJB_ITEMS; ++it) {
        s_off[threadIdx.x * MJB_ITEMS + it] = ex;
        ex += mine[it];
    }
    if (threadIdx.x == 0) {
        tile_total[blockIdx.x] = tot;
        if (tot >> 32) atomicOr((unsigned long long *)wide, 1ull);  // offsets beyond 32 bits
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < MJB_ITEMS; ++it)
        loc[row0 + (uint64_t)it * MJB_BLOCK + threadIdx.x] =
            (uint32_t)s_off[it * MJB_BLOCK + threadIdx.x];
}

// Output size of every pair from the global offsets.
__global__ void k_pair_totals(const JPairDev *P, int np, TileOff off, uint64_t nl,
                              uint64_t ntiles, uint64_t *tot) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= np) return;
    auto at = [&](uint64_t i) { return i < nl ? off(i) : off.tpre[ntiles]; };
    tot[j] = at(P[j].lbase + P[j].nl) - at(P[j].lbase);
}

__global__ __launch_bounds__(256) void k_join_partition_multi(TileOff off, uint64_t nl,
                                                              uint64_t total, int nv,
                                                              uint64_t nblocks, uint64_t *mp) {
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nblocks) return;
    const uint64_t diag = min(b * (uint64_t)nv, nl + total);
    uint64_t lo = 0, hi = nl;
    while (lo < hi) {
        const uint64_t mid = lo + ((hi - lo) >> 1);
        if (mid + off(mid) < diag) lo = mid + 1; else hi = mid;
    }
    mp[b] = lo;
}

// NV: merge-path diagonal per block (the join_nv tunable).
template <int MJ_NV>
__global__ __launch_bounds__(MJ_BLOCK) void k_join_expand_multi(
    TileOff off, const uint32_t *lower, const uint64_t *mp, uint64_t nl, uint64_t total,
    const JPairDev *P, int np, const uint64_t *lbase) {
    constexpr int ITEMS = MJ_NV / MJ_BLOCK;
    static_assert(MJ_NV < 65536, "offsets within a diagonal in 16 bits");
    // Offsets relative to the block's first output b0 (the first staged
    // row, which may start earlier, is staged as 0 with its exact offset in
    // s_off0), as the single join's expansion stages them.
    __shared__ uint16_t s_rel[MJ_NV + 2];
    __shared__ uint32_t s_low[MJ_NV + 1];
    __shared__ uint16_t s_pair[MJ_NV + 1];  // pair of each staged left row
    __shared__ uint64_t s_off0;
    const uint64_t blk = blockIdx.x;
    const uint64_t d0 = blk * MJ_NV;
    const uint64_t d1 = min(d0 + MJ_NV, nl + total);
    const uint64_t a0 = mp[blk], a1 = mp[blk + 1];
    const uint64_t b0 = d0 - a0, b1 = d1 - a1;
    if (b1 <= b0) return;
    const uint64_t sa = a0 > 0 ? a0 - 1 : 0;
    const int m = (int)(a1 - sa);
    // Staging loads unconditional (indices clamped to the last row), so all
    // of a batch are in flight together.
    constexpr int SB = 4;
    for (int t0 = 0; t0 < m; t0 += SB * MJ_BLOCK) {
        uint64_t vpre[SB];
        uint32_t vloc[SB], vl[SB];
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            const uint64_t i = sa + (uint64_t)min(t0 + u * MJ_BLOCK + (int)threadIdx.x, m - 1);
            vloc[u] = off.loc[i];
            vpre[u] = off.tpre[i / OFF_TILE];
            vl[u] = lower[i];
        }
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            const int t = t0 + u * MJ_BLOCK + threadIdx.x;
            if (t < m) {
                const uint64_t vo = vpre[u] + vloc[u];
                s_rel[t] = t == 0 ? (uint16_t)0 : (uint16_t)(vo - b0);
                s_low[t] = vl[u];
                s_pair[t] = (uint16_t)find_seg(lbase, np, sa + t);
                if (t == 0) s_off0 = vo;
            }
        }
    }
    __syncthreads();
    const int nout = (int)(b1 - b0);
    // Right row of output b0 + q, whose left row is staged row k.
    auto right_row = [&](int k, int q) -> uint32_t {
        return k == 0 ? s_low[0] + (uint32_t)(b0 + (uint64_t)q - s_off0)
                      : s_low[k] + (uint32_t)(q - (int)s_rel[k]);
    };
    if (s_pair[0] == s_pair[m - 1]) {
        // One join for the whole block (the common case): its descriptor is
        // uniform, columns are written one after the other, coalesced.
        const JPairDev &D = P[uniform(s_pair[0])];
        uint32_t rrow[ITEMS];
        int kk[ITEMS];  // last staged row with s_rel <= q
        dev::lockstep_last_le<ITEMS>(s_rel, m,
                                     [&](int it) { return it * MJ_BLOCK + (int)threadIdx.x; }, kk);
        // every item's rows without branches; past the outputs: row 0 of
        // the pair's inputs (nothing is stored for them)
#pragma unroll
        for (int it = 0; it < ITEMS; ++it) {
            const int q = it * MJ_BLOCK + threadIdx.x;
            const bool in = q < nout;
            const uint32_t r = right_row(kk[it], q);
            rrow[it] = in ? r : 0u;
            if (!in) kk[it] = -1;
        }
        const uint64_t ob = b0 - D.obase;
        // Output columns [L columns, R columns after the key], one at a time
        // (fewer registers, more waves per CU, as in the single join).
        constexpr int CB = 1;
        const int ncl = D.ncl, nco = D.ncl + D.ncr - 1;
        for (int c0 = 0; c0 < nco; c0 += CB) {
            uint32_t v[CB][ITEMS];
#pragma unroll
            for (int cb = 0; cb < CB; ++cb) {
                const int c = c0 + cb;
                if (c < nco) {
                    const bool from_l = c < ncl;
                    const uint32_t *src = from_l ? D.lc[c] : D.rc[c - ncl + 1];
#pragma unroll
                    for (int it = 0; it < ITEMS; ++it) {
                        const uint32_t lr = kk[it] < 0 ? 0u : (uint32_t)(sa + kk[it] - D.lbase);
                        v[cb][it] = src[from_l ? lr : rrow[it]];
                    }
                }
            }
#pragma unroll
            for (int cb = 0; cb < CB; ++cb) {
                const int c = c0 + cb;
                if (c < nco) {
                    uint32_t *dst = D.oc[c] + ob;
#pragma unroll
                    for (int it = 0; it < ITEMS; ++it) {
                        const int q = it * MJ_BLOCK + threadIdx.x;
                        if (q < nout) dst[q] = v[cb][it];
                    }
                }
            }
        }
        return;
    }
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
        const int q = it * MJ_BLOCK + threadIdx.x;
        if (q >= nout) continue;
        const uint64_t o = b0 + q;
        int lo = 0, hi = m;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if ((int)s_rel[mid] <= q) lo = mid + 1; else hi = mid;
        }
        const int k = lo - 1;
        const uint64_t grow = sa + k;  // global left row
        const JPairDev &D = P[s_pair[k]];
        const uint64_t lrow = grow - D.lbase;
        const uint64_t rrow = right_row(k, q);
        const uint64_t loc = o - D.obase;
        int oc = 0;
        for (int c = 0; c < D.ncl; ++c) D.oc[oc++][loc] = D.lc[c][lrow];
        for (int c = 1; c < D.ncr; ++c) D.oc[oc++][loc] = D.rc[c][rrow];
    }
}

// ---- sort by column 0 -----------------------------------------------------------
struct SortRelDev {
    const uint32_t *in[VD_MAX_COLS];
    uint32_t *out[VD_MAX_COLS];
    uint64_t n, base;
    int ncols;
};

__global__ __launch_bounds__(MB_BLOCK) void k_max_col0_multi(const SortRelDev *S, int ns,
                                                             const uint64_t *tbase,
                                                             uint32_t *out) {
    __shared__ int s_j;
    __shared__ uint32_t s_m[MB_BLOCK / 64];
    if (threadIdx.x == 0) s_j = find_seg(tbase, ns, blockIdx.x);
    __syncthreads();
    const int j = uniform(s_j);
    const SortRelDev &D = S[j];
    const uint64_t base = (blockIdx.x - tbase[j]) * (uint64_t)MB_TILE;
    uint32_t m = 0;
    for (int it = 0; it < MB_ITEMS; ++it) {
        const uint64_t i = base + (uint64_t)it * MB_BLOCK + threadIdx.x;
        if (i < D.n) m = max(m, D.in[0][i]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor(m, off, 64));
    if (dev::lane_id() == 0) s_m[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t r = 0;
        for (int w = 0; w < MB_BLOCK / 64; ++w) r = max(r, s_m[w]);
        atomicMax(out, r);
    }
}

// key = seg << (b + R) | col0 << R | global row
__global__ __launch_bounds__(MB_BLOCK) void k_pack_multi(const SortRelDev *S, int ns,
                                                         const uint64_t *tbase, int rbits,
                                                         int cbits, uint64_t *keys, int lo_bit,
                                                         int hi_bit, int npass, uint32_t *ghist) {
    __shared__ radix::LdsHist h;
    __shared__ int s_j;
    radix::hist_zero(&h);
    if (threadIdx.x == 0) s_j = find_seg(tbase, ns, blockIdx.x);
    __syncthreads();
    const int j = uniform(s_j);
    const SortRelDev &D = S[j];
    const uint64_t base = (blockIdx.x - tbase[j]) * (uint64_t)MB_TILE;
    for (int it = 0; it < MB_ITEMS; ++it) {
        const uint64_t i = base + (uint64_t)it * MB_BLOCK + threadIdx.x;
        if (i < D.n) {
            const uint64_t r = D.base + i;
            const uint64_t key = ((uint64_t)j << (cbits + rbits)) |
                                 ((uint64_t)D.in[0][i] << rbits) | r;
            keys[r] = key;
            radix::hist_add<true>(&h, key, lo_bit, hi_bit, npass);
        }
    }
    __syncthreads();
    radix::hist_flush(&h, ghist, npass);
}

__global__ __launch_bounds__(256) void k_gather_sort_multi(const uint64_t *keys, uint64_t total,
                                                           const SortRelDev *S, int rbits,
                                                           int cbits) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t rmask = (rbits >= 64) ? ~0ull : ((1ull << rbits) - 1);
    const uint64_t cmask = (1ull << cbits) - 1;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const uint64_t k = keys[i];
        const int j = (int)(k >> (cbits + rbits));
        const SortRelDev &D = S[j];
        const uint64_t src = (k & rmask) - D.base;
        const uint64_t dst = i - D.base;
        D.out[0][dst] = (uint32_t)((k >> rbits) & cmask);
        for (int c = 1; c < D.ncols; ++c) D.out[c][dst] = D.in[c][src];
    }
}

// ---- unique -------------------------------------------------------------------
__global__ __launch_bounds__(MB_BLOCK) void k_unique_count_multi(const SortRelDev *S, int ns,
                                                                 const uint64_t *tbase,
                                                                 uint32_t *cnt) {
    __shared__ int s_j;
    __shared__ uint32_t s_tmp[MB_BLOCK / 64];
    if (threadIdx.x == 0) s_j = find_seg(tbase, ns, blockIdx.x);
    __syncthreads();
    const int j = uniform(s_j);
    const SortRelDev &D = S[j];
    const uint64_t base = (blockIdx.x - tbase[j]) * (uint64_t)MB_TILE;
    const uint32_t c =
        (uint32_t)__popc(dev::unique_keep_bits<MB_BLOCK, MB_ITEMS>(D.in, D.ncols, D.n, base));
    const uint32_t t = dev::block_reduce<MB_BLOCK>(c, s_tmp);
    if (threadIdx.x == 0) cnt[blockIdx.x] = t;
}

__global__ void k_unique_totals(const uint64_t *tbase, int ns, const uint64_t *off,
                                uint64_t *kept) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ns) return;
    kept[j] = off[tbase[j + 1]] - off[tbase[j]];
}

__global__ __launch_bounds__(MB_BLOCK) void k_unique_scatter_multi(const SortRelDev *S, int ns,
                                                                   const uint64_t *tbase,
                                                                   const uint64_t *off) {
    __shared__ int s_j;
    __shared__ uint32_t s_cnt[MB_ITEMS][MB_BLOCK / 64];
    if (threadIdx.x == 0) s_j = find_seg(tbase, ns, blockIdx.x);
    __syncthreads();
    const int j = uniform(s_j);
    const SortRelDev &D = S[j];
    const uint64_t base = (blockIdx.x - tbase[j]) * (uint64_t)MB_TILE;
    constexpr int KC = 6;  // columns kept in registers between the two phases
    uint32_t vals[KC][MB_ITEMS];
    const uint32_t keep =
        dev::unique_keep_bits<MB_BLOCK, MB_ITEMS, KC>(D.in, D.ncols, D.n, base, vals);
    dev::unique_scatter_tile<MB_BLOCK, MB_ITEMS, KC>(keep, D.in, D.out, D.ncols, base,
                                                     off[blockIdx.x] - off[tbase[j]], s_cnt,
                                                     vals);
}

// ---- host drivers ---------------------------------------------------------------
template <typename T>
DBuf<T> to_device(vd_ctx *ctx, const std::vector<T> &h) {
    DBuf<T> d(ctx, h.size());
    upload(ctx, d.p, h.data(), h.size() * sizeof(T));
    return d;
}

// Tile base prefix (n + 1 entries) for row counts.
std::vector<uint64_t> tile_prefix(const std::vector<uint64_t> &rows) {
    std::vector<uint64_t> t(rows.size() + 1, 0);
    for (size_t j = 0; j < rows.size(); ++j) t[j + 1] = t[j] + ceil_div(rows[j], MB_TILE);
    return t;
}

// Row ranges of many lookups (any permutations): hr[2j], hr[2j+1].
std::vector<uint64_t> lookup_multi(vd_ctx *ctx, const vd_index *idx, const vd_select_job_t *jobs,
                                   int nj) {
    std::vector<SelJobDev> hj(nj);
    for (int j = 0; j < nj; ++j) {
        const vd_perm_arrays &pa = get_perm(idx, jobs[j].perm);
        hj[j] = SelJobDev{perm_dev(pa), jobs[j].k1, jobs[j].k2};
    }
    DBuf<SelJobDev> dj = to_device(ctx, hj);
    DBuf<uint64_t> range(ctx, 2 * (uint64_t)nj);
    std::vector<uint64_t> hr(2 * (size_t)nj);
    // one block: the kernel publishes the words itself
    if (nj == 0) return {};
    const bool direct = nj <= LK_JOBS;
    const Readback rb = direct ? rb_next(ctx) : Readback{};
    launch(ctx, "select_lookup", 24.0 * nj, dim3((unsigned)ceil_div(nj, LK_JOBS)),
           dim3(LK_BLOCK), 0, k_sel_lookup_multi, (const SelJobDev *)dj.p, nj, range.p, rb);
    if (direct) rb_wait(ctx, rb, hr.data(), 2 * nj);
    else download(ctx, hr.data(), range.p, hr.size() * sizeof(uint64_t));
    return hr;
}

// Ranges and column end values of many lookups: hr[3j..3j+2] as
// k_sel_bounds_multi writes them.
std::vector<uint64_t> bounds_multi(vd_ctx *ctx, const vd_index *idx, const vd_select_job_t *jobs,
                                   int nj) {
    std::vector<BoundJobDev> hj(nj);
    for (int j = 0; j < nj; ++j) {
        const vd_perm_arrays &pa = get_perm(idx, jobs[j].perm);
        hj[j] = BoundJobDev{perm_dev(pa), jobs[j].k2 == VD_UNBOUND ? pa.y : pa.z, jobs[j].k1,
                            jobs[j].k2};
    }
    DBuf<BoundJobDev> dj = to_device(ctx, hj);
    DBuf<uint64_t> out(ctx, 3 * (uint64_t)nj);
    std::vector<uint64_t> hr(3 * (size_t)nj);
    // one block: the kernel publishes the words itself
    if (nj == 0) return {};
    const bool direct = nj <= LK_JOBS;
    const Readback rb = direct ? rb_next(ctx) : Readback{};
    launch(ctx, "select_bounds", 32.0 * nj, dim3((unsigned)ceil_div(nj, LK_JOBS)),
           dim3(LK_BLOCK), 0, k_sel_bounds_multi, (const BoundJobDev *)dj.p, nj, out.p, rb);
    if (direct) rb_wait(ctx, rb, hr.data(), 3 * nj);
    else download(ctx, hr.data(), out.p, hr.size() * sizeof(uint64_t));
    return hr;
}

void select_multi(vd_ctx *ctx, const vd_index *idx, const vd_select_job_t *jobs, int nj,
                  vd_rel **outs) {
    for (int j = 0; j < nj; ++j)
        VD_REQUIRE(jobs[j].cols == 1 || jobs[j].cols == 3, VD_EINVAL,
                   "batched select supports cols 1 and 3");
    const std::vector<uint64_t> hr = lookup_multi(ctx, idx, jobs, nj);
    // Each result is the contiguous row range of its permutation: a view.
    for (int j = 0; j < nj; ++j) outs[j] = nullptr;
    try {
        for (int j = 0; j < nj; ++j) {
            const vd_perm_arrays &pa = get_perm(idx, jobs[j].perm);
            const uint64_t b = hr[2 * j], e = hr[2 * j + 1];
            uint32_t *vc[2];
            int nc = 0;
            if (jobs[j].cols == 3) vc[nc++] = pa.y + b;
            vc[nc++] = pa.z + b;
            outs[j] = rel_view(nc, e - b, vc, idx->hold);
            outs[j]->distinct = idx->distinct();
            if (pa.bounded) {
                if (jobs[j].cols == 3) rel_set_bound(outs[j], 0, pa.ymax);
                rel_set_bound(outs[j], nc - 1, pa.zmax);
            }
        }
    } catch (...) {
        for (int j = 0; j < nj; ++j) rel_free(ctx, outs[j]), outs[j] = nullptr;
        throw;
    }
}

// One chunk of a batched join (at most JM_CHUNK pairs).
void join_multi_chunk(vd_ctx *ctx, const vd_rel *const *L, const vd_rel *const *R, int np,
                      vd_rel **outs) {
    VD_REQUIRE(np <= 65535, VD_ERANGE, "at most 65535 joins per chunk");
    std::vector<JPairDev> hp(np);
    std::vector<uint64_t> lrows(np);
    for (int j = 0; j < np; ++j) {
        VD_REQUIRE(L[j] && R[j] && L[j]->ncols >= 1 && R[j]->ncols >= 1, VD_EINVAL,
                   "bad join input");
        VD_REQUIRE(L[j]->ncols + R[j]->ncols - 1 <= VD_MAX_COLS, VD_EINVAL,
                   "join output has too many columns");
        // right rows are addressed by 32-bit first-match indices
        VD_REQUIRE(R[j]->nrows < (1ull << 32), VD_ERANGE, "batched join right input too large");
        // an empty right side contributes no left rows (all counts 0)
        lrows[j] = R[j]->nrows ? L[j]->nrows : 0;
    }
    // Row space of the batch: input j's left rows start on tile tb[j], so a
    // bounds block (one tile) scans its own counts; rows past an input's end
    // up to its next tile are padding with no matches.
    const std::vector<uint64_t> tb = tile_prefix(lrows);
    const uint64_t ntiles = tb[np];
    const uint64_t nrow = ntiles * MB_TILE;
    VD_REQUIRE(nrow < (1ull << 32), VD_ERANGE, "batched join inputs too large");
    std::vector<uint64_t> lbase(np + 1);
    for (int j = 0; j <= np; ++j) lbase[j] = tb[j] * MB_TILE;
    for (int j = 0; j < np; ++j) {
        JPairDev d{};
        for (int c = 0; c < L[j]->ncols; ++c) d.lc[c] = L[j]->cols[c];
        for (int c = 0; c < R[j]->ncols; ++c) d.rc[c] = R[j]->cols[c];
        d.ncl = L[j]->ncols;
        d.ncr = R[j]->ncols;
        d.nl = lrows[j];
        d.nr = R[j]->nrows;
        d.lbase = lbase[j];
        hp[j] = d;
    }
    for (int j = 0; j < np; ++j) outs[j] = nullptr;
    std::vector<uint64_t> tot(np + 1, 0);  // + the flag of a tile beyond 32-bit offsets
    DBuf<JPairDev> dp = to_device(ctx, hp);
    DBuf<uint64_t> dlbase = to_device(ctx, lbase);
    DBuf<uint32_t> lower(ctx, nrow), loc(ctx, nrow);
    DBuf<uint64_t> ttot(ctx, ntiles), tpre(ctx, ntiles + 1), dtot(ctx, np + 1);
    const TileOff off{(const uint32_t *)loc.p, tpre.p};
    if (ntiles > 0) {
        DBuf<uint64_t> dtb = to_device(ctx, tb);
        double bbytes = 0;
        for (int j = 0; j < np; ++j) bbytes += 4.0 * (hp[j].nl + hp[j].nr) + 8.0 * hp[j].nl;
        DBuf<uint64_t> ranges(ctx, 2 * ntiles);
        fill_u64(ctx, dtot.p + np, 0, 1);
        launch(ctx, "join_ranges", 0.0, dim3((unsigned)ceil_div(ntiles * 2 * MR_GROUP, 256)),
               dim3(256), 0, k_join_ranges_multi, (const JPairDev *)dp.p, np,
               (const uint64_t *)dtb.p, ntiles, ranges.p);
        launch(ctx, "join_bounds", bbytes, dim3((unsigned)ntiles), dim3(MJB_BLOCK), 0,
               k_join_bounds_multi, (const JPairDev *)dp.p, np, (const uint64_t *)dtb.p,
               (const uint64_t *)ranges.p, lower.p, loc.p, ttot.p, dtot.p + np);
        scan_tile_totals(ctx, ttot.p, ntiles, tpre.p);
        launch(ctx, "join_totals", 0.0, dim3((unsigned)ceil_div(np, 256)), dim3(256), 0,
               k_pair_totals, (const JPairDev *)dp.p, np, off, nrow, ntiles, dtot.p);
        download(ctx, tot.data(), dtot.p, (np + 1) * sizeof(uint64_t));
        VD_REQUIRE(tot[np] == 0, VD_ERANGE,
                   "batched join: 2048 consecutive left rows match more than 2^32-1 right rows");
    }
    uint64_t total = 0;
    try {
        for (int j = 0; j < np; ++j) {
            outs[j] = rel_new(ctx, L[j]->ncols + R[j]->ncols - 1, tot[j]);
            outs[j]->distinct = L[j]->distinct && R[j]->distinct;
            for (int c = 0; c < outs[j]->ncols; ++c) {
                hp[j].oc[c] = outs[j]->cols[c];
                const bool from_l = c < L[j]->ncols;
                rel_copy_bound(outs[j], c, from_l ? L[j] : R[j], from_l ? c : c - L[j]->ncols + 1);
            }
            hp[j].obase = total;  // outputs of the pairs are laid end to end
            total += tot[j];
        }
        VD_REQUIRE(total < (1ull << 40), VD_ERANGE, "join output too large");
        if (total > 0) {
            upload(ctx, dp.p, hp.data(), np * sizeof(JPairDev));
            const int nv = ctx->join_nv;
            const uint64_t nblocks = ceil_div(nrow + total, (uint64_t)nv);
            VD_REQUIRE(nblocks < (1ull << 31), VD_ERANGE, "join expansion grid too large");
            DBuf<uint64_t> mp(ctx, nblocks + 1);
            launch(ctx, "join_partition", 0.0, dim3((unsigned)ceil_div(nblocks + 1, 256)),
                   dim3(256), 0, k_join_partition_multi, off, nrow, total,
                   nv, nblocks, mp.p);
            double xbytes = 0;
            for (int j = 0; j < np; ++j)
                xbytes += 4.0 * tot[j] * outs[j]->ncols + 4.0 * hp[j].nl * hp[j].ncl +
                          4.0 * hp[j].nr * (hp[j].ncr - 1);
            auto go = [&](auto kern) {
                launch(ctx, "join_expand", xbytes, dim3((unsigned)nblocks), dim3(MJ_BLOCK), 0,
                       kern, off, (const uint32_t *)lower.p,
                       (const uint64_t *)mp.p, nrow, total, (const JPairDev *)dp.p, np,
                       (const uint64_t *)dlbase.p);
            };
            if (nv == 1024) go(k_join_expand_multi<1024>);
            else if (nv == 4096) go(k_join_expand_multi<4096>);
            else go(k_join_expand_multi<2048>);
        }
    } catch (...) {
        for (int j = 0; j < np; ++j) rel_free(ctx, outs[j]), outs[j] = nullptr;
        throw;
    }
}

// Pairs in chunks of JM_CHUNK: every input's left rows start on a tile of
// the chunk's row space, so the padding up to a tile (at most 2047 rows
// per pair) stays below JM_CHUNK * 2048 rows per chunk (64 MiB of row
// offsets), whatever the number of tiny joins in the call.
constexpr int JM_CHUNK = 4096;
void join_multi(vd_ctx *ctx, const vd_rel *const *L, const vd_rel *const *R, int np,
                vd_rel **outs) {
    for (int j = 0; j < np; ++j) outs[j] = nullptr;
    try {
        for (int a = 0; a < np; a += JM_CHUNK)
            join_multi_chunk(ctx, L + a, R + a, std::min(JM_CHUNK, np - a), outs + a);
    } catch (...) {
        for (int j = 0; j < np; ++j) rel_free(ctx, outs[j]), outs[j] = nullptr;
        throw;
    }
}

void sort_col0_range(vd_ctx *ctx, vd_rel *const *rels, int n);

// Relations of at least this many rows are sorted one by one (sort_col0):
// its launch and read-back costs are small against their rows, and it moves
// the columns with the keys instead of gathering them by row id afterwards.
constexpr uint64_t MULTI_SORT_SINGLE_ROWS = 1ull << 20;

void sort_col0_multi(vd_ctx *ctx, vd_rel *const *rels, int n) {
    std::vector<vd_rel *> work;
    for (int j = 0; j < n; ++j) {
        if (!rels[j] || rels[j]->nrows <= 1) continue;
        if (rels[j]->nrows >= MULTI_SORT_SINGLE_ROWS) {
            const int rc = vd_sort_col0(ctx, rels[j]);
            if (rc != VD_OK) throw Error(rc, ctx->last_error);
        } else {
            work.push_back(rels[j]);
        }
    }
    if (!work.empty()) sort_col0_range(ctx, work.data(), (int)work.size());
}

void sort_col0_range(vd_ctx *ctx, vd_rel *const *rels, int n) {
    std::vector<SortRelDev> hs(n);
    std::vector<uint64_t> rows(n);
    uint64_t total = 0;
    for (int j = 0; j < n; ++j) {
        SortRelDev d{};
        d.n = rels[j]->nrows;
        d.base = total;
        d.ncols = rels[j]->ncols;
        for (int c = 0; c < d.ncols; ++c) d.in[c] = rels[j]->cols[c];
        hs[j] = d;
        rows[j] = d.n;
        total += d.n;
    }
    VD_REQUIRE(total < (1ull << 30), VD_ERANGE, "batched sort supports < 2^30 rows");
    const std::vector<uint64_t> tb = tile_prefix(rows);
    DBuf<SortRelDev> ds = to_device(ctx, hs);
    DBuf<uint64_t> dtb = to_device(ctx, tb);
    DBuf<uint64_t> dmax(ctx, 1);
    fill_u64(ctx, dmax.p, 0, 1);
    launch(ctx, "sort_max", 4.0 * total, dim3((unsigned)tb[n]), dim3(MB_BLOCK), 0,
           k_max_col0_multi, (const SortRelDev *)ds.p, n, (const uint64_t *)dtb.p,
           (uint32_t *)dmax.p);
    const uint32_t mx = (uint32_t)read_u64(ctx, dmax.p);
    const int rbits = std::max(1, bits_for(total - 1));
    const int cbits = std::max(1, bits_for(mx));
    const int sbits = bits_for((uint64_t)n - 1);
    if (rbits + cbits + sbits > 64) {
        VD_REQUIRE(n > 1, VD_ERANGE, "sort key does not fit 64 bits");
        sort_col0_range(ctx, rels, n / 2);
        sort_col0_range(ctx, rels + n / 2, n - n / 2);
        return;
    }
    const int lo = rbits, hi = rbits + cbits + sbits;
    const int npass = radix_passes(lo, hi);
    DBuf<uint64_t> keys(ctx, total), tmp(ctx, total);
    DBuf<uint32_t> hist(ctx, radix_hist_words(lo, hi));
    fill_u32(ctx, hist.p, 0, radix_hist_words(lo, hi));
    launch(ctx, "sort_pack", 12.0 * total, dim3((unsigned)tb[n]), dim3(MB_BLOCK), 0,
           k_pack_multi, (const SortRelDev *)ds.p, n, (const uint64_t *)dtb.p, rbits, cbits,
           keys.p, lo, hi, npass, hist.p);
    uint64_t *sorted = radix_sort_u64(ctx, keys.p, tmp.p, total, lo, hi, hist.p, true);
    std::vector<uint32_t *> fresh;
    try {
        for (int j = 0; j < n; ++j)
            for (int c = 0; c < hs[j].ncols; ++c) {
                fresh.push_back((uint32_t *)dalloc(ctx, hs[j].n * 4));
                hs[j].out[c] = fresh.back();
            }
    } catch (...) {
        for (auto p : fresh) dfree(ctx, p);
        throw;
    }
    upload(ctx, ds.p, hs.data(), n * sizeof(SortRelDev));
    double gbytes = 0;
    for (int j = 0; j < n; ++j) gbytes += (8.0 + 4.0 + 8.0 * (hs[j].ncols - 1)) * hs[j].n;
    launch(ctx, "sort_gather", gbytes, dim3(grid_cap(total)), dim3(256), 0, k_gather_sort_multi,
           (const uint64_t *)sorted, total, (const SortRelDev *)ds.p, rbits, cbits);
    for (int j = 0; j < n; ++j)
        for (int c = 0; c < hs[j].ncols; ++c) {
            rel_set_col(ctx, rels[j], c, hs[j].out[c]);
        }
}

void unique_multi(vd_ctx *ctx, vd_rel *const *rels_in, int n_in) {
    std::vector<vd_rel *> rels;
    // (rows known to be distinct have no adjacent duplicates)
    for (int j = 0; j < n_in; ++j)
        if (rels_in[j] && rels_in[j]->nrows > 1 && !rels_in[j]->distinct)
            rels.push_back(rels_in[j]);
    const int n = (int)rels.size();
    if (n == 0) return;
    std::vector<SortRelDev> hs(n);
    std::vector<uint64_t> rows(n);
    double bytes = 0;
    for (int j = 0; j < n; ++j) {
        SortRelDev d{};
        d.n = rels[j]->nrows;
        d.ncols = rels[j]->ncols;
        for (int c = 0; c < d.ncols; ++c) d.in[c] = rels[j]->cols[c];
        hs[j] = d;
        rows[j] = d.n;
        bytes += 4.0 * d.n * d.ncols;
    }
    const std::vector<uint64_t> tb = tile_prefix(rows);
    const uint64_t tiles = tb[n];
    DBuf<SortRelDev> ds = to_device(ctx, hs);
    DBuf<uint64_t> dtb = to_device(ctx, tb);
    DBuf<uint32_t> cnt(ctx, tiles);
    DBuf<uint64_t> off(ctx, tiles + 1), kept(ctx, n);
    launch(ctx, "unique_count", bytes, dim3((unsigned)tiles), dim3(MB_BLOCK), 0,
           k_unique_count_multi, (const SortRelDev *)ds.p, n, (const uint64_t *)dtb.p, cnt.p);
    scan_u32_to_u64(ctx, cnt.p, off.p, tiles, off.p + tiles);
    launch(ctx, "unique_totals", 0.0, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
           k_unique_totals, (const uint64_t *)dtb.p, n, (const uint64_t *)off.p, kept.p);
    std::vector<uint64_t> hk(n);
    download(ctx, hk.data(), kept.p, n * sizeof(uint64_t));
    std::vector<uint32_t *> fresh;
    try {
        for (int j = 0; j < n; ++j)
            for (int c = 0; c < hs[j].ncols; ++c) {
                fresh.push_back((uint32_t *)dalloc(ctx, std::max<uint64_t>(hk[j], 1) * 4));
                hs[j].out[c] = fresh.back();
            }
    } catch (...) {
        for (auto p : fresh) dfree(ctx, p);
        throw;
    }
    upload(ctx, ds.p, hs.data(), n * sizeof(SortRelDev));
    launch(ctx, "unique_scatter", 2.0 * bytes, dim3((unsigned)tiles), dim3(MB_BLOCK), 0,
           k_unique_scatter_multi, (const SortRelDev *)ds.p, n, (const uint64_t *)dtb.p,
           (const uint64_t *)off.p);
    for (int j = 0; j < n; ++j) {
        for (int c = 0; c < hs[j].ncols; ++c) {
            rel_set_col(ctx, rels[j], c, hs[j].out[c]);
        }
        rels[j]->nrows = hk[j];
    }
}

}  // namespace
}  // namespace vd

using namespace vd;

extern "C" {

int vd_select_multi(vd_ctx *ctx, const vd_index *idx, const vd_select_job_t *jobs, int njobs,
                    vd_rel **outs) {
    if (!ctx || njobs < 0 || (njobs && (!jobs || !outs))) return VD_EINVAL;
    if (njobs == 0) return VD_OK;
    VD_TRY(ctx, select_multi(ctx, idx, jobs, njobs, outs));
}

int vd_index_lookup_multi(vd_ctx *ctx, const vd_index *idx, const vd_select_job_t *jobs,
                          int njobs, uint64_t *begin, uint64_t *end) {
    if (!ctx || njobs < 0 || (njobs && (!jobs || !begin || !end))) return VD_EINVAL;
    if (njobs == 0) return VD_OK;
    VD_TRY(ctx, {
        const std::vector<uint64_t> hr = lookup_multi(ctx, idx, jobs, njobs);
        for (int j = 0; j < njobs; ++j) {
            begin[j] = hr[2 * j];
            end[j] = hr[2 * j + 1];
        }
    });
}

int vd_index_bounds_multi(vd_ctx *ctx, const vd_index *idx, const vd_select_job_t *jobs,
                          int njobs, uint64_t *begin, uint64_t *end, uint32_t *first,
                          uint32_t *last) {
    if (!ctx || njobs < 0 || (njobs && (!jobs || !begin || !end || !first || !last)))
        return VD_EINVAL;
    if (njobs == 0) return VD_OK;
    VD_TRY(ctx, {
        const std::vector<uint64_t> hr = bounds_multi(ctx, idx, jobs, njobs);
        for (int j = 0; j < njobs; ++j) {
            begin[j] = hr[3 * j];
            end[j] = hr[3 * j + 1];
            first[j] = (uint32_t)hr[3 * j + 2];
            last[j] = (uint32_t)(hr[3 * j + 2] >> 32);
        }
    });
}

int vd_join_multi(vd_ctx *ctx, const vd_rel *const *L, const vd_rel *const *R, int npairs,
                  vd_rel **outs) {
    if (!ctx || npairs < 0 || (npairs && (!L || !R || !outs))) return VD_EINVAL;
    if (npairs == 0) return VD_OK;
    VD_TRY(ctx, join_multi(ctx, L, R, npairs, outs));
}

int vd_sort_col0_multi(vd_ctx *ctx, vd_rel *const *rels, int n) {
    if (!ctx || n < 0 || (n && !rels)) return VD_EINVAL;
    VD_TRY(ctx, sort_col0_multi(ctx, rels, n));
}

int vd_unique_adjacent_multi(vd_ctx *ctx, vd_rel *const *rels, int n) {
    if (!ctx || n < 0 || (n && !rels)) return VD_EINVAL;
    VD_TRY(ctx, unique_multi(ctx, rels, n));
}

}  // extern "C"
