// Sort-merge bag inner join of two relations sorted on their key columns.
//
//  1. k_join_bounds : for every left row i, lower_i = first right row with
//                     key >= L[i] and cnt_i = number of right rows with key
//                     == L[i].  A block takes 2048 consecutive left rows;
//                     the right range they can touch (k_join_ranges) is
//                     staged in LDS, or a directory of every k-th key of it
//                     when it is longer than the LDS capacity.  The same
//                     kernel scans the counts of its tile: loc_i = exclusive
//                     scan of cnt_i within the tile, plus the tile total.
//  2. k_join_scan_total : one block scans the tile totals into tpre (tpre[t]
//                     = rows before tile t; tpre[ntiles] = output size, read
//                     back once to allocate, with the first and last output
//                     keys).  The output offset of left row
//                     i is off_i = tpre[i / 2048] + loc_i (join_off below).
//  3. k_join_partition / k_join_expand : load-balanced expansion.  The merge
//                     of the left rows (row i at merge position i + off_i)
//                     with the output indices is cut into diagonals of NV
//                     items, so every block gets at most NV rows + outputs
//                     whatever the duplicate-run lengths.  A block stages
//                     off/lower of its rows in LDS, maps each output index
//                     to (left row, right row) by binary search in LDS, and
//                     gathers the output columns with coalesced writes; the
//                     (left, right) pair list is never written to HBM.
//
// Output rows are left-major (left row order, then right row order), with
// columns [keys, left payload, right payload] -- the order the reference's
// ModernGPU inner_join + gather produce (JoinQueryJob.cu:68-101, 262-380).

#include "vd_internal.hpp"

namespace vd {
namespace {

// 512 threads x 4 rows per 2048-row tile: the 32 KB of staged right keys
// allow 4 blocks per CU, so 512-thread blocks keep 32 waves per CU where
// 256-thread ones kept 16 (the searches are latency-bound: 0.47 -> 0.40 ms
// of bounds per join-set step).
constexpr int JB_BLOCK = 512;
constexpr int JB_ITEMS = 4;
constexpr int JB_TILE = JB_BLOCK * JB_ITEMS;
constexpr int JX_BLOCK = 256;
constexpr int JB_RCAP_BYTES = 32768;  // LDS for staged right keys

struct KeyCols {
    const uint32_t *c0, *c1;
};

template <int NK>
struct KeyT;
template <>
struct KeyT<1> {
    using T = uint32_t;
    static __device__ __forceinline__ T at(const KeyCols &k, uint64_t i) {
        return k.c0[i];
    }
};
template <>
struct KeyT<2> {
    using T = uint64_t;
    static __device__ __forceinline__ T at(const KeyCols &k, uint64_t i) {
        return ((uint64_t)k.c0[i] << 32) | k.c1[i];
    }
};

// Right-row range [lo, hi) each bounds block can touch: two groups of 16
// lanes per block, one for each end, all searches in parallel (a serial
// search inside every block would put ~2 log2(|R|) dependent global loads
// in front of its work).
constexpr int JR_GROUP = 16;
template <int NK>
__global__ __launch_bounds__(256) void k_join_ranges(KeyCols L, uint64_t nl, KeyCols R,
                                                     uint64_t nr, uint64_t nblocks,
                                                     uint64_t *__restrict__ ranges) {
    using T = typename KeyT<NK>::T;
    const uint64_t g = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / JR_GROUP;
    const uint64_t b = g / 2;
    if (b >= nblocks) return;  // whole groups leave together
    const uint64_t base = b * JB_TILE;
    // even group: first right row >= the tile's first key; odd group: first
    // right row > its last key (one code path, so both run at once)
    const bool upper = g % 2 != 0;
    const T key = KeyT<NK>::at(L, upper ? min(base + JB_TILE, nl) - 1 : base);
    const uint64_t end = dev::group_bound<JR_GROUP>(0, nr, [&](uint64_t q) {
        const T r = KeyT<NK>::at(R, q);
        return upper ? key < r : !(r < key);
    });
    if (threadIdx.x % JR_GROUP == 0) ranges[2 * b + (upper ? 1 : 0)] = end;
}

// The right range of a tile goes to LDS whole when it has at most rcap
// keys.  A longer range (dense right side, hot keys) is staged as a
// directory of every k-th key (k = ceil(m / rcap)): a search runs in LDS
// over the directory, then over at most k - 1 keys in global memory.
template <int NK>
__global__ __launch_bounds__(JB_BLOCK) void k_join_bounds(
    KeyCols L, uint64_t nl, KeyCols R, uint64_t nr, int rcap,
    const uint64_t *__restrict__ ranges, uint32_t *__restrict__ lower,
    uint32_t *__restrict__ loc, uint64_t *__restrict__ tile_total,
    uint64_t *__restrict__ tile_keys) {
    using T = typename KeyT<NK>::T;
    constexpr int SR_WORDS = JB_RCAP_BYTES / sizeof(T);
    static_assert(JB_RCAP_BYTES >= JB_TILE * 12, "counts and offsets reuse the staging buffer");
    // Right keys; once the searches are done, the tile's offsets (u64, row
    // order) and counts (u32, row order) in the same space.
    __shared__ T s_r[SR_WORDS];
    __shared__ uint64_t s_tmp[JB_BLOCK / 64];
    const uint64_t tile = blockIdx.x;
    const uint64_t base = tile * JB_TILE;
    const uint64_t rlo = ranges[2 * tile], rhi = ranges[2 * tile + 1];
    const uint64_t m = rhi - rlo;
    const uint64_t k = m <= (uint64_t)rcap ? 1 : (m + rcap - 1) / rcap;  // directory stride
    const uint64_t nd = (m + k - 1) / k;                                 // directory entries
    // Stage the right keys in batches of SB per thread: the loads of a
    // batch are all in flight before the first LDS store.
    constexpr int SB = 8;
    for (uint64_t t0 = 0; t0 < nd; t0 += (uint64_t)SB * JB_BLOCK) {
        T v[SB];
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            const uint64_t t = t0 + (uint64_t)u * JB_BLOCK + threadIdx.x;
            if (t < nd) v[u] = KeyT<NK>::at(R, rlo + t * k);
        }
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            const uint64_t t = t0 + (uint64_t)u * JB_BLOCK + threadIdx.x;
            if (t < nd) s_r[t] = v[u];
        }
    }
    // The tile's left keys, all loads in flight together.
    T lkey[JB_ITEMS];
#pragma unroll
    for (int j = 0; j < JB_ITEMS; ++j) {
        const uint64_t i = base + (uint64_t)j * JB_BLOCK + threadIdx.x;
        lkey[j] = i < nl ? KeyT<NK>::at(L, i) : T(0);
    }
    __syncthreads();
    // Every row's first staged entry >= key (t0) and > key (t1).
    uint32_t t0[JB_ITEMS], t1[JB_ITEMS];
    dev::lockstep_bounds<JB_ITEMS>(s_r, (uint32_t)nd, lkey, t0, t1);
    // Bounds relative to rlo: the staged entries found, or with a directory
    // (k > 1) the exact positions within the gaps before them.
    uint32_t lo[JB_ITEMS], hi[JB_ITEMS];
    if (k > 1) {
        bool valid[JB_ITEMS];
#pragma unroll
        for (int j = 0; j < JB_ITEMS; ++j) valid[j] = base + (uint64_t)j * JB_BLOCK + threadIdx.x < nl;
        dev::lockstep_gap_bounds<JB_ITEMS>(
            [&](uint32_t q) { return KeyT<NK>::at(R, rlo + q); }, k, m, lkey, valid, t0, t1, lo,
            hi);
    } else {
#pragma unroll
        for (int j = 0; j < JB_ITEMS; ++j) {
            lo[j] = t0[j];
            hi[j] = t1[j];
        }
    }
    uint32_t cnt[JB_ITEMS];
#pragma unroll
    for (int j = 0; j < JB_ITEMS; ++j) {
        const uint64_t i = base + (uint64_t)j * JB_BLOCK + threadIdx.x;
        uint32_t c = 0;
        if (i < nl) {
            lower[i] = (uint32_t)(rlo + lo[j]);
            c = (uint32_t)(hi[j] - lo[j]);
        }
        cnt[j] = c;
    }
    // Smallest and largest first key column of the tile's rows with matches
    // (the first and last output keys of the tile).
    uint32_t kmn = ~0u, kmx = 0;
#pragma unroll
    for (int j = 0; j < JB_ITEMS; ++j) {
        const uint32_t c0 = NK == 1 ? (uint32_t)lkey[j] : (uint32_t)((uint64_t)lkey[j] >> 32);
        if (cnt[j]) {
            kmn = min(kmn, c0);
            kmx = max(kmx, c0);
        }
    }
    // (its barriers also end every read of the staged right keys)
    dev::block_min_max<JB_BLOCK>(kmn, kmx, s_tmp);
    uint64_t *s_off = reinterpret_cast<uint64_t *>(s_r);
    uint32_t *s_cnt = reinterpret_cast<uint32_t *>(s_off + JB_TILE);
#pragma unroll
    for (int j = 0; j < JB_ITEMS; ++j) s_cnt[j * JB_BLOCK + threadIdx.x] = cnt[j];
    __syncthreads();
    // Thread t scans rows t*ITEMS .. t*ITEMS+ITEMS-1 of the tile.
    uint64_t run = 0;
    uint32_t mine[JB_ITEMS];
#pragma unroll
    for (int j = 0; j < JB_ITEMS; ++j) {
        mine[j] = s_cnt[threadIdx.x * JB_ITEMS + j];
        run += mine[j];
    }
    uint64_t tot;
    uint64_t ex = dev::block_excl_scan<JB_BLOCK>(run, s_tmp, &tot);
#pragma unroll
    for (int j = 0; j < JB_ITEMS; ++j) {
        s_off[threadIdx.x * JB_ITEMS + j] = ex;
        ex += mine[j];
    }
    if (threadIdx.x == 0) {
        tile_total[tile] = tot;
        tile_keys[tile] = ((uint64_t)kmx << 32) | kmn;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < JB_ITEMS; ++j) {
        const uint64_t i = base + (uint64_t)j * JB_BLOCK + threadIdx.x;
        if (i < nl) loc[i] = (uint32_t)s_off[j * JB_BLOCK + threadIdx.x];
    }
}

static_assert(JB_TILE == OFF_TILE, "bounds tiles are the offset tiles");
using JoinOff = TileOff;

// mp[b] = number of left rows i with i + off[i] < b * nv (clamped to the
// end of the diagonal), for b = 0 .. nblocks: a group of lanes per b.
__global__ __launch_bounds__(256) void k_join_partition(
    JoinOff off, uint64_t nl, uint64_t total, int nv,
    uint64_t nblocks, uint64_t *__restrict__ mp) {
    const uint64_t b = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / JR_GROUP;
    if (b > nblocks) return;  // whole groups leave together
    const uint64_t diag = min(b * (uint64_t)nv, nl + total);
    const uint64_t lo =
        dev::group_bound<JR_GROUP>(0, nl, [&](uint64_t i) { return !(i + off(i) < diag); });
    if (threadIdx.x % JR_GROUP == 0) mp[b] = lo;
}

// The same cuts made before the host reads the output size back: the size
// comes from *total_p (tpre[ntiles], written by the scan kernel before this
// one on the stream), the groups stride over b = 0 .. min(nblocks, cap - 1),
// so the cuts are ready when the host launches the expansion.
__global__ __launch_bounds__(256) void k_join_partition_ahead(
    JoinOff off, uint64_t nl, const uint64_t *__restrict__ total_p, int nv, uint64_t cap,
    uint64_t *__restrict__ mp) {
    const uint64_t total = *total_p;
    const uint64_t last = min((nl + total + (uint64_t)nv - 1) / (uint64_t)nv, cap - 1);
    const uint64_t ngroups = (uint64_t)gridDim.x * blockDim.x / JR_GROUP;
    for (uint64_t b = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / JR_GROUP; b <= last;
         b += ngroups) {
        const uint64_t diag = min(b * (uint64_t)nv, nl + total);
        const uint64_t lo =
            dev::group_bound<JR_GROUP>(0, nl, [&](uint64_t i) { return !(i + off(i) < diag); });
        if (threadIdx.x % JR_GROUP == 0) mp[b] = lo;
    }
}

// Cuts made ahead of the size read-back for outputs up to 32 times their
// left rows (a larger output gets the partition kernel after it), by groups
// enough for outputs up to 8 times the left rows in one pass.
inline uint64_t cut_capacity(uint64_t nl, int nv) { return ceil_div(33 * nl, (uint64_t)nv) + 2; }

void partition_ahead(vd_ctx *ctx, const JoinOff &joff, uint64_t nl, const uint64_t *total_p,
                     int nv, DBuf<uint64_t> &mp, uint64_t &cap) {
    cap = cut_capacity(nl, nv);
    mp = DBuf<uint64_t>(ctx, cap);
    const uint64_t groups = std::min(cap, ceil_div(9 * nl, (uint64_t)nv) + 2);
    launch(ctx, "join_partition", 0.0, dim3((unsigned)ceil_div(groups * JR_GROUP, 256)),
           dim3(256), 0, k_join_partition_ahead, joff, nl, total_p, nv, cap, mp.p);
}

// The expansion's cuts: those made ahead when they cover the output, else
// the partition kernel's into `spare`.
const uint64_t *join_cuts(vd_ctx *ctx, const JoinOff &joff, uint64_t nl, uint64_t total, int nv,
                          const DBuf<uint64_t> &mp, uint64_t cap, DBuf<uint64_t> &spare) {
    const uint64_t nblocks = ceil_div(nl + total, (uint64_t)nv);
    if (nblocks + 1 <= cap) return mp.p;
    spare = DBuf<uint64_t>(ctx, nblocks + 1);
    launch(ctx, "join_partition", 0.0, dim3((unsigned)ceil_div((nblocks + 1) * JR_GROUP, 256)),
           dim3(256), 0, k_join_partition, joff, nl, total, nv, nblocks, spare.p);
    return spare.p;
}

// One block: scans the tile totals into the tile prefixes, then reads back
// {total, first output key, last output key} (the output's column 0 holds
// the left keys of matched rows in order: the smallest and largest of the
// tiles' matched keys).
__global__ __launch_bounds__(dev::TT_BLOCK) void k_join_scan_total(
    const uint64_t *__restrict__ ttot, const uint64_t *__restrict__ tkeys, uint64_t ntiles,
    uint64_t *__restrict__ tpre, Readback rb, const uint32_t *__restrict__ flag) {
    __shared__ uint64_t s_tmp[dev::TT_BLOCK / 64];
    // Key bounds and the high halves of the totals first, from coalesced
    // loads all in flight together (one round trip for up to TT_BLOCK *
    // TT_PER tiles); the scan then reads the totals from the cache.
    constexpr int PER = dev::TT_PER;
    uint32_t mn = ~0u, mx = 0;
    uint64_t wide = 0;  // high halves of the tile totals
    for (uint64_t b = 0; b < ntiles; b += (uint64_t)dev::TT_BLOCK * PER) {
        uint64_t v[PER], w[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const uint64_t t = b + (uint64_t)j * dev::TT_BLOCK + threadIdx.x;
            const bool in = t < ntiles;
            v[j] = in ? ttot[t] : 0;
            w[j] = in ? tkeys[t] : 0xFFFFFFFFull;  // (max 0, min ~0): no keys
        }
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            wide |= v[j] >> 32;
            mn = min(mn, (uint32_t)w[j]);
            mx = max(mx, (uint32_t)(w[j] >> 32));
        }
    }
    dev::scan_tile_totals_block(ttot, ntiles, tpre, s_tmp);
    dev::block_min_max<dev::TT_BLOCK>(mn, mx, s_tmp);
    const int overflow = __syncthreads_or(wide != 0);
    if (threadIdx.x >= 64) return;
    if (threadIdx.x == 0) {
        const uint64_t total = tpre[ntiles];  // written by this thread
        rb.dst[0] = total;
        rb.dst[1] = total ? mn : 0;
        rb.dst[2] = total ? mx : 0;
        // bit 0: a tile's offsets do not fit 32 bits; bit 1: the caller's
        // flag word (driven joins: repeated driver rows) is set
        rb.dst[3] = (overflow ? 1u : 0u) | (flag && *flag ? 2u : 0u);
    }
    dev::rb_publish(rb);
}

// Output column c is gathered from src[c] at the output's left (driving)
// row, or at its matched right row when bit c of from_r is set.  With
// records (the expansion's AW > 0), the right rows' columns come from
// rec, AW words per right row, column c in word slot[c]: one load per
// output row for all of them.
struct OutCols {
    const uint32_t *src[VD_MAX_COLS];
    uint32_t *o[VD_MAX_COLS];
    uint32_t from_r;
    int n;
    const uint32_t *rec;
    int slot[VD_MAX_COLS];
};

template <int AW>
struct RecT;
template <>
struct RecT<2> {
    using T = uint2;
    static __device__ __forceinline__ uint32_t word(const T &v, int s) { return s == 0 ? v.x : v.y; }
};
template <>
struct RecT<4> {
    using T = uint4;
    static __device__ __forceinline__ uint32_t word(const T &v, int s) {
        return s == 0 ? v.x : s == 1 ? v.y : s == 2 ? v.z : v.w;
    }
};

// rec[i] = (c[0][i], .., c[nc - 1][i], 0 ..) as AW-word records.
struct RecCols {
    const uint32_t *c[4];
    int nc;
};
template <int AW>
__global__ __launch_bounds__(256) void k_interleave(RecCols rc, uint64_t n,
                                                    typename RecT<AW>::T *__restrict__ rec) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint32_t w[AW];
#pragma unroll
    for (int s = 0; s < AW; ++s) w[s] = s < rc.nc ? rc.c[s][i] : 0u;
    typename RecT<AW>::T v;
    if constexpr (AW == 2) v = make_uint2(w[0], w[1]);
    else v = make_uint4(w[0], w[1], w[2], w[3]);
    rec[i] = v;
}

template <int NV, int AW = 0>
__global__ __launch_bounds__(JX_BLOCK) void k_join_expand(
    JoinOff off, const uint32_t *__restrict__ lower,
    const uint64_t *__restrict__ mp, uint64_t nl, uint64_t total, OutCols oc) {
    constexpr int ITEMS = NV / JX_BLOCK;
    static_assert(NV < 65536, "offsets within a diagonal in 16 bits");
    // Offsets of the diagonal's left rows relative to its first output b0
    // (rows after the first start at or after b0, so they fit 16 bits; the
    // first row, which may start earlier, is staged as 0 and its exact
    // offset kept in s_off0), and the rows' first right rows.  12 KB for
    // NV = 2048, so LDS allows 12 blocks per CU.
    __shared__ uint16_t s_rel[NV + 2];
    __shared__ uint32_t s_low[NV + 1];
    __shared__ uint64_t s_off0;
    const uint64_t blk = blockIdx.x;
    const uint64_t d0 = blk * NV;
    const uint64_t d1 = min(d0 + NV, nl + total);
    const uint64_t a0 = mp[blk], a1 = mp[blk + 1];
    const uint64_t b0 = d0 - a0, b1 = d1 - a1;
    if (b1 <= b0) return;  // no outputs on this diagonal
    const uint64_t sa = a0 > 0 ? a0 - 1 : 0;
    const int m = (int)(a1 - sa);
    // Stage offsets and lower bounds of the diagonal's left rows.  The
    // loads of a batch are unconditional (indices clamped to the last row)
    // so all of them are in flight before the first is used.
    constexpr int SB = 4;
    for (int t0 = 0; t0 < m; t0 += SB * JX_BLOCK) {
        uint64_t vpre[SB];
        uint32_t vloc[SB], vl[SB];
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            const uint64_t i = sa + (uint64_t)min(t0 + u * JX_BLOCK + (int)threadIdx.x, m - 1);
            vloc[u] = off.loc[i];
            vpre[u] = off.tpre[i / OFF_TILE];
            vl[u] = lower[i];
        }
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            const int t = t0 + u * JX_BLOCK + threadIdx.x;
            if (t < m) {
                const uint64_t vo = vpre[u] + vloc[u];
                s_rel[t] = t == 0 ? (uint16_t)0 : (uint16_t)(vo - b0);
                s_low[t] = vl[u];
                if (t == 0) s_off0 = vo;
            }
        }
    }
    __syncthreads();

    // Output b0 + q's left row: the last staged k with s_rel[k] <= q
    // (s_rel[0] = 0 <= every q).
    uint32_t rrow[ITEMS];
    const int nout = (int)(b1 - b0);
    int kk[ITEMS];
    dev::lockstep_last_le<ITEMS>(s_rel, m, [&](int j) { return j * JX_BLOCK + (int)threadIdx.x; },
                                 kk);
    // Every item's rows, with no branch: k is a valid staged index for
    // every item, and an item past the diagonal's outputs takes row 0 of
    // both inputs (it stores nothing), so the LDS reads and the gathers of
    // a batch are all in flight together.
    uint32_t low[ITEMS], rel[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        low[j] = s_low[kk[j]];
        rel[j] = s_rel[kk[j]];
    }
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const int q = j * JX_BLOCK + threadIdx.x;
        const int k = kk[j];
        const bool in = q < nout;
        const uint32_t r = k == 0 ? low[j] + (uint32_t)(b0 + (uint64_t)q - s_off0)
                                  : low[j] + (uint32_t)(q - (int)rel[j]);
        rrow[j] = in ? r : 0u;
        if (!in) kk[j] = -1;  // past the outputs: left row 0
    }
    // Output columns gathered CB columns at a time: the loads of a batch
    // are all issued before its stores (the output never aliases the
    // inputs).  One column per batch: fewer registers, more waves per CU
    // (measured: 1, 2, 4 columns per batch give 1.61 / 1.68 / 1.78 ms of
    // expansion per join-set step).
    constexpr int CB = 1;
    const int nco = oc.n;
    if constexpr (AW > 0) {
        // every right-row column from one record load per output row
        using R = RecT<AW>;
        typename R::T rv[ITEMS];
        const typename R::T *rec = reinterpret_cast<const typename R::T *>(oc.rec);
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) rv[j] = rec[rrow[j]];
        for (int c = 0; c < nco; ++c) {
            if (((oc.from_r >> c) & 1u) == 0) continue;
            const int sl = oc.slot[c];
            uint32_t *dst = oc.o[c];
#pragma unroll
            for (int j = 0; j < ITEMS; ++j) {
                const int q = j * JX_BLOCK + threadIdx.x;
                if (q < nout) dst[b0 + q] = R::word(rv[j], sl);
            }
        }
    }
    for (int c0 = 0; c0 < nco; c0 += CB) {
        uint32_t v[CB][ITEMS];
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) {
            const int c = c0 + cb;
            if (AW > 0 && ((oc.from_r >> c) & 1u)) continue;  // written from the records
            if (c < nco) {
                const bool from_l = ((oc.from_r >> c) & 1u) == 0;
                const uint32_t *src = oc.src[c];
#pragma unroll
                for (int j = 0; j < ITEMS; ++j) {
                    const uint32_t lr = kk[j] < 0 ? 0u : (uint32_t)(sa + kk[j]);
                    v[cb][j] = src[from_l ? lr : rrow[j]];
                }
            }
        }
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) {
            const int c = c0 + cb;
            if (AW > 0 && ((oc.from_r >> c) & 1u)) continue;
            if (c < nco) {
                uint32_t *dst = oc.o[c];
#pragma unroll
                for (int j = 0; j < ITEMS; ++j) {
                    const int q = j * JX_BLOCK + threadIdx.x;
                    if (q < nout) dst[b0 + q] = v[cb][j];
                }
            }
        }
    }
}


// ---- joins driven by a relation in output order ------------------------------
//
// vd_join_driven: for every row of D, in D's order, the rows of Y (sorted on
// column 0) whose key equals D's key column, in Y's order.  Y's runs come
// from a dense table over the key values (first row, end row per key), so
// D may be in any order; the counts, offsets, partition and expansion are
// the sort-merge join's.

// tab[k] = {first row, end row} of key k's run in a sorted key column, for
// the keys <= kmax the column holds (the table is zero-filled first, so an
// absent key has an empty run).  A block takes RT_ROWS consecutive rows,
// stages their keys in LDS and writes every table entry of its key span
// (key[r0 - 1], key[r1 - 1]] in key order, absent keys included, so its
// writes cover whole cache lines: a key's entry comes from two searches in
// the staged keys.  The span's last key may run on past the block: its
// entry gets its first row here and its end from the block holding the
// run's last row (which starts no entry of its own for it).  A block whose
// span is longer than RT_SPAN_MAX keys (sparse keys) writes only the
// entries of the keys it holds, on the zero fill.
constexpr int RT_BLOCK = 256;
constexpr int RT_ROWS = 1024;
constexpr uint32_t RT_SPAN_MAX = 16 * RT_ROWS;
__global__ __launch_bounds__(RT_BLOCK) void k_run_table(const uint32_t *__restrict__ key,
                                                        uint64_t n, uint32_t kmax,
                                                        uint2 *__restrict__ tab) {
    __shared__ uint32_t s_k[RT_ROWS];
    const uint64_t r0 = (uint64_t)blockIdx.x * RT_ROWS;
    const int cnt = (int)min<uint64_t>(RT_ROWS, n - r0);
    for (int t = threadIdx.x; t < RT_ROWS; t += RT_BLOCK) s_k[t] = t < cnt ? key[r0 + t] : ~0u;
    // the key before the block and the one after it
    const uint32_t kprev = r0 > 0 ? key[r0 - 1] : 0u;
    const bool has_prev = r0 > 0;
    const uint32_t knext = r0 + cnt < n ? key[r0 + cnt] : ~0u;
    const bool has_next = r0 + cnt < n;
    __syncthreads();
    const uint32_t kfirst = s_k[0], klast = s_k[cnt - 1];
    // keys of this block's span: (kprev, min(klast, kmax)]; the first
    // block's span starts at its first key
    const uint64_t lo = has_prev ? (uint64_t)kprev + 1 : (uint64_t)kfirst;
    const uint64_t hi = min<uint64_t>(klast, kmax);
    // A run that starts before the block and ends inside it: its end.
    if (has_prev && kprev == kfirst && kfirst <= kmax && threadIdx.x == 0) {
        int e = 0;  // first staged key past the run
        int len = cnt;
        while (len > 0) {
            const int half = len >> 1;
            if (s_k[e + half] <= kfirst) { e += half + 1; len -= half + 1; }
            else len = half;
        }
        if (e < cnt || !has_next || knext != kfirst) tab[kfirst].y = (uint32_t)(r0 + e);
    }
    if (lo > hi) return;
    auto lower = [&](uint32_t q) {  // first staged index with s_k >= q
        int p = 0, len = cnt;
        while (len > 0) {
            const int half = len >> 1;
            if (s_k[p + half] < q) { p += half + 1; len -= half + 1; }
            else len = half;
        }
        return p;
    };
    if (hi - lo + 1 <= RT_SPAN_MAX) {
        for (uint64_t q = lo + threadIdx.x; q <= hi; q += RT_BLOCK) {
            const int a = lower((uint32_t)q);
            uint2 ent = make_uint2(0u, 0u);
            if (a < cnt && s_k[a] == (uint32_t)q) {
                const int b = q < 0xFFFFFFFFull ? lower((uint32_t)q + 1) : cnt;
                ent.x = (uint32_t)(r0 + a);
                ent.y = (uint32_t)(r0 + b);
                // the run goes on in the next block: that block writes the end
                if (b == cnt && has_next && knext == (uint32_t)q) {
                    tab[q].x = ent.x;
                    continue;
                }
            }
            tab[q] = ent;
        }
    } else {
        // sparse keys: the entries of the keys held, on the zero fill
        for (int t = threadIdx.x; t < cnt; t += RT_BLOCK) {
            const uint32_t q = s_k[t];
            if ((uint64_t)q < lo || q > hi) continue;
            if (t > 0 && s_k[t - 1] == q) continue;  // not the run's first row here
            const int b = q < 0xFFFFFFFFu ? lower(q + 1) : cnt;
            if (b == cnt && has_next && knext == q) tab[q].x = (uint32_t)(r0 + t);
            else tab[q] = make_uint2((uint32_t)(r0 + t), (uint32_t)(r0 + b));
        }
    }
}

// Driver columns the count kernel reads: the key column, and (check > 0)
// every column, to flag a row equal to the row before it that matches more
// than one row.
struct DriveCols {
    const uint32_t *key;
    const uint32_t *all[VD_MAX_COLS];
    int check;  // columns compared for repeated rows (0: no check)
};

// Per driver row: first matched row and match count (keys outside [klo,
// khi] match nothing; khi <= the table's last key), then the tile's scan
// of the counts (32-bit tile-local offsets + tile total) and the smallest
// and largest key with matches, as k_join_bounds writes them.
__global__ __launch_bounds__(JB_BLOCK) void k_drive_bounds(
    DriveCols D, uint64_t nd, const uint2 *__restrict__ tab, uint32_t klo, uint32_t khi,
    uint32_t *__restrict__ lower, uint32_t *__restrict__ loc, uint64_t *__restrict__ tile_total,
    uint64_t *__restrict__ tile_keys, uint32_t *__restrict__ flag) {
    __shared__ uint64_t s_off[JB_TILE];
    __shared__ uint32_t s_cnt[JB_TILE];
    __shared__ uint64_t s_tmp[JB_BLOCK / 64];
    const uint64_t tile = blockIdx.x;
    const uint64_t base = tile * JB_TILE;
    uint32_t key[JB_ITEMS];
#pragma unroll
    for (int j = 0; j < JB_ITEMS; ++j) {
        const uint64_t i = base + (uint64_t)j * JB_BLOCK + threadIdx.x;
        key[j] = i < nd ? D.key[i] : 0u;
    }
    uint2 run[JB_ITEMS];
#pragma unroll
    for (int j = 0; j < JB_ITEMS; ++j) {
        const uint64_t i = base + (uint64_t)j * JB_BLOCK + threadIdx.x;
        const bool in = i < nd && key[j] >= klo && key[j] <= khi;
        run[j] = in ? tab[key[j]] : make_uint2(0u, 0u);
    }
    uint32_t cnt[JB_ITEMS], kmn = ~0u, kmx = 0;
    bool repeated = false;
#pragma unroll
    for (int j = 0; j < JB_ITEMS; ++j) {
        const uint64_t i = base + (uint64_t)j * JB_BLOCK + threadIdx.x;
        cnt[j] = run[j].y - run[j].x;
        if (i < nd) lower[i] = run[j].x;
        if (cnt[j]) {
            kmn = min(kmn, key[j]);
            kmx = max(kmx, key[j]);
        }
        if (D.check && cnt[j] > 1 && i > 0) {
            bool same = true;
            for (int c = 0; c < D.check; ++c) same = same && D.all[c][i] == D.all[c][i - 1];
            repeated = repeated || same;
        }
    }
    if (repeated) atomicOr(flag, 1u);
    dev::block_min_max<JB_BLOCK>(kmn, kmx, s_tmp);
#pragma unroll
    for (int j = 0; j < JB_ITEMS; ++j) s_cnt[j * JB_BLOCK + threadIdx.x] = cnt[j];
    __syncthreads();
    // Thread t scans rows t*ITEMS .. t*ITEMS+ITEMS-1 of the tile.
    uint64_t sum = 0;
    uint32_t mine[JB_ITEMS];
#pragma unroll
    for (int j = 0; j < JB_ITEMS; ++j) {
        mine[j] = s_cnt[threadIdx.x * JB_ITEMS + j];
        sum += mine[j];
    }
    uint64_t tot;
    uint64_t ex = dev::block_excl_scan<JB_BLOCK>(sum, s_tmp, &tot);
#pragma unroll
    for (int j = 0; j < JB_ITEMS; ++j) {
        s_off[threadIdx.x * JB_ITEMS + j] = ex;
        ex += mine[j];
    }
    if (threadIdx.x == 0) {
        tile_total[tile] = tot;
        tile_keys[tile] = ((uint64_t)kmx << 32) | kmn;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < JB_ITEMS; ++j) {
        const uint64_t i = base + (uint64_t)j * JB_BLOCK + threadIdx.x;
        if (i < nd) loc[i] = (uint32_t)s_off[j * JB_BLOCK + threadIdx.x];
    }
}

// The count phase of a join: per left row its first match (lower) and
// tile-local offset (off), per tile the prefix (tpre, in `ranges`), and the
// read-back {total, first key, last key}.
struct JoinCount {
    DBuf<uint32_t> lower, off;
    DBuf<uint64_t> ranges, mp;  // mp: the cuts made ahead (mp_cap entries, for nv)
    uint64_t *tpre = nullptr;
    uint64_t total = 0, mp_cap = 0;
    uint32_t key_lo = 0, key_hi = 0;
    int nv = 0;
};

template <int NK>
void join_count(vd_ctx *ctx, const vd_rel *L, const vd_rel *R, JoinCount &jc) {
    const uint64_t nl = L->nrows, nr = R->nrows;
    KeyCols lk{L->cols[0], NK == 2 ? L->cols[1] : nullptr};
    KeyCols rk{R->cols[0], NK == 2 ? R->cols[1] : nullptr};
    jc.lower = DBuf<uint32_t>(ctx, nl);
    jc.off = DBuf<uint32_t>(ctx, nl);  // offsets within each bounds tile
    const uint64_t jb_tiles = ceil_div(nl, JB_TILE);
    const int rcap = std::min<int>(ctx->join_rcap, JB_RCAP_BYTES / (NK == 2 ? 8 : 4));
    // ranges (2 per tile), tile totals (1 per tile), tile prefixes (1 per
    // tile + the total), tile key bounds (1 per tile).
    jc.ranges = DBuf<uint64_t>(ctx, 5 * jb_tiles + 1);
    uint64_t *ranges = jc.ranges.p;
    uint64_t *ttot = ranges + 2 * jb_tiles, *tpre = ranges + 3 * jb_tiles;
    uint64_t *tkeys = ranges + 4 * jb_tiles + 1;
    jc.tpre = tpre;
    launch(ctx, "join_ranges", 0.0, dim3((unsigned)ceil_div(jb_tiles * 2 * JR_GROUP, 256)),
           dim3(256), 0, k_join_ranges<NK>, lk, nl, rk, nr, jb_tiles, ranges);
    launch(ctx, "join_bounds", 4.0 * NK * (double)(nl + nr) + 8.0 * nl,
           dim3((unsigned)jb_tiles), dim3(JB_BLOCK), 0, k_join_bounds<NK>, lk,
           nl, rk, nr, rcap, (const uint64_t *)ranges, jc.lower.p, jc.off.p, ttot, tkeys);
    const Readback rb = rb_next(ctx);
    launch(ctx, "scan_tiles", 16.0 * (double)jb_tiles, dim3(1), dim3(dev::TT_BLOCK), 0,
           k_join_scan_total, (const uint64_t *)ttot, (const uint64_t *)tkeys, jb_tiles, tpre,
           rb, (const uint32_t *)nullptr);
    jc.nv = ctx->join_nv;
    partition_ahead(ctx, JoinOff{(const uint32_t *)jc.off.p, tpre}, nl, tpre + jb_tiles, jc.nv,
                    jc.mp, jc.mp_cap);
    uint64_t h[4];
    rb_wait(ctx, rb, h, 4);
    jc.total = h[0];
    jc.key_lo = (uint32_t)h[1];
    jc.key_hi = (uint32_t)h[2];
    VD_REQUIRE(jc.total < (1ull << 40), VD_ERANGE, "join output too large");
    VD_REQUIRE(h[3] == 0, VD_ERANGE,
               "join: 2048 consecutive left rows match more than 2^32-1 right rows");
}


// The join; with max_rows, nullptr (and *rows = the output size) when the
// output would have more rows than that.
template <int NK>
vd_rel *join_impl(vd_ctx *ctx, const vd_rel *L, const vd_rel *R,
                  uint64_t max_rows = ~0ull, uint64_t *rows = nullptr) {
    const uint64_t nl = L->nrows, nr = R->nrows;
    const int out_cols = L->ncols + R->ncols - NK;
    VD_REQUIRE(out_cols <= VD_MAX_COLS, VD_EINVAL, "join output has too many columns");
    if (rows) *rows = 0;
    if (nl == 0 || nr == 0) return rel_new(ctx, out_cols, 0);
    VD_REQUIRE(nl < (1ull << 32) && nr < (1ull << 32), VD_ERANGE,
               "join inputs limited to 2^32-1 rows");
    JoinCount jc;
    join_count<NK>(ctx, L, R, jc);
    const uint64_t total = jc.total;
    if (rows) *rows = total;
    if (total > max_rows) return nullptr;
    const JoinOff joff{(const uint32_t *)jc.off.p, jc.tpre};
    DBuf<uint32_t> &lower = jc.lower;

    vd_rel *out = rel_new(ctx, out_cols, total);
    // columns [L's, R's after the keys]: their values come from those columns
    for (int c = 0; c < out_cols; ++c)
        rel_copy_bound(out, c, c < L->ncols ? L : R, c < L->ncols ? c : c - L->ncols + NK);
    out->key_range = true;  // output column 0 = left keys of matched rows
    out->distinct = L->distinct && R->distinct;  // one row per pair of input rows
    out->key_lo = jc.key_lo;
    out->key_hi = jc.key_hi;
    if (total == 0) return out;
    try {
        // columns [L's, R's after the keys]
        OutCols oc{};
        for (int c = 0; c < out_cols; ++c) {
            const bool from_l = c < L->ncols;
            oc.src[c] = from_l ? L->cols[c] : R->cols[c - L->ncols + NK];
            if (!from_l) oc.from_r |= 1u << c;
            oc.o[c] = out->cols[c];
        }
        oc.n = out_cols;
        const int nv = jc.nv;
        const uint64_t nblocks = ceil_div(nl + total, (uint64_t)nv);
        VD_REQUIRE(nblocks < (1ull << 31), VD_ERANGE, "join expansion grid too large");
        DBuf<uint64_t> spare;
        const uint64_t *mp = join_cuts(ctx, joff, nl, total, nv, jc.mp, jc.mp_cap, spare);
        const double xbytes = 4.0 * (double)total * out_cols + 8.0 * nl +
                              4.0 * (double)nl * (L->ncols) + 4.0 * (double)nr * (R->ncols - NK);
        auto go = [&](auto kern) {
            launch(ctx, "join_expand", xbytes, dim3((unsigned)nblocks), dim3(JX_BLOCK),
                   0, kern, joff, (const uint32_t *)lower.p, mp, nl, total, oc);
        };
        if (nv == 1024) go(k_join_expand<1024>);
        else if (nv == 4096) go(k_join_expand<4096>);
        else go(k_join_expand<2048>);
    } catch (...) {
        rel_free(ctx, out);
        throw;
    }
    return out;
}


// Largest key value a driven join's run table covers (8 bytes per value).
constexpr uint32_t DRIVE_MAX_KEY = (1u << 28) - 1;
// The driven form pays for a zero-filled table of every key value up to the
// bound and one random table read per driver row; it replaces a re-sort of
// the join output.  It is taken for inputs of at least ctx->drive_min_rows
// rows together (2^20 by default) whose key bound is at most
// DRIVE_KEYS_PER_ROW times that.
constexpr uint64_t DRIVE_KEYS_PER_ROW = 2;

// Whether a driven join of D (key column dkey) and Y applies, and the
// largest key its run table covers: keys above khi or a known bound of
// either key column match nothing.
bool drive_applies(const vd_ctx *ctx, const vd_rel *D, int dkey, const vd_rel *Y, uint32_t khi,
                   uint32_t *kmax_out) {
    const uint64_t nd = D->nrows, ny = Y->nrows;
    uint32_t kmax = khi;
    if (Y->bounded & 1u) kmax = std::min(kmax, Y->colmax[0]);
    if ((D->bounded >> dkey) & 1u) kmax = std::min(kmax, D->colmax[dkey]);
    *kmax_out = kmax;
    return kmax <= DRIVE_MAX_KEY && nd + ny >= (uint64_t)ctx->drive_min_rows &&
           (uint64_t)kmax + 1 <= DRIVE_KEYS_PER_ROW * (nd + ny);
}

// vd_join_driven (see the C ABI).  Returns nullptr, with nothing left on the
// stream, when the driven form does not apply: inputs too small or a key
// bound too large for the run table, or (check_dups) a repeated driver row
// matching several rows.
vd_rel *join_driven(vd_ctx *ctx, const vd_rel *D, int dkey, const vd_rel *Y, uint32_t klo,
                    uint32_t khi, int nout, const int *src, int check_dups,
                    uint32_t *key_lo, uint32_t *key_hi) {
    VD_REQUIRE(dkey >= 0 && dkey < D->ncols, VD_EINVAL, "driver key column out of range");
    VD_REQUIRE(nout >= 1 && nout <= VD_MAX_COLS, VD_EINVAL, "driven join output columns");
    for (int c = 0; c < nout; ++c)
        VD_REQUIRE(src[c] >= 0 ? src[c] < Y->ncols : -1 - src[c] < D->ncols, VD_EINVAL,
                   "driven join column source out of range");
    const uint64_t nd = D->nrows, ny = Y->nrows;
    VD_REQUIRE(nd < (1ull << 32) && ny < (1ull << 32), VD_ERANGE,
               "join inputs limited to 2^32-1 rows");
    uint32_t kmax = 0;
    if (!drive_applies(ctx, D, dkey, Y, khi, &kmax)) return nullptr;
    *key_lo = *key_hi = 0;
    auto new_out = [&](uint64_t rows) {
        vd_rel *o = rel_new(ctx, nout, rows);
        // One row per (D row, Y row) pair: distinct when both inputs are and
        // the output holds every column of D and every column of Y but its
        // key (equal to D's key in every output row).
        uint32_t dcov = 0, ycov = 0;
        for (int c = 0; c < nout; ++c) {
            if (src[c] >= 0) ycov |= 1u << src[c];
            else dcov |= 1u << (-1 - src[c]);
        }
        const uint32_t dall = (D->ncols >= 32 ? ~0u : (1u << D->ncols) - 1u);
        const uint32_t yall = (Y->ncols >= 32 ? ~0u : (1u << Y->ncols) - 1u) & ~1u;
        o->distinct = D->distinct && Y->distinct && dcov == dall && (ycov | 1u) == (yall | 1u);
        for (int c = 0; c < nout; ++c) {
            if (src[c] >= 0) rel_copy_bound(o, c, Y, src[c]);
            else rel_copy_bound(o, c, D, -1 - src[c]);
        }
        return o;
    };
    if (nd == 0 || ny == 0 || klo > kmax) return new_out(0);

    // one zero fill: the run table, then the repeated-row flag word
    DBuf<uint2> tab(ctx, (uint64_t)kmax + 2);
    fill_u32(ctx, reinterpret_cast<uint32_t *>(tab.p), 0, 2 * ((uint64_t)kmax + 2));
    uint32_t *flag = reinterpret_cast<uint32_t *>(tab.p + kmax + 1);
    launch(ctx, "run_table", 4.0 * (double)ny, dim3((unsigned)ceil_div(ny, RT_ROWS)),
           dim3(RT_BLOCK), 0, k_run_table, (const uint32_t *)Y->cols[0], ny, kmax, tab.p);
    DBuf<uint32_t> lower(ctx, nd), off(ctx, nd);
    const uint64_t jb_tiles = ceil_div(nd, JB_TILE);
    // tile totals, tile prefixes (+ the total), tile key bounds
    DBuf<uint64_t> tt(ctx, 3 * jb_tiles + 1);
    uint64_t *ttot = tt.p, *tpre = tt.p + jb_tiles, *tkeys = tt.p + 2 * jb_tiles + 1;
    DriveCols dc{};
    dc.key = D->cols[dkey];
    if (check_dups) {
        dc.check = D->ncols;
        for (int c = 0; c < D->ncols; ++c) dc.all[c] = D->cols[c];
    }
    launch(ctx, "drive_bounds", 4.0 * (double)nd * (1 + (check_dups ? D->ncols : 0)) +
                                    16.0 * (double)nd,
           dim3((unsigned)jb_tiles), dim3(JB_BLOCK), 0, k_drive_bounds, dc, nd,
           (const uint2 *)tab.p, klo, kmax, lower.p, off.p, ttot, tkeys, flag);
    const JoinOff joff{(const uint32_t *)off.p, tpre};
    const Readback rb = rb_next(ctx);
    launch(ctx, "scan_tiles", 16.0 * (double)jb_tiles, dim3(1), dim3(dev::TT_BLOCK), 0,
           k_join_scan_total, (const uint64_t *)ttot, (const uint64_t *)tkeys, jb_tiles, tpre,
           rb, (const uint32_t *)flag);
    const int nv = ctx->join_nv;
    DBuf<uint64_t> mp;
    uint64_t mp_cap = 0;
    partition_ahead(ctx, joff, nd, tpre + jb_tiles, nv, mp, mp_cap);
    uint64_t h[4];
    rb_wait(ctx, rb, h, 4);
    if (h[3] & 2u) return nullptr;  // a repeated driver row: the order would differ
    const uint64_t total = h[0];
    VD_REQUIRE(total < (1ull << 40), VD_ERANGE, "join output too large");
    VD_REQUIRE((h[3] & 1u) == 0, VD_ERANGE,
               "join: 2048 consecutive driver rows match more than 2^32-1 rows");
    *key_lo = (uint32_t)h[1];
    *key_hi = (uint32_t)h[2];
    vd_rel *out = new_out(total);
    if (total == 0) return out;
    try {
        OutCols oc{};
        // Y's columns the output takes, each once
        int ycol[VD_MAX_COLS], nyc = 0;
        for (int c = 0; c < nout; ++c) {
            if (src[c] >= 0) {
                oc.src[c] = Y->cols[src[c]];
                oc.from_r |= 1u << c;
                int k = 0;
                while (k < nyc && ycol[k] != src[c]) ++k;
                if (k == nyc) ycol[nyc++] = src[c];
                oc.slot[c] = k;
            } else {
                oc.src[c] = D->cols[-1 - src[c]];
            }
            oc.o[c] = out->cols[c];
        }
        oc.n = nout;
        // Y's runs sit at random rows, so each of its columns costs a cache
        // line per run: two to four of them go into one array of records
        // first (one line per run) when the lines saved, one per column
        // after the first for each driver row, outweigh the record pass.
        const int aw = nyc == 2 ? 2 : (nyc == 3 || nyc == 4) ? 4 : 0;
        DBuf<uint32_t> rec;
        if (aw && 64.0 * (nyc - 1) * (double)nd > 4.0 * (double)ny * (nyc + aw)) {
            rec = DBuf<uint32_t>(ctx, ny * (uint64_t)aw);
            RecCols rc{};
            for (int k = 0; k < nyc; ++k) rc.c[k] = Y->cols[ycol[k]];
            rc.nc = nyc;
            const dim3 g((unsigned)ceil_div(ny, 256));
            if (aw == 2)
                launch(ctx, "interleave", 4.0 * (double)ny * (nyc + aw), g, dim3(256), 0,
                       k_interleave<2>, rc, ny, reinterpret_cast<uint2 *>(rec.p));
            else
                launch(ctx, "interleave", 4.0 * (double)ny * (nyc + aw), g, dim3(256), 0,
                       k_interleave<4>, rc, ny, reinterpret_cast<uint4 *>(rec.p));
            oc.rec = rec.p;
        }
        const uint64_t nblocks = ceil_div(nd + total, (uint64_t)nv);
        VD_REQUIRE(nblocks < (1ull << 31), VD_ERANGE, "join expansion grid too large");
        DBuf<uint64_t> spare;
        const uint64_t *cuts = join_cuts(ctx, joff, nd, total, nv, mp, mp_cap, spare);
        int nfd = 0, nfy = 0;
        for (int c = 0; c < nout; ++c) (src[c] >= 0 ? nfy : nfd) += 1;
        const double xbytes = 4.0 * (double)total * nout + 8.0 * nd + 4.0 * (double)nd * nfd +
                              4.0 * (double)ny * nfy;
        auto go = [&](auto kern) {
            launch(ctx, "join_expand", xbytes, dim3((unsigned)nblocks), dim3(JX_BLOCK), 0, kern,
                   joff, (const uint32_t *)lower.p, cuts, nd, total, oc);
        };
        if (!rec.p) {
            if (nv == 1024) go(k_join_expand<1024>);
            else if (nv == 4096) go(k_join_expand<4096>);
            else go(k_join_expand<2048>);
        } else if (aw == 2) {
            if (nv == 1024) go(k_join_expand<1024, 2>);
            else if (nv == 4096) go(k_join_expand<4096, 2>);
            else go(k_join_expand<2048, 2>);
        } else {
            if (nv == 1024) go(k_join_expand<1024, 4>);
            else if (nv == 4096) go(k_join_expand<4096, 4>);
            else go(k_join_expand<2048, 4>);
        }
    } catch (...) {
        rel_free(ctx, out);
        throw;
    }
    return out;
}

}  // namespace
}  // namespace vd

using namespace vd;

extern "C" int vd_join(vd_ctx *ctx, const vd_rel *L, const vd_rel *R, int nkeys,
                       vd_rel **out) {
    if (!ctx || !out) return VD_EINVAL;
    *out = nullptr;
    VD_TRY(ctx, {
        VD_REQUIRE(L && R, VD_EINVAL, "null relation");
        VD_REQUIRE(nkeys == 1 || nkeys == 2, VD_EINVAL, "nkeys must be 1 or 2");
        VD_REQUIRE(L->ncols >= nkeys && R->ncols >= nkeys, VD_EINVAL,
                   "relation has fewer columns than keys");
        *out = nkeys == 1 ? join_impl<1>(ctx, L, R) : join_impl<2>(ctx, L, R);
    });
}

extern "C" int vd_join_driven(vd_ctx *ctx, const vd_rel *D, int dkey, const vd_rel *Y,
                              uint32_t key_lo, uint32_t key_hi, int nout, const int *src,
                              int check_dups, vd_rel **out, uint32_t *match_lo,
                              uint32_t *match_hi) {
    if (!ctx || !out || !match_lo || !match_hi) return VD_EINVAL;
    *out = nullptr;
    VD_TRY(ctx, {
        VD_REQUIRE(D && Y && src, VD_EINVAL, "null argument");
        VD_REQUIRE(Y->ncols >= 1, VD_EINVAL, "relation without columns");
        *out = join_driven(ctx, D, dkey, Y, key_lo, key_hi, nout, src, check_dups != 0,
                           match_lo, match_hi);
    });
}

extern "C" int vd_join_capped(vd_ctx *ctx, const vd_rel *L, const vd_rel *R, int nkeys,
                              uint64_t max_rows, vd_rel **out, uint64_t *rows) {
    if (!ctx || !out || !rows) return VD_EINVAL;
    *out = nullptr;
    *rows = 0;
    VD_TRY(ctx, {
        VD_REQUIRE(L && R, VD_EINVAL, "null relation");
        VD_REQUIRE(nkeys == 1 || nkeys == 2, VD_EINVAL, "nkeys must be 1 or 2");
        VD_REQUIRE(L->ncols >= nkeys && R->ncols >= nkeys, VD_EINVAL,
                   "relation has fewer columns than keys");
        *out = nkeys == 1 ? join_impl<1>(ctx, L, R, max_rows, rows)
                          : join_impl<2>(ctx, L, R, max_rows, rows);
    });
}

extern "C" int vd_join_driven_applies(const vd_ctx *ctx, const vd_rel *D, int dkey,
                                      const vd_rel *Y, uint32_t key_hi) {
    if (!ctx || !D || !Y || dkey < 0 || dkey >= D->ncols || Y->ncols < 1) return 0;
    uint32_t kmax = 0;
    return vd::drive_applies(ctx, D, dkey, Y, key_hi, &kmax) ? 1 : 0;
}
