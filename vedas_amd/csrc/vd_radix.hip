// LSD radix sort of 64-bit keys, one kernel launch per 8-bit digit.
//
// Each pass is a single read and a single write of the keys ("onesweep"):
//  * a block takes the next tile from an atomic counter, so tiles are
//    handed out in the order blocks start;
//  * it ranks its tile's keys with 64-lane ballots (peers with the same
//    digit) and per-wave digit counters in LDS, which keeps the sort stable;
//  * it publishes its per-digit counts and walks back over earlier tiles'
//    published counts (decoupled look-back) to find its global digit
//    offsets;
//  * it stages the tile in LDS in digit order and writes it out, so
//    consecutive lanes write consecutive addresses of a digit bucket.
// The global digit histograms of all passes are counted once up front
// (fused into the producer of the keys where possible).
//
// Look-back termination: tile t only waits on tiles < t.  Tile ids come
// from an atomic counter as blocks start, so every tile < t belongs to a
// block that is already resident, and a block publishes its aggregate
// before it waits on anyone.  Tile 0 publishes an inclusive value without
// waiting.  Hence every wait ends.  The look-back reads a window of
// predecessors per round trip and stops at the first unpublished one.
//
// With tiles grouped per XCD (the radix_group tunable) each of the 8
// counters hands out its own tiles in increasing order as the blocks of
// that counter start.  Let u be the lowest tile no block has taken yet, on
// counter v.  The resident blocks of counter v all hold tiles below u, so
// they wait only on tiles below u, all of which are taken: the lowest
// unfinished taken tile waits on nothing unfinished and completes, and so
// on, until a block of counter v finishes and the next block of counter v
// takes u.  Every wait still ends.

#include "vd_radix_dev.hpp"

namespace vd {
namespace {

using radix::RBINS;
using radix::RBITS;


constexpr uint32_t LB_AGG = 1u << 30;   // tile aggregate published
constexpr uint32_t LB_INC = 2u << 30;   // inclusive prefix published
constexpr uint32_t LB_VAL = (1u << 30) - 1u;

__device__ __forceinline__ void lb_store(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t lb_load(const uint32_t *p) {
    return __hip_atomic_load(const_cast<uint32_t *>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(radix::HIST_BLOCK) void k_rs_hist(
    const uint64_t *__restrict__ keys, uint64_t n, int lo_bit, int hi_bit,
    int npass, uint32_t *ghist) {
    __shared__ radix::LdsHist s;
    radix::hist_zero(&s);
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += stride)
        radix::hist_add(&s, keys[i], lo_bit, hi_bit, npass);
    __syncthreads();
    radix::hist_flush(&s, ghist, npass);
}

// Exclusive scan of each pass's 256-bin histogram: one block per pass.
__global__ __launch_bounds__(RBINS) void k_rs_scan_hist(const uint32_t *ghist,
                                                        uint32_t *gofs) {
    __shared__ uint32_t s_tmp[RBINS / 64];
    const uint32_t v = ghist[blockIdx.x * RBINS + threadIdx.x];
    uint32_t tot;
    const uint32_t ex = dev::block_excl_scan<RBINS>(v, s_tmp, &tot);
    gofs[blockIdx.x * RBINS + threadIdx.x] = ex;
}

// Look-back window: predecessor tiles read per round trip.  The published
// values cross XCDs (agent scope), so each round trip is long; reading two
// tiles at once halves the trips, while wider windows cost more in extra
// loads than they save (measured on 100M keys: windows of 1/2/4/8 tiles give
// 2.99/3.19/3.18/3.10 TB/s per pass).
constexpr int LB_WINDOW = 2;

// Occupancy target (waves per SIMD) of each tile shape: two 512-thread
// blocks per CU, which needs <= 128 VGPRs and <= 80 KB of LDS per block.
template <int BLOCK, int ITEMS, int M>
struct OsShape {
    static constexpr int WAVES_PER_EU = (BLOCK == 512 && ITEMS * (2 + M) <= 48) ? 4 : 1;
};

// Digit width of a pass: 8 bits (256 bins), or 9 bits (512 bins) where one
// 9-bit pass replaces a nearly empty 8-bit one (keys of 25-27 bits in 3
// passes instead of 4).  Threads 0..BINS-1 own one digit each.

// M 32-bit payload columns travel with the keys (pay.in[c] -> pay.out[c]);
// they are loaded with the keys and staged, column after column, through the
// same LDS buffer as the keys once those are written out, so every global
// access stays coalesced and the block keeps two-per-CU occupancy.
// BLOCK threads (a multiple of 256: threads 0..255 own one digit each in the
// count and look-back phases), ITEMS keys per thread.
template <int BLOCK, int ITEMS, int M, int DB = RBITS>
__global__ __launch_bounds__(BLOCK)
__attribute__((amdgpu_waves_per_eu(OsShape<BLOCK, ITEMS, M>::WAVES_PER_EU, 8))) void k_rs_onesweep(
    const uint64_t *__restrict__ in, uint64_t *__restrict__ out, RadixPayload pay,
    RadixKeyIO kio, uint64_t n, int shift, int width, const uint32_t *__restrict__ gofs,
    uint32_t *lookback, uint32_t *tile_counter, uint32_t group) {
    constexpr int WAVES = BLOCK / 64;
    constexpr int TILE = BLOCK * ITEMS;
    constexpr int RBINS = 1 << DB;
    static_assert(BLOCK % RBINS == 0, "one digit per thread in the count phase");
    static_assert(M >= 0 && M <= RADIX_MAX_PAYLOAD, "payload columns");
    static_assert(TILE <= 65535, "16-bit wave counters");
    __shared__ uint64_t s_keys[TILE];  // keys, then each payload column (as u32)
    __shared__ uint16_t s_wcnt[WAVES][RBINS];
    __shared__ uint32_t s_lstart[RBINS];
    __shared__ uint32_t s_gbase[RBINS];
    __shared__ uint32_t s_tmp[WAVES];
    __shared__ uint32_t s_tile;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = tid >> 6;

    for (int i = tid; i < WAVES * RBINS / 2; i += BLOCK)
        reinterpret_cast<uint32_t *>(&s_wcnt[0][0])[i] = 0;
    if (tid == 0) {
        if (group == 0) {
            s_tile = atomicAdd(tile_counter, 1u);
        } else {
            // Block b runs on XCD b mod 8 (workgroups are dispatched to the
            // XCDs round-robin).  Each XCD takes its own tiles from its own
            // counter, `group` consecutive tiles at a time: groups g = v,
            // v + 8, v + 16, ... belong to XCD v.  Neighbouring tiles of a
            // group then run together on one XCD, whose L2 merges the
            // partial cache lines they write at the ends of each digit run.
            const uint32_t v = blockIdx.x & 7u;
            const uint32_t m = atomicAdd(tile_counter + v, 1u);
            s_tile = ((m / group) * 8u + v) * group + m % group;
        }
    }
    __syncthreads();

    const uint32_t tile = s_tile;
    const uint64_t base = (uint64_t)tile * TILE;
    if (base >= n) return;  // a grouped grid rounds up to whole groups per XCD
    const uint64_t left = n - base;
    const uint32_t nvalid = left < (uint64_t)TILE ? (uint32_t)left : (uint32_t)TILE;
    const uint32_t dmask = (1u << width) - 1u;

    // Load: position within the tile = (w * ITEMS + j) * 64 + lane, which is
    // also the ranking order below.  Missing keys are all-ones: they carry
    // the largest digit and rank after every real key of the tile.  Keys
    // come from the key buffer, from two columns, or from another key
    // buffer with their fields moved; a payload column
    // without a source is the row id.
    uint64_t k[ITEMS];
    uint32_t v[M ? M : 1][ITEMS];
    if (kio.in_hi) {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const uint64_t idx = base + (uint64_t)((w * ITEMS + j) * 64 + lane);
            k[j] = idx < n ? ((uint64_t)kio.in_hi[idx] << 32) | (kio.in_lo ? kio.in_lo[idx] : 0u)
                           : ~0ull;
        }
    } else if (kio.in_k) {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const uint64_t idx = base + (uint64_t)((w * ITEMS + j) * 64 + lane);
            const uint64_t v = idx < n ? kio.in_k[idx] : 0ull;
            uint64_t key = 0;
#pragma unroll
            for (int c = 0; c < 3; ++c)
                key |= (uint64_t)((uint32_t)(v >> kio.rp.from_sh[c]) & kio.rp.mask[c])
                       << kio.rp.to_sh[c];
            k[j] = idx < n ? key : ~0ull;
        }
    } else {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const uint64_t idx = base + (uint64_t)((w * ITEMS + j) * 64 + lane);
            k[j] = idx < n ? in[idx] : ~0ull;
        }
    }
#pragma unroll
    for (int c = 0; c < M; ++c) {
        const uint32_t *src = pay.in[c];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const uint64_t idx = base + (uint64_t)((w * ITEMS + j) * 64 + lane);
            v[c][j] = idx < n ? (src ? src[idx] : (uint32_t)idx) : 0u;
        }
    }

    // Rank within the wave: peers = lanes holding the same digit (digit bits
    // above `width` are zero for every lane, so their ballots change nothing).
    // Per digit bit: m = 0 where the lane's bit is set, all-ones where it is
    // clear (sign-extended bit of ~d), so the lane keeps `bal ^ m` = the
    // lanes agreeing with it; both 32-bit halves of the mask use the same m.
    uint32_t rk[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t d = (uint32_t)(k[j] >> shift) & dmask;
        const uint32_t nd = ~d;
        uint32_t plo = ~0u, phi = ~0u;
#pragma unroll
        for (int b = 0; b < DB; ++b) {
            const uint32_t m = (uint32_t)((int32_t)(nd << (31 - b)) >> 31);
            const uint64_t bal = __ballot(m == 0u);
            plo &= (uint32_t)bal ^ m;
            phi &= (uint32_t)(bal >> 32) ^ m;
        }
        const uint32_t before = s_wcnt[w][d];
        const uint32_t r = __builtin_amdgcn_mbcnt_hi(phi, __builtin_amdgcn_mbcnt_lo(plo, 0u));
        const uint32_t cnt = (uint32_t)(__builtin_popcount(plo) + __builtin_popcount(phi));
        __builtin_amdgcn_wave_barrier();
        if (r == cnt - 1) s_wcnt[w][d] = (uint16_t)(before + cnt);
        __builtin_amdgcn_wave_barrier();
        rk[j] = before + r;
    }
    __syncthreads();

    // Threads 0..255, one digit each: wave-exclusive offsets, tile count,
    // publish, look back.
    uint32_t tcount = 0, excl = 0;
    if (tid < RBINS) {
        uint32_t sum = 0;
#pragma unroll
        for (int ww = 0; ww < WAVES; ++ww) {
            const uint32_t c = s_wcnt[ww][tid];
            s_wcnt[ww][tid] = (uint16_t)sum;
            sum += c;
        }
        tcount = sum;
        uint32_t pub = tcount;
        if ((uint32_t)tid == dmask) pub -= (uint32_t)TILE - nvalid;  // padding
        uint32_t *mine = lookback + (uint64_t)tile * RBINS + tid;
        // A digit no key of the whole input holds (bins above a narrow last
        // digit, the top bins of the highest digit) is neither published
        // nor looked up: no tile ever reads it.
        const uint32_t gcount =
            (tid + 1 < RBINS ? gofs[tid + 1] : (uint32_t)n) - gofs[tid];
        if (gcount == 0) {
            // nothing to publish: excl stays 0 and no staged key uses it
        } else if (tile == 0) {
            lb_store(mine, LB_INC | pub);
        } else {
            lb_store(mine, LB_AGG | pub);
            int64_t j = (int64_t)tile - 1;
            for (;;) {
                // Read tiles j, j-1, ..., j-LB_WINDOW+1 at once; take them in
                // order up to the first inclusive value or unpublished tile.
                uint32_t lv[LB_WINDOW];
#pragma unroll
                for (int q = 0; q < LB_WINDOW; ++q)
                    lv[q] = j - q >= 0 ? lb_load(lookback + (uint64_t)(j - q) * RBINS + tid)
                                       : LB_INC;
                bool done = false, stop = false;
                int used = 0;
#pragma unroll
                for (int q = 0; q < LB_WINDOW; ++q) {
                    const uint32_t st = lv[q] & ~LB_VAL;
                    const bool take = !done && !stop && st != 0;
                    stop = stop || st == 0;
                    excl += take ? (lv[q] & LB_VAL) : 0u;
                    used += take ? 1 : 0;
                    done = done || (take && st == LB_INC);
                }
                if (done) break;
                j -= used;
                if (used == 0) __builtin_amdgcn_s_sleep(1);
            }
            lb_store(mine, LB_INC | (excl + pub));
        }
    }

    uint32_t tot;
    const uint32_t lstart = dev::block_excl_scan<BLOCK>(tid < RBINS ? tcount : 0u, s_tmp, &tot);
    if (tid < RBINS) {
        s_lstart[tid] = lstart;
        s_gbase[tid] = gofs[tid] + excl - lstart;
    }
    __syncthreads();

    // Staged positions (< TILE <= 65535), two per register: they stay live
    // until the payload columns are staged.
    uint32_t pos2[(ITEMS + 1) / 2];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t d = (uint32_t)(k[j] >> shift) & dmask;
        const uint32_t p = s_lstart[d] + s_wcnt[w][d] + rk[j];
        if (j % 2 == 0) pos2[j / 2] = p; else pos2[j / 2] |= p << 16;
        s_keys[p] = k[j];
    }
    auto pos = [&](int j) { return (pos2[j / 2] >> (16 * (j % 2))) & 0xFFFFu; };
    __syncthreads();

    // Staged position i (i = tid + BLOCK * r) goes to dst[r].  Positions at
    // or past nvalid hold the padding keys and are not written.
    // The sorted keys go to the key buffer or, split, to two columns.
    uint32_t dst[ITEMS];
    if (kio.out_hi) {
#pragma unroll
        for (int r = 0; r < ITEMS; ++r) {
            const uint32_t i = (uint32_t)(tid + r * BLOCK);
            const uint64_t key = s_keys[i];
            const uint32_t d = (uint32_t)(key >> shift) & dmask;
            dst[r] = s_gbase[d] + i;
            if (i < nvalid) {
                kio.out_hi[dst[r]] = (uint32_t)(key >> 32);
                if (kio.out_lo) kio.out_lo[dst[r]] = (uint32_t)key;
            }
        }
    } else {
#pragma unroll
        for (int r = 0; r < ITEMS; ++r) {
            const uint32_t i = (uint32_t)(tid + r * BLOCK);
            const uint64_t key = s_keys[i];
            const uint32_t d = (uint32_t)(key >> shift) & dmask;
            dst[r] = s_gbase[d] + i;
            if (i < nvalid) out[dst[r]] = key;
        }
    }
    uint32_t *s_v = reinterpret_cast<uint32_t *>(s_keys);
#pragma unroll
    for (int c = 0; c < M; ++c) {
        __syncthreads();  // the previous reads of the staging buffer are done
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) s_v[pos(j)] = v[c][j];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < ITEMS; ++r) {
            const uint32_t i = (uint32_t)(tid + r * BLOCK);
            if (i < nvalid) pay.out[c][dst[r]] = s_v[i];
        }
    }
}

// Pass p reads the keys from a and writes b (then swaps), and moves the
// payload from src (first pass) or the buffer the previous pass wrote, into
// pa (odd passes) or pb (even passes); *final receives the last buffers.
// Pass p sorts on the digit at shifts[p] (widths[p] bits) with the 256
// exclusive digit offsets at gofs[p].  kio (may be null): the first pass
// reads, the last writes, split key columns instead of a and b.
template <int BLOCK, int ITEMS, int M, int DB>
void run_passes(vd_ctx *ctx, uint64_t *&a, uint64_t *&b, const RadixCols &src,
                const RadixCols &pa, const RadixCols &pb, RadixCols *final,
                uint64_t n, const int *shifts, const int *widths, int npass,
                const uint32_t *const *gofs, const RadixKeyIO *kio) {
    constexpr uint64_t TILE = (uint64_t)BLOCK * ITEMS;
    constexpr int BINS = 1 << DB;
    const uint64_t ntiles = ceil_div(n, TILE);
    const uint64_t lb_words = (uint64_t)npass * ntiles * BINS;
    DBuf<uint32_t> lb(ctx, lb_words + 8 * npass);
    fill_u32(ctx, lb.p, 0, lb_words + 8 * npass);
    uint32_t *counters = lb.p + lb_words;
    // Grouped tile order: 8 counters per pass, and a grid of whole groups
    // per XCD (blocks past the last tile exit at once).
    const uint32_t group = (uint32_t)ctx->radix_group;
    const uint64_t grid = group ? 8 * ceil_div(ceil_div(ntiles, group), 8) * group : ntiles;
    RadixCols cur = src;
    for (int p = 0; p < npass; ++p) {
        const RadixCols &dstc = (p % 2 == 0) ? pb : pa;
        RadixPayload pay{};
        for (int c = 0; c < M; ++c) {
            pay.in[c] = cur.p[c];
            pay.out[c] = dstc.p[c];
        }
        RadixKeyIO io{};
        double kbytes = 16.0;
        if (kio && p == 0) {
            io.in_hi = kio->in_hi;
            io.in_lo = kio->in_lo;
            io.in_k = kio->in_k;
            io.rp = kio->rp;
            if (io.in_hi) kbytes -= io.in_lo ? 0.0 : 4.0;
        }
        if (kio && p == npass - 1) {
            io.out_hi = kio->out_hi;
            io.out_lo = kio->out_lo;
            if (io.out_hi) kbytes -= io.out_lo ? 0.0 : 4.0;
        }
        launch(ctx, M ? "radix_onesweep_kv" : "radix_onesweep",
               (double)n * (kbytes + 8.0 * M), dim3((unsigned)grid),
               dim3(BLOCK), 0, k_rs_onesweep<BLOCK, ITEMS, M, DB>, (const uint64_t *)a, b, pay,
               io, n, shifts[p], widths[p], gofs[p],
               lb.p + (uint64_t)p * ntiles * BINS, counters + 8 * p, group);
        std::swap(a, b);
        cur = dstc;
    }
    if (final) *final = cur;
}

// Runs the passes with the tile shape the radix_items tunable selects:
// 16 -> 512 x 16 (default), 8 -> 512 x 8, 24 -> 256 x 24 (keys only).  With
// payload columns the default shape narrows as columns are added (512 x 16 /
// 12 / 8 for 1 / 2 / 3 columns), which keeps the registers of a thread
// within two blocks per CU.
void dispatch_passes(vd_ctx *ctx, uint64_t *&a, uint64_t *&b, int m, const RadixCols &src,
                     const RadixCols &pa, const RadixCols &pb, RadixCols *final, uint64_t n,
                     const int *shifts, const int *widths, int npass,
                     const uint32_t *const *g, int digit_bits = RBITS,
                     const RadixKeyIO *kio = nullptr) {
    const int items = ctx->radix_items;
    if (digit_bits == 9) {
        // 512-thread shapes only (one digit per thread), as the default.
#define VD_RUN9(IT, MM) \
    run_passes<512, IT, MM, 9>(ctx, a, b, src, pa, pb, final, n, shifts, widths, npass, g, kio)
        switch (m) {
            case 0: VD_RUN9(16, 0); break;
            case 1: VD_RUN9(16, 1); break;
            case 2: VD_RUN9(12, 2); break;
            default: VD_RUN9(8, 3); break;
        }
#undef VD_RUN9
        return;
    }
#define VD_RUN(BL, IT, MM) \
    run_passes<BL, IT, MM, RBITS>(ctx, a, b, src, pa, pb, final, n, shifts, widths, npass, g, \
                                  kio)
    if (m > 0) {
        if (items == 8) {
            switch (m) {
                case 1: VD_RUN(512, 8, 1); break;
                case 2: VD_RUN(512, 8, 2); break;
                default: VD_RUN(512, 8, 3); break;
            }
        } else {
            switch (m) {
                case 1: VD_RUN(512, 16, 1); break;
                case 2: VD_RUN(512, 12, 2); break;
                default: VD_RUN(512, 8, 3); break;
            }
        }
        return;
    }
    switch (items) {
        case 8: VD_RUN(512, 8, 0); break;
        case 24: VD_RUN(256, 24, 0); break;
        default: VD_RUN(512, 16, 0); break;
    }
#undef VD_RUN
}

}  // namespace

int radix_passes(int lo_bit, int hi_bit) {
    if (hi_bit <= lo_bit) return 0;
    return (hi_bit - lo_bit + RBITS - 1) / RBITS;
}

static uint64_t *sort_impl(vd_ctx *ctx, uint64_t *a, uint64_t *b, int m,
                           const RadixCols &src, const RadixCols &pa, const RadixCols &pb,
                           RadixCols *final, uint64_t n, int lo_bit, int hi_bit,
                           uint32_t *d_hist, bool hist_ready);

uint64_t *radix_sort_u64(vd_ctx *ctx, uint64_t *keys, uint64_t *tmp,
                         uint64_t n, int lo_bit, int hi_bit, uint32_t *d_hist,
                         bool hist_ready) {
    RadixCols none{};
    return sort_impl(ctx, keys, tmp, 0, none, none, none, nullptr, n, lo_bit, hi_bit, d_hist,
                     hist_ready);
}

void radix_sort_u64_kv(vd_ctx *ctx, uint64_t *&keys, uint64_t *&ktmp, uint32_t *&vals,
                       uint32_t *&vtmp, uint64_t n, int lo_bit, int hi_bit,
                       uint32_t *d_hist, bool hist_ready) {
    RadixCols src{}, pa{}, pb{}, fin{};
    src.p[0] = pa.p[0] = vals;
    pb.p[0] = vtmp;
    uint64_t *k0 = keys, *k1 = ktmp;
    uint64_t *sorted = sort_impl(ctx, k0, k1, 1, src, pa, pb, &fin, n, lo_bit, hi_bit, d_hist,
                                 hist_ready);
    if (sorted != keys) std::swap(keys, ktmp);
    if (fin.p[0] && fin.p[0] != vals) std::swap(vals, vtmp);
}

uint64_t *radix_sort_u64_payload(vd_ctx *ctx, uint64_t *keys, uint64_t *tmp, uint64_t n,
                                 int lo_bit, int hi_bit, uint32_t *d_hist, bool hist_ready,
                                 int m, const RadixCols &src, const RadixCols &pa,
                                 const RadixCols &pb, RadixCols *final) {
    VD_REQUIRE(m >= 0 && m <= RADIX_MAX_PAYLOAD, VD_EINVAL, "payload columns");
    return sort_impl(ctx, keys, tmp, m, src, pa, pb, final, n, lo_bit, hi_bit, d_hist,
                     hist_ready);
}

static uint64_t *sort_impl(vd_ctx *ctx, uint64_t *a, uint64_t *b, int m,
                           const RadixCols &src, const RadixCols &pa, const RadixCols &pb,
                           RadixCols *final, uint64_t n, int lo_bit, int hi_bit,
                           uint32_t *d_hist, bool hist_ready) {
    uint64_t *keys = a;
    const int npass = radix_passes(lo_bit, hi_bit);
    VD_REQUIRE(lo_bit >= 0 && hi_bit <= 64, VD_EINVAL, "radix bit range");
    VD_REQUIRE(npass <= radix::MAX_PASSES, VD_EINVAL, "too many radix passes");
    if (final) *final = src;
    if (n <= 1 || npass == 0) return a;
    VD_REQUIRE(n <= (uint64_t)LB_VAL, VD_ERANGE,
               "radix sort supports at most 2^30-1 keys per call");
    if (!hist_ready) {
        fill_u32(ctx, d_hist, 0, (uint64_t)npass * RBINS);
        uint64_t g = ceil_div(n, 256 * 16);
        if (g > 2048) g = 2048;
        launch(ctx, "radix_hist", (double)n * 8.0, dim3((unsigned)g),
               dim3(radix::HIST_BLOCK), 0, k_rs_hist, (const uint64_t *)keys, n,
               lo_bit, hi_bit, npass, d_hist);
    }
    DBuf<uint32_t> gofs(ctx, (uint64_t)npass * RBINS);
    launch(ctx, "radix_scan_hist", 0.0, dim3(npass), dim3(RBINS), 0,
           k_rs_scan_hist, (const uint32_t *)d_hist, gofs.p);
    int shifts[radix::MAX_PASSES], widths[radix::MAX_PASSES];
    const uint32_t *g[radix::MAX_PASSES];
    for (int p = 0; p < npass; ++p) {
        shifts[p] = lo_bit + p * RBITS;
        widths[p] = std::min(RBITS, hi_bit - shifts[p]);
        g[p] = gofs.p + (uint64_t)p * RBINS;
    }
    dispatch_passes(ctx, a, b, m, src, pa, pb, final, n, shifts, widths, npass, g);
    return a;
}

uint64_t *radix_sort_shifts(vd_ctx *ctx, uint64_t *keys, uint64_t *tmp, uint64_t n,
                            const int *shifts, int npass, const uint32_t *const *d_gofs, int m,
                            const RadixCols &src, const RadixCols &pa, const RadixCols &pb,
                            RadixCols *final, int digit_bits, const RadixKeyIO *kio) {
    VD_REQUIRE(m >= 0 && m <= RADIX_MAX_PAYLOAD, VD_EINVAL, "payload columns");
    VD_REQUIRE(npass >= 0 && npass <= radix::MAX_PASSES, VD_EINVAL, "radix pass count");
    VD_REQUIRE(digit_bits == 8 || digit_bits == 9, VD_EINVAL, "radix digit width");
    if (final) *final = src;
    if (npass == 0 || n <= 1) return keys;
    VD_REQUIRE(n <= (uint64_t)LB_VAL, VD_ERANGE,
               "radix sort supports at most 2^30-1 keys per call");
    int widths[radix::MAX_PASSES];
    for (int p = 0; p < npass; ++p) widths[p] = std::min(digit_bits, 64 - shifts[p]);
    uint64_t *a = keys, *b = tmp;
    dispatch_passes(ctx, a, b, m, src, pa, pb, final, n, shifts, widths, npass, d_gofs,
                    digit_bits, kio);
    return a;
}


}  // namespace vd
