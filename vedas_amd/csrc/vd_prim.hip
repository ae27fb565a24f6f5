// Device primitives: fills, max-reduce and multi-level exclusive scans.
//
// The scans are reduce-then-scan over tiles of SCAN_TILE elements: one
// kernel reduces each tile, the tile sums are scanned recursively, and a
// second kernel scans each tile again adding its tile prefix.  No kernel
// waits on another block, so every launch terminates independently.

#include "vd_internal.hpp"

namespace vd {
namespace {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

template <typename T>
__global__ __launch_bounds__(256) void k_fill(T *p, T v, uint64_t n) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += stride)
        p[i] = v;
}

// 16-byte loads over the aligned body; the < 4 elements before it and the
// < 4 after it are read one by one.
__global__ __launch_bounds__(256) void k_max_u32(const uint32_t *p, uint64_t n,
                                                 uint32_t *out) {
    __shared__ uint32_t s[4];
    uint32_t m = 0;
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t mis = (reinterpret_cast<uintptr_t>(p) & 15u) / 4u;
    const uint64_t head = min<uint64_t>(n, (4u - mis) & 3u);
    if (tid < head) m = p[tid];
    const uint4 *q = reinterpret_cast<const uint4 *>(p + head);
    const uint64_t nq = (n - head) / 4;
#pragma unroll 4
    for (uint64_t i = tid; i < nq; i += stride) {
        const uint4 v = q[i];
        m = max(m, max(max(v.x, v.y), max(v.z, v.w)));
    }
    const uint64_t t0 = head + nq * 4;
    if (t0 + tid < n) m = max(m, p[t0 + tid]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor(m, off, 64));
    if (dev::lane_id() == 0) s[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t r = max(max(s[0], s[1]), max(s[2], s[3]));
        atomicMax(out, r);
    }
}

// Sum of each tile -> partial[tile].
template <typename InT>
__global__ __launch_bounds__(SCAN_BLOCK) void k_tile_reduce(const InT *in,
                                                            uint64_t n,
                                                            uint64_t *partial) {
    __shared__ uint64_t s_tmp[SCAN_BLOCK / 64];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    uint64_t acc = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        uint64_t i = base + (uint64_t)j * SCAN_BLOCK + threadIdx.x;
        if (i < n) acc += (uint64_t)in[i];
    }
    uint64_t tot = dev::block_reduce<SCAN_BLOCK>(acc, s_tmp);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Exclusive scan of each tile plus its prefix (prefix may be null: single
// tile).  The total of the last tile is written to *d_total when given.
template <typename InT>
__global__ __launch_bounds__(SCAN_BLOCK) void k_tile_scan(
    const InT *in, uint64_t *out, uint64_t n, const uint64_t *prefix,
    uint64_t *d_total, uint64_t *tile_tot) {
    __shared__ uint64_t s_val[SCAN_TILE];
    __shared__ uint64_t s_tmp[SCAN_BLOCK / 64];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        int li = j * SCAN_BLOCK + threadIdx.x;
        uint64_t i = base + li;
        s_val[li] = (i < n) ? (uint64_t)in[i] : 0;
    }
    __syncthreads();
    uint64_t v[SCAN_ITEMS];
    uint64_t acc = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        v[j] = s_val[threadIdx.x * SCAN_ITEMS + j];
        acc += v[j];
    }
    uint64_t tot;
    uint64_t run = dev::block_excl_scan<SCAN_BLOCK>(acc, s_tmp, &tot);
    const uint64_t pre = prefix ? prefix[blockIdx.x] : 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        s_val[threadIdx.x * SCAN_ITEMS + j] = pre + run;
        run += v[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        int li = j * SCAN_BLOCK + threadIdx.x;
        uint64_t i = base + li;
        if (i < n) out[i] = s_val[li];
    }
    if (d_total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
        *d_total = pre + tot;
    if (tile_tot && threadIdx.x == 0) tile_tot[blockIdx.x] = tot;
}

// tpre[t] = sum of tot[0..t), t = 0..ntiles.
__global__ __launch_bounds__(dev::TT_BLOCK) void k_scan_tile_totals(
    const uint64_t *__restrict__ tot, uint64_t ntiles, uint64_t *__restrict__ tpre) {
    __shared__ uint64_t s_tmp[dev::TT_BLOCK / 64];
    dev::scan_tile_totals_block(tot, ntiles, tpre, s_tmp);
}

template <typename InT>
void scan_impl(vd_ctx *ctx, const InT *in, uint64_t *out, uint64_t n,
               uint64_t *d_total) {
    if (n == 0) {
        if (d_total) fill_u64(ctx, d_total, 0, 1);
        return;
    }
    const uint64_t tiles = ceil_div(n, SCAN_TILE);
    VD_REQUIRE(tiles < (1ull << 31), VD_ERANGE, "scan too large");
    const double bytes = (double)n * (sizeof(InT) * 2 + sizeof(uint64_t));
    if (tiles == 1) {
        launch(ctx, "scan_tile", bytes, dim3(1), dim3(SCAN_BLOCK), 0,
               k_tile_scan<InT>, in, out, n, (const uint64_t *)nullptr, d_total,
               (uint64_t *)nullptr);
        return;
    }
    DBuf<uint64_t> partial(ctx, tiles);
    launch(ctx, "scan_reduce", (double)n * sizeof(InT), dim3((unsigned)tiles),
           dim3(SCAN_BLOCK), 0, k_tile_reduce<InT>, in, n, partial.p);
    scan_impl<uint64_t>(ctx, partial.p, partial.p, tiles, d_total);
    launch(ctx, "scan_tile", bytes, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0,
           k_tile_scan<InT>, in, out, n, (const uint64_t *)partial.p,
           (uint64_t *)nullptr, (uint64_t *)nullptr);
}

}  // namespace

void scan_u32_to_u64(vd_ctx *ctx, const uint32_t *in, uint64_t *out,
                     uint64_t n, uint64_t *d_total) {
    scan_impl<uint32_t>(ctx, in, out, n, d_total);
}

void scan_u64(vd_ctx *ctx, const uint64_t *in, uint64_t *out, uint64_t n,
              uint64_t *d_total) {
    scan_impl<uint64_t>(ctx, in, out, n, d_total);
}

void scan_u64_ws(vd_ctx *ctx, const uint64_t *in, uint64_t *out, uint64_t n,
                 uint64_t *d_total, uint64_t *ws, uint64_t ws_words) {
    if (n == 0) {
        if (d_total) fill_u64(ctx, d_total, 0, 1);
        return;
    }
    const uint64_t tiles = ceil_div(n, SCAN_TILE);
    VD_REQUIRE(tiles <= ws_words && tiles <= (uint64_t)SCAN_TILE, VD_ERANGE,
               "scan larger than its workspace");
    const double bytes = (double)n * 24.0;
    if (tiles == 1) {
        launch(ctx, "scan_tile", bytes, dim3(1), dim3(SCAN_BLOCK), 0, k_tile_scan<uint64_t>, in,
               out, n, (const uint64_t *)nullptr, d_total, (uint64_t *)nullptr);
        return;
    }
    launch(ctx, "scan_reduce", (double)n * 8.0, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0,
           k_tile_reduce<uint64_t>, in, n, ws);
    launch(ctx, "scan_tile", (double)tiles * 24.0, dim3(1), dim3(SCAN_BLOCK), 0,
           k_tile_scan<uint64_t>, (const uint64_t *)ws, ws, tiles, (const uint64_t *)nullptr,
           d_total, (uint64_t *)nullptr);
    launch(ctx, "scan_tile", bytes, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0,
           k_tile_scan<uint64_t>, in, out, n, (const uint64_t *)ws, (uint64_t *)nullptr,
           (uint64_t *)nullptr);
}

static_assert(SCAN_TILE == OFF_TILE, "tile offsets use the scan tile");

void scan_tile_totals(vd_ctx *ctx, const uint64_t *tot, uint64_t ntiles, uint64_t *tpre) {
    launch(ctx, "scan_tiles", 8.0 * (double)ntiles, dim3(1), dim3(dev::TT_BLOCK), 0,
           k_scan_tile_totals,
           tot, ntiles, tpre);
}

static unsigned grid_for(uint64_t n) {
    uint64_t g = ceil_div(n, 256);
    if (g > 8192) g = 8192;
    if (g == 0) g = 1;
    return (unsigned)g;
}

void fill_u32(vd_ctx *ctx, uint32_t *p, uint32_t v, uint64_t n) {
    if (n == 0) return;
    launch(ctx, "fill", (double)n * 4, dim3(grid_for(n)), dim3(256), 0,
           k_fill<uint32_t>, p, v, n);
}

void fill_u64(vd_ctx *ctx, uint64_t *p, uint64_t v, uint64_t n) {
    if (n == 0) return;
    launch(ctx, "fill", (double)n * 8, dim3(grid_for(n)), dim3(256), 0,
           k_fill<uint64_t>, p, v, n);
}

void max_u32_dev(vd_ctx *ctx, const uint32_t *p, uint64_t n, uint32_t *d_out) {
    launch(ctx, "max_u32", (double)n * 4, dim3(grid_for(n) > 2048 ? 2048 : grid_for(n)),
           dim3(256), 0, k_max_u32, p, n, d_out);
}

uint32_t max_u32(vd_ctx *ctx, const uint32_t *p, uint64_t n) {
    if (n == 0) return 0;
    DBuf<uint64_t> d(ctx, 1);
    fill_u64(ctx, d.p, 0, 1);
    launch(ctx, "max_u32", (double)n * 4, dim3(grid_for(n) > 2048 ? 2048 : grid_for(n)),
           dim3(256), 0, k_max_u32, p, n, (uint32_t *)d.p);
    return (uint32_t)read_u64(ctx, d.p);
}

}  // namespace vd
