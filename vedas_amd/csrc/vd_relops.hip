// Relation re-sorts and finalisation: sort by column 0 (IndexSwap),
// lexicographic sort, removal of adjacent duplicate rows, and the
// semi-join filter that may precede an IndexSwap.
//
// sort_col0: key = col0 << 32 | col1, radix-sorted on the bits of col0 only
// (LSD passes are stable, so equal keys keep their original order), with a
// 32-bit value for a third column or the row id of wider relations.

#include <vector>

#include "vd_radix_dev.hpp"
#include "vd_swap_dev.hpp"

namespace vd {
namespace {

constexpr int RB = radix::HIST_BLOCK;
constexpr int R_ITEMS = 16;
constexpr int U_BLOCK = 256;
constexpr int U_ITEMS = 8;
constexpr int U_TILE = U_BLOCK * U_ITEMS;
constexpr int U_KEEP_COLS = 6;  // columns the unique scatter keeps in registers

struct Cols {
    const uint32_t *in[VD_MAX_COLS];
    uint32_t *out[VD_MAX_COLS];
    int n;
};

// key = col[perm[i]] << 32 | i (perm may be null = identity).
__global__ __launch_bounds__(RB) void k_pack_rowid(
    const uint32_t *__restrict__ col, const uint32_t *__restrict__ perm,
    uint64_t n, uint64_t *__restrict__ keys, int hi_bit, int npass,
    uint32_t *ghist) {
    __shared__ radix::LdsHist s;
    radix::hist_zero(&s);
    __syncthreads();
    for (uint64_t base = (uint64_t)blockIdx.x * (RB * R_ITEMS); base < n;
         base += (uint64_t)gridDim.x * (RB * R_ITEMS)) {
#pragma unroll 4
        for (int j = 0; j < R_ITEMS; ++j) {
            const uint64_t i = base + (uint64_t)j * RB + threadIdx.x;
            if (i < n) {
                const uint32_t v = perm ? col[perm[i]] : col[i];
                const uint64_t key = ((uint64_t)v << 32) | (uint64_t)i;
                keys[i] = key;
                radix::hist_add(&s, key, 32, hi_bit, npass);
            }
        }
    }
    __syncthreads();
    radix::hist_flush(&s, ghist, npass);
}

// Histograms of column 0's digits under the plan of *gmax (column 0's
// maximum, computed before) into g8 (8-bit plan) or g9 (9-bit plan).
// 16-byte loads over the aligned body (column 0 of a view starts at any
// 4-byte offset); the < 4 values before it and the < 4 after it are counted
// by the first threads of the grid.
__global__ __launch_bounds__(RB) void k_swap_hist(const uint32_t *__restrict__ c0, uint64_t n,
                                                  uint32_t *g8, uint32_t *g9,
                                                  const uint32_t *gmax, int bits) {
    __shared__ uint32_t s_h[radix::HIST_WAVES][swap::HW];
    static_assert(RB == radix::HIST_BLOCK, "one histogram copy per wave");
    for (int i = threadIdx.x; i < radix::HIST_WAVES * swap::HW; i += RB) (&s_h[0][0])[i] = 0;
    const swap::Plan plan = swap::plan_for(*gmax, bits);
    __syncthreads();
    uint32_t *h = s_h[threadIdx.x / 64];
    const uint64_t tid = (uint64_t)blockIdx.x * RB + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * RB;
    const uint64_t mis = (reinterpret_cast<uintptr_t>(c0) & 15u) / 4u;
    const uint64_t head = min<uint64_t>(n, (4u - mis) & 3u);
    const uint4 *q = reinterpret_cast<const uint4 *>(c0 + head);
    const uint64_t nq = (n - head) / 4;
    const uint64_t t0 = head + nq * 4;
    if (tid < head) swap::hist_add(h, c0[tid], plan);
    if (t0 + tid < n) swap::hist_add(h, c0[t0 + tid], plan);
#pragma unroll 2
    for (uint64_t i = tid; i < nq; i += stride) {
        const uint4 v = q[i];
        swap::hist_add(h, v.x, plan);
        swap::hist_add(h, v.y, plan);
        swap::hist_add(h, v.z, plan);
        swap::hist_add(h, v.w, plan);
    }
    __syncthreads();
    swap::hist_flush(&s_h[0][0], radix::HIST_WAVES, plan, g8, g9);
}

// One block: the exclusive scans of the digits the producer counted (the
// plan of the maximum *gmax) into gofs8 or gofs9, then (publish) a
// read-back of {bit d set: digit d varies (no bin holds all n rows), the
// maximum}.
constexpr int SP_BLOCK = 512;
__global__ __launch_bounds__(SP_BLOCK) void k_swap_plan_scan(
    const uint32_t *__restrict__ g8, const uint32_t *__restrict__ g9, const uint32_t *gmax,
    uint64_t n, uint32_t *__restrict__ gofs8, uint32_t *__restrict__ gofs9, Readback rb,
    int bits, int publish) {
    __shared__ uint32_t s_tmp[SP_BLOCK / 64];
    const uint32_t mx = *gmax;
    const swap::Plan plan = swap::plan_for(mx, bits);
    const int bins = plan.use9 ? 512 : 256;
    const uint32_t *h = plan.use9 ? g9 : g8;
    uint32_t *go = plan.use9 ? gofs9 : gofs8;
    const int tid = (int)threadIdx.x;
    uint64_t varies = 0;
    for (int d = 0; d < plan.ndig; ++d) {
        const uint32_t v = tid < bins ? h[d * bins + tid] : 0u;
        uint32_t tot;
        const uint32_t ex = dev::block_excl_scan<SP_BLOCK>(v, s_tmp, &tot);
        if (tid < bins) go[d * bins + tid] = ex;
        const int one_bin = __syncthreads_or(tid < bins && (uint64_t)v == n);
        varies |= (one_bin ? 0ull : 1ull) << d;
    }
    if (publish && tid < 64) {
        if (tid == 0) {
            rb.dst[0] = varies;
            rb.dst[1] = mx;
        }
        dev::rb_publish(rb);
    }
}

// perm2[i] = perm[key_i & 0xffffffff] (perm null = identity).
__global__ __launch_bounds__(256) void k_compose_perm(
    const uint64_t *__restrict__ keys, const uint32_t *__restrict__ perm,
    uint64_t n, uint32_t *__restrict__ perm2) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += stride) {
        const uint32_t r = (uint32_t)keys[i];
        perm2[i] = perm ? perm[r] : r;
    }
}

__global__ __launch_bounds__(256) void k_gather_perm(
    const uint32_t *__restrict__ perm, uint64_t n, Cols cols) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += stride) {
        const uint32_t r = perm[i];
        for (int c = 0; c < cols.n; ++c) cols.out[c][i] = cols.in[c][r];
    }
}

__global__ __launch_bounds__(U_BLOCK) void k_unique_count(Cols cols, uint64_t n,
                                                          uint32_t *cnt) {
    __shared__ uint32_t s_tmp[U_BLOCK / 64];
    const uint64_t base = (uint64_t)blockIdx.x * U_TILE;
    const uint32_t c =
        (uint32_t)__popc(dev::unique_keep_bits<U_BLOCK, U_ITEMS>(cols.in, cols.n, n, base));
    const uint32_t t = dev::block_reduce<U_BLOCK>(c, s_tmp);
    if (threadIdx.x == 0) cnt[blockIdx.x] = t;
}

__global__ __launch_bounds__(U_BLOCK) void k_unique_scatter(
    Cols cols, uint64_t n, const uint64_t *__restrict__ off) {
    __shared__ uint32_t s_cnt[U_ITEMS][U_BLOCK / 64];
    const uint64_t base = (uint64_t)blockIdx.x * U_TILE;
    uint32_t vals[U_KEEP_COLS][U_ITEMS];
    const uint32_t keep =
        dev::unique_keep_bits<U_BLOCK, U_ITEMS, U_KEEP_COLS>(cols.in, cols.n, n, base, vals);
    dev::unique_scatter_tile<U_BLOCK, U_ITEMS, U_KEEP_COLS>(keep, cols.in, cols.out, cols.n,
                                                            base, off[blockIdx.x], s_cnt, vals);
}

unsigned sgrid(uint64_t n) {
    uint64_t g = ceil_div(n, 256 * 4);
    if (g > 16384) g = 16384;
    return g ? (unsigned)g : 1u;
}

// Sort keys (col[perm[i]] << 32 | i) on the bits of col; returns sorted keys
// (owned by one of the two buffers).
uint64_t *sort_by_col(vd_ctx *ctx, const uint32_t *col, const uint32_t *perm,
                      uint64_t n, DBuf<uint64_t> &keys, DBuf<uint64_t> &tmp) {
    const int b = std::max(1, bits_for(max_u32(ctx, col, n)));
    const int hi = 32 + b;
    const int npass = radix_passes(32, hi);
    DBuf<uint32_t> hist(ctx, radix_hist_words(32, hi));
    fill_u32(ctx, hist.p, 0, radix_hist_words(32, hi));
    launch(ctx, "sort_pack", (double)n * (perm ? 16.0 : 12.0),
           dim3(radix::hist_grid(ceil_div(n, RB * R_ITEMS))), dim3(RB), 0, k_pack_rowid,
           col, perm, n, keys.p, hi, npass, hist.p);
    return radix_sort_u64(ctx, keys.p, tmp.p, n, 32, hi, hist.p, true);
}

// Stable sort by column 0 (the reference's sortByFirstColumn).  The sort
// key is col0 << 32 | col1, sorted on the digits of column 0 only (LSD
// passes are stable, so equal keys keep their order).  Column 0's maximum
// picks the digit plan; one kernel counts its digits' histograms, which are
// read back as one flag per digit, and the digits that are equal in every
// row (the top digits of small IDs) get no pass.  The first pass reads the
// keys from columns 0 and 1 and the last writes them back as columns, so
// the keys exist in memory only between passes.  Columns 2..4 travel
// through the passes as payload columns (read from the relation by the
// first pass); wider relations carry the row id and gather columns 2.. once
// at the end.  Owned columns are reused as buffers; borrowed (view) columns
// are replaced, never written.
void sort_col0(vd_ctx *ctx, vd_rel *rel) {
    const uint64_t n = rel->nrows;
    const int nc = rel->ncols;
    if (n <= 1) return;
    rel->key_range = false;
    VD_REQUIRE(n < (1ull << 30), VD_ERANGE, "sort supports fewer than 2^30 rows");
    const bool by_row = nc > 2 + RADIX_MAX_PAYLOAD;
    const int m = nc <= 2 ? 0 : (by_row ? 1 : nc - 2);
    DBuf<uint32_t> hist(ctx, swap::HBUF), gofs(ctx, swap::HBUF), dmax(ctx, 1);
    uint32_t *h8 = hist.p, *h9 = hist.p + swap::H8;
    fill_u32(ctx, hist.p, 0, swap::HBUF);
    int shifts[4], dbits = 8, npass;
    const uint32_t *g[4];
    if (rel->bounded & 1u) {
        // The value bound fixes the plan: no maximum and no read-back; every
        // digit below its bit length gets a pass (one that no row changes
        // leaves the order as it is).
        fill_u32(ctx, dmax.p, rel->colmax[0], 1);
        swap_hist(ctx, rel->cols[0], n, h8, h9, dmax.p);
        npass = swap_passes_bounded(ctx, h8, h9, dmax.p, rel->colmax[0], n, gofs.p, 32, shifts,
                                    g, &dbits);
    } else {
        fill_u32(ctx, dmax.p, 0, 1);
        max_u32_dev(ctx, rel->cols[0], n, dmax.p);
        swap_hist(ctx, rel->cols[0], n, h8, h9, dmax.p);
        npass = swap_passes(ctx, h8, h9, dmax.p, n, gofs.p, 32, shifts, g, &dbits);
    }
    if (npass == 0) return;  // column 0 is constant: already in order

    // Key buffers between passes: none for one pass, one for two.
    DBuf<uint64_t> keys(ctx, npass >= 3 ? n : 0), ktmp(ctx, npass >= 2 ? n : 0);
    DBuf<uint32_t> rows(ctx, by_row ? n : 0);
    // Payload buffers: src = the relation's columns (null for the row ids);
    // pa aliases src where the relation owns the column, pb is fresh.
    std::vector<uint32_t *> fresh;
    auto fresh_col = [&]() {
        fresh.push_back(nullptr);
        fresh.back() = (uint32_t *)dalloc(ctx, n * 4);
        return fresh.back();
    };
    RadixCols src{}, pa{}, pb{}, fin{};
    uint32_t *out01[2] = {nullptr, nullptr};
    uint32_t *gathered[VD_MAX_COLS] = {};
    try {
        for (int c = 0; c < m; ++c) {
            if (by_row) {
                src.p[0] = nullptr;
                pa.p[0] = rows.p;
            } else {
                src.p[c] = rel->cols[2 + c];
                pa.p[c] = rel_borrowed(rel, 2 + c) ? fresh_col() : rel->cols[2 + c];
            }
            pb.p[c] = fresh_col();
        }
        // A single pass reads and writes columns 0 and 1 in one launch: it
        // writes fresh ones.
        for (int c = 0; c < std::min(nc, 2); ++c)
            out01[c] = (npass == 1 || rel_borrowed(rel, c)) ? fresh_col() : rel->cols[c];
        if (by_row)
            for (int c = 2; c < nc; ++c) gathered[c] = fresh_col();
        RadixKeyIO kio{rel->cols[0], nc >= 2 ? rel->cols[1] : nullptr, out01[0], out01[1]};
        radix_sort_shifts(ctx, keys.p, ktmp.p, n, shifts, npass, g, m, src, pa, pb, &fin, dbits,
                          &kio);
    } catch (...) {
        for (auto p : fresh) dfree(ctx, p);
        throw;
    }
    if (by_row) {
        Cols cols{};
        cols.n = nc - 2;
        for (int c = 2; c < nc; ++c) {
            cols.in[c - 2] = rel->cols[c];
            cols.out[c - 2] = gathered[c];
        }
        launch(ctx, "sort_gather", (double)n * (4.0 + 8.0 * (nc - 2)), dim3(sgrid(n)),
               dim3(256), 0, k_gather_perm, (const uint32_t *)fin.p[0], n, cols);
    }
    // Install the new columns; free the scratch that is not installed.
    auto installed = [&](uint32_t *p) {
        for (int c = 0; c < nc; ++c)
            if (rel->cols[c] == p) return true;
        return false;
    };
    for (int c = 0; c < std::min(nc, 2); ++c)
        if (out01[c] != rel->cols[c]) rel_set_col(ctx, rel, c, out01[c]);
    if (by_row) {
        for (int c = 2; c < nc; ++c) rel_set_col(ctx, rel, c, gathered[c]);
    } else {
        for (int c = 0; c < m; ++c)
            if (fin.p[c] != rel->cols[2 + c]) rel_set_col(ctx, rel, 2 + c, fin.p[c]);
    }
    for (auto p : fresh)
        if (!installed(p)) dfree(ctx, p);
}

void sort_lex(vd_ctx *ctx, vd_rel *rel) {
    const uint64_t n = rel->nrows;
    if (n <= 1) return;
    rel->key_range = false;
    VD_REQUIRE(n < (1ull << 30), VD_ERANGE, "sort supports fewer than 2^30 rows");
    DBuf<uint64_t> keys(ctx, n), tmp(ctx, n);
    DBuf<uint32_t> perm(ctx, n), perm2(ctx, n);
    bool have_perm = false;
    for (int c = rel->ncols - 1; c >= 0; --c) {
        uint64_t *sorted = sort_by_col(ctx, rel->cols[c],
                                       have_perm ? perm.p : nullptr, n, keys, tmp);
        launch(ctx, "sort_compose", (double)n * 16.0, dim3(sgrid(n)), dim3(256),
               0, k_compose_perm, (const uint64_t *)sorted,
               have_perm ? (const uint32_t *)perm.p : (const uint32_t *)nullptr,
               n, perm2.p);
        std::swap(perm.p, perm2.p);
        have_perm = true;
    }
    Cols cols{};
    cols.n = rel->ncols;
    uint32_t *fresh[VD_MAX_COLS] = {};
    try {
        for (int c = 0; c < rel->ncols; ++c) {
            fresh[c] = (uint32_t *)dalloc(ctx, n * 4);
            cols.in[c] = rel->cols[c];
            cols.out[c] = fresh[c];
        }
    } catch (...) {
        for (auto p : fresh) dfree(ctx, p);
        throw;
    }
    launch(ctx, "sort_gather", (double)n * (4.0 + 8.0 * rel->ncols),
           dim3(sgrid(n)), dim3(256), 0, k_gather_perm, (const uint32_t *)perm.p,
           n, cols);
    for (int c = 0; c < rel->ncols; ++c) {
        rel_set_col(ctx, rel, c, fresh[c]);
    }
}

void unique_adjacent(vd_ctx *ctx, vd_rel *rel) {
    const uint64_t n = rel->nrows;
    if (n <= 1 || rel->distinct) return;  // distinct rows: nothing adjacent repeats
    Cols cols{};
    cols.n = rel->ncols;
    for (int c = 0; c < rel->ncols; ++c) cols.in[c] = rel->cols[c];
    const uint64_t tiles = ceil_div(n, U_TILE);
    DBuf<uint32_t> cnt(ctx, tiles);
    DBuf<uint64_t> off(ctx, tiles + 1);
    launch(ctx, "unique_count", (double)n * 4.0 * rel->ncols, dim3((unsigned)tiles),
           dim3(U_BLOCK), 0, k_unique_count, cols, n, cnt.p);
    scan_u32_to_u64(ctx, cnt.p, off.p, tiles, off.p + tiles);
    const uint64_t kept = read_u64(ctx, off.p + tiles);
    if (kept == n) return;
    uint32_t *fresh[VD_MAX_COLS] = {};
    try {
        for (int c = 0; c < rel->ncols; ++c) {
            fresh[c] = (uint32_t *)dalloc(ctx, (kept ? kept : 1) * 4);
            cols.out[c] = fresh[c];
        }
    } catch (...) {
        for (auto p : fresh) dfree(ctx, p);
        throw;
    }
    launch(ctx, "unique_scatter", 4.0 * rel->ncols * (double)(n + kept),
           dim3((unsigned)tiles), dim3(U_BLOCK), 0, k_unique_scatter, cols, n,
           (const uint64_t *)off.p);
    for (int c = 0; c < rel->ncols; ++c) {
        rel_set_col(ctx, rel, c, fresh[c]);
    }
    rel->nrows = kept;
}

// ---- semi-join filter ------------------------------------------------------------
//
// The rows of X whose key (column xcol) occurs in column ycol of Y, in X's
// order.  Y's keys set bits of a bitmap over the key values; X's rows are
// kept by their bit, counted per tile, scanned, and scattered as the
// adjacent unique scatters (one ballot per round).  A join after a stable
// sort of X on the key keeps exactly the rows this filter keeps, in the
// same order, so the filter may run before the sort.

// Largest key value the bitmap covers (32 MB of bits).
constexpr uint32_t SEMI_MAX_KEY = (1u << 28) - 1;

// Bit k of bits (word k / 32, bit k % 32) for each value k <= kmax of col.
// Only a row whose value differs from the row before it sets a bit (a
// sorted column's runs of one value, hot keys included, set it once), and
// a wave merges the bits of its lanes that fall in one word, so one lane
// per word sets them.
__global__ __launch_bounds__(256) void k_semi_mark(const uint32_t *__restrict__ col, uint64_t n,
                                                   uint32_t kmax, uint32_t *__restrict__ bits) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const int lane = (int)dev::lane_id();
    const uint32_t k = i < n ? col[i] : ~0u;
    const uint32_t before = lane == 0 && i > 0 && i <= n ? col[i - 1] : 0u;
    const uint32_t up = (uint32_t)__shfl_up((int)k, 1, dev::WAVE);
    const uint32_t prev = lane == 0 ? before : up;
    const bool in = k <= kmax;
    const uint32_t w = in ? k >> 5 : ~0u;
    uint32_t b = in && (i == 0 || prev != k) ? 1u << (k & 31u) : 0u;
    // after the step of offset o, a lane holds the bits of the lanes
    // [lane, lane + 2o) of its run of one word
#pragma unroll
    for (int o = 1; o < dev::WAVE; o <<= 1) {
        const uint32_t w2 = (uint32_t)__shfl_down((int)w, o, dev::WAVE);
        const uint32_t b2 = (uint32_t)__shfl_down((int)b, o, dev::WAVE);
        if (lane + o < dev::WAVE && w2 == w) b |= b2;
    }
    const uint32_t wp = (uint32_t)__shfl_up((int)w, 1, dev::WAVE);
    if (b && (lane == 0 || wp != w)) atomicOr(bits + w, b);
}

// Keep flags of a tile (bit j: the row of round j exists and its key's bit
// is set), all key loads then all bitmap loads in flight together.
__device__ __forceinline__ uint32_t semi_keep_bits(const uint32_t *__restrict__ key, uint64_t n,
                                                   uint64_t base, uint32_t kmax,
                                                   const uint32_t *__restrict__ bits) {
    uint32_t k[U_ITEMS], w[U_ITEMS];
#pragma unroll
    for (int j = 0; j < U_ITEMS; ++j) {
        const uint64_t i = base + (uint64_t)j * U_BLOCK + threadIdx.x;
        k[j] = i < n ? key[i] : ~0u;  // ~0u > kmax: not kept
    }
#pragma unroll
    for (int j = 0; j < U_ITEMS; ++j) w[j] = k[j] <= kmax ? bits[k[j] >> 5] : 0u;
    uint32_t keep = 0;
#pragma unroll
    for (int j = 0; j < U_ITEMS; ++j) keep |= ((w[j] >> (k[j] & 31u)) & 1u) << j;
    return keep;
}

__global__ __launch_bounds__(U_BLOCK) void k_semi_count(const uint32_t *__restrict__ key,
                                                        uint64_t n, uint32_t kmax,
                                                        const uint32_t *__restrict__ bits,
                                                        uint32_t *__restrict__ cnt) {
    __shared__ uint32_t s_tmp[U_BLOCK / 64];
    const uint64_t base = (uint64_t)blockIdx.x * U_TILE;
    const uint32_t c = (uint32_t)__popc(semi_keep_bits(key, n, base, kmax, bits));
    const uint32_t t = dev::block_reduce<U_BLOCK>(c, s_tmp);
    if (threadIdx.x == 0) cnt[blockIdx.x] = t;
}

__global__ __launch_bounds__(U_BLOCK) void k_semi_scatter(Cols cols, const uint32_t *__restrict__ key,
                                                          uint64_t n, uint32_t kmax,
                                                          const uint32_t *__restrict__ bits,
                                                          const uint64_t *__restrict__ off) {
    __shared__ uint32_t s_cnt[U_ITEMS][U_BLOCK / 64];
    const uint64_t base = (uint64_t)blockIdx.x * U_TILE;
    const uint32_t keep = semi_keep_bits(key, n, base, kmax, bits);
    dev::unique_scatter_tile<U_BLOCK, U_ITEMS, 0>(keep, cols.in, cols.out, cols.n, base,
                                                  off[blockIdx.x], s_cnt, nullptr);
}

// vd_semi_filter (see the C ABI).
vd_rel *semi_filter(vd_ctx *ctx, const vd_rel *X, int xcol, const vd_rel *Y, int ycol,
                    uint32_t key_hi, uint64_t max_rows, uint64_t *rows) {
    uint32_t kmax = key_hi;
    if ((X->bounded >> xcol) & 1u) kmax = std::min(kmax, X->colmax[xcol]);
    if ((Y->bounded >> ycol) & 1u) kmax = std::min(kmax, Y->colmax[ycol]);
    *rows = ~0ull;
    if (kmax > SEMI_MAX_KEY) return nullptr;  // no bitmap that large
    const uint64_t n = X->nrows, ny = Y->nrows;
    auto new_out = [&](uint64_t kept) {
        vd_rel *o = rel_new(ctx, X->ncols, kept);
        for (int c = 0; c < X->ncols; ++c) rel_copy_bound(o, c, X, c);
        o->distinct = X->distinct;  // rows of X, each once
        return o;
    };
    if (n == 0 || ny == 0) {
        *rows = 0;
        return new_out(0);
    }
    const uint64_t words = (uint64_t)kmax / 32 + 1;
    DBuf<uint32_t> bits(ctx, words);
    fill_u32(ctx, bits.p, 0, words);
    launch(ctx, "semi_mark", 4.0 * (double)ny, dim3((unsigned)ceil_div(ny, 256)), dim3(256), 0,
           k_semi_mark, (const uint32_t *)Y->cols[ycol], ny, kmax, bits.p);
    const uint64_t tiles = ceil_div(n, U_TILE);
    DBuf<uint32_t> cnt(ctx, tiles);
    DBuf<uint64_t> off(ctx, tiles + 1);
    const uint32_t *key = X->cols[xcol];
    launch(ctx, "semi_count", 4.0 * (double)n, dim3((unsigned)tiles), dim3(U_BLOCK), 0,
           k_semi_count, key, n, kmax, (const uint32_t *)bits.p, cnt.p);
    scan_u32_to_u64(ctx, cnt.p, off.p, tiles, off.p + tiles);
    const uint64_t kept = read_u64(ctx, off.p + tiles);
    *rows = kept;
    if (kept > max_rows) return nullptr;
    vd_rel *out = new_out(kept);
    if (kept == 0) return out;
    Cols cols{};
    cols.n = X->ncols;
    for (int c = 0; c < X->ncols; ++c) {
        cols.in[c] = X->cols[c];
        cols.out[c] = out->cols[c];
    }
    launch(ctx, "semi_scatter", 4.0 * (double)n + 4.0 * X->ncols * (double)kept,
           dim3((unsigned)tiles), dim3(U_BLOCK), 0, k_semi_scatter, cols, key, n, kmax,
           (const uint32_t *)bits.p, (const uint64_t *)off.p);
    return out;
}

}  // namespace

int swap_passes_bounded(vd_ctx *ctx, const uint32_t *h8, const uint32_t *h9,
                        const uint32_t *d_max, uint32_t mx, uint64_t n, uint32_t *gofs,
                        int lo_bit, int *shifts, const uint32_t **g, int *digit_bits) {
    launch(ctx, "radix_scan_hist", 0.0, dim3(1), dim3(SP_BLOCK), 0, k_swap_plan_scan, h8, h9,
           d_max, n, gofs, gofs + swap::H8, Readback{}, ctx->radix_bits, 0);
    const swap::Plan plan = swap::plan_for(mx, ctx->radix_bits);
    const int W = plan.use9 ? 9 : 8;
    *digit_bits = W;
    for (int d = 0; d < plan.ndig; ++d) {
        shifts[d] = lo_bit + W * d;
        g[d] = plan.use9 ? gofs + swap::H8 + d * 512 : gofs + d * 256;
    }
    return plan.ndig;
}

void swap_hist(vd_ctx *ctx, const uint32_t *col, uint64_t n, uint32_t *h8, uint32_t *h9,
               const uint32_t *d_max) {
    launch(ctx, "sort_hist", (double)n * 4.0, dim3(radix::hist_grid(ceil_div(n, RB * R_ITEMS))),
           dim3(RB), 0, k_swap_hist, col, n, h8, h9, d_max, ctx->radix_bits);
}

int swap_passes(vd_ctx *ctx, const uint32_t *h8, const uint32_t *h9, const uint32_t *d_max,
                uint64_t n, uint32_t *gofs, int lo_bit, int *shifts, const uint32_t **g,
                int *digit_bits) {
    const Readback rb = rb_next(ctx);
    launch(ctx, "radix_scan_hist", 0.0, dim3(1), dim3(SP_BLOCK), 0, k_swap_plan_scan, h8, h9,
           d_max, n, gofs, gofs + swap::H8, rb, ctx->radix_bits, 1);
    uint64_t hv[2];
    rb_wait(ctx, rb, hv, 2);
    // The digits of the plan that vary (the producer counted only the
    // plan's digits below the maximum's bit length).
    const swap::Plan plan = swap::plan_for((uint32_t)hv[1], ctx->radix_bits);
    const int W = plan.use9 ? 9 : 8;
    *digit_bits = W;
    int npass = 0;
    for (int d = 0; d < plan.ndig; ++d)
        if ((hv[0] >> d) & 1u) {
            shifts[npass] = lo_bit + W * d;
            g[npass] = plan.use9 ? gofs + swap::H8 + d * 512 : gofs + d * 256;
            ++npass;
        }
    return npass;
}

}  // namespace vd

using namespace vd;

extern "C" {

int vd_sort_col0(vd_ctx *ctx, vd_rel *rel) {
    if (!ctx || !rel) return VD_EINVAL;
    VD_TRY(ctx, sort_col0(ctx, rel));
}

int vd_sort_lex(vd_ctx *ctx, vd_rel *rel) {
    if (!ctx || !rel) return VD_EINVAL;
    VD_TRY(ctx, sort_lex(ctx, rel));
}

int vd_semi_filter(vd_ctx *ctx, const vd_rel *X, int xcol, const vd_rel *Y, int ycol,
                   uint32_t key_hi, uint64_t max_rows, vd_rel **out, uint64_t *rows) {
    if (!ctx || !out || !rows) return VD_EINVAL;
    *out = nullptr;
    *rows = 0;
    VD_TRY(ctx, {
        VD_REQUIRE(X && Y, VD_EINVAL, "null relation");
        VD_REQUIRE(xcol >= 0 && xcol < X->ncols && ycol >= 0 && ycol < Y->ncols, VD_EINVAL,
                   "semi-join filter column out of range");
        *out = semi_filter(ctx, X, xcol, Y, ycol, key_hi, max_rows, rows);
    });
}

int vd_unique_adjacent(vd_ctx *ctx, vd_rel *rel) {
    if (!ctx || !rel) return VD_EINVAL;
    VD_TRY(ctx, unique_adjacent(ctx, rel));
}

}  // extern "C"
