// Internal definitions shared by the HIP translation units of libvedas_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "vedas_hip.h"

namespace vd {

// Error carried from the point of failure to the C-ABI boundary, where it
// becomes a status code plus the context's last-error message.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define VD_HIP(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess)                                                 \
            throw ::vd::Error(VD_EHIP, std::string(#call) + ": " +           \
                                           hipGetErrorString(e_));           \
    } while (0)

#define VD_REQUIRE(cond, code, msg)                                           \
    do {                                                                      \
        if (!(cond)) throw ::vd::Error((code), (msg));                        \
    } while (0)

// Body of a C-ABI entry point: maps exceptions to status codes and records
// the message on the context.
#define VD_TRY(ctx, ...)                                                      \
    try {                                                                     \
        __VA_ARGS__;                                                          \
        return VD_OK;                                                         \
    } catch (const ::vd::Error &e_) {                                         \
        if (ctx) (ctx)->last_error = e_.what();                               \
        return e_.code;                                                       \
    } catch (const std::bad_alloc &) {                                        \
        if (ctx) (ctx)->last_error = "host allocation failed";               \
        return VD_ENOMEM;                                                     \
    } catch (const std::exception &e_) {                                      \
        if (ctx) (ctx)->last_error = e_.what();                               \
        return VD_EINVAL;                                                     \
    }

struct ProfStat {
    uint64_t launches = 0;
    double ms = 0.0;
    double bytes = 0.0;
};

struct PendingLaunch {
    const char *name;
    double bytes;
    hipEvent_t start, stop;
};

}  // namespace vd

struct vd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;

    // Caching allocator: blocks are rounded to size classes and reused in
    // stream order (every kernel of the context runs on `stream`, so a block
    // freed after its last use may be handed to the next allocation at once).
    std::multimap<size_t, void *> free_blocks;
    std::unordered_map<void *, size_t> live_blocks;
    uint64_t bytes_live = 0, bytes_cached = 0;
    uint64_t bytes_peak = 0;  // highest bytes_live since the last reset
    // Guards the pool: a relation may be freed from another host thread
    // than the one issuing work on the context.
    std::recursive_mutex alloc_mu;

    // Read-back slot: RB_WORDS words plus a sequence word in coherent
    // pinned host memory that kernels write directly (d_rb is its device
    // address).  A read-back kernel stores its words, then the sequence
    // number rb_seq of the read-back; the host waits for that number.
    uint64_t *h_rb = nullptr, *d_rb = nullptr;
    uint64_t rb_seq = 0;
    // Pinned staging for descriptor uploads; reused once the previous
    // upload's event has completed.
    void *h_stage = nullptr;
    size_t h_stage_bytes = 0;
    hipEvent_t stage_done = nullptr;
    bool stage_pending = false;

    // Graph recording (vd_ctx_record, vd_ctx_capture_begin/end): a recorded
    // run logs every word the host reads back; the same run captured into
    // a HIP graph gets those words from the log instead of waiting (nothing
    // executes during a capture).
    int rb_mode = 0;  // 0: live, 1: live and logged, 2: answered from rb_log
    std::vector<uint64_t> rb_log;
    size_t rb_pos = 0;
    bool capturing = false;
    bool prof_before_capture = false;
    // Device blocks a capture allocates or frees (its graph owns them) and
    // the capture's own free list; pinned host sources of its uploads.
    std::vector<std::pair<void *, size_t>> cap_blocks;
    std::multimap<size_t, void *> cap_free;
    std::vector<void *> cap_host;

    // Profiling (of the launches whose label is in prof_only, or of all
    // launches when it is empty).
    bool prof = false;
    std::vector<std::string> prof_only;
    std::vector<vd::PendingLaunch> pending;
    std::vector<hipEvent_t> event_pool;
    std::vector<std::string> prof_names;
    std::map<std::string, vd::ProfStat> prof_stats;

    // Tunables.
    int radix_items = 16;   // keys per thread in a radix tile
    int join_nv = 2048;     // merge-path diagonal length per expansion block
    int join_rcap = 4096;   // right keys staged in LDS per bounds block
    int radix_bits = 0;     // IndexSwap digit width: 0 = by rule, 8, 9
    int radix_group = 8;    // radix tiles per XCD group: 0 = one global order
    int drive_min_rows = 1 << 20;  // smallest inputs of a driven join
};

// A relation: ncols device columns of nrows uint32 each.  A column whose bit
// is set in `borrowed` is a read-only window into storage the relation does
// not own (a range of an index permutation); `hold` keeps that storage alive.
// Operators never write into a borrowed column: they replace it.
struct vd_rel {
    int ncols = 0;
    uint64_t nrows = 0;
    uint32_t *cols[VD_MAX_COLS] = {};
    uint32_t borrowed = 0;
    std::shared_ptr<void> hold;
    // First and last value of column 0 when the operator that produced the
    // relation knew them (a join); cleared when column 0 changes.
    bool key_range = false;
    uint32_t key_lo = 0, key_hi = 0;
    // Value bounds: with bit c of `bounded` set, every value of column c is
    // <= colmax[c].  Selects take them from the index's column maxima and
    // joins from their inputs; they follow columns through swaps, drops and
    // the operators that only reorder or drop rows, and a column handed out
    // writable (vd_rel_col) loses its bound.
    uint32_t bounded = 0;
    uint32_t colmax[VD_MAX_COLS] = {};
    // The rows are pairwise distinct: selects of an index without repeated
    // triples, joins of distinct inputs (each output row is one pair of
    // input rows), and anything that only reorders or drops rows of such a
    // relation.  Dropping a column or handing one out writable clears it.
    bool distinct = false;
};

namespace vd {

// ---- allocator -------------------------------------------------------------
void *dalloc(vd_ctx *ctx, size_t bytes);
void dfree(vd_ctx *ctx, void *p);

// RAII device buffer from the context pool.
template <typename T>
struct DBuf {
    vd_ctx *ctx = nullptr;
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(vd_ctx *c, size_t count) : ctx(c), n(count) {
        p = static_cast<T *>(dalloc(c, count ? count * sizeof(T) : sizeof(T)));
    }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : ctx(o.ctx), p(o.p), n(o.n) { o.p = nullptr; }
    DBuf &operator=(DBuf &&o) noexcept {
        reset();
        ctx = o.ctx; p = o.p; n = o.n; o.p = nullptr;
        return *this;
    }
    ~DBuf() { reset(); }
    void reset() {
        if (p) dfree(ctx, p);
        p = nullptr;
    }
    T *release() { T *q = p; p = nullptr; return q; }
};

// ---- launches --------------------------------------------------------------
hipEvent_t take_event(vd_ctx *ctx);
// A launch timed by events its own dispatch recorded.
void prof_push(vd_ctx *ctx, const char *name, double bytes, hipEvent_t ev0, hipEvent_t ev1);
// After a stream sync, or (done_only) the launches whose stop event is done.
void prof_collect(vd_ctx *ctx, bool done_only);
inline bool prof_selected(const vd_ctx *ctx, const char *name) {
    if (ctx->prof_only.empty()) return true;
    for (const auto &l : ctx->prof_only)
        if (l == name) return true;
    return false;
}

// Launch `kernel` on the context stream; with profiling on, start and stop
// events are recorded on the stream around it (events without system-scope
// fences, read at the vd_prof_* calls) and it is charged `bytes`
// algorithmic HBM bytes.
template <typename... KArgs, typename... Args>
inline void launch(vd_ctx *ctx, const char *name, double bytes, dim3 grid,
                   dim3 block, size_t shmem, void (*kernel)(KArgs...),
                   Args &&...args) {
    if (grid.x == 0 || grid.y == 0 || grid.z == 0) return;
    if (ctx->prof && prof_selected(ctx, name)) {
        hipEvent_t ev0 = take_event(ctx), ev1 = take_event(ctx);
        VD_HIP(hipEventRecord(ev0, ctx->stream));
        hipLaunchKernelGGL(kernel, grid, block, shmem, ctx->stream, std::forward<Args>(args)...);
        VD_HIP(hipGetLastError());
        VD_HIP(hipEventRecord(ev1, ctx->stream));
        prof_push(ctx, name, bytes, ev0, ev1);
        return;
    }
    hipLaunchKernelGGL(kernel, grid, block, shmem, ctx->stream,
                       std::forward<Args>(args)...);
    VD_HIP(hipGetLastError());
}

void sync(vd_ctx *ctx);

// ---- small read-backs ----------------------------------------------------------
// The host learns that a read-back is complete from the sequence word the
// kernel writes into pinned memory, not from the stream's completion
// signal: the word lands a few microseconds after the store, the signal
// tens of microseconds later.  Every kernel before the read-back kernel on
// the stream has finished when its words are visible.
constexpr int RB_WORDS = 4096;
struct Readback {
    uint64_t *dst;  // RB_WORDS words, then the sequence word
    uint64_t seq;
};
Readback rb_next(vd_ctx *ctx);
// Waits for rb's sequence word, then copies n words of the slot to dst.
void rb_wait(vd_ctx *ctx, const Readback &rb, uint64_t *dst, int n);
uint64_t read_u64(vd_ctx *ctx, const uint64_t *dptr);          // syncs
// Stream-ordered host -> device copy through pinned staging (the host
// buffer may be reused as soon as this returns).
void upload(vd_ctx *ctx, void *dst, const void *src, size_t bytes);
// Device -> host copy of `bytes`, synchronising the stream.
void download(vd_ctx *ctx, void *dst, const void *src, size_t bytes);
// n <= RB_WORDS device words to the host through the read-back slot.
void read_words(vd_ctx *ctx, uint64_t *dst, const uint64_t *dptr, int n);

// ---- primitives ------------------------------------------------------------
// Two-level offsets of per-row counts, in tiles of OFF_TILE rows: loc[i] =
// exclusive scan of the counts within i's tile (32 bits: a tile's total is
// checked to fit), tpre[t] = the counts before tile t (tpre[ntiles] = the
// total); the offset of row i < n is tpre[i / OFF_TILE] + loc[i] (TileOff).
// No pass adds the two.
constexpr int OFF_TILE = 2048;
struct TileOff {
    const uint32_t *loc;
    const uint64_t *tpre;
    __device__ __forceinline__ uint64_t operator()(uint64_t i) const {
        return tpre[i / OFF_TILE] + loc[i];
    }
};
// tpre from the tile totals (one block).
void scan_tile_totals(vd_ctx *ctx, const uint64_t *tot, uint64_t ntiles, uint64_t *tpre);
// Exclusive scan of n uint32 counts into uint64 offsets; *d_total (device)
// receives the sum.  Arbitrary n.
void scan_u32_to_u64(vd_ctx *ctx, const uint32_t *in, uint64_t *out,
                     uint64_t n, uint64_t *d_total);
// Exclusive scan of n uint64 in place-capable (in may equal out).
void scan_u64(vd_ctx *ctx, const uint64_t *in, uint64_t *out, uint64_t n,
              uint64_t *d_total);
// scan_u64 with its tile totals in the caller's workspace of ws_words words
// (no allocation): n up to min(ws_words, 2048) * 2048 elements.
void scan_u64_ws(vd_ctx *ctx, const uint64_t *in, uint64_t *out, uint64_t n,
                 uint64_t *d_total, uint64_t *ws, uint64_t ws_words);
// Resolves, once per device, this GPU's address of the no-allocation select
// scratch (device `device` current); called by vd_ctx_create.
void noalloc_init(int device);
void fill_u32(vd_ctx *ctx, uint32_t *p, uint32_t v, uint64_t n);
void fill_u64(vd_ctx *ctx, uint64_t *p, uint64_t v, uint64_t n);
uint32_t max_u32(vd_ctx *ctx, const uint32_t *p, uint64_t n);  // syncs
// atomicMax of p[0, n) into *d_out (which the caller has set), no sync.
void max_u32_dev(vd_ctx *ctx, const uint32_t *p, uint64_t n, uint32_t *d_out);
void copy_d2d(vd_ctx *ctx, void *dst, const void *src, size_t bytes);

// Radix sort of 64-bit keys on bits [lo_bit, hi_bit), stable (LSD).
// keys and tmp hold n keys each; the sorted keys end up in the returned
// pointer (either keys or tmp).  If hist_ready, d_hist already holds the
// per-pass digit histograms (8-bit digits from lo_bit; npass x 256 uint32),
// counted by a fused producer kernel.
uint64_t *radix_sort_u64(vd_ctx *ctx, uint64_t *keys, uint64_t *tmp,
                         uint64_t n, int lo_bit, int hi_bit,
                         uint32_t *d_hist, bool hist_ready);
// Same with a 32-bit value per key; on return keys/vals hold the sorted
// data (the four pointers are swapped as the passes alternate buffers).
void radix_sort_u64_kv(vd_ctx *ctx, uint64_t *&keys, uint64_t *&ktmp, uint32_t *&vals,
                       uint32_t *&vtmp, uint64_t n, int lo_bit, int hi_bit,
                       uint32_t *d_hist, bool hist_ready);
// Payload columns moved with the keys by radix_sort_u64_payload.
constexpr int RADIX_MAX_PAYLOAD = 3;
struct RadixCols {
    uint32_t *p[RADIX_MAX_PAYLOAD];
};
struct RadixPayload {  // kernel argument of one pass
    const uint32_t *in[RADIX_MAX_PAYLOAD];  // null: the row id
    uint32_t *out[RADIX_MAX_PAYLOAD];
};
// Same with m <= RADIX_MAX_PAYLOAD uint32 columns: the first pass reads
// src, pass p writes pb (p even) or pa (p odd); *final gets the buffers
// holding the sorted columns (src itself if no pass ran).  src may alias pa.
// Returns the buffer holding the sorted keys (keys or tmp).
uint64_t *radix_sort_u64_payload(vd_ctx *ctx, uint64_t *keys, uint64_t *tmp, uint64_t n,
                                 int lo_bit, int hi_bit, uint32_t *d_hist, bool hist_ready,
                                 int m, const RadixCols &src, const RadixCols &pa,
                                 const RadixCols &pb, RadixCols *final);
// Keys held as two 32-bit columns, key = hi << 32 | lo (lo null: 0): with
// in_hi set, the first pass reads the keys from (in_hi, in_lo) instead of
// the key buffer; with out_hi set, the last pass writes the sorted keys'
// halves to (out_hi, out_lo) (out_lo null: the low half is dropped) instead
// of a key buffer.  The columns written must not alias those read when a
// single pass does both.
// Field move of a packed key: field c (mask[c]) goes from bit from_sh[c]
// to bit to_sh[c].
struct KeyRepack {
    int from_sh[3];
    uint32_t mask[3];
    int to_sh[3];
};
struct RadixKeyIO {
    const uint32_t *in_hi, *in_lo;
    uint32_t *out_hi, *out_lo;
    // With in_k set (and in_hi not), the first pass reads its keys from
    // in_k, each with its fields moved by rp, instead of the key buffer.
    const uint64_t *in_k;
    KeyRepack rp;
};
// Stable sort of n keys on the digits of digit_bits (8 or 9) bits at the
// given shifts (LSD order; digits not listed must be equal in every key).
// d_gofs[p] points to the 2^digit_bits exclusive digit offsets of pass p
// (device).  Payload columns as for radix_sort_u64_payload; a null src
// column is the row id (0, 1, ...).  Returns the buffer holding the sorted
// keys (keys or tmp; meaningless when kio->out_hi is set).  With kio, a
// first pass reading columns does not read `keys` and a last pass writing
// columns does not write its key buffer: 1 pass needs neither buffer, 2
// passes only tmp.
uint64_t *radix_sort_shifts(vd_ctx *ctx, uint64_t *keys, uint64_t *tmp, uint64_t n,
                            const int *shifts, int npass, const uint32_t *const *d_gofs, int m,
                            const RadixCols &src, const RadixCols &pa, const RadixCols &pb,
                            RadixCols *final, int digit_bits = 8,
                            const RadixKeyIO *kio = nullptr);
// Digit plan of a stable sort on a 32-bit value held at bits [lo_bit,
// lo_bit + 32) of its keys, from the histograms h8 / h9 (the swap::Plan of
// the value's maximum *d_max under the radix_bits tunable, counted by the
// producer with swap::hist_add): writes their exclusive offsets into gofs
// (swap::HBUF words) and reads back which digits vary.  Fills shifts[] and
// g[] (at most 4) for those digits, sets *digit_bits (8 or 9) and returns
// the number of passes (0: the value is the same in every row).
int swap_passes(vd_ctx *ctx, const uint32_t *h8, const uint32_t *h9, const uint32_t *d_max,
                uint64_t n, uint32_t *gofs, int lo_bit, int *shifts, const uint32_t **g,
                int *digit_bits);
// swap_passes for a value bound mx (the value stored at d_max): the scans
// run without a read-back and every digit of mx's plan gets a pass.
int swap_passes_bounded(vd_ctx *ctx, const uint32_t *h8, const uint32_t *h9,
                        const uint32_t *d_max, uint32_t mx, uint64_t n, uint32_t *gofs,
                        int lo_bit, int *shifts, const uint32_t **g, int *digit_bits);
// Histograms h8 / h9 (swap::HBUF words, zeroed by the caller) of the
// digits of col[0, n) under the plan of *d_max (col's maximum or an upper
// bound of it) and the radix_bits tunable, as swap_passes reads them.
void swap_hist(vd_ctx *ctx, const uint32_t *col, uint64_t n, uint32_t *h8, uint32_t *h9,
               const uint32_t *d_max);
int radix_passes(int lo_bit, int hi_bit);
// Device histogram buffer size (uint32 words) needed for radix_sort_u64.
inline size_t radix_hist_words(int lo_bit, int hi_bit) {
    return (size_t)radix_passes(lo_bit, hi_bit) * 256;
}

inline int bits_for(uint64_t maxv) {
    int b = 0;
    while (b < 64 && (maxv >> b) != 0) ++b;
    return b;
}

inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Relation helpers.
vd_rel *rel_new(vd_ctx *ctx, int ncols, uint64_t nrows);
// A relation of borrowed columns (no device allocation).
vd_rel *rel_view(int ncols, uint64_t nrows, uint32_t *const *cols, std::shared_ptr<void> hold);
void rel_free(vd_ctx *ctx, vd_rel *r);
// Replace column c by `fresh` (a pool allocation the relation now owns),
// freeing the old column if the relation owned it.
void rel_set_col(vd_ctx *ctx, vd_rel *r, int c, uint32_t *fresh);
inline bool rel_borrowed(const vd_rel *r, int c) { return (r->borrowed >> c) & 1u; }
inline void rel_set_bound(vd_rel *r, int c, uint32_t mx) {
    r->bounded |= 1u << c;
    r->colmax[c] = mx;
}
// Column d of `to` takes the bound (or its absence) of column s of `from`.
inline void rel_copy_bound(vd_rel *to, int d, const vd_rel *from, int s) {
    if ((from->bounded >> s) & 1u)
        rel_set_bound(to, d, from->colmax[s]);
    else
        to->bounded &= ~(1u << d);
}

}  // namespace vd

// ---- device helpers ----------------------------------------------------------
namespace vd {
namespace dev {

constexpr int WAVE = 64;

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & (WAVE - 1); }

// Ends a read-back: called by every lane of the one wave that wrote the
// words of rb.dst (after those stores), it makes them visible to the host,
// then stores the sequence word.
__device__ __forceinline__ void rb_publish(const Readback &rb) {
    __threadfence_system();
    if (lane_id() == 0)
        __hip_atomic_store(rb.dst + RB_WORDS, rb.seq, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_SYSTEM);
}

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    const int lane = lane_id();
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        T u = __shfl_up(v, off, WAVE);
        if (lane >= off) v += u;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_reduce(T v) {
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// Block-wide exclusive scan; returns the block total through *total.
// s_tmp must hold BLOCK/64 entries.  All threads must call it.
template <int BLOCK, typename T>
__device__ __forceinline__ T block_excl_scan(T v, T *s_tmp, T *total) {
    constexpr int NW = BLOCK / WAVE;
    const int lane = lane_id();
    const int w = threadIdx.x / WAVE;
    T inc = wave_incl_scan(v);
    if (lane == WAVE - 1) s_tmp[w] = inc;
    __syncthreads();
    T base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        T x = s_tmp[i];
        if (i < w) base += x;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return base + inc - v;
}

// tpre[t] = sum of tot[0..t), t = 0..ntiles, by one block of TT_BLOCK
// threads (s_tmp: TT_BLOCK / 64 words): a thread adds TT_PER consecutive
// totals (independent loads), then one block scan per TT_BLOCK * TT_PER
// totals.
constexpr int TT_BLOCK = 1024, TT_PER = 16;
__device__ __forceinline__ void scan_tile_totals_block(const uint64_t *__restrict__ tot,
                                                       uint64_t ntiles,
                                                       uint64_t *__restrict__ tpre,
                                                       uint64_t *s_tmp) {
    uint64_t carry = 0;
    for (uint64_t b = 0; b < ntiles; b += (uint64_t)TT_BLOCK * TT_PER) {
        const uint64_t t0 = b + (uint64_t)threadIdx.x * TT_PER;
        uint64_t v[TT_PER];
        uint64_t run = 0;
#pragma unroll
        for (int j = 0; j < TT_PER; ++j) {
            v[j] = t0 + j < ntiles ? tot[t0 + j] : 0;
            run += v[j];
        }
        uint64_t sum;
        uint64_t ex = carry + block_excl_scan<TT_BLOCK>(run, s_tmp, &sum);
#pragma unroll
        for (int j = 0; j < TT_PER; ++j) {
            if (t0 + j < ntiles) tpre[t0 + j] = ex;
            ex += v[j];
        }
        carry += sum;
    }
    if (threadIdx.x == 0) tpre[ntiles] = carry;
}

template <int BLOCK, typename T>
__device__ __forceinline__ T block_reduce(T v, T *s_tmp) {
    constexpr int NW = BLOCK / WAVE;
    T r = wave_reduce(v);
    if (lane_id() == 0) s_tmp[threadIdx.x / WAVE] = r;
    __syncthreads();
    T tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) tot += s_tmp[i];
    __syncthreads();
    return tot;
}

// Block-wide min and max of u32 values (every thread ends with both);
// s_tmp must hold BLOCK/64 words.  All threads must call it.
template <int BLOCK>
__device__ __forceinline__ void block_min_max(uint32_t &mn, uint32_t &mx, uint64_t *s_tmp) {
    constexpr int NW = BLOCK / WAVE;
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        mn = min(mn, (uint32_t)__shfl_xor(mn, off, WAVE));
        mx = max(mx, (uint32_t)__shfl_xor(mx, off, WAVE));
    }
    if (lane_id() == 0) s_tmp[threadIdx.x / WAVE] = ((uint64_t)mx << 32) | mn;
    __syncthreads();
    mn = ~0u;
    mx = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        const uint64_t w = s_tmp[i];
        mn = min(mn, (uint32_t)w);
        mx = max(mx, (uint32_t)(w >> 32));
    }
    __syncthreads();
}

// For ITEMS keys at once: lo[j] = first index in a[0, n) with a >= key[j],
// hi[j] = first index with a > key[j] (a sorted ascending).  Branchless and
// in lockstep over the keys, so the 2 * ITEMS loads of a step are
// independent and their latencies overlap (n is uniform over the block).
template <int ITEMS, typename T, typename K>
__device__ __forceinline__ void lockstep_bounds(const T *a, uint32_t n, const K (&key)[ITEMS],
                                                uint32_t (&lo)[ITEMS], uint32_t (&hi)[ITEMS]) {
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) lo[j] = hi[j] = 0;
    if (n == 0) return;
    for (uint32_t len = n; len > 1;) {
        const uint32_t half = len >> 1;
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            lo[j] += a[lo[j] + half] < key[j] ? half : 0u;
            hi[j] += key[j] < a[hi[j] + half] ? 0u : half;
        }
        len -= half;
    }
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        lo[j] += a[lo[j]] < key[j] ? 1u : 0u;
        hi[j] += key[j] < a[hi[j]] ? 0u : 1u;
    }
}

// For ITEMS values v(j) at once: k[j] = last index in a[0, n) with
// a[k] <= v(j), where a is ascending and a[0] <= every value (n >= 1).
// Branchless and in lockstep, as lockstep_bounds.
template <int ITEMS, typename T, typename F>
__device__ __forceinline__ void lockstep_last_le(const T *a, int n, F v, int (&k)[ITEMS]) {
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) k[j] = 0;
    for (int len = n; len > 1;) {
        const int half = len >> 1;
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) k[j] += a[k[j] + half] <= v(j) ? half : 0;
        len -= half;
    }
}

// Directory searches finished in global memory.  A range of m sorted keys
// was staged as every k-th key (k > 1); for ITEMS keys, t0 / t1 are the
// first staged entries >= / > the key.  The exact first position >= key
// (lo) and > key (hi), relative to the range, lie within the k - 1 keys
// before those entries: both searches of every item run branch by branch
// in lockstep, so their 2 x ITEMS dependent loads per step overlap (a hot
// key's gap holds thousands of keys).  at(q) reads key q of the range.
// Items with valid[j] false are left at lo = hi = 0.  m < 2^32.
template <int ITEMS, typename T, typename At>
__device__ __forceinline__ void lockstep_gap_bounds(At at, uint64_t k, uint64_t m,
                                                    const T (&key)[ITEMS],
                                                    const bool (&valid)[ITEMS],
                                                    const uint32_t (&t0)[ITEMS],
                                                    const uint32_t (&t1)[ITEMS],
                                                    uint32_t (&lo)[ITEMS],
                                                    uint32_t (&hi)[ITEMS]) {
    uint32_t ll[ITEMS], ul[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint64_t d0 = t0[j], d1 = t1[j];
        lo[j] = (uint32_t)(d0 ? (d0 - 1) * k + 1 : 0);
        hi[j] = (uint32_t)(d1 ? (d1 - 1) * k + 1 : 0);
        ll[j] = valid[j] ? (uint32_t)min(d0 * k, m) - lo[j] : 0u;
        ul[j] = valid[j] ? (uint32_t)min(d1 * k, m) - hi[j] : 0u;
        if (!valid[j]) lo[j] = hi[j] = 0;
    }
    // every step at least halves each remaining length (< k)
    for (uint64_t left = k; left > 0; left >>= 1) {
        T lv[ITEMS], uv[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            lv[j] = ll[j] ? at(lo[j] + (ll[j] >> 1)) : T(0);
            uv[j] = ul[j] ? at(hi[j] + (ul[j] >> 1)) : T(0);
        }
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            if (ll[j]) {
                const uint32_t h = ll[j] >> 1;
                if (lv[j] < key[j]) {
                    lo[j] += h + 1;
                    ll[j] -= h + 1;
                } else {
                    ll[j] = h;
                }
            }
            if (ul[j]) {
                const uint32_t h = ul[j] >> 1;
                if (!(key[j] < uv[j])) {
                    hi[j] += h + 1;
                    ul[j] -= h + 1;
                } else {
                    ul[j] = h;
                }
            }
        }
    }
}

// First index in [lo, hi) with a[i] >= v (a sorted ascending).
template <typename T, typename V>
__device__ __forceinline__ uint64_t lower_bound(const T *a, uint64_t lo,
                                                uint64_t hi, V v) {
    while (lo < hi) {
        uint64_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// First index in [lo, hi) with a[i] > v.
template <typename T, typename V>
__device__ __forceinline__ uint64_t upper_bound(const T *a, uint64_t lo,
                                                uint64_t hi, V v) {
    while (lo < hi) {
        uint64_t mid = lo + ((hi - lo) >> 1);
        if (v < a[mid]) hi = mid; else lo = mid + 1;
    }
    return lo;
}

// Wave-cooperative search over [lo, hi): the first index whose element is
// >= v (UPPER = false) or > v (UPPER = true).  Each round the 64 lanes probe
// the last element of 64 equal segments and a ballot picks the segment the
// bound lies in, so a search of n elements takes about log64(n) dependent
// loads instead of log2(n).  Every lane of the wave must call it with the
// same arguments; every lane gets the result.
template <bool UPPER, typename T, typename V>
__device__ __forceinline__ uint64_t wave_bound(const T *a, uint64_t lo, uint64_t hi, V v) {
    const int lane = lane_id();
    auto before = [&](uint64_t i) {  // the bound lies after index i
        const T x = a[i];
        return UPPER ? !(v < x) : (x < v);
    };
    while (hi - lo > (uint64_t)WAVE) {
        const uint64_t step = (hi - lo + WAVE - 1) / WAVE;
        const uint64_t q = lo + (uint64_t)(lane + 1) * step - 1;  // last of segment `lane`
        const bool b = q < hi && before(q);
        const int c = __popcll(__ballot(b));  // segments wholly before the bound
        const uint64_t qc = lo + (uint64_t)(c + 1) * step - 1;  // last of segment c
        hi = qc < hi ? qc + 1 : hi;
        lo = min(lo + (uint64_t)c * step, hi);  // c == 64: every element before
    }
    const uint64_t q = lo + (uint64_t)lane;
    const bool b = q < hi && before(q);
    return lo + (uint64_t)__popcll(__ballot(b));
}

// First index q in [lo, hi) with pred(q), for pred false then true over
// the range, searched by a group of G lanes (G divides 64; a group's lanes
// hold the same lo/hi and run together): each round samples G evenly
// spaced indices, so the search takes log_G of the range in dependent
// loads instead of log_2.
template <int G, typename Pred>
__device__ __forceinline__ uint64_t group_bound(uint64_t lo, uint64_t hi, Pred pred) {
    static_assert(64 % G == 0, "groups tile the wave");
    const int lane = lane_id();
    const int g = lane % G;
    const int shift = lane - g;
    const uint64_t gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
    while (hi - lo > (uint64_t)G) {
        const uint64_t step = (hi - lo + G - 1) / G;
        const uint64_t q = lo + (uint64_t)(g + 1) * step - 1;  // last of segment g
        const bool before = q < hi && !pred(q);
        const int c = __popcll((__ballot(before) >> shift) & gmask);  // segments before the bound
        const uint64_t qc = lo + (uint64_t)(c + 1) * step - 1;
        hi = qc < hi ? qc + 1 : hi;
        lo = min(lo + (uint64_t)c * step, hi);
    }
    const uint64_t q = lo + (uint64_t)g;
    const bool before = q < hi && !pred(q);
    return lo + (uint64_t)__popcll((__ballot(before) >> shift) & gmask);
}

// Adjacent-unique over a tile of ITEMS rounds of BLOCK consecutive rows
// (row = base + round * BLOCK + thread).  Keep flags: bit j = the row of
// round j exists (< n) and differs from the row before it in some column.
// Column by column, all rounds at once, so a thread's loads are
// independent; the previous row's value comes from the lane before (a wave
// holds consecutive rows), and lane 0 loads it.
// The values of columns < KC are left in kept[][] for the scatter, so it
// does not read them again (a tile's rows outlive the L2 between the two).
template <int BLOCK, int ITEMS, int KC = 0>
__device__ __forceinline__ uint32_t unique_keep_bits(const uint32_t *const *in, int ncols,
                                                     uint64_t n, uint64_t base,
                                                     uint32_t (*kept)[ITEMS] = nullptr) {
    static_assert(ITEMS <= 32, "keep flags in one word");
    const int lane = lane_id();
    uint32_t diff = 0, valid = 0;
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint64_t i = base + (uint64_t)j * BLOCK + threadIdx.x;
        valid |= (i < n ? 1u : 0u) << j;
        diff |= (i == 0 ? 1u : 0u) << j;
    }
    for (int c = 0; c < ncols; ++c) {
        const uint32_t *col = in[c];
        uint32_t v[ITEMS], pv[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const uint64_t i = base + (uint64_t)j * BLOCK + threadIdx.x;
            v[j] = i < n ? col[i] : 0u;
            pv[j] = (lane == 0 && i > 0 && i <= n) ? col[i - 1] : 0u;
        }
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const uint32_t up = __shfl_up(v[j], 1, WAVE);
            const uint32_t prev = lane == 0 ? pv[j] : up;
            diff |= (v[j] != prev ? 1u : 0u) << j;
        }
#pragma unroll
        for (int kc = 0; kc < KC; ++kc)
            if (kc == c) {
#pragma unroll
                for (int j = 0; j < ITEMS; ++j) kept[kc][j] = v[j];
            }
    }
    return diff & valid;
}

// Writes the kept rows of the tile (flags from unique_keep_bits) to out at
// o + (rows kept before them in row order).  One ballot per round gives
// each wave its mask; s_cnt (ITEMS x BLOCK/64 words of LDS) carries the
// per-(round, wave) counts across one block barrier.  All threads call it.
// Columns < KC come from vals (filled by unique_keep_bits), the others
// are read again.
template <int BLOCK, int ITEMS, int KC>
__device__ __forceinline__ void unique_scatter_tile(uint32_t keep, const uint32_t *const *in,
                                                    uint32_t *const *out, int ncols,
                                                    uint64_t base, uint64_t o,
                                                    uint32_t (*s_cnt)[BLOCK / WAVE],
                                                    const uint32_t (*vals)[ITEMS]) {
    constexpr int WAVES = BLOCK / WAVE;
    const int lane = lane_id();
    const int w = threadIdx.x / WAVE;
    uint64_t mask[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        mask[j] = __ballot((keep >> j) & 1u);
        if (lane == 0) s_cnt[j][w] = (uint32_t)__popcll(mask[j]);
    }
    __syncthreads();
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        uint32_t before = 0, round_total = 0;
#pragma unroll
        for (int ww = 0; ww < WAVES; ++ww) {
            const uint32_t c = s_cnt[j][ww];
            before += ww < w ? c : 0u;
            round_total += c;
        }
        if ((keep >> j) & 1u) {
            const uint64_t i = base + (uint64_t)j * BLOCK + threadIdx.x;
            const uint64_t pos = o + before + (uint64_t)__popcll(mask[j] & lt);
#pragma unroll
            for (int cc = 0; cc < KC; ++cc)
                if (cc < ncols) out[cc][pos] = vals[cc][j];
            for (int cc = KC; cc < ncols; ++cc) out[cc][pos] = in[cc][i];
        }
        o += round_total;
    }
}

}  // namespace dev
}  // namespace vd
