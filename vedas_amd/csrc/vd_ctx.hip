// Context, allocator, profiler, buffers and relation handles of libvedas_hip.

#include <algorithm>
#include <cstring>

#include <chrono>

#include "vd_internal.hpp"

namespace vd {

// ---- allocator ---------------------------------------------------------------
// Size classes: powers of two up to 64 MiB, then 32 MiB multiples.  Blocks
// are kept per context and handed out again in stream order.
static size_t size_class(size_t bytes) {
    if (bytes < 256) bytes = 256;
    const size_t big = size_t(64) << 20;
    if (bytes <= big) {
        size_t c = 256;
        while (c < bytes) c <<= 1;
        return c;
    }
    const size_t step = size_t(32) << 20;
    return (bytes + step - 1) / step * step;
}

static void release_cached(vd_ctx *ctx) {
    std::lock_guard<std::recursive_mutex> lk(ctx->alloc_mu);
    if (ctx->free_blocks.empty()) return;
    VD_HIP(hipStreamSynchronize(ctx->stream));
    for (auto &kv : ctx->free_blocks) (void)hipFree(kv.second);
    ctx->free_blocks.clear();
    ctx->bytes_cached = 0;
}

void *dalloc(vd_ctx *ctx, size_t bytes) {
    std::lock_guard<std::recursive_mutex> lk(ctx->alloc_mu);
    const size_t c = size_class(bytes);
    if (ctx->capturing) {
        // a block the capture freed before: stream order within the graph
        auto ci = ctx->cap_free.find(c);
        if (ci != ctx->cap_free.end()) {
            void *q = ci->second;
            ctx->cap_free.erase(ci);
            ctx->live_blocks[q] = c;
            return q;
        }
    }
    auto it = ctx->free_blocks.find(c);
    void *p = nullptr;
    if (it != ctx->free_blocks.end()) {
        p = it->second;
        ctx->free_blocks.erase(it);
        ctx->bytes_cached -= c;
    } else {
        // the calling host thread may not have selected the context's device
        VD_HIP(hipSetDevice(ctx->device));
        hipError_t e = hipMalloc(&p, c);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            release_cached(ctx);
            e = hipMalloc(&p, c);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                throw Error(VD_ENOMEM, "device allocation of " +
                                           std::to_string(c) + " bytes failed");
            }
        }
    }
    ctx->live_blocks[p] = c;
    ctx->bytes_live += c;
    if (ctx->bytes_live > ctx->bytes_peak) ctx->bytes_peak = ctx->bytes_live;
    if (ctx->capturing) ctx->cap_blocks.emplace_back(p, c);
    return p;
}

void dfree(vd_ctx *ctx, void *p) {
    if (!p) return;
    std::lock_guard<std::recursive_mutex> lk(ctx->alloc_mu);
    auto it = ctx->live_blocks.find(p);
    if (it == ctx->live_blocks.end()) return;  // not ours: ignore
    const size_t c = it->second;
    ctx->live_blocks.erase(it);
    if (ctx->capturing) {
        // The graph may read the block at any replay: it stays out of the
        // pool.  A block the capture allocated may serve a later
        // allocation of the same capture (stream order holds in the graph).
        bool mine = false;
        for (const auto &b : ctx->cap_blocks) mine = mine || b.first == p;
        if (mine) ctx->cap_free.emplace(c, p);
        else ctx->cap_blocks.emplace_back(p, c);
        return;
    }
    ctx->bytes_live -= c;
    ctx->free_blocks.emplace(c, p);
    ctx->bytes_cached += c;
}

// ---- profiling -----------------------------------------------------------------
hipEvent_t take_event(vd_ctx *ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    // Timing events only: no system-scope release or acquire around them
    // (the host never reads device memory through them).
    hipEvent_t e;
    VD_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    return e;
}

void prof_push(vd_ctx *ctx, const char *name, double bytes, hipEvent_t ev0, hipEvent_t ev1) {
    ctx->pending.push_back({name, bytes, ev0, ev1});
}

// Timed launches whose events are read only when this many are pending, or
// at the vd_prof_* calls: reading an event pair costs host time that would
// otherwise sit between a read-back and the next launch.
constexpr size_t PROF_PENDING_MAX = 1024;

void prof_collect(vd_ctx *ctx, bool done_only) {
    size_t k = 0;
    for (; k < ctx->pending.size(); ++k) {
        auto &p = ctx->pending[k];
        if (done_only) {
            const hipError_t q = hipEventQuery(p.stop);
            if (q == hipErrorNotReady) break;
            VD_HIP(q);
        }
        float ms = 0.f;
        VD_HIP(hipEventElapsedTime(&ms, p.start, p.stop));
        auto &st = ctx->prof_stats[p.name];
        st.launches += 1;
        st.ms += ms;
        st.bytes += p.bytes;
        ctx->event_pool.push_back(p.start);
        ctx->event_pool.push_back(p.stop);
    }
    ctx->pending.erase(ctx->pending.begin(), ctx->pending.begin() + (std::ptrdiff_t)k);
}

// Most waits here are short: a size read-back right after the kernel that
// computes it.  Polling the stream for a while before the blocking wait
// keeps the host thread's wake-up latency out of each of them.
void sync(vd_ctx *ctx) {
    if (ctx->capturing) return;  // nothing runs until the graph is launched
    constexpr double kSpinMicros = 200.0;
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        const hipError_t e = hipStreamQuery(ctx->stream);
        if (e == hipSuccess) break;
        if (e != hipErrorNotReady) VD_HIP(e);
        const double us = std::chrono::duration<double, std::micro>(
                              std::chrono::steady_clock::now() - t0).count();
        if (us > kSpinMicros) {
            VD_HIP(hipStreamSynchronize(ctx->stream));
            break;
        }
    }
    // timed launches are read at vd_prof_* calls (off the launch path),
    // here only when many are pending
    if (ctx->pending.size() >= PROF_PENDING_MAX) prof_collect(ctx, false);
}

// ---- small read-backs -------------------------------------------------------------
Readback rb_next(vd_ctx *ctx) { return Readback{ctx->d_rb, ++ctx->rb_seq}; }

// Words a captured run reads back: the recorded run's, in the same order.
static void rb_replay(vd_ctx *ctx, uint64_t *dst, size_t n) {
    VD_REQUIRE(ctx->rb_pos + n <= ctx->rb_log.size(), VD_EINVAL,
               "captured run reads back more than the recorded run did");
    std::memcpy(dst, ctx->rb_log.data() + ctx->rb_pos, n * sizeof(uint64_t));
    ctx->rb_pos += n;
}

void rb_wait(vd_ctx *ctx, const Readback &rb, uint64_t *dst, int n) {
    if (ctx->rb_mode == 2) {
        rb_replay(ctx, dst, (size_t)n);
        return;
    }
    // A read-back usually follows the kernels it waits for by well under a
    // millisecond; after kSpinMicros the thread blocks on the stream, which
    // also surfaces a failed kernel (its sequence word never comes).
    constexpr double kSpinMicros = 20000.0;
    const uint64_t *word = ctx->h_rb + RB_WORDS;
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t spin = 1; __atomic_load_n(word, __ATOMIC_ACQUIRE) != rb.seq; ++spin) {
#if defined(__x86_64__)
        __builtin_ia32_pause();
#endif
        if ((spin & 1023) != 0) continue;
        const double us = std::chrono::duration<double, std::micro>(
                              std::chrono::steady_clock::now() - t0).count();
        if (us > kSpinMicros) {
            VD_HIP(hipStreamSynchronize(ctx->stream));
            VD_REQUIRE(__atomic_load_n(word, __ATOMIC_ACQUIRE) == rb.seq, VD_EHIP,
                       "read-back kernel did not publish its words");
            break;
        }
    }
    std::memcpy(dst, ctx->h_rb, (size_t)n * sizeof(uint64_t));
    if (ctx->rb_mode == 1) ctx->rb_log.insert(ctx->rb_log.end(), dst, dst + n);
    if (ctx->pending.size() >= PROF_PENDING_MAX) prof_collect(ctx, true);
}

__global__ __launch_bounds__(64) void k_rb_copy(const uint64_t *__restrict__ src, int n,
                                                Readback rb) {
    for (int i = (int)threadIdx.x; i < n; i += 64) rb.dst[i] = src[i];
    dev::rb_publish(rb);
}

void read_words(vd_ctx *ctx, uint64_t *dst, const uint64_t *dptr, int n) {
    VD_REQUIRE(n >= 0 && n <= RB_WORDS, VD_EINVAL, "read_words: too many words");
    const Readback rb = rb_next(ctx);
    launch(ctx, "read_back", 0.0, dim3(1), dim3(64), 0, k_rb_copy, dptr, n, rb);
    rb_wait(ctx, rb, dst, n);
}

uint64_t read_u64(vd_ctx *ctx, const uint64_t *dptr) {
    uint64_t v = 0;
    read_words(ctx, &v, dptr, 1);
    return v;
}

void upload(vd_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return;
    if (ctx->capturing) {
        // the graph copies from a host buffer of its own at every replay
        void *h = nullptr;
        VD_HIP(hipHostMalloc(&h, bytes, hipHostMallocDefault));
        ctx->cap_host.push_back(h);
        std::memcpy(h, src, bytes);
        VD_HIP(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, ctx->stream));
        return;
    }
    if (ctx->stage_pending) {
        VD_HIP(hipEventSynchronize(ctx->stage_done));
        ctx->stage_pending = false;
    }
    if (bytes > ctx->h_stage_bytes) {
        if (ctx->h_stage) VD_HIP(hipHostFree(ctx->h_stage));
        ctx->h_stage = nullptr;
        size_t cap = 1 << 16;
        while (cap < bytes) cap <<= 1;
        VD_HIP(hipHostMalloc(&ctx->h_stage, cap, hipHostMallocDefault));
        ctx->h_stage_bytes = cap;
    }
    if (!ctx->stage_done) VD_HIP(hipEventCreateWithFlags(&ctx->stage_done, hipEventDisableTiming));
    std::memcpy(ctx->h_stage, src, bytes);
    VD_HIP(hipMemcpyAsync(dst, ctx->h_stage, bytes, hipMemcpyHostToDevice, ctx->stream));
    VD_HIP(hipEventRecord(ctx->stage_done, ctx->stream));
    ctx->stage_pending = true;
}

void download(vd_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return;
    if (bytes % sizeof(uint64_t) == 0 && bytes <= RB_WORDS * sizeof(uint64_t) &&
        reinterpret_cast<uintptr_t>(src) % alignof(uint64_t) == 0) {
        uint64_t w[RB_WORDS];
        const int n = (int)(bytes / sizeof(uint64_t));
        read_words(ctx, w, static_cast<const uint64_t *>(src), n);
        std::memcpy(dst, w, bytes);
        return;
    }
    const size_t nw = (bytes + sizeof(uint64_t) - 1) / sizeof(uint64_t);
    if (ctx->rb_mode == 2) {
        std::vector<uint64_t> w(nw);
        rb_replay(ctx, w.data(), nw);
        std::memcpy(dst, w.data(), bytes);
        return;
    }
    VD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    if (ctx->rb_mode == 1) {
        std::vector<uint64_t> w(nw, 0);
        std::memcpy(w.data(), dst, bytes);
        ctx->rb_log.insert(ctx->rb_log.end(), w.begin(), w.end());
    }
}

void copy_d2d(vd_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return;
    VD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
}

// ---- relations -----------------------------------------------------------------
vd_rel *rel_new(vd_ctx *ctx, int ncols, uint64_t nrows) {
    VD_REQUIRE(ncols >= 1 && ncols <= VD_MAX_COLS, VD_EINVAL,
               "relation column count out of range");
    vd_rel *r = new vd_rel();
    r->ncols = ncols;
    r->nrows = nrows;
    try {
        for (int c = 0; c < ncols; ++c)
            r->cols[c] = static_cast<uint32_t *>(
                dalloc(ctx, (nrows ? nrows : 1) * sizeof(uint32_t)));
    } catch (...) {
        rel_free(ctx, r);
        throw;
    }
    return r;
}

vd_rel *rel_view(int ncols, uint64_t nrows, uint32_t *const *cols, std::shared_ptr<void> hold) {
    VD_REQUIRE(ncols >= 1 && ncols <= VD_MAX_COLS, VD_EINVAL,
               "relation column count out of range");
    vd_rel *r = new vd_rel();
    r->ncols = ncols;
    r->nrows = nrows;
    for (int c = 0; c < ncols; ++c) r->cols[c] = cols[c];
    r->borrowed = (1u << ncols) - 1u;
    r->hold = std::move(hold);
    return r;
}

void rel_free(vd_ctx *ctx, vd_rel *r) {
    if (!r) return;
    for (int c = 0; c < r->ncols; ++c)
        if (!rel_borrowed(r, c)) dfree(ctx, r->cols[c]);
    delete r;
}

void rel_set_col(vd_ctx *ctx, vd_rel *r, int c, uint32_t *fresh) {
    if (c == 0) r->key_range = false;
    if (!rel_borrowed(r, c)) dfree(ctx, r->cols[c]);
    r->cols[c] = fresh;
    r->borrowed &= ~(1u << c);
    if (!r->borrowed) r->hold.reset();
}

}  // namespace vd

// ---- C ABI ---------------------------------------------------------------------
using vd::Error;

extern "C" {

int vd_ctx_create(int device, vd_ctx **out) {
    if (!out) return VD_EINVAL;
    *out = nullptr;
    vd_ctx *ctx = new (std::nothrow) vd_ctx();
    if (!ctx) return VD_ENOMEM;
    try {
        int ndev = 0;
        VD_HIP(hipGetDeviceCount(&ndev));
        VD_REQUIRE(device >= 0 && device < ndev, VD_EINVAL, "no such device");
        ctx->device = device;
        VD_HIP(hipSetDevice(device));
        vd::noalloc_init(device);  // this GPU's address of the select scratch
        VD_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        // Coherent: kernel stores reach host memory without a cache flush
        // of their own, so the sequence word is seen as soon as it is written.
        VD_HIP(hipHostMalloc((void **)&ctx->h_rb, (vd::RB_WORDS + 1) * sizeof(uint64_t),
                             hipHostMallocCoherent | hipHostMallocMapped));
        std::memset(ctx->h_rb, 0, (vd::RB_WORDS + 1) * sizeof(uint64_t));
        VD_HIP(hipHostGetDevicePointer((void **)&ctx->d_rb, ctx->h_rb, 0));
    } catch (const Error &e) {
        std::fprintf(stderr, "vd_ctx_create: %s\n", e.what());
        int code = e.code;
        if (ctx->h_rb) (void)hipHostFree(ctx->h_rb);
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return code;
    }
    *out = ctx;
    return VD_OK;
}

int vd_ctx_destroy(vd_ctx *ctx) {
    if (!ctx) return VD_EINVAL;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &p : ctx->pending) {
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->stage_done) (void)hipEventDestroy(ctx->stage_done);
    if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
    for (auto &kv : ctx->free_blocks) (void)hipFree(kv.second);
    for (auto &kv : ctx->live_blocks) (void)hipFree(kv.first);
    if (ctx->h_rb) (void)hipHostFree(ctx->h_rb);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return VD_OK;
}

const char *vd_ctx_last_error(const vd_ctx *ctx) {
    return ctx ? ctx->last_error.c_str() : "null context";
}

// A captured run: the instantiated graph and the memory it reads and writes
// (device blocks from the context's pool, pinned host sources of copies).
struct vd_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<std::pair<void *, size_t>> blocks;
    std::vector<void *> host;
};

int vd_ctx_record(vd_ctx *ctx, int on) {
    if (!ctx) return VD_EINVAL;
    if (ctx->capturing) return VD_EINVAL;
    ctx->rb_mode = on ? 1 : 0;
    if (on) ctx->rb_log.clear();
    return VD_OK;
}

int vd_ctx_capture_begin(vd_ctx *ctx) {
    if (!ctx) return VD_EINVAL;
    VD_TRY(ctx, {
        VD_REQUIRE(!ctx->capturing, VD_EINVAL, "a capture is already open");
        vd::sync(ctx);
        VD_HIP(hipSetDevice(ctx->device));
        VD_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
        ctx->capturing = true;
        ctx->rb_mode = 2;
        ctx->rb_pos = 0;
        ctx->prof_before_capture = ctx->prof;
        ctx->prof = false;
        ctx->cap_blocks.clear();
        ctx->cap_free.clear();
        ctx->cap_host.clear();
    });
}

int vd_ctx_capture_end(vd_ctx *ctx, vd_graph **out) {
    if (!ctx) return VD_EINVAL;
    if (out) *out = nullptr;
    if (!ctx->capturing) return VD_EINVAL;
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    ctx->capturing = false;
    ctx->prof = ctx->prof_before_capture;
    const bool all_read = ctx->rb_pos == ctx->rb_log.size();
    ctx->rb_mode = 0;
    auto *g = new vd_graph();
    g->graph = graph;
    {
        std::lock_guard<std::recursive_mutex> lk(ctx->alloc_mu);
        for (const auto &b : ctx->cap_blocks) {
            ctx->live_blocks.erase(b.first);  // the graph owns it now
            ctx->bytes_live -= b.second;
        }
        g->blocks.swap(ctx->cap_blocks);
        ctx->cap_free.clear();
    }
    g->host.swap(ctx->cap_host);
    int rc = VD_OK;
    if (e != hipSuccess) {
        (void)hipGetLastError();
        ctx->last_error = std::string("stream capture failed: ") + hipGetErrorString(e);
        rc = VD_EHIP;
    } else if (!all_read) {
        ctx->last_error = "captured run read back fewer words than the recorded run";
        rc = VD_EINVAL;
    } else if (out) {
        const hipError_t ie = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            (void)hipGetLastError();
            ctx->last_error = std::string("graph instantiation failed: ") + hipGetErrorString(ie);
            rc = VD_EHIP;
        }
    }
    if (rc != VD_OK || !out) {
        vd_graph_free(ctx, g);
        return rc;
    }
    *out = g;
    return VD_OK;
}

int vd_graph_launch(vd_ctx *ctx, vd_graph *g) {
    if (!ctx || !g || !g->exec) return VD_EINVAL;
    VD_TRY(ctx, {
        VD_REQUIRE(!ctx->capturing, VD_EINVAL, "graph launch during a capture");
        VD_HIP(hipGraphLaunch(g->exec, ctx->stream));
    });
}

int vd_graph_free(vd_ctx *ctx, vd_graph *g) {
    if (!ctx || !g) return VD_EINVAL;
    if (!ctx->capturing) (void)hipStreamSynchronize(ctx->stream);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    {
        std::lock_guard<std::recursive_mutex> lk(ctx->alloc_mu);
        for (const auto &b : g->blocks) {
            ctx->free_blocks.emplace(b.second, b.first);
            ctx->bytes_cached += b.second;
        }
    }
    for (void *h : g->host) (void)hipHostFree(h);
    delete g;
    return VD_OK;
}

int vd_ctx_sync(vd_ctx *ctx) {
    if (!ctx) return VD_EINVAL;
    VD_TRY(ctx, vd::sync(ctx));
}

void *vd_ctx_stream(vd_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int vd_ctx_trim(vd_ctx *ctx) {
    if (!ctx) return VD_EINVAL;
    VD_TRY(ctx, vd::release_cached(ctx));
}

int vd_ctx_mem(vd_ctx *ctx, uint64_t *in_use, uint64_t *cached) {
    if (!ctx) return VD_EINVAL;
    if (in_use) *in_use = ctx->bytes_live;
    if (cached) *cached = ctx->bytes_cached;
    return VD_OK;
}

int vd_ctx_mem_peak(vd_ctx *ctx, uint64_t *peak, int reset) {
    if (!ctx) return VD_EINVAL;
    if (peak) *peak = ctx->bytes_peak;
    if (reset) ctx->bytes_peak = ctx->bytes_live;
    return VD_OK;
}

int vd_ctx_tune(vd_ctx *ctx, const char *key, int value) {
    if (!ctx || !key) return VD_EINVAL;
    std::string k(key);
    if (k == "radix_items") {
        if (value != 8 && value != 16 && value != 24) return VD_EINVAL;
        ctx->radix_items = value;
    } else if (k == "join_nv") {
        if (value != 1024 && value != 2048 && value != 4096) return VD_EINVAL;
        ctx->join_nv = value;
    } else if (k == "join_rcap") {
        if (value != 2048 && value != 4096 && value != 8192) return VD_EINVAL;
        ctx->join_rcap = value;
    } else if (k == "radix_bits") {
        if (value != 0 && value != 8 && value != 9) return VD_EINVAL;
        ctx->radix_bits = value;
    } else if (k == "drive_min_rows") {
        if (value < 0) return VD_EINVAL;
        ctx->drive_min_rows = value;
    } else if (k == "radix_group") {
        if (value != 0 && value != 4 && value != 8 && value != 16 && value != 32)
            return VD_EINVAL;
        ctx->radix_group = value;
    } else {
        ctx->last_error = "unknown tunable " + k;
        return VD_EINVAL;
    }
    return VD_OK;
}

int vd_dev_alloc(vd_ctx *ctx, size_t bytes, void **dptr) {
    if (!ctx || !dptr) return VD_EINVAL;
    VD_TRY(ctx, *dptr = vd::dalloc(ctx, bytes));
}

int vd_dev_free(vd_ctx *ctx, void *dptr) {
    if (!ctx) return VD_EINVAL;
    VD_TRY(ctx, vd::dfree(ctx, dptr));
}

int vd_copy_h2d(vd_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx) return VD_EINVAL;
    if (bytes == 0) return VD_OK;
    VD_TRY(ctx, {
        VD_REQUIRE(!ctx->capturing, VD_EINVAL, "host copies are not captured");
        VD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        vd::sync(ctx);
    });
}

int vd_copy_d2h(vd_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx) return VD_EINVAL;
    if (bytes == 0) return VD_OK;
    VD_TRY(ctx, {
        VD_REQUIRE(!ctx->capturing, VD_EINVAL, "host copies are not captured");
        VD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        vd::sync(ctx);
    });
}

int vd_copy_d2d(vd_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx) return VD_EINVAL;
    VD_TRY(ctx, vd::copy_d2d(ctx, dst, src, bytes));
}

int vd_rel_create(vd_ctx *ctx, int ncols, uint64_t nrows, vd_rel **out) {
    if (!ctx || !out) return VD_EINVAL;
    VD_TRY(ctx, *out = vd::rel_new(ctx, ncols, nrows));
}

int vd_rel_destroy(vd_ctx *ctx, vd_rel *rel) {
    if (!ctx) return VD_EINVAL;
    VD_TRY(ctx, vd::rel_free(ctx, rel));
}

int vd_rel_ncols(const vd_rel *rel) { return rel ? rel->ncols : 0; }
uint64_t vd_rel_nrows(const vd_rel *rel) { return rel ? rel->nrows : 0; }

uint32_t *vd_rel_col(vd_rel *rel, int c) {
    if (!rel || c < 0 || c >= rel->ncols) return nullptr;
    rel->bounded &= ~(1u << c);  // the caller may write through the pointer
    rel->distinct = false;
    return rel->cols[c];
}

int vd_rel_swap_cols(vd_rel *rel, int i, int j) {
    if (!rel || i < 0 || j < 0 || i >= rel->ncols || j >= rel->ncols)
        return VD_EINVAL;
    std::swap(rel->cols[i], rel->cols[j]);
    std::swap(rel->colmax[i], rel->colmax[j]);
    const uint32_t mi = (rel->bounded >> i) & 1u, mj = (rel->bounded >> j) & 1u;
    rel->bounded &= ~((1u << i) | (1u << j));
    rel->bounded |= (mi << j) | (mj << i);
    if (i == 0 || j == 0) rel->key_range = false;
    const uint32_t bi = (rel->borrowed >> i) & 1u, bj = (rel->borrowed >> j) & 1u;
    rel->borrowed &= ~((1u << i) | (1u << j));
    rel->borrowed |= (bi << j) | (bj << i);
    return VD_OK;
}

int vd_rel_drop_col(vd_ctx *ctx, vd_rel *rel, int i) {
    if (!ctx || !rel || i < 0 || i >= rel->ncols || rel->ncols == 1)
        return VD_EINVAL;
    if (!vd::rel_borrowed(rel, i)) vd::dfree(ctx, rel->cols[i]);
    if (i == 0) rel->key_range = false;
    rel->distinct = false;  // rows may agree on the columns left
    for (int c = i; c + 1 < rel->ncols; ++c) {
        rel->cols[c] = rel->cols[c + 1];
        rel->colmax[c] = rel->colmax[c + 1];
    }
    rel->cols[rel->ncols - 1] = nullptr;
    rel->bounded = (rel->bounded & ((1u << i) - 1u)) | ((rel->bounded >> (i + 1)) << i);
    const uint32_t low = rel->borrowed & ((1u << i) - 1u);
    rel->borrowed = low | ((rel->borrowed >> (i + 1)) << i);
    rel->ncols -= 1;
    if (!rel->borrowed) rel->hold.reset();
    return VD_OK;
}

int vd_rel_materialize(vd_ctx *ctx, vd_rel *rel) {
    if (!ctx || !rel) return VD_EINVAL;
    VD_TRY(ctx, {
        for (int c = 0; c < rel->ncols; ++c) {
            if (!vd::rel_borrowed(rel, c)) continue;
            const size_t bytes = (rel->nrows ? rel->nrows : 1) * sizeof(uint32_t);
            uint32_t *fresh = static_cast<uint32_t *>(vd::dalloc(ctx, bytes));
            if (rel->nrows) vd::copy_d2d(ctx, fresh, rel->cols[c], rel->nrows * sizeof(uint32_t));
            vd::rel_set_col(ctx, rel, c, fresh);
        }
    });
}

int vd_rel_key_range(const vd_rel *rel, uint32_t *lo, uint32_t *hi) {
    if (!rel || !rel->key_range || rel->nrows == 0) return 0;
    if (lo) *lo = rel->key_lo;
    if (hi) *hi = rel->key_hi;
    return 1;
}

int vd_rel_distinct(const vd_rel *rel) { return rel && rel->distinct ? 1 : 0; }

int vd_rel_is_view(const vd_rel *rel, int c) {
    if (!rel || c < 0 || c >= rel->ncols) return 0;
    return vd::rel_borrowed(rel, c) ? 1 : 0;
}

int vd_prof_enable(vd_ctx *ctx, int on) {
    if (!ctx) return VD_EINVAL;
    ctx->prof = on != 0;
    return VD_OK;
}

int vd_prof_only(vd_ctx *ctx, const char *labels) {
    if (!ctx) return VD_EINVAL;
    VD_TRY(ctx, {
        std::vector<std::string> v;
        std::string cur;
        for (const char *c = labels ? labels : ""; ; ++c) {
            if (*c == ',' || *c == 0) {
                if (!cur.empty()) v.push_back(cur);
                cur.clear();
                if (*c == 0) break;
            } else {
                cur += *c;
            }
        }
        ctx->prof_only = v;
    });
}

int vd_prof_reset(vd_ctx *ctx) {
    if (!ctx) return VD_EINVAL;
    VD_TRY(ctx, {
        vd::sync(ctx);
        vd::prof_collect(ctx, false);
        ctx->prof_stats.clear();
        ctx->prof_names.clear();
    });
}

int vd_prof_count(vd_ctx *ctx) {
    if (!ctx) return VD_EINVAL;
    try {
        vd::sync(ctx);
        vd::prof_collect(ctx, false);
    } catch (const Error &e) {
        ctx->last_error = e.what();
        return e.code;
    }
    ctx->prof_names.clear();
    for (auto &kv : ctx->prof_stats) ctx->prof_names.push_back(kv.first);
    return (int)ctx->prof_names.size();
}

int vd_prof_get(vd_ctx *ctx, int i, const char **name, uint64_t *launches,
                double *ms, double *bytes) {
    if (!ctx || i < 0 || i >= (int)ctx->prof_names.size()) return VD_EINVAL;
    const auto &nm = ctx->prof_names[i];
    const auto &st = ctx->prof_stats[nm];
    if (name) *name = nm.c_str();
    if (launches) *launches = st.launches;
    if (ms) *ms = st.ms;
    if (bytes) *bytes = st.bytes;
    return VD_OK;
}

}  // extern "C"
