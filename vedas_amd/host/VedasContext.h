// Per-request execution context: replaces mgpu::standard_context_t
// (VedasQuery.cu:86) and the reference's static QueryExecutor counters and
// exe_log (QueryExecutor.h:60-75): everything a query mutates lives here,
// so concurrent requests on different contexts share no state.
#ifndef VEDAS_HOST_CONTEXT_H
#define VEDAS_HOST_CONTEXT_H

#include <string>
#include <vector>

#include "vedas.h"

class VedasContext {
  public:
    explicit VedasContext(int device = 0);
    ~VedasContext();
    VedasContext(const VedasContext &) = delete;
    VedasContext &operator=(const VedasContext &) = delete;

    vd_ctx *raw() const { return ctx_; }
    int device() const { return device_; }
    // Throws VedasError when rc != VD_OK.
    void check(int rc, const char *what) const;
    void sync() const;

    // The context FullRelationIR(cols, rows) allocates on (thread local,
    // like the CUDA current device the reference relies on).
    static VedasContext *current();
    static void setCurrent(VedasContext *c);

    // Per-request statistics (the reference keeps these as statics).
    double upload_ns = 0, join_ns = 0, swap_index_ns = 0,
           eliminate_duplicate_ns = 0, eif_ns = 0;
    int eif_count = 0;
    uint64_t join_rows = 0, select_rows = 0, swap_rows = 0;
    uint64_t driven_joins = 0;  // joins that delivered their IndexSwap order
    uint64_t filtered_rows = 0;  // rows dropped before a join's deferred sort
    std::vector<ExecuteLogRecord> exe_log;
    void resetStats();

  private:
    vd_ctx *ctx_ = nullptr;
    int device_ = 0;
};

// Makes `c` current for the lifetime of the guard.
class ContextGuard {
  public:
    explicit ContextGuard(VedasContext *c) : prev_(VedasContext::current()) {
        VedasContext::setCurrent(c);
    }
    ~ContextGuard() { VedasContext::setCurrent(prev_); }
  private:
    VedasContext *prev_;
};

#endif
