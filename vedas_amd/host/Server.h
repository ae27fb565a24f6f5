// Serving loop with continuous batching (no counterpart in the reference,
// whose vdQuery reads and runs one query at a time, VedasQuery.cu:136-205).
//
// Requests arrive open loop at given times.  A dispatcher keeps a FIFO of
// arrived requests; whenever one of `slots` workers is idle it admits the
// queue head and the requests behind it while their estimated peak memory
// (PlanEstimator over the store statistics) fits the budget left by the
// batches in flight, up to max_batch, and hands that batch to the worker.
// Each worker owns a context (its own HIP stream and memory pool) and runs
// its batch through the BatchExecutor, so batches of different slots
// overlap on the device; requests arriving meanwhile wait for the next free
// slot.  Plain mode runs every request alone through QueryExecutor (the
// drop-in single-query path) on one worker, as the baseline.
#ifndef VEDAS_HOST_SERVER_H
#define VEDAS_HOST_SERVER_H

#include <memory>
#include <set>
#include <string>
#include <vector>

#include "PlanTreeNode.h"
#include "SparqlResult.h"
#include "Statistics.h"
#include "VedasStorage.h"

struct ServeRequest {
    std::vector<TriplePattern> patterns;
    std::set<std::string> selected;
    PlanTreeNode *root = nullptr;
    double arrival_ms = 0;
};

struct ServeConfig {
    int slots = 4;
    int max_batch = 64;
    uint64_t mem_budget = 64ull << 30;
    bool plain = false;
    bool keep_results = false;
    // While two or more batches run, a free slot waits for min_fill queued
    // requests or for the oldest to have waited max_wait_ms (min_fill 0: no
    // waiting).
    int min_fill = 0;
    double max_wait_ms = 0.0;
};

struct ServeRecord {
    double arrival_ms = 0, start_ms = 0, done_ms = 0;
    double est_bytes = 0;
    uint64_t rows = 0;
    int batch = -1, slot = -1;
    int mem_held = 0;  // 1: waited at the head of the queue for memory
};

class Server {
  public:
    // slot_ctx: one context per worker slot (at least cfg.slots, or one in
    // plain mode); results live in these contexts, so they must outlive the
    // results.
    Server(VedasStorage *storage, const Statistics *stats, const ServeConfig &cfg,
           std::vector<VedasContext *> slot_ctx);
    // Serves all requests; recs[i] / results[i] describe request i.
    void run(std::vector<ServeRequest> &reqs, std::vector<ServeRecord> &recs,
             std::vector<SparqlResult> *results);

  private:
    VedasStorage *storage;
    const Statistics *stats;
    ServeConfig cfg;
    std::vector<VedasContext *> slot_ctx;
};

#endif
