#include "Server.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>

#include "BatchExecutor.h"
#include "QueryExecutor.h"

Server::Server(VedasStorage *s, const Statistics *st, const ServeConfig &c,
               std::vector<VedasContext *> slots)
    : storage(s), stats(st), cfg(c), slot_ctx(std::move(slots)) {}

namespace {

using Clock = std::chrono::steady_clock;

struct Slot {
    std::thread th;
    std::vector<int> batch;  // request indices; empty = idle
    bool busy = false;
    int batch_id = -1;
};

}  // namespace

void Server::run(std::vector<ServeRequest> &reqs, std::vector<ServeRecord> &recs,
                 std::vector<SparqlResult> *results) {
    const int n = (int)reqs.size();
    recs.assign(n, ServeRecord());
    if (results) {
        results->clear();
        results->resize(n);
    }
    if (n == 0) return;
    const int nslots = cfg.plain ? 1 : std::max(1, cfg.slots);
    const int max_batch = cfg.plain ? 1 : std::max(1, cfg.max_batch);
    if ((int)slot_ctx.size() < nslots) throw VedasError(VD_EINVAL, "too few slot contexts");

    // Memory estimate of every request (admission).
    PlanEstimator est(*stats, cfg.plain);
    std::vector<double> need(n);
    for (int i = 0; i < n; ++i) {
        need[i] = est.estimate(reqs[i].patterns, reqs[i].root).peak_bytes;
        recs[i].arrival_ms = reqs[i].arrival_ms;
        recs[i].est_bytes = need[i];
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return reqs[a].arrival_ms < reqs[b].arrival_ms; });

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<Slot> slots(nslots);
    bool stop = false;
    double in_flight = 0;   // estimated bytes of the requests on the GPU
    int in_flight_reqs = 0; // their number (the byte sum carries rounding)
    int finished = 0;
    std::string failure;
    const auto t0 = Clock::now();
    auto now_ms = [&]() {
        return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    };

    auto worker = [&](int si) {
        VedasContext &ctx = *slot_ctx[si];
        for (;;) {
            std::vector<int> batch;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || !slots[si].batch.empty(); });
                if (slots[si].batch.empty()) return;  // stop
                batch = slots[si].batch;
            }
            const double start = now_ms();
            std::vector<SparqlResult> out;
            std::string err;
            try {
                if (cfg.plain) {
                    out.resize(1);
                    ServeRequest &r = reqs[batch[0]];
                    QueryExecutor ex(storage, &ctx);
                    ex.query(r.patterns, r.selected, r.root, out[0]);
                } else {
                    std::vector<BatchQuery> bq(batch.size());
                    for (size_t k = 0; k < batch.size(); ++k) {
                        ServeRequest &r = reqs[batch[k]];
                        bq[k].patterns = &r.patterns;
                        bq[k].selected = r.selected;
                        bq[k].root = r.root;
                    }
                    BatchExecutor bx(storage, &ctx);
                    bx.run(bq, out, nullptr);
                }
                ctx.sync();
            } catch (const std::exception &e) {
                err = e.what();
            }
            const double done = now_ms();
            {
                std::lock_guard<std::mutex> lk(mu);
                for (size_t k = 0; k < batch.size(); ++k) {
                    ServeRecord &rc = recs[batch[k]];
                    rc.start_ms = start;
                    rc.done_ms = done;
                    rc.slot = si;
                    rc.batch = slots[si].batch_id;
                    if (err.empty()) {
                        FullRelationIR *ir = out[k].getResultIR();
                        rc.rows = ir ? ir->size() : 0;
                        if (results && cfg.keep_results) (*results)[batch[k]].setResult(out[k].release());
                    }
                    in_flight -= need[batch[k]];
                }
                in_flight_reqs -= (int)batch.size();
                if (in_flight_reqs == 0) in_flight = 0;
                if (!err.empty() && failure.empty()) failure = err;
                finished += (int)batch.size();
                slots[si].batch.clear();
                slots[si].busy = false;
            }
            cv_done.notify_all();
        }
    };
    for (int si = 0; si < nslots; ++si) slots[si].th = std::thread(worker, si);

    // Dispatcher: FIFO of arrived requests, admitted by memory while a slot
    // is free.
    std::deque<int> queue;
    size_t next = 0;
    int batch_id = 0;
    {
        std::unique_lock<std::mutex> lk(mu);
        while (finished < n && failure.empty()) {
            const double t = now_ms();
            while (next < order.size() && reqs[order[next]].arrival_ms <= t) queue.push_back(order[next++]);
            bool dispatched = false, held = false;
            int busy = 0;
            for (int si = 0; si < nslots; ++si) busy += slots[si].busy ? 1 : 0;
            for (int si = 0; si < nslots && !queue.empty(); ++si) {
                if (slots[si].busy) continue;
                // Fill policy: with two batches already on the GPU (a load
                // that keeps it busy), a small queue waits to become a fuller
                // batch, for up to max_wait_ms.
                if (!cfg.plain && cfg.min_fill > 0 && busy >= 2 &&
                    (int)queue.size() < cfg.min_fill &&
                    t - reqs[queue.front()].arrival_ms < cfg.max_wait_ms) {
                    held = true;
                    break;
                }
                std::vector<int> batch;
                double add = 0;
                while (!queue.empty() && (int)batch.size() < max_batch) {
                    const int r = queue.front();
                    const bool fits = in_flight + add + need[r] <= (double)cfg.mem_budget;
                    // nothing in flight and an empty batch: admit even an
                    // oversized request alone rather than stall
                    if (!fits && !(batch.empty() && in_flight_reqs == 0)) {
                        recs[r].mem_held = 1;  // waits for memory to free up
                        break;
                    }
                    batch.push_back(r);
                    add += need[r];
                    queue.pop_front();
                }
                if (batch.empty()) break;  // head does not fit yet
                in_flight += add;
                in_flight_reqs += (int)batch.size();
                slots[si].batch = batch;
                slots[si].busy = true;
                slots[si].batch_id = batch_id++;
                dispatched = true;
                ++busy;
            }
            if (dispatched) {
                cv_work.notify_all();
                continue;
            }
            // wait for a completion, the next arrival or the end of a hold
            double wake = next < order.size() ? reqs[order[next]].arrival_ms : -1.0;
            if (held) {
                const double hold_end = reqs[queue.front()].arrival_ms + cfg.max_wait_ms;
                wake = wake < 0 ? hold_end : std::min(wake, hold_end);
            }
            if (wake >= 0) {
                const double dt = wake - now_ms();
                if (dt > 0)
                    cv_done.wait_for(lk, std::chrono::microseconds((long long)(dt * 1000.0) + 1));
            } else {
                cv_done.wait(lk);
            }
        }
        stop = true;
    }
    cv_work.notify_all();
    for (auto &s : slots) s.th.join();
    if (!failure.empty()) throw VedasError(VD_EINVAL, "serving failed: " + failure);
}
