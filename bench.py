#!/usr/bin/env python3
"""Benchmark of the VEDAS query path on one MI355X.

Store (synthetic, seeded): 100M distinct triples -- 10M entities with
Zipf(0.8) subjects, 256 predicates Zipf(1.1), objects 60% entities / 40%
literals -- indexed into SPO, PSO, POS and OSP, all resident in HBM.

Headline (`value`): join output rows per second of the single-query join
set run through the drop-in operator API (QueryExecutor -> Select / Join /
IndexSwap jobs -> QueryPlan), one step = the three queries once:

  star4      ?s 30 ?a . ?s 31 ?b . ?x 1 ?s . ?y 2 ?s
  chain4     ?a 1 ?b . ?b 1 ?c . ?c 1 ?d . ?d 1 ?e      (index swaps between joins)
  snowflake6 ?x 50 ?y . ?x 51 ?u . ?w 1 ?x . ?y 30 ?v . ?y 31 ?t . ?z 1 ?y

Also reported: the index build (4 permutations), 10k range selects, a
256-query mixed batch run one query at a time vs batched, and an open-loop
Poisson arrival stream served plain vs with continuous batching (p50/p99).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from vedas_amd import workload as W  # noqa: E402
from vedas_amd.workload import JOIN_SET, STORE_100M, STORE_500M, query_mix  # noqa: E402

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
GEN = {k: v for k, v in STORE_100M.items() if k != "n"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


PERM_COLS = {"SPO": (0, 1, 2), "SOP": (0, 2, 1), "PSO": (1, 0, 2), "POS": (1, 2, 0),
             "OSP": (2, 0, 1), "OPS": (2, 1, 0)}


def build_order(perms):
    """(permutation, sibling it is derived from or None) in build order, as
    the index build plans it: a permutation XYZ whose Y comes before its Z in
    a built sibling's key is derived from that sibling (a stable sort on X),
    the others are sorted on all bits (vd_index.hip build_packed)."""
    order, todo = [], [q for q in PERM_COLS if q in perms]
    while todo:
        pick = None
        for q in todo:
            for b, _ in order:
                pos = {c: k for k, c in enumerate(PERM_COLS[b])}
                if pos[PERM_COLS[q][1]] < pos[PERM_COLS[q][2]]:
                    pick = (q, b)
                    break
            if pick:
                break
        pick = pick or (todo[0], None)
        order.append(pick)
        todo.remove(pick[0])
    return order


def x_passes(xbits: int) -> int:
    """Passes of a stable sort on an X of xbits bits (radix_bits 0: 9-bit
    digits when they need fewer passes than bytes, up to 27 bits)."""
    n8, n9 = -(-xbits // 8), -(-xbits // 9)
    return n9 if xbits <= 27 and n9 < n8 else n8


def index_bytes_full_sort(n: int, kbits: int, nperms: int, rec: int = 8) -> float:
    """The scope table's unit for the index build, per permutation of n
    records of `rec` bytes (the packed 64-bit key) sorted on a K-bit key in
    8-bit digits: ceil(K/8) scatter passes reading and writing every record
    plus one histogram read, B = ceil(K/8)*2*n*rec + n*rec.  It prices every
    permutation as a full sort, whichever way it is built."""
    return nperms * (-(-kbits // 8) * 2 * n * rec + n * rec)


def index_bytes(n: int, bits, perms) -> float:
    """Algorithmic HBM bytes of the index build when the triple fits one
    64-bit key (bits = bit lengths of s, p, o).  A permutation sorted on all
    K bits packs the triple (12 B read, 8 B written) and runs ceil(K/8)
    passes (8 B read, 8 B written); a derived one runs the passes of its X
    bits, the first reading its sibling's sorted keys, after one histogram
    read (4 B) of each input column that is a derived X.  Every permutation
    writes Y and Z from its sorted keys (8 B + 8 B) and reads the keys once
    more for its L1/L2 heads (8 B; the head entries written are not
    counted)."""
    kbits = sum(bits)
    assert kbits <= 64
    total = 0.0
    xcols = set()
    for q, src in build_order(perms):
        if src is None:
            total += 20 * n + -(-kbits // 8) * 16 * n
        else:
            xcols.add(PERM_COLS[q][0])
            total += x_passes(bits[PERM_COLS[q][0]]) * 16 * n
        total += 24 * n
    return total + 4 * n * len(xcols)


_M1 = np.uint64(0xFF51AFD7ED558CCD)
_M2 = np.uint64(0xC4CEB9FE1A85EC53)


def _mix(h):
    h ^= h >> np.uint64(33)
    h *= _M1
    h ^= h >> np.uint64(33)
    h *= _M2
    h ^= h >> np.uint64(33)
    return h


def result_digest(headers, cols, chunk=1 << 22):
    """Order-independent digest of a result's rows (columns ordered by
    variable name): (names, rows, sum of a 64-bit row hash).  Used to check
    that two runs gave the same bag of rows."""
    order = sorted(range(len(headers)), key=lambda i: headers[i])
    names = tuple(headers[i] for i in order)
    n = len(cols[0]) if cols else 0
    total = 0
    with np.errstate(over="ignore"):
        for a in range(0, n, chunk):
            h = np.full(min(chunk, n - a), 0x243F6A8885A308D3, np.uint64)
            for c, i in enumerate(order):
                h = _mix(h ^ (cols[i][a:a + chunk].astype(np.uint64) +
                              np.uint64((0x9E3779B97F4A7C15 * (c + 1)) & 0xFFFFFFFFFFFFFFFF)))
            total = (total + int(h.sum(dtype=np.uint64))) & 0xFFFFFFFFFFFFFFFF
    return names, n, total


def kernel_table(prof):
    return {k: {"launches": v["launches"], "ms": round(v["ms"], 3),
                "GB/s": round(v["bytes"] / (v["ms"] / 1e3) / 1e9, 1) if v["ms"] > 0 else None}
            for k, v in sorted(prof.items(), key=lambda kv: -kv[1]["ms"])}


def select_bench(ctx, idx, s, p, o, nq=W.N_SELECTS, seed=W.SEED_SELECTS,
                 max_chunk_rows=400_000_000):
    """10k range selects with one or two constants taken from uniformly drawn
    existing triples: S -> SPO, P -> PSO, O -> OSP, (S,P) -> SPO, (P,O) -> POS,
    (O,S) -> OSP.  Each permutation's queries run as batched selects whose
    output (Y, Z) rows are materialised in HBM, in chunks of at most
    max_chunk_rows rows."""
    import ctypes as C
    from vedas_amd import _native
    groups = W.select_mix(nq, seed, s, p, o)
    plans_ = []
    for perm, (k1, k2) in groups.items():
        d1, d2 = ctx.upload(k1), ctx.upload(k2)
        db, de = ctx.alloc(8 * len(k1)), ctx.alloc(8 * len(k1))
        ctx.check(ctx.L.vd_index_lookup_batch(ctx.h, idx.h, _native.PERMS[perm], C.c_void_p(d1),
                                              C.c_void_p(d2), len(k1), C.c_void_p(db),
                                              C.c_void_p(de)))
        sz = ctx.download(de, len(k1), np.uint64) - ctx.download(db, len(k1), np.uint64)
        for d in (d1, d2, db, de):
            ctx.free(d)
        start = 0
        while start < len(k1):
            acc, end = 0, start
            while end < len(k1) and (acc + int(sz[end]) <= max_chunk_rows or end == start):
                acc += int(sz[end])
                end += 1
            plans_.append((perm, k1[start:end].copy(), k2[start:end].copy(), acc))
            start = end
    cap = max(max(a for *_, a in plans_), 1)
    dy, dz = ctx.alloc(4 * cap), ctx.alloc(4 * cap)
    staged = [(perm, ctx.upload(a), ctx.upload(b), len(a), ctx.alloc(8 * (len(a) + 1)), rows)
              for perm, a, b, rows in plans_]
    ctx.sync()
    total = C.c_uint64()
    rows = 0
    t0 = time.perf_counter()
    for perm, d1, d2, m, doff, r in staged:
        ctx.check(ctx.L.vd_select_batch(ctx.h, idx.h, _native.PERMS[perm], C.c_void_p(d1),
                                        C.c_void_p(d2), m, C.c_void_p(doff), C.c_void_p(dy),
                                        C.c_void_p(dz), cap, C.byref(total)))
        assert total.value == r
        rows += total.value
    ctx.sync()
    dt = time.perf_counter() - t0
    for _, d1, d2, _, doff, _ in staged:
        for d in (d1, d2, doff):
            ctx.free(d)
    ctx.free(dy)
    ctx.free(dz)
    return {"selects": nq, "ms": dt * 1e3, "selects_per_s": nq / dt, "rows": rows,
            "rows_per_s": rows / dt, "GB/s": 16.0 * rows / dt / 1e9, "chunks": len(staged)}


def batch_bench(eng, texts, reps=2, budget=96 << 30):
    """The same query mix one query at a time (plain drop-in path) and in
    batches (as many queries per batch as the memory estimate allows);
    results compared canonically."""
    groups = W.chunk_by_memory(eng, texts, budget)
    plain_t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        for t in texts:
            eng.query(t).free()
        plain_t.append(time.perf_counter() - t0)
    batch_t = []
    rounds = calls = join_rows = 0
    from vedas_amd import _native
    ctx = _native.Ctx.wrap(eng.ctx_handle())
    prof = None
    for rep in range(reps + 1):
        if rep == reps:  # one more run with per-kernel timing, not counted
            ctx.prof(True)
            ctx.prof_reset()
        t0 = time.perf_counter()
        rounds = calls = join_rows = 0
        for g in groups:
            res, st = eng.batch(g)
            rounds += st["rounds"]
            calls += st["calls"]
            join_rows += st["join_rows"]
            for r in res:
                r.free()
        if rep < reps:
            batch_t.append(time.perf_counter() - t0)
        else:
            prof = kernel_table(ctx.prof_stats())
            ctx.prof(False)
    # The same batches captured as HIP graphs (vdq_graph_create) and
    # replayed: every kernel of each batch again, no host work between them.
    graphs = [eng.graph(g) for g in groups]
    graph_t = []
    try:
        for _ in range(reps + 1):
            t0 = time.perf_counter()
            for gr in graphs:
                gr.replay()
            graph_t.append(time.perf_counter() - t0)
        # equality of results, checked outside the timed runs: every query
        # alone, batched and from the graph's last replay
        same = same_graph = 0
        for g, gr in zip(groups, graphs):
            res, _ = eng.batch(g)
            for i, (t, r) in enumerate(zip(g, res)):
                a = eng.query(t)
                d = result_digest(r.headers(), r.columns())
                same += result_digest(a.headers(), a.columns()) == d
                gres = gr.result(i)
                same_graph += result_digest(gres.headers(), gres.columns()) == d
                a.free()
                r.free()
    finally:
        for gr in graphs:
            gr.free()
    pt, bt, gt = min(plain_t), min(batch_t), min(graph_t[1:])
    return {"queries": len(texts), "batches": len(groups), "plain_ms": pt * 1e3,
            "batched_ms": bt * 1e3, "plain_qps": len(texts) / pt, "batched_qps": len(texts) / bt,
            "speedup": pt / bt, "graph_ms": gt * 1e3, "graph_qps": len(texts) / gt,
            "graph_speedup_vs_batched": bt / gt, "graph_speedup_vs_plain": pt / gt,
            "graph_note": "each batch captured once as a HIP graph (two runs: one logs the "
                          "read-back sizes, one is captured), then replayed: the same kernels "
                          "over the same store, no host work between them",
            "rounds": rounds, "calls": calls, "join_rows": join_rows,
            "results_equal": same, "graph_results_equal": same_graph, "kernels": prof}


def serve_bench(eng, texts, qps, seed=1, slots=4, max_batch=128, budget=96 << 30,
                min_fill=48, max_wait_ms=1.0):
    """Open-loop Poisson arrivals at `qps` over the mix, served plain (one
    query at a time) and with continuous batching; p50/p99 latency.  Four
    slots of up to 128 requests (a sweep of 2-8 slots and batches of 32-256
    on the 500M store gives the same throughput within 3%), and the fill
    policy: while two batches run, a free slot waits for 48 queued requests
    or 1 ms (1-3% more throughput at saturation, p50 at 2k q/s 0.44 ->
    0.46 ms in a sweep)."""
    rng = np.random.default_rng(seed)
    arr = list(np.cumsum(rng.exponential(1000.0 / qps, len(texts))))

    def summarise(recs):
        lat = np.array([r["done_ms"] - r["arrival_ms"] for r in recs])
        span = max(r["done_ms"] for r in recs) - min(r["arrival_ms"] for r in recs)
        return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "qps": len(recs) / (span / 1e3), "batches": len({r["batch"] for r in recs}),
                "mem_held": int(sum(r["mem_held"] for r in recs)),
                "est_bytes_max": float(max(r["est_bytes"] for r in recs))}

    # Each mode runs the stream three times and keeps the median run: the
    # served stream is short (tens of ms at saturation) and the first run of
    # a mode pays its warm-up (the slot contexts' pools, for example).
    def median_run(**kw):
        runs = []
        for k in range(3):
            runs.append(summarise(eng.serve(texts, arr, **kw)[0]))
            log(f"[bench]     {'plain' if kw.get('plain') else 'served'} run {k}: "
                f"{runs[-1]['qps']:.0f} q/s")
        runs.sort(key=lambda r: r["qps"])
        return runs[1], [r["qps"] for r in runs]
    plain, plain_runs = median_run(plain=True)
    served, served_runs = median_run(slots=slots, max_batch=max_batch, mem_budget=budget,
                                     min_fill=min_fill, max_wait_ms=max_wait_ms)
    out = {"offered_qps": qps, "requests": len(texts), "slots": slots, "max_batch": max_batch,
           "min_fill": min_fill, "max_wait_ms": max_wait_ms,
           "mem_budget": budget, "plain": plain, "served": served,
           "plain_qps_runs": plain_runs, "served_qps_runs": served_runs}
    out["qps_speedup"] = out["served"]["qps"] / out["plain"]["qps"]
    return out


def serve_section(store: dict, loads, n_queries: int, seed: int, mix_seed: int):
    """The served form: a fresh engine over its own store (distinct triples,
    all permutations resident), the seeded query mix drawn on that store,
    served at each offered load; then every request's served result
    compared (bag digest) with the same query run alone."""
    from vedas_amd import _native
    from vedas_amd.engine import Engine, query_text
    prm = {k: v for k, v in store.items() if k != "n"}
    log(f"[bench] serving store: generating {store['n']} triples")
    s, p, o = _native.gen_triples(store["n"], seed=seed, distinct=True, **prm)
    eng = Engine(0)
    ctx = _native.Ctx.wrap(eng.ctx_handle())
    ds, dp, do = ctx.upload(s), ctx.upload(p), ctx.upload(o)
    t0 = time.perf_counter()
    eng.load_device(ds, dp, do, len(s), ("SPO", "PSO", "POS", "OSP"))
    ctx.sync()
    build_ms = (time.perf_counter() - t0) * 1e3
    for d in (ds, dp, do):
        ctx.free(d)
    kbits = int(s.max()).bit_length() + int(p.max()).bit_length() + int(o.max()).bit_length()
    log(f"[bench] serving store built in {build_ms:.0f} ms ({kbits}-bit keys)")
    mix = query_mix(n_queries, mix_seed, store, s, p, o)
    del s, p, o
    texts = [query_text(pats, node) for _, pats, node in mix]
    for r in eng.batch(texts[:16])[0]:  # warm-up
        r.free()
    loads_out = []
    for q in loads:
        log(f"[bench] serving at {q:.0f} offered q/s")
        sv = serve_bench(eng, texts, q)
        loads_out.append(sv)
        log(f"[bench]   plain {sv['plain']['qps']:.0f} q/s p50 {sv['plain']['p50_ms']:.1f} "
            f"p99 {sv['plain']['p99_ms']:.1f} ms; served {sv['served']['qps']:.0f} q/s "
            f"p50 {sv['served']['p50_ms']:.1f} p99 {sv['served']['p99_ms']:.1f} ms")
    # the highest load again under intermediate-memory budgets that bind:
    # admission holds requests until the estimates of those in flight leave
    # room (median of three runs each)
    budget_runs = []
    rng = np.random.default_rng(seed + 3)
    arr = list(np.cumsum(rng.exponential(1000.0 / max(loads), len(texts))))
    for budget in (4 << 30, 1 << 30):
        runs = []
        for _ in range(3):
            recs, _ = eng.serve(texts, arr, slots=4, max_batch=128, mem_budget=budget,
                                min_fill=48, max_wait_ms=1.0)
            lat = np.array([r["done_ms"] - r["arrival_ms"] for r in recs])
            span = max(r["done_ms"] for r in recs) - min(r["arrival_ms"] for r in recs)
            runs.append({"qps": len(recs) / (span / 1e3), "p50_ms": float(np.percentile(lat, 50)),
                         "p99_ms": float(np.percentile(lat, 99)),
                         "mem_held": int(sum(r["mem_held"] for r in recs)),
                         "batches": len({r["batch"] for r in recs})})
        runs.sort(key=lambda r: r["qps"])
        budget_runs.append({"mem_budget": budget, "offered_qps": max(loads), **runs[1]})
        log(f"[bench]   budget {budget >> 20} MiB: {runs[1]['qps']:.0f} q/s p50 "
            f"{runs[1]['p50_ms']:.1f} p99 {runs[1]['p99_ms']:.1f} ms, "
            f"{runs[1]['mem_held']} requests held for memory")
    # same result alone and in the served mix
    rng = np.random.default_rng(seed + 2)
    arr = list(np.cumsum(rng.exponential(1000.0 / max(loads), len(texts))))
    _, outs = eng.serve(texts, arr, keep_results=True)
    same = 0
    for i in range(len(texts)):
        a = eng.query(texts[i])
        same += result_digest(a.headers(), a.columns()) == result_digest(outs[i].headers(),
                                                                          outs[i].columns())
        a.free()
    for r in outs:
        r.free()
    eng.close()
    return {"store": {**store, "seed": seed, "distinct": True, "key_bits": kbits},
            "build_ms": build_ms, "queries": len(texts), "loads": loads_out,
            "budget_runs": budget_runs,
            "results_equal_alone": f"{same}/{len(texts)}"}


def host_info() -> dict:
    """Cores, CPU model and NUMA layout of the host the baseline runs on."""
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    try:
        with open("/proc/cpuinfo") as f:
            models = [ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")]
        info["cpu_model"] = models[0] if models else None
    except OSError:
        info["cpu_model"] = None
    node_dir = "/sys/devices/system/node"
    try:
        nodes = sorted(d for d in os.listdir(node_dir) if d.startswith("node") and d[4:].isdigit())
        info["numa_nodes"] = len(nodes)
        info["numa_cpus"] = {d: open(os.path.join(node_dir, d, "cpulist")).read().strip()
                             for d in nodes}
    except OSError:
        info["numa_nodes"] = None
    info["OMP_NUM_THREADS"] = os.environ.get("OMP_NUM_THREADS")
    return info


def cpu_baseline(idx, names, gpu_results: dict, runs: int = 5):
    """The join set on the host cores: the multi-threaded CPU port of the
    operators (select with the bound clip, join, stable re-sort, adjacent
    unique; bit-equal to the oracle by tests/test_cpu_port.py) walking the
    same plans with the reference's join-variable bounds, over the PSO and
    POS arrays of the same store (downloaded from the device build: the
    index build is not part of the timed step).  One warm-up run, whose
    results are checked against the GPU's (bag digest, join rows, final
    rows), then `runs` timed runs of the whole join set; the value is from
    the median run, the minimum is reported beside it (vdBench reports the
    minimum, VedasBench.cu:182-189)."""
    from oracle import cpu_port as P
    ops = P.make_ops()
    ix = {perm: idx.arrays(perm) for perm in ("PSO", "POS")}
    parity, rows = {}, 0
    for q in names:
        pats, node = JOIN_SET[q]
        st = {}
        t0 = time.perf_counter()
        h, cols = P.run_query(ix, pats, node, ops, st)
        dt = time.perf_counter() - t0
        g = gpu_results[q]
        d = result_digest(h, cols)
        parity[q] = {"digest_equal": d == g["digest"], "rows": d[1], "gpu_rows": g["rows"],
                     "join_rows": st["join_rows"], "gpu_join_rows": g["join_rows"],
                     "select_rows": st["select_rows"], "gpu_select_rows": g["select_rows"],
                     "join_rows_equal": st["join_rows"] == g["join_rows"]}
        rows += st["join_rows"]
        del cols
        log(f"[bench] cpu baseline warm-up: {q} {dt:.2f} s, equal to the GPU's: "
            f"{parity[q]['digest_equal'] and parity[q]['join_rows_equal']}")
    times = []
    for k in range(runs):
        t0 = time.perf_counter()
        for q in names:
            pats, node = JOIN_SET[q]
            P.run_query(ix, pats, node, ops, {})
        times.append(time.perf_counter() - t0)
        log(f"[bench] cpu baseline run {k}: {times[-1]:.2f} s")
    med, mn = float(np.median(times)), min(times)
    return {"value": rows / med, "unit": "join rows/s", "cores": ops.threads, "kind": "port",
            "sample": f"the join set ({', '.join(names)}; {rows} join rows) on the same "
                      f"100M-triple store by the multi-threaded CPU port on {ops.threads} "
                      f"threads; median of {runs} runs after one warm-up",
            "runs_s": times, "median_s": med, "min_s": mn, "value_at_min": rows / mn,
            "host": host_info(),
            "parity_vs_gpu": parity,
            "all_equal": all(v["digest_equal"] and v["join_rows_equal"] for v in parity.values())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--triples", type=int, default=STORE_100M["n"])
    ap.add_argument("--seed", type=int, default=W.SEED_100M)
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed runs of the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default kernel tunables")
    ap.add_argument("--queries", default="star4,chain4,snowflake6")
    ap.add_argument("--selects", type=int, default=W.N_SELECTS)
    ap.add_argument("--mix", type=int, default=W.N_MIX, help="queries in the batched mix (0: skip)")
    ap.add_argument("--serve-qps", type=float, nargs="*", default=[2000.0, 50000.0],
                    help="offered loads of the open-loop arrival stream (none: skip)")
    ap.add_argument("--serve-requests", type=int, default=W.N_SERVE,
                    help="requests in the served arrival stream")
    ap.add_argument("--serve-store", choices=["500M", "100M"], default="500M",
                    help="store of the served form")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the query path does not shard across GPUs.", file=sys.stderr)
        sys.exit(2)

    from vedas_amd import _native
    from vedas_amd.engine import Engine, get_option, query_text, set_option

    log(f"[bench] generating {args.triples} distinct triples")
    t = time.perf_counter()
    s, p, o = _native.gen_triples(args.triples, seed=args.seed, distinct=True, **GEN)
    log(f"[bench] generated in {time.perf_counter() - t:.1f}s")

    eng = Engine(0)
    ctx = _native.Ctx.wrap(eng.ctx_handle())
    ds, dp, do = ctx.upload(s), ctx.upload(p), ctx.upload(o)
    n = len(s)
    perms = ("SPO", "PSO", "POS", "OSP")
    log("[bench] index build")
    # Warm-up build, a build with every launch timed (the per-kernel table),
    # then the timed build without events (each timed launch adds two to the
    # stream).
    eng.load_device(ds, dp, do, n, perms)
    ctx.sync()
    ctx.prof(True)
    ctx.prof_reset()
    eng.load_device(ds, dp, do, n, perms)
    ctx.sync()
    build_prof = ctx.prof_stats()
    ctx.prof(False)
    t0 = time.perf_counter()
    eng.load_device(ds, dp, do, n, perms)
    ctx.sync()
    build_s = time.perf_counter() - t0
    for d in (ds, dp, do):
        ctx.free(d)
    bits = (int(s.max()).bit_length(), int(p.max()).bit_length(), int(o.max()).bit_length())
    kb = {name: sum(bits) for name in perms}
    build_alg = index_bytes(n, bits, perms)
    full_b = index_bytes_full_sort(n, sum(bits), len(perms))
    log(f"[bench] index build {build_s * 1e3:.1f} ms")

    idx_view = _native.Index.wrap(ctx, eng.index_handle(), n)
    sel = None
    if args.selects > 0:
        log("[bench] range selects")
        select_bench(ctx, idx_view, s, p, o, nq=min(args.selects, 500))  # warm-up
        sel = select_bench(ctx, idx_view, s, p, o, nq=args.selects)
        log(f"[bench] selects: {sel['selects_per_s']:.0f}/s, {sel['GB/s']:.0f} GB/s")

    names = [q for q in args.queries.split(",") if q]
    texts = {q: query_text(JOIN_SET[q][0], JOIN_SET[q][1]) for q in names}

    def step(collect=None):
        rows = 0
        for q in names:
            r = eng.query(texts[q])
            st = r.stats()
            r.free()
            rows += st["join_rows"]
            if collect is not None:
                collect.setdefault(q, []).append(st)
        return rows

    tune = None
    if not args.no_tune:
        log("[bench] autotuning kernel tunables on the join set")
        tune = eng.autotune([texts[q] for q in names], reps=2)
        log(f"[bench] autotune: radix_items {tune['radix_items']} join_nv {tune['join_nv']} "
            f"join_rcap {tune['join_rcap']} radix_bits {tune['radix_bits']} "
            f"radix_group {tune['radix_group']} "
            f"({tune['base_ms']:.2f} -> {tune['tuned_ms']:.2f} ms)")
    for w in range(args.warmup):
        log(f"[bench] warmup {w}")
        step()
    # One more untimed step with every launch timed, query by query: the
    # per-kernel table of a step, each query's algorithmic bytes, and the
    # dominant kernel.  The timed steps then time only the dominant kernel's
    # launches with HIP events (each timed launch adds two events to the
    # stream, about 1 ms per step when every launch is timed).
    ctx.sync()
    ctx.prof(True)
    step_prof, q_bytes = {}, {}
    for q in names:
        ctx.prof_reset()
        r = eng.query(texts[q])
        r.free()
        ctx.sync()
        qp = ctx.prof_stats()
        q_bytes[q] = sum(v["bytes"] for v in qp.values())
        for k, v in qp.items():
            acc = step_prof.setdefault(k, {"launches": 0, "ms": 0.0, "bytes": 0.0})
            for f in ("launches", "ms", "bytes"):
                acc[f] += v[f]
    dom_name = max(step_prof.items(), key=lambda kv: kv[1]["ms"])[0]
    ctx.prof_only([dom_name])
    ctx.prof_reset()
    per_q = {}
    total_rows = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        total_rows += step(per_q)
        log(f"[bench] step {k} done")
    ctx.sync()
    dt = time.perf_counter() - t0
    prof = ctx.prof_stats()
    ctx.prof(False)
    ctx.prof_only(None)

    # dominant kernel by device time, timed over the timed steps
    dom = prof[dom_name]
    achieved = dom["bytes"] / (dom["ms"] / 1e3) if dom["ms"] > 0 else 0.0
    alg_per_launch = dom["bytes"] / max(dom["launches"], 1)
    roofline = {"bound": "hbm", "kernel": dom_name, "achieved": achieved / 1e9,
                "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                "traffic": None, "launches": dom["launches"],
                "avg_launch_ms": dom["ms"] / max(dom["launches"], 1),
                "alg_bytes_per_launch": alg_per_launch}
    # HBM bytes per launch of the same kernel on the same workload, measured
    # by two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) and kept as text.
    traffic_file = os.path.join(ROOT, "profiles", "hbm_traffic_joinset_100M.md")
    if os.path.exists(traffic_file) and n == STORE_100M["n"]:
        from vedas_amd.prof_summary import read_traffic
        measured = read_traffic(traffic_file).get(dom_name)
        if measured:
            roofline["traffic"] = measured
            roofline["traffic_over_alg"] = measured / alg_per_launch if alg_per_launch else None
            roofline["traffic_source"] = os.path.relpath(traffic_file, ROOT)

    # Outside the timed steps: each query's result digest (checked against
    # the CPU baseline's), its intermediates without the bound pre-scan,
    # and the same query planned by the query graph planner (no PLAN).
    gpu_results, planned, prescan = {}, {}, {}
    for q in names:
        r = eng.query(texts[q])
        st = r.stats()
        gpu_results[q] = {"digest": result_digest(r.headers(), r.columns()), "rows": st["rows"],
                          "join_rows": st["join_rows"], "select_rows": st["select_rows"]}
        r.free()
        set_option("bound_prescan", 0)
        try:
            r = eng.query(texts[q])
            off = r.stats()
            off_digest = result_digest(r.headers(), r.columns())
            r.free()
        finally:
            set_option("bound_prescan", 1)
        prescan[q] = {"select_rows": st["select_rows"], "join_rows": st["join_rows"],
                      "select_rows_without": off["select_rows"],
                      "join_rows_without": off["join_rows"], "ms_without": off["elapsed_ms"],
                      "same_result": off_digest == gpu_results[q]["digest"]}
        pq = query_text(JOIN_SET[q][0], None)
        r = eng.query(pq)  # warm
        r.free()
        r = eng.query(pq)
        pst = r.stats()
        planned[q] = {"ms": pst["elapsed_ms"], "rows": pst["rows"], "join_rows": pst["join_rows"],
                      "rows_given_plan": st["rows"],
                      "same_result": result_digest(r.headers(), r.columns()) ==
                      gpu_results[q]["digest"]}
        r.free()
    log("[bench] planner: " + ", ".join(f"{q} {v['ms']:.1f} ms" for q, v in planned.items()))

    # The same steps with the plan rewrites off, one at a time and both
    # (every IndexSwap then re-sorts its whole input, the reference's order
    # of operations): step time, rows re-sorted, identical results.
    def steps_with(opts):
        old = {k: get_option(k) for k in opts}
        for k, v in opts.items():
            set_option(k, v)
        try:
            step()
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(3):
                step()
            ctx.sync()
            ms = (time.perf_counter() - t0) / 3 * 1e3
            qs = {}
            for q in names:
                r = eng.query(texts[q])
                st0 = r.stats()
                qs[q] = {"swap_rows": st0["swap_rows"], "join_rows": st0["join_rows"],
                         "filtered_rows": st0["filtered_rows"],
                         "same_result": result_digest(r.headers(), r.columns()) ==
                         gpu_results[q]["digest"]}
                r.free()
        finally:
            for k, v in old.items():
                set_option(k, v)
        return ms, qs

    def compare(ms_without, qs):
        return {"ms_per_step_with": dt / args.steps * 1e3, "ms_per_step_without": ms_without,
                "queries": {q: {"driven_joins": per_q[q][0]["driven_joins"],
                                "filtered_rows_with": per_q[q][0]["filtered_rows"],
                                "swap_rows_with": per_q[q][0]["swap_rows"],
                                "swap_rows_without": qs[q]["swap_rows"],
                                "same_join_rows": qs[q]["join_rows"] == per_q[q][0]["join_rows"],
                                "same_result": qs[q]["same_result"]} for q in names}}

    driven = compare(*steps_with({"driven_swap": 0}))
    filtered = compare(*steps_with({"filtered_swap": 0}))
    ref_ms, ref_q = steps_with({"driven_swap": 0, "filtered_swap": 0})
    reference_order = {"ms_per_step": ref_ms,
                       "same_result": all(v["same_result"] for v in ref_q.values()),
                       "swap_rows": {q: v["swap_rows"] for q, v in ref_q.items()}}
    log(f"[bench] driven joins: {driven['ms_per_step_with']:.2f} ms per step with, "
        f"{driven['ms_per_step_without']:.2f} ms without; filtered IndexSwaps: "
        f"{filtered['ms_per_step_without']:.2f} ms without; both off {ref_ms:.2f} ms")

    batch = serve = None
    if args.mix > 0:
        mix = query_mix(args.mix, W.SEED_MIX, STORE_100M, s, p, o)
        mtexts = [query_text(pats, node) for _, pats, node in mix]
        log(f"[bench] batched mix of {len(mtexts)} queries")
        for r in eng.batch(mtexts[:16])[0]:  # warm-up
            r.free()
        batch = batch_bench(eng, mtexts)
        log(f"[bench] mix: plain {batch['plain_qps']:.0f} q/s, batched {batch['batched_qps']:.0f} q/s, "
            f"graph replay {batch['graph_qps']:.0f} q/s")

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline")
        cpu = cpu_baseline(idx_view, names, gpu_results, runs=args.cpu_runs)

    if args.serve_qps:
        eng.close()  # the served form runs on its own store
        big = args.serve_store == "500M"
        serve = serve_section(STORE_500M if big else STORE_100M, args.serve_qps,
                              args.serve_requests, W.SEED_500M if big else args.seed,
                              W.SEED_SERVE_MIX)

    queries = {}
    for q, sts in per_q.items():
        ms = float(np.median([x["elapsed_ms"] for x in sts]))
        jr = sts[0]["join_rows"]
        queries[q] = {"ms_median": ms, "join_rows": jr, "rows": sts[0]["rows"],
                      "select_rows": sts[0]["select_rows"], "swap_rows": sts[0]["swap_rows"],
                      "filtered_rows": sts[0]["filtered_rows"],
                      "join_rows_per_s": jr / (ms / 1e3), "alg_bytes": q_bytes[q],
                      "GB/s": q_bytes[q] / (ms / 1e3) / 1e9}
    out = {
        "metric": "join rows/s (single-query join set on the 100M-triple store)",
        "value": total_rows / dt,
        "unit": "join rows/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic (seeded generator, distinct triples)",
        "config": {"workload": "single-query joins (star4, chain4, snowflake6) on 100M triples",
                   "triples": n, "distinct": True, "entities": GEN["n_entities"],
                   "predicates": GEN["n_predicates"], "subject_zipf": GEN["subject_zipf"],
                   "pred_zipf": GEN["pred_zipf"], "literal_frac": GEN["literal_frac"],
                   "queries": names, "seed": args.seed, "bound_prescan": True,
                   "driven_joins": True, "filtered_swaps": True},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "queries": queries,
        "queries_note": "GB/s = algorithmic operator bytes of the query / its median wall time",
        "bound_prescan": prescan,
        "driven_joins": driven,
        "filtered_swaps": filtered,
        "reference_order": reference_order,
        "plan_rewrites_note": "driven joins: a join and the IndexSwap after it in one pass; "
                              "filtered IndexSwaps: the rows a swapped relation's join cannot "
                              "match dropped before the sort; reference_order: both off",
        "index_build": {"ms": build_s * 1e3, "permutations": list(perms), "key_bits": kb,
                        "alg_bytes": build_alg, "GB/s": build_alg / build_s / 1e9,
                        "frac": build_alg / build_s / HBM_PEAK,
                        "keys_per_s": len(perms) * n / build_s,
                        "order": [f"{q}<-{b}" if b else f"{q} (all bits)"
                                  for q, b in build_order(perms)],
                        "alg_note": "bytes of the derived-permutation build (index_bytes)",
                        "full_sort_bytes": full_b, "full_sort_GB/s": full_b / build_s / 1e9,
                        "full_sort_frac": full_b / build_s / HBM_PEAK,
                        "one_touch_bytes": 2 * 8 * n * len(perms),
                        "one_touch_frac": 2 * 8 * n * len(perms) / build_s / HBM_PEAK,
                        "full_sort_note": "the scope table's unit: ceil(K/8)*2*N*R + N*R per "
                                          "permutation, R = 8 B packed record, K = key bits; "
                                          "one-touch = 2*N*R per permutation",
                        "kernels": kernel_table(build_prof)},
        "autotune": tune,
        "planned_queries": planned,
        "range_selects": sel,
        "batched_mix": batch,
        "serving": serve,
        "kernels": kernel_table(step_prof),
        "kernels_note": "one untimed step with every launch timed",
    }
    print(json.dumps(out), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
