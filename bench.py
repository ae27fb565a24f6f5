#!/usr/bin/env python3
"""Benchmark of the VEDAS query path on one MI355X.

Workload (synthetic, seeded): the 100M-triple store -- 10M entities with
Zipf(0.8) subjects, 256 predicates Zipf(1.1), objects 60% entities / 40%
literals -- indexed into SPO, PSO, POS and OSP in HBM, then the single-query
join set run through the drop-in operator API (QueryExecutor ->
Select/Join/IndexSwap jobs -> QueryPlan):

  star4      ?s 30 ?a . ?s 31 ?b . ?x 1 ?s . ?y 2 ?s
  chain4     ?a 1 ?b . ?b 1 ?c . ?c 1 ?d . ?d 1 ?e      (index swaps between joins)
  snowflake6 ?x 50 ?y . ?x 51 ?u . ?w 1 ?x . ?y 30 ?v . ?y 31 ?t . ?z 1 ?y

A step runs the three queries once (triples and permutations already
resident in HBM).  `value` = join output rows per second over the timed
steps.  The index build (4 permutations) is timed once beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--triples N]
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

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s

QUERIES = {
    "star4": ([("?s", 30, "?a"), ("?s", 31, "?b"), ("?x", 1, "?s"), ("?y", 2, "?s")],
              ("JOIN", ("JOIN", ("JOIN", ("UPLOAD", 0, "?s"), ("UPLOAD", 1, "?s"), "?s"),
                        ("UPLOAD", 2, "?s"), "?s"), ("UPLOAD", 3, "?s"), "?s")),
    "chain4": ([("?a", 1, "?b"), ("?b", 1, "?c"), ("?c", 1, "?d"), ("?d", 1, "?e")],
               ("JOIN", ("INDEXSWAP", ("JOIN", ("INDEXSWAP", ("JOIN", ("UPLOAD", 0, "?b"),
                                                              ("UPLOAD", 1, "?b"), "?b"), "?c"),
                                                 ("UPLOAD", 2, "?c"), "?c"), "?d"),
                ("UPLOAD", 3, "?d"), "?d")),
    "snowflake6": ([("?x", 50, "?y"), ("?x", 51, "?u"), ("?w", 1, "?x"),
                    ("?y", 30, "?v"), ("?y", 31, "?t"), ("?z", 1, "?y")],
                   ("JOIN",
                    ("INDEXSWAP", ("JOIN", ("JOIN", ("UPLOAD", 0, "?x"), ("UPLOAD", 1, "?x"), "?x"),
                                   ("UPLOAD", 2, "?x"), "?x"), "?y"),
                    ("JOIN", ("JOIN", ("UPLOAD", 3, "?y"), ("UPLOAD", 4, "?y"), "?y"),
                     ("UPLOAD", 5, "?y"), "?y"),
                    "?y")),
}

GEN = dict(n_entities=10_000_000, subject_zipf=0.8, n_predicates=256, pred_zipf=1.1,
           literal_frac=0.4, literal_base=50_000_000, n_literals=50_000_000, distinct=False)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def index_bytes(n: int, kbits: int) -> float:
    """Algorithmic HBM bytes of one permutation build: ceil(K/8) radix
    passes reading and writing 8-byte packed keys, one pass packing
    (12 B read, 8 B written) and one unpacking (8 B read, 8 B written)."""
    passes = -(-kbits // 8)
    return passes * 2 * 8 * n + 20 * n + 16 * n


def select_bench(ctx, idx, s, p, o, nq=10_000, seed=7, max_chunk_rows=400_000_000):
    """10k range selects with one or two constants taken from uniformly drawn
    existing triples: S -> SPO, P -> PSO, O -> OSP, (S,P) -> SPO, (P,O) -> POS,
    (O,S) -> OSP.  Each permutation's queries run as batched selects whose
    output (Y, Z) rows are materialised in HBM, in chunks of at most
    max_chunk_rows rows.  Returns selects/s, rows/s and GB/s."""
    import ctypes as C
    from vedas_amd import _native
    rng = np.random.default_rng(seed)
    pick = rng.integers(0, len(s), nq)
    two = rng.random(nq) < 0.5
    pos = rng.integers(0, 3, nq)
    groups = {"SPO": ([], []), "PSO": ([], []), "POS": ([], []), "OSP": ([], [])}
    U = _native.VD_UNBOUND
    for i, t in enumerate(pick):
        S, P, O = int(s[t]), int(p[t]), int(o[t])
        if not two[i]:
            perm, k1, k2 = [("SPO", S, U), ("PSO", P, U), ("OSP", O, U)][pos[i]]
        else:
            perm, k1, k2 = [("SPO", S, P), ("POS", P, O), ("OSP", O, S)][pos[i]]
        groups[perm][0].append(k1)
        groups[perm][1].append(k2)
    plans_ = []
    for perm, (k1, k2) in groups.items():
        if not k1:
            continue
        k1 = np.array(k1, np.uint32)
        k2 = np.array(k2, np.uint32)
        # sizes on the host to cut the batch into chunks that fit
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
            "rows_per_s": rows / dt, "GB/s": 16.0 * rows / dt / 1e9,
            "chunks": len(staged)}


def cpu_baseline(engine_ctx, sample_seconds: float):
    """The oracle's join (C, one thread) on the first left rows of the
    star4 query's first join (?s 30 ?a ⋈ ?s 31 ?b on the same store), so the
    CPU does the same kind of work; rows/s extrapolate linearly in left rows."""
    from oracle import pyoracle as O
    from vedas_amd import _native
    ctx, idx = engine_ctx
    rels = []
    for pred in (30, 31):
        b, e = idx.lookup("PSO", pred)
        r = idx.select("PSO", b, e, 3)
        rels.append(r.columns())
        r.free()
    L, R = rels if len(rels[0][0]) <= len(rels[1][0]) else rels[::-1]
    m = min(len(L[0]), 2000)
    rows = 0
    t0 = time.perf_counter()
    elapsed = 0.0
    while elapsed < sample_seconds and m <= len(L[0]):
        out = O.join([c[:m] for c in L], R, 1)
        rows = len(out[0])
        elapsed = time.perf_counter() - t0
        if elapsed < sample_seconds / 4 and m < len(L[0]):
            m = min(len(L[0]), m * 4)
            t0 = time.perf_counter()
            elapsed = 0.0
            continue
        break
    return {"value": rows / max(elapsed, 1e-9), "unit": "join rows/s", "cores": 1,
            "kind": "port",
            "sample": f"oracle or_join (1 thread) of the first {m} left rows of "
                      f"star4's first join ({len(L[0])} x {len(R[0])} rows); {rows} output rows"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--triples", type=int, default=100_000_000)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--queries", default="star4,chain4,snowflake6")
    ap.add_argument("--selects", type=int, default=10_000)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the query path does not shard across GPUs.", file=sys.stderr)
        sys.exit(2)

    from vedas_amd import _native
    from vedas_amd.engine import Engine, query_text

    log(f"[bench] generating {args.triples} triples")
    t = time.perf_counter()
    s, p, o = _native.gen_triples(args.triples, seed=args.seed, **GEN)
    log(f"[bench] generated in {time.perf_counter() - t:.1f}s")

    eng = Engine(0)
    ctx = _native.Ctx.wrap(eng.ctx_handle())
    ds, dp, do = ctx.upload(s), ctx.upload(p), ctx.upload(o)
    n = len(s)
    perms = ("SPO", "PSO", "POS", "OSP")
    # index build: one untimed warm-up build, then a timed one
    log("[bench] index build")
    eng.load_device(ds, dp, do, n, perms)
    ctx.sync()
    ctx.prof(True)
    ctx.prof_reset()
    t0 = time.perf_counter()
    eng.load_device(ds, dp, do, n, perms)
    ctx.sync()
    build_s = time.perf_counter() - t0
    build_prof = ctx.prof_stats()
    ctx.prof(False)
    for d in (ds, dp, do):
        ctx.free(d)
    kb = {name: int(s.max()).bit_length() + int(p.max()).bit_length() + int(o.max()).bit_length()
          for name in perms}
    build_alg = sum(index_bytes(n, kb[x]) for x in perms)
    log(f"[bench] index build {build_s * 1e3:.1f} ms")

    idx_view = _native.Index.wrap(ctx, eng.index_handle(), n)
    sel = None
    if args.selects > 0:
        log("[bench] range selects")
        select_bench(ctx, idx_view, s, p, o, nq=min(args.selects, 500))  # warm-up
        sel = select_bench(ctx, idx_view, s, p, o, nq=args.selects)
        log(f"[bench] selects: {sel['selects_per_s']:.0f}/s, {sel['GB/s']:.0f} GB/s")

    names = [q for q in args.queries.split(",") if q]
    texts = {q: query_text(QUERIES[q][0], QUERIES[q][1]) for q in names}

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

    for w in range(args.warmup):
        log(f"[bench] warmup {w}")
        step()
    ctx.sync()
    ctx.prof(True)
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

    # dominant kernel by device time in the timed region
    dom_name, dom = max(prof.items(), key=lambda kv: kv[1]["ms"])
    achieved = dom["bytes"] / (dom["ms"] / 1e3) if dom["ms"] > 0 else 0.0
    roofline = {"bound": "hbm", "kernel": dom_name, "achieved": achieved / 1e9,
                "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                "traffic": None, "launches": dom["launches"],
                "avg_launch_ms": dom["ms"] / max(dom["launches"], 1),
                "alg_bytes_per_launch": dom["bytes"] / max(dom["launches"], 1)}

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline")
        cpu = cpu_baseline((ctx, idx_view), args.cpu_seconds)

    queries = {}
    for q, sts in per_q.items():
        ms = float(np.median([x["elapsed_ms"] for x in sts]))
        jr = sts[0]["join_rows"]
        queries[q] = {"ms_median": ms, "join_rows": jr, "rows": sts[0]["rows"],
                      "join_rows_per_s": jr / (ms / 1e3), "swap_rows": sts[0]["swap_rows"]}
    kernels = {k: {"launches": v["launches"], "ms": round(v["ms"], 3),
                   "GB/s": round(v["bytes"] / (v["ms"] / 1e3) / 1e9, 1) if v["ms"] > 0 else None}
               for k, v in sorted(prof.items(), key=lambda kv: -kv[1]["ms"])}
    out = {
        "metric": "join rows/s (config-2 query set on the 100M-triple store)",
        "value": total_rows / dt,
        "unit": "join rows/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic (seeded generator, duplicates kept)",
        "config": {"workload": "single-query joins (star4, chain4, snowflake6) on 100M triples",
                   "triples": n, "entities": GEN["n_entities"], "predicates": GEN["n_predicates"],
                   "subject_zipf": GEN["subject_zipf"], "pred_zipf": GEN["pred_zipf"],
                   "literal_frac": GEN["literal_frac"], "queries": names, "seed": args.seed},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "queries": queries,
        "index_build": {"ms": build_s * 1e3, "permutations": list(perms),
                        "key_bits": kb, "alg_bytes": build_alg,
                        "GB/s": build_alg / build_s / 1e9,
                        "frac": build_alg / build_s / HBM_PEAK,
                        "keys_per_s": 4 * n / build_s,
                        "kernels": {k: {"launches": v["launches"], "ms": round(v["ms"], 3),
                                        "GB/s": round(v["bytes"] / (v["ms"] / 1e3) / 1e9, 1)
                                        if v["ms"] > 0 else None}
                                    for k, v in sorted(build_prof.items(), key=lambda kv: -kv[1]["ms"])}},
        "kernels": kernels,
        "range_selects": sel,
    }
    print(json.dumps(out), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
