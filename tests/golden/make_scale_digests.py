"""Writes tests/golden/scale_digests.json: what the oracle computes on the
benchmark stores at their full sizes, for the GPU scale tests to compare
with.

For each store (seeded generator, distinct triples) the oracle builds every
permutation the GPU builds with its own sort (or_build_perm: qsort of
tuples), and records the SHA-256 of each of the six arrays, then runs the
workloads over its own PSO / POS arrays:
  * the 10k range selects (lookups: SHA-256 of the begin / end arrays per
    permutation, total rows);
  * the single-query join set, given plans and planner plans, with and
    without the reference's join-variable bounds (rows, join rows, select
    rows, bag digest of the result);
  * the 256-query batched mix and the 1024-request served mix (bag digest
    and sizes per query).
Nothing here touches a GPU.  Run (about 20 min, 40 GB of RAM for 500M):
  python tests/golden/make_scale_digests.py 100M
  python tests/golden/make_scale_digests.py 500M
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import pyoracle as O  # noqa: E402
from vedas_amd import workload as W  # noqa: E402
from vedas_amd._native import gen_triples  # noqa: E402
from vedas_amd.engine import plan, parse_plan, query_text  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scale_digests.json")
BUILT = ("SPO", "PSO", "POS", "OSP")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def build(s, p, o):
    out, keep = {}, {}
    for perm in BUILT:
        t = time.time()
        ix = O.build_perm(s, p, o, perm)
        out[perm] = {"n": len(ix["y"]), "n_l1": len(ix["l1v"]), "n_l2": len(ix["l2v"]),
                     "sha256": {k: sha(ix[k]) for k in ("l1v", "l1o", "l2v", "l2o", "y", "z")}}
        log(f"  {perm} built in {time.time() - t:.0f} s")
        if perm in ("PSO", "POS"):
            keep[perm] = ix
        else:
            del ix
    return out, keep


def run_query(ix, pats, node):
    rec = {}
    for bounds in (True, False):
        st = {}
        h, cols = O.run_plan(ix, pats, node, st, bounds=bounds)
        d = O.bag_digest(h, cols)
        sfx = "" if bounds else "_unbounded"
        rec["join_rows" + sfx] = st["join_rows"]
        rec["select_rows" + sfx] = st["select_rows"]
        if bounds:
            rec["rows"] = d[1]
            rec["digest"] = [list(d[0]), d[1], str(d[2]), str(d[3])]
        else:
            assert d == (tuple(rec["digest"][0]), rec["digest"][1], int(rec["digest"][2]),
                         int(rec["digest"][3])), "bounds changed a result"
        del cols
    return rec


def planner_plan(ix, pats):
    """The plan the query-graph planner makes from the patterns' range rows
    (PSO for a bound or free subject, POS for a bound object)."""
    est = []
    for sv, pv, ov in pats:
        if isinstance(sv, str) and not isinstance(ov, str):
            b, e = O.lookup(ix["POS"], pv, ov)
        elif not isinstance(sv, str):
            b, e = O.lookup(ix["PSO"], pv, sv)
        else:
            b, e = O.lookup(ix["PSO"], pv)
        est.append(e - b)
    return parse_plan(plan(query_text(pats, None), est))


def store_100m(res):
    prm = {k: v for k, v in W.STORE_100M.items() if k != "n"}
    t = time.time()
    s, p, o = gen_triples(W.STORE_100M["n"], seed=W.SEED_100M, distinct=True, **prm)
    log(f"100M store generated in {time.time() - t:.0f} s")
    rec = {"n": len(s), "seed": W.SEED_100M, "triples_digest": [str(x) for x in O.triples_digest(s, p, o)]}
    rec["perms"], ix = build(s, p, o)
    # range selects
    sel = {}
    for perm, (k1, k2) in W.select_mix(W.N_SELECTS, W.SEED_SELECTS, s, p, o).items():
        ar = O.build_perm(s, p, o, perm) if perm not in ix else ix[perm]
        be = np.array([O.lookup(ar, int(a), None if b == 0xFFFFFFFF else int(b))
                       for a, b in zip(k1, k2)], np.uint64).reshape(-1, 2)
        sel[perm] = {"count": len(k1), "rows": int((be[:, 1] - be[:, 0]).sum()),
                     "begin_sha256": sha(be[:, 0].copy()), "end_sha256": sha(be[:, 1].copy())}
        del ar
    rec["selects"] = sel
    log("selects done")
    js = {}
    for name, (pats, node) in W.JOIN_SET.items():
        t = time.time()
        js[name] = run_query(ix, pats, node)
        js[name]["text"] = query_text(pats, node)
        pn = planner_plan(ix, pats)
        js[name]["planned"] = run_query(ix, pats, pn)
        js[name]["planned"]["text"] = query_text(pats, pn)
        log(f"{name}: {js[name]['rows']} rows, {js[name]['join_rows']} join rows "
            f"({time.time() - t:.0f} s)")
    rec["join_set"] = js
    mix = []
    t = time.time()
    for _, pats, node in W.query_mix(W.N_MIX, W.SEED_MIX, W.STORE_100M, s, p, o):
        r = run_query(ix, pats, node)
        r["text"] = query_text(pats, node)
        mix.append(r)
    log(f"mix done in {time.time() - t:.0f} s")
    rec["mix"] = mix
    res["store_100M"] = rec


def store_500m(res):
    prm = {k: v for k, v in W.STORE_500M.items() if k != "n"}
    t = time.time()
    s, p, o = gen_triples(W.STORE_500M["n"], seed=W.SEED_500M, distinct=True, **prm)
    log(f"500M store generated in {time.time() - t:.0f} s")
    rec = {"n": len(s), "seed": W.SEED_500M, "triples_digest": [str(x) for x in O.triples_digest(s, p, o)]}
    rec["perms"], ix = build(s, p, o)
    mix = W.query_mix(W.N_SERVE, W.SEED_SERVE_MIX, W.STORE_500M, s, p, o)
    del s, p, o
    out = []
    t = time.time()
    for k, (_, pats, node) in enumerate(mix):
        r = run_query(ix, pats, node)
        r["text"] = query_text(pats, node)
        out.append(r)
        if k % 128 == 127:
            log(f"  served mix {k + 1} queries ({time.time() - t:.0f} s)")
    rec["serve_mix"] = out
    res["store_500M"] = rec


def main():
    which = sys.argv[1:] or ["100M", "500M"]
    res = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for w in which:
        {"100M": store_100m, "500M": store_500m}[w](res)
        with open(OUT + ".tmp", "w") as f:
            json.dump(res, f, indent=0, sort_keys=True)
        os.replace(OUT + ".tmp", OUT)
        log(f"wrote {OUT} ({w})")


if __name__ == "__main__":
    main()
