"""Helpers of the full-size GPU parity tests: the benchmark stores built
through the engine, their permutations checked at full size, and the
oracle's digests of the same stores (tests/golden/scale_digests.json,
written by tests/golden/make_scale_digests.py)."""
import hashlib
import json
import os
import time

import numpy as np

from oracle import pyoracle as O

GOLDEN_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                           "scale_digests.json")
PERMS = ("SPO", "PSO", "POS", "OSP")


def note(msg: str) -> None:
    print(f"[scale] {msg}", flush=True)


def golden(key: str) -> dict:
    with open(GOLDEN_PATH) as f:
        return json.load(f)[key]


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def digest_of(rec: dict):
    names, n, a, b = rec["digest"]
    return tuple(names), n, int(a), int(b)


def build_store(prm: dict, seed: int):
    """The store of `prm` (distinct triples, `seed`) in a fresh engine with
    SPO, PSO, POS and OSP resident.  Returns a dict with the engine, its
    context and index views and the host triples."""
    from vedas_amd import _native
    from vedas_amd.engine import Engine
    gen = {k: v for k, v in prm.items() if k != "n"}
    t = time.time()
    s, p, o = _native.gen_triples(prm["n"], seed=seed, distinct=True, **gen)
    note(f"generated {len(s)} distinct triples in {time.time() - t:.1f} s")
    eng = Engine(0)
    ctx = _native.Ctx.wrap(eng.ctx_handle())
    ds, dp, do = ctx.upload(s), ctx.upload(p), ctx.upload(o)
    eng.load_device(ds, dp, do, len(s), PERMS)
    ctx.sync()
    for d in (ds, dp, do):
        ctx.free(d)
    idx = _native.Index.wrap(ctx, eng.index_handle(), len(s))
    note("index built")
    return {"eng": eng, "ctx": ctx, "idx": idx, "s": s, "p": p, "o": o}


def checked_arrays(st: dict, gold: dict, perms=PERMS, keep=PERMS) -> dict:
    """Every permutation's arrays downloaded and checked at full size:
    sorted (X, Y, Z) order with L1 / L2 at the run heads, the same multiset
    of triples as the input, L1 run lengths equal to the input's histogram
    of the leading column, and -- when the generator reproduced the
    oracle's store -- every array equal (SHA-256) to the oracle's own
    build.  Returns {perm: arrays} for the permutations in `keep`, and
    "_same_store"."""
    s, p, o = st["s"], st["p"], st["o"]
    td = O.triples_digest(s, p, o)
    same_store = [str(x) for x in td] == gold["triples_digest"]
    note(f"store equals the oracle's: {same_store}")
    cols = {"S": s, "P": p, "O": o}
    out = {}
    for perm in perms:
        a = st["idx"].arrays(perm)
        O.check_perm(a)
        assert O.perm_digest(a, perm) == td, f"{perm}: not the input's triples"
        hist = np.bincount(cols[perm[0]])
        vals = np.flatnonzero(hist)
        counts = hist[vals]
        del hist
        runs = np.diff(np.append(a["l1o"].astype(np.int64), len(a["y"])))
        assert np.array_equal(a["l1v"], vals) and np.array_equal(runs, counts), perm
        if same_store:
            g = gold["perms"][perm]
            assert (len(a["y"]), len(a["l1v"]), len(a["l2v"])) == (g["n"], g["n_l1"], g["n_l2"])
            for k in ("l1v", "l1o", "l2v", "l2o", "y", "z"):
                assert sha(a[k]) == g["sha256"][k], (perm, k)
        note(f"{perm} checked ({len(a['y'])} rows, {len(a['l1v'])} L1, {len(a['l2v'])} L2)")
        if perm in keep:
            out[perm] = a
        del a
    out["_same_store"] = same_store
    return out
