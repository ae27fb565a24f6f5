"""HIP graph capture: a run recorded (read-backs logged), captured (read-backs
answered from the log) and replayed; every replay, with eager work of the
program between replays, must equal the eager run."""
import numpy as np
import pytest

from oracle import pyoracle as O
from tests import plans

pytestmark = pytest.mark.gpu


def triples(n, seed):
    rng = np.random.default_rng(seed)
    t = np.stack([rng.integers(1, 3000, n), rng.integers(1, 6, n), rng.integers(1, 4000, n)], 1)
    t = np.unique(t, axis=0)
    rng.shuffle(t)
    return [np.ascontiguousarray(t[:, i], np.uint32) for i in range(3)]


def test_operator_run_captured_and_replayed():
    """Select, join, IndexSwap and unique on one context: recorded, captured
    and replayed four times, with an eager join on the same context between
    replays; every replay equals the eager run and the oracle."""
    from vedas_amd._native import Ctx, Index
    ctx = Ctx(0)
    s, p, o = triples(60_000, 5)
    s = np.append(s, s[:50])  # repeated triples: the unique has work
    p = np.append(p, p[:50])
    o = np.append(o, o[:50])
    idx = Index(ctx, s, p, o, ["PSO", "POS"])
    ix = {q: O.build_perm(s, p, o, q) for q in ("PSO", "POS")}
    b1, e1 = O.lookup(ix["POS"], 1)
    b2, e2 = O.lookup(ix["PSO"], 2)
    lw = O.select(ix["POS"], b1, e1, 3)
    rw = O.select(ix["PSO"], b2, e2, 3)
    jw = O.join(lw, rw, 1)
    want = O.unique_adjacent(O.sort_col0([jw[2], jw[1], jw[0]]))

    def run():
        L = idx.select("POS", b1, e1, 3)
        R = idx.select("PSO", b2, e2, 3)
        out = L.join(R, 1)
        out.swap_cols(0, 2)
        out.sort_col0()
        out.unique_adjacent()
        L.free()
        R.free()
        return out

    try:
        ctx.record(True)
        first = run()
        ctx.sync()
        ctx.record(False)
        eager = first.columns()
        for g, w in zip(eager, want):
            assert np.array_equal(g, w)
        first.free()
        ctx.capture_begin()
        try:
            out = run()
        except Exception:
            ctx.capture_end(keep=False)
            raise
        g = ctx.capture_end()
        try:
            for k in range(4):
                print(f"replay {k}", flush=True)
                ctx.graph_launch(g)
                ctx.sync()
                got = out.columns()
                for a, w in zip(got, want):
                    assert np.array_equal(a, w), k
                # eager work of the program between replays
                L = idx.select("PSO", b2, e2, 3)
                R = idx.select("PSO", b2, e2, 3)
                x = L.join(R, 1)
                assert x.nrows > 0
                for r in (L, R, x):
                    r.free()
        finally:
            out.free()
            ctx.graph_free(g)
    finally:
        idx.free()
        ctx.close()


def test_batch_graph_equals_batch():
    """A batch of star, chain and snowflake queries captured as one graph
    (vdq_graph_create) and replayed three times, with eager queries and
    batches on the engine between replays: every replay's results equal
    the eager batch's, column for column and row for row."""
    from vedas_amd.engine import Engine, query_text
    s, p, o = triples(120_000, 9)
    shapes = [plans.star_plan([1, 2, 3]), plans.chain_plan([1, 2, 3]),
              plans.chain_plan([2, 2, 1, 4]), plans.snowflake_plan([1, 2], [3, 4]),
              plans.star_plan([4, 5]), plans.chain_plan([5, 3])]
    texts = [query_text(pats, node) for pats, node in shapes]
    e = Engine(0)
    try:
        e.load(s, p, o, ("PSO", "POS"))
        res, _ = e.batch(texts)
        want = [(r.headers(), r.columns()) for r in res]
        for r in res:
            r.free()
        g = e.graph(texts)
        try:
            for k in range(3):
                if k:
                    g.replay()
                for i in range(len(texts)):
                    r = g.result(i)
                    h, cols = r.headers(), r.columns()
                    assert h == want[i][0], (k, i)
                    for a, b in zip(cols, want[i][1]):
                        assert np.array_equal(a, b), (k, i)
                # eager work between replays
                q = e.query(texts[k])
                assert q.shape()[1] == len(want[k][1][0])
                q.free()
                more, _ = e.batch(texts[::-1])
                for r in more:
                    r.free()
        finally:
            g.free()
        ex = plans.OracleExecutor(s, p, o)
        for (pats, node), (h, cols) in zip(shapes, want):
            rel = ex.run(node)
            assert O.bag_digest(h, cols) == O.bag_digest(rel.headers, ex.columns(rel))
    finally:
        e.close()
