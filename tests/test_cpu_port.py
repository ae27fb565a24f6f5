"""The multi-threaded CPU port (the benchmark's CPU baseline) gives the
oracle's results bit for bit, row order included, for any thread count.
CPU only."""
import numpy as np
import pytest

from oracle import cpu_port as P
from oracle import pyoracle as O
from tests import plans


def rel(rng, n, ncols, keymax):
    return O.sort_col0([rng.integers(0, keymax, n).astype(np.uint32) for _ in range(ncols)])


@pytest.mark.parametrize("nt", [1, 3, 8])
def test_join_sort_select_unique_equal_oracle(nt):
    ops = P.make_ops(nt)
    rng = np.random.default_rng(nt)
    for nl, nr, km in [(0, 5, 3), (5, 0, 3), (1, 1, 1), (3000, 5000, 40), (70_000, 20, 9000),
                       (5, 90_000, 3), (200_000, 150_000, 500_000), (40_000, 40_000, 2)]:
        L = rel(rng, nl, int(rng.integers(1, 4)), km)
        R = rel(rng, nr, int(rng.integers(1, 4)), km)
        want = O.join(L, R, 1)
        got = ops.join(L, R, 1)
        assert len(got) == len(want)
        for g, w in zip(got, want):
            assert np.array_equal(g, w), (nl, nr, km)
    for n, nc, km in [(0, 2, 5), (1, 1, 5), (70_000, 3, 7), (300_000, 2, 1 << 31), (100_000, 4, 300)]:
        cols = [rng.integers(0, km, n).astype(np.uint32) for _ in range(nc)]
        for g, w in zip(ops.sort_col0(cols), O.sort_col0(cols)):
            assert np.array_equal(g, w)
        s = O.sort_col0(cols)
        for g, w in zip(ops.unique_adjacent(s), O.unique_adjacent(s)):
            assert np.array_equal(g, w)
    s, p, o = (rng.integers(1, 60, 200_000).astype(np.uint32) for _ in range(3))
    ix = O.build_perm(s, p, o, "PSO")
    for k1 in (1, 7, 59, 61):
        b, e = O.lookup(ix, k1)
        for clip in (None, (10, 40), (30, 20), (0, 0xFFFFFFFF)):
            for cols in (1, 3):
                g = ops.select(ix, b, e, cols, clip=clip)
                w = O.select(ix, b, e, cols, clip=clip)
                for gg, ww in zip(g, w):
                    assert np.array_equal(gg, ww)


@pytest.mark.parametrize("shape", ["star3", "chain3", "snow"])
def test_plan_walk_equals_oracle(shape):
    rng = np.random.default_rng(2)
    trip = sorted(set(zip(rng.integers(1, 400, 30_000).tolist(), rng.integers(1, 6, 30_000).tolist(),
                          rng.integers(1, 400, 30_000).tolist())))
    s, p, o = (np.array(c, np.uint32) for c in zip(*trip))
    pats, node = {"star3": plans.star_plan([1, 2, 3]), "chain3": plans.chain_plan([1, 2, 3]),
                  "snow": plans.snowflake_plan([1, 2], [3, 4])}[shape]
    ix = {perm: O.build_perm(s, p, o, perm) for perm in ("PSO", "POS")}
    st0, st1 = {}, {}
    h0, c0 = O.run_plan(ix, pats, node, st0, bounds=True)
    h1, c1 = P.run_query(ix, pats, node, P.make_ops(4), st1)
    assert h0 == h1
    for a, b in zip(c0, c1):
        assert np.array_equal(a, b)  # distinct triples: the unique keeps every row
    assert st0 == st1
    assert len(c0[0]) > 0
