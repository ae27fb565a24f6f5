"""The semi-join filter (vd_semi_filter) and the IndexSwap it serves: an
IndexSwapJob whose output feeds a join on its variable leaves a large
relation unsorted, and the join drops the rows whose key the other input
lacks before sorting the rest.  The filter is checked against numpy
membership on the same inputs, and whole plans are checked identical, row
for row, with and without it (and against the oracle's plan walk)."""
import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from vedas_amd._native import Ctx
    c = Ctx(0)
    yield c
    c.close()


def check_equal(got, want):
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


@pytest.mark.parametrize("nx,ny,nk", [(1, 1, 3), (5000, 3000, 4000), (200_000, 50_000, 1 << 20),
                                      (300_000, 1000, 100)])
@pytest.mark.parametrize("sorted_y", [True, False])
def test_semi_filter_keeps_rows_with_keys_in_y(ctx, nx, ny, nk, sorted_y):
    from vedas_amd._native import Rel
    rng = np.random.default_rng(nx + ny)
    x = [rng.integers(0, nk, nx).astype(np.uint32), rng.integers(0, 1 << 32, nx, dtype=np.uint64).astype(np.uint32),
         np.arange(nx, dtype=np.uint32)]
    yk = rng.integers(0, nk, ny).astype(np.uint32)
    if sorted_y:
        yk = np.sort(yk)
    X = Rel.from_columns(ctx, x)
    Y = Rel.from_columns(ctx, [yk])
    print(f"semi filter nx={nx} ny={ny} nk={nk} sorted_y={sorted_y}", flush=True)
    keep = np.isin(x[0], yk)
    want = [c[keep] for c in x]
    out, rows = X.semi_filter(0, Y, key_hi=nk - 1)
    assert rows == int(keep.sum())
    check_equal(out.columns(), want)
    out.free()
    # over max_rows: nothing built, the count returned
    if rows > 0:
        none, r2 = X.semi_filter(0, Y, key_hi=nk - 1, max_rows=rows - 1)
        assert none is None and r2 == rows
    # the filter column may be any column; keys above key_hi match nothing
    out, rows = X.semi_filter(0, Y, key_hi=nk // 2)
    keep2 = keep & (x[0] <= nk // 2)
    assert rows == int(keep2.sum())
    check_equal(out.columns(), [c[keep2] for c in x])
    out.free()
    for r in (X, Y):
        r.free()


def test_semi_filter_declines_and_empty(ctx):
    from vedas_amd._native import Rel
    X = Rel.from_columns(ctx, [np.array([5, 1 << 30, 7], np.uint32)])
    Y = Rel.from_columns(ctx, [np.array([7], np.uint32)])
    # no bound below 2^28: no bitmap, nothing built
    assert X.semi_filter(0, Y) == (None, None)
    out, rows = X.semi_filter(0, Y, key_hi=100)
    assert rows == 1 and out.columns()[0].tolist() == [7]
    out.free()
    E = Rel.from_columns(ctx, [np.zeros(0, np.uint32)])
    out, rows = X.semi_filter(0, E, key_hi=100)
    assert rows == 0 and out.nrows == 0
    out.free()
    for r in (X, Y, E):
        r.free()


def distinct_triples(n, seed, ns, npred, no):
    rng = np.random.default_rng(seed)
    t = np.stack([rng.integers(1, ns, n), rng.integers(1, npred + 1, n),
                  rng.integers(1, no, n)], 1)
    t = np.unique(t, axis=0)
    rng.shuffle(t)
    return [np.ascontiguousarray(t[:, i], np.uint32) for i in range(3)]


@pytest.mark.parametrize("shape", ["snow", "chain4"])
def test_plans_with_and_without_filtered_swaps(shape):
    """Whole plans through the operator API with the IndexSwap left to its
    join (every size) and without: identical columns and rows, the same
    join rows, fewer rows sorted, and the oracle's plan walk."""
    from tests import plans
    from vedas_amd.engine import Engine, get_option, query_text, set_option
    s, p, o = distinct_triples(120_000, 31, ns=8000, npred=4, no=9000)
    keep = (p != 4) | (np.arange(len(p)) % 25 == 0)  # predicate 4 rare: selective last joins
    s, p, o = s[keep], p[keep], o[keep]
    if shape == "snow":
        pats, node = plans.snowflake_plan([1, 2], [3, 4])
    else:
        pats, node = plans.chain_plan([1, 2, 3, 4])
    text = query_text(pats, node)
    default = get_option("filtered_swap")
    e = Engine(0)
    try:
        e.load(s, p, o, ("PSO", "POS"))
        got = {}
        for on in (1, 0):
            set_option("filtered_swap", on)  # 1: every swapped relation, 0: none
            set_option("driven_swap", 0)     # the chain's swaps as IndexSwapJobs
            r = e.query(text)
            got[on] = (r.headers(), r.columns(), r.stats())
            r.free()
    finally:
        set_option("filtered_swap", default)
        set_option("driven_swap", 1)
        e.close()
    (h1, c1, st1), (h0, c0, st0) = got[1], got[0]
    print(f"{shape}: rows {st1['rows']} join rows {st1['join_rows']} swap rows "
          f"{st1['swap_rows']} / {st0['swap_rows']} filtered {st1['filtered_rows']}", flush=True)
    assert h1 == h0
    for a, b in zip(c1, c0):
        assert np.array_equal(a, b)
    assert st1["join_rows"] == st0["join_rows"] and st1["rows"] == st0["rows"]
    assert st0["filtered_rows"] == 0
    assert st1["filtered_rows"] > 0 and st1["swap_rows"] < st0["swap_rows"]
    assert st1["swap_rows"] + st1["filtered_rows"] == st0["swap_rows"]
    ex = plans.OracleExecutor(s, p, o)
    rel = ex.run(node)
    assert O.bag_digest(h1, c1) == O.bag_digest(rel.headers, ex.columns(rel))
    assert get_option("filtered_swap") == default


def test_batch_with_and_without_filtered_swaps():
    """The batched executor with every IndexSwap left to its join (filter
    forced on) and with none: each query's result equals the other run's
    and the oracle's plan walk (canonical order), and fewer rows are
    re-sorted."""
    from tests import plans
    from vedas_amd.engine import Engine, get_option, query_text, set_option
    s, p, o = distinct_triples(120_000, 37, ns=8000, npred=4, no=9000)
    keep = (p != 4) | (np.arange(len(p)) % 25 == 0)
    s, p, o = s[keep], p[keep], o[keep]
    shapes = [plans.snowflake_plan([1, 2], [3, 4]), plans.chain_plan([1, 2, 3, 4]),
              plans.chain_plan([2, 1, 4]), plans.snowflake_plan([2, 3], [4, 1])]
    texts = [query_text(pats, node) for pats, node in shapes]
    default = get_option("filtered_swap")
    e = Engine(0)
    try:
        e.load(s, p, o, ("PSO", "POS"))
        got, stats = {}, {}
        for on in (1, 0):
            set_option("filtered_swap", on)
            res, st = e.batch(texts)
            got[on] = [O.bag_digest(r.headers(), r.columns()) for r in res]
            stats[on] = st
            for r in res:
                r.free()
    finally:
        set_option("filtered_swap", default)
        e.close()
    print(f"batched swap rows {stats[1]['swap_rows']} / {stats[0]['swap_rows']}", flush=True)
    assert got[1] == got[0]
    assert stats[1]["swap_rows"] < stats[0]["swap_rows"]
    assert stats[1]["join_rows"] == stats[0]["join_rows"]
    ex = plans.OracleExecutor(s, p, o)
    for (pats, node), g in zip(shapes, got[1]):
        rel = ex.run(node)
        assert O.bag_digest(rel.headers, ex.columns(rel)) == g
