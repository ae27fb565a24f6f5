"""Joins driven by a relation in output order (vd_join_driven) against the
oracle's join followed by its stable sort on the swapped column (the
reference's JoinQueryJob::join then IndexSwapJob::startJob): the driver is
the other permutation's range of one input's pattern (POS for a PSO
select), so the driven output must equal join + IndexSwap bit for bit."""
import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from vedas_amd._native import Ctx
    c = Ctx(0)
    c.tune("drive_min_rows", 0)  # the stores here are small
    yield c
    c.close()


def distinct_triples(n, seed, ns=3000, npred=4, no=5000, hot_s=7, hot_o=11):
    rng = np.random.default_rng(seed)
    t = np.stack([rng.integers(1, ns, n), rng.integers(1, npred + 1, n),
                  rng.integers(1, no, n)], 1)
    t[::hot_s, 0] = 5  # a hot subject
    t[::hot_o, 2] = 9  # a hot object
    t = np.unique(t, axis=0)
    rng.shuffle(t)
    return [np.ascontiguousarray(t[:, i], np.uint32) for i in range(3)]


@pytest.fixture(scope="module")
def store(ctx):
    from vedas_amd._native import Index
    s, p, o = distinct_triples(120_000, 3)
    idx = Index(ctx, s, p, o, ["PSO", "POS"])
    ix = {q: O.build_perm(s, p, o, q) for q in ("PSO", "POS")}
    yield idx, ix
    idx.free()


def check_equal(got, want):
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


@pytest.mark.parametrize("p1,p2", [(1, 2), (2, 2), (3, 1), (4, 4)])
@pytest.mark.parametrize("clip", [None, (40, 2500)])
def test_driven_equals_join_then_swap_right(ctx, store, p1, p2, clip):
    """?a p1 ?b . ?b p2 ?c joined on ?b, then swapped to ?c (held by the
    right input): the driver is p2's POS range (c, b)."""
    idx, ix = store
    print(f"driven right p1={p1} p2={p2} clip={clip}", flush=True)
    b1, e1 = O.lookup(ix["POS"], p1)
    b2, e2 = O.lookup(ix["PSO"], p2)
    lw = O.select(ix["POS"], b1, e1, 3, clip=clip)        # (b, a)
    rw = O.select(ix["PSO"], b2, e2, 3, clip=clip)        # (b, c)
    jw = O.join(lw, rw, 1)                                # [b, a, c]
    want = O.sort_col0([jw[2], jw[1], jw[0]])             # swap(0, 2), stable sort
    L = idx.select("POS", b1, e1, 3, clip=clip)
    D = idx.select("POS", b2, e2, 3)                      # (c, b), unclipped
    lo, hi = clip if clip else (0, 0xFFFFFFFF)
    r = D.join_driven(1, L, [-1, 1, -2], lo, hi, check_dups=True)
    assert r is not None
    out, mlo, mhi = r
    check_equal(out.columns(), want)
    if len(jw[0]):
        assert (mlo, mhi) == (int(jw[0].min()), int(jw[0].max()))
    for x in (L, D, out):
        x.free()


@pytest.mark.parametrize("p1,p2", [(1, 2), (3, 3)])
def test_driven_equals_join_then_swap_left(ctx, store, p1, p2):
    """The swapped variable held by the left input: ?b p2 ?c (left, PSO)
    joined with ?a p1 ?b (right, POS) on ?b, then swapped to ?c; the driver
    is p2's POS range and the right input's runs are looked up."""
    idx, ix = store
    print(f"driven left p1={p1} p2={p2}", flush=True)
    b1, e1 = O.lookup(ix["POS"], p1)
    b2, e2 = O.lookup(ix["PSO"], p2)
    lw = O.select(ix["PSO"], b2, e2, 3)                   # (b, c)
    rw = O.select(ix["POS"], b1, e1, 3)                   # (b, a)
    jw = O.join(lw, rw, 1)                                # [b, c, a]
    want = O.sort_col0([jw[1], jw[0], jw[2]])             # swap(0, 1)
    R = idx.select("POS", b1, e1, 3)
    D = idx.select("POS", b2, e2, 3)                      # (c, b)
    out, _, _ = D.join_driven(1, R, [-1, -2, 1])
    check_equal(out.columns(), want)
    for x in (R, D, out):
        x.free()


def test_driven_over_a_join_output(ctx, store):
    """Y is itself a join output sorted on its key (the second join of a
    chain): ((?a p1 ?b . ?b p2 ?c) swapped to ?c) . ?c p3 ?d, then swapped
    to ?d, with both swaps done by driven joins."""
    idx, ix = store
    p1, p2, p3 = 1, 2, 3
    b1, e1 = O.lookup(ix["POS"], p1)
    b2, e2 = O.lookup(ix["PSO"], p2)
    b3, e3 = O.lookup(ix["PSO"], p3)
    j1 = O.join(O.select(ix["POS"], b1, e1, 3), O.select(ix["PSO"], b2, e2, 3), 1)
    s1 = O.sort_col0([j1[2], j1[1], j1[0]])               # [c, a, b]
    j2 = O.join(s1, O.select(ix["PSO"], b3, e3, 3), 1)    # [c, a, b, d]
    want = O.sort_col0([j2[3], j2[1], j2[2], j2[0]])      # swap(0, 3)
    L = idx.select("POS", b1, e1, 3)
    D1 = idx.select("POS", b2, e2, 3)
    o1, _, _ = D1.join_driven(1, L, [-1, 1, -2], check_dups=True)
    check_equal(o1.columns(), s1)
    D2 = idx.select("POS", b3, e3, 3)                     # (d, c)
    o2, _, _ = D2.join_driven(1, o1, [-1, 1, 2, -2], check_dups=True)
    check_equal(o2.columns(), want)
    for x in (L, D1, o1, D2, o2):
        x.free()


def test_driven_declines(ctx):
    """Adjacent equal driver rows matching several rows (with check_dups)
    and key values beyond the run table's range return None; without
    check_dups the output is driver-major."""
    from vedas_amd._native import Rel
    y = Rel.from_columns(ctx, [np.array([1, 1, 2, 5, 5, 5], np.uint32),
                               np.array([10, 11, 20, 50, 51, 52], np.uint32)])
    d = Rel.from_columns(ctx, [np.array([7, 7, 8, 9], np.uint32),
                               np.array([5, 5, 2, 1], np.uint32)])
    # host-made columns carry no value bound: key_hi bounds the run table
    assert d.join_driven(1, y, [-1, 1], key_hi=9, check_dups=True) is None
    out, lo, hi = d.join_driven(1, y, [-1, 1, -2], key_hi=9)
    got = out.columns()
    assert got[0].tolist() == [7, 7, 7, 7, 7, 7, 8, 9, 9]
    assert got[1].tolist() == [50, 51, 52, 50, 51, 52, 20, 10, 11]
    assert got[2].tolist() == [5, 5, 5, 5, 5, 5, 2, 1, 1]
    assert (lo, hi) == (1, 5)
    out.free()
    # keys outside [key_lo, key_hi] match nothing
    out, lo, hi = d.join_driven(1, y, [-1, 1], key_lo=2, key_hi=4)
    assert out.columns()[1].tolist() == [20] and (lo, hi) == (2, 2)
    out.free()
    big = Rel.from_columns(ctx, [np.array([1 << 29], np.uint32)])
    assert big.join_driven(0, y, [-1, 1]) is None  # no bound at all: too many key values
    assert d.join_driven(1, y, [-1, 1], key_hi=100) is None  # 101 key values for 10 rows
    ctx.tune("drive_min_rows", 11)
    try:
        assert d.join_driven(1, y, [-1, 1], key_hi=9) is None  # 10 rows < 11
    finally:
        ctx.tune("drive_min_rows", 0)
    for x in (y, d, big):
        x.free()


@pytest.mark.parametrize("shape", ["chain4", "chain3", "snow"])
def test_plans_with_and_without_driven_swaps(shape):
    """Whole plans through the operator API: the results with driven joins
    (the default) are identical, column for column and row for row, to the
    plain join + IndexSwap path, and equal the oracle's plan walk."""
    from tests import plans
    from vedas_amd.engine import Engine, query_text, set_option, get_option
    # hot keys milder than the operator tests': the snowflake joins two
    # stars on ?y, and a hot object makes that join quadratic (about 1M
    # rows per result here)
    s, p, o = distinct_triples(80_000, 11, ns=6000, npred=3, no=8000, hot_s=997, hot_o=1009)
    if shape == "chain4":
        pats, node = plans.chain_plan([1, 1, 1, 2])
    elif shape == "chain3":
        pats, node = plans.chain_plan([2, 3, 1])
    else:
        pats, node = plans.snowflake_plan([1, 2], [3, 1])
    text = query_text(pats, node)
    from vedas_amd._native import Ctx
    e = Engine(0)
    try:
        Ctx.wrap(e.ctx_handle()).tune("drive_min_rows", 0)  # a small store
        e.load(s, p, o, ("PSO", "POS"))
        got = {}
        for on in (1, 0):
            set_option("driven_swap", on)
            r = e.query(text)
            got[on] = (r.headers(), r.columns(), r.stats())
            r.free()
    finally:
        set_option("driven_swap", 1)
        e.close()
    (h1, c1, st1), (h0, c0, st0) = got[1], got[0]
    assert h1 == h0
    for a, b in zip(c1, c0):
        assert np.array_equal(a, b)
    assert st1["join_rows"] == st0["join_rows"] and st1["rows"] == st0["rows"]
    assert st0["driven_joins"] == 0
    if shape.startswith("chain"):
        assert st1["driven_joins"] >= 1 and st1["swap_rows"] < st0["swap_rows"]
    ex = plans.OracleExecutor(s, p, o)
    rel = ex.run(node)
    assert O.bag_digest(h1, c1) == O.bag_digest(rel.headers, ex.columns(rel))
    assert get_option("driven_swap") == 1


def test_plan_with_resorted_left_input():
    """A join whose left input is an intermediate holding the next swap
    variable, with an output several times that input: the input is
    re-sorted on the swap variable and drives (vd_join_size decides), and
    the plan's result is identical to the plain path's and equals the
    oracle's plan walk."""
    from tests import plans
    from vedas_amd._native import Ctx
    from vedas_amd.engine import Engine, query_text, set_option
    s, p, o = distinct_triples(120_000, 21, ns=4000, npred=4, no=6000, hot_s=997, hot_o=1009)
    keep = (p != 4) | (np.arange(len(p)) % 20 == 0)  # predicate 4 rare
    s, p, o = s[keep], p[keep], o[keep]
    x, y, a, z, b = "?x", "?y", "?a", "?z", "?b"
    pats = [(x, 4, y), (x, 2, a), (z, 3, x), (y, 1, b)]
    inner = ("JOIN", ("JOIN", ("UPLOAD", pats[0], x), ("UPLOAD", pats[1], x), x),
             ("UPLOAD", pats[2], x), x)
    node = ("JOIN", ("INDEXSWAP", inner, y), ("UPLOAD", pats[3], y), y)
    text = query_text(pats, node)
    e = Engine(0)
    try:
        Ctx.wrap(e.ctx_handle()).tune("drive_min_rows", 0)
        e.load(s, p, o, ("PSO", "POS"))
        got = {}
        for on in (1, 0):
            set_option("driven_swap", on)
            r = e.query(text)
            got[on] = (r.headers(), r.columns(), r.stats())
            r.free()
    finally:
        set_option("driven_swap", 1)
        e.close()
    (h1, c1, st1), (h0, c0, st0) = got[1], got[0]
    print(f"rows {st1['rows']} join rows {st1['join_rows']} driven {st1['driven_joins']} "
          f"swap rows {st1['swap_rows']} / {st0['swap_rows']}", flush=True)
    assert st1["driven_joins"] == 1 and st1["swap_rows"] < st0["swap_rows"]
    assert h1 == h0
    for u, w in zip(c1, c0):
        assert np.array_equal(u, w)
    ex = plans.OracleExecutor(s, p, o)
    rel = ex.run(node)
    assert O.bag_digest(h1, c1) == O.bag_digest(rel.headers, ex.columns(rel))


def test_join_capped(ctx, store):
    """vd_join_capped: the join when its output fits max_rows (equal to
    vd_join's and the oracle's), nothing but the size beyond."""
    idx, ix = store
    b1, e1 = O.lookup(ix["POS"], 1)
    b2, e2 = O.lookup(ix["PSO"], 2)
    want = O.join(O.select(ix["POS"], b1, e1, 3), O.select(ix["PSO"], b2, e2, 3), 1)
    n = len(want[0])
    L = idx.select("POS", b1, e1, 3)
    R = idx.select("PSO", b2, e2, 3)
    out, rows = L.join_capped(R, n - 1)
    assert out is None and rows == n
    out, rows = L.join_capped(R, n)
    assert rows == n
    check_equal(out.columns(), want)
    assert L.join_capped(R, 0)[1] == n
    for x in (L, R, out):
        x.free()


@pytest.mark.parametrize("ny_cols", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("repeat", [False, True])
def test_driven_record_gather(ctx, ny_cols, repeat):
    """Y with 1 to 5 gathered columns (records of 2 or 4 words for 2 to 4 of
    them, one of them taken twice when `repeat`; plain columns for 1 and
    5): the output equals the oracle's join of D's key with Y followed by
    the stable sort on D's first column."""
    from vedas_amd._native import Rel
    rng = np.random.default_rng(100 + ny_cols)
    ny, nd, nk = 40_000, 30_000, 20_000
    ykey = np.sort(rng.integers(0, nk, ny)).astype(np.uint32)
    ycols = [ykey] + [rng.integers(0, 1 << 32, ny, dtype=np.uint64).astype(np.uint32)
                      for _ in range(ny_cols)]
    # D: (v, key) sorted on v then key, distinct rows
    dv = rng.integers(0, 5000, nd).astype(np.uint32)
    dk = rng.integers(0, nk, nd).astype(np.uint32)
    dd = np.unique(np.stack([dv, dk], 1), axis=0)
    dcols = [np.ascontiguousarray(dd[:, 0]), np.ascontiguousarray(dd[:, 1])]
    src = [-1, -2] + list(range(1, ny_cols + 1))
    if repeat:
        src.append(1)
    # the oracle: join D's (key, v) with Y on the key, then the output's
    # columns [v, key, Y...] sorted stably on v (the join is key-major)
    jw = O.join(O.sort_col0([dcols[1], dcols[0]]), ycols, 1)   # [key, v, y1..]
    cols = {-1: jw[1], -2: jw[0]}
    for c in range(1, ny_cols + 1):
        cols[c] = jw[1 + c]
    want = O.sort_col0([cols[c] for c in src])
    D = Rel.from_columns(ctx, dcols)
    Y = Rel.from_columns(ctx, ycols)
    print(f"record gather ny_cols={ny_cols} repeat={repeat}", flush=True)
    out, lo, hi = D.join_driven(1, Y, src, key_hi=nk - 1)
    check_equal(out.columns(), want)
    for x in (D, Y, out):
        x.free()
