"""Pins the oracle (C restatement) with checks that do not share its code:
Python sorting, linear scans, nested-loop joins, a naive BGP evaluator and
the committed golden fixtures.  CPU only."""
import itertools
import os

import numpy as np
import pytest

from oracle import pyoracle as O
from tests import plans

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
PERM_NAMES = ["SPO", "SOP", "PSO", "POS", "OSP", "OPS"]


def rand_triples(n, seed, ns=30, npred=5, no=40):
    rng = np.random.default_rng(seed)
    s = rng.integers(1, ns + 1, n).astype(np.uint32)
    p = rng.integers(1, npred + 1, n).astype(np.uint32)
    o = rng.integers(1, no + 1, n).astype(np.uint32)
    return s, p, o


@pytest.mark.parametrize("perm", PERM_NAMES)
@pytest.mark.parametrize("seed", [0, 1])
def test_build_perm_is_sorted_tuples(perm, seed):
    s, p, o = rand_triples(500, seed)  # duplicates included
    ix = O.build_perm(s, p, o, perm)
    cols = (s, p, o)
    xi, yi, zi = O.PERM_COLS[perm]
    tup = sorted(zip(cols[xi].tolist(), cols[yi].tolist(), cols[zi].tolist()))
    assert ix["y"].tolist() == [t[1] for t in tup]
    assert ix["z"].tolist() == [t[2] for t in tup]
    l1 = O.groupby_starts([t[0] for t in tup])
    assert ix["l1v"].tolist() == [k for k, _ in l1]
    assert ix["l1o"].tolist() == [i for _, i in l1]
    l2 = O.groupby_starts([(t[0], t[1]) for t in tup])
    assert ix["l2v"].tolist() == [k[1] for k, _ in l2]
    assert ix["l2o"].tolist() == [i for _, i in l2]


def test_build_perm_empty():
    e = np.zeros(0, np.uint32)
    ix = O.build_perm(e, e, e, "PSO")
    assert all(len(v) == 0 for v in ix.values())


@pytest.mark.parametrize("perm", ["PSO", "POS", "SPO", "OSP"])
def test_lookup_matches_linear_scan(perm):
    s, p, o = rand_triples(400, 7)
    ix = O.build_perm(s, p, o, perm)
    cols = (s, p, o)
    xi, yi, _ = O.PERM_COLS[perm]
    tup = sorted(zip(cols[xi].tolist(), cols[yi].tolist(), cols[O.PERM_COLS[perm][2]].tolist()))
    xs = sorted(set(cols[xi].tolist()))
    probes = xs + [0, max(xs) + 1, 10_000]
    for k1 in probes:
        rows = [i for i, t in enumerate(tup) if t[0] == k1]
        b, e = O.lookup(ix, k1)
        if rows:
            assert (b, e) == (rows[0], rows[-1] + 1)
        else:
            assert b == e
        for k2 in list(range(0, 45, 3)):
            rows2 = [i for i, t in enumerate(tup) if t[0] == k1 and t[1] == k2]
            b2, e2 = O.lookup(ix, k1, k2)
            if rows2:
                assert (b2, e2) == (rows2[0], rows2[-1] + 1)
            else:
                assert b2 == e2


def test_select_matches_filter():
    s, p, o = rand_triples(600, 3)
    ix = O.build_perm(s, p, o, "PSO")
    tup = sorted(zip(p.tolist(), s.tolist(), o.tolist()))
    for pred in range(1, 6):
        b, e = O.lookup(ix, pred)
        rng = tup[b:e]
        y, z = O.select(ix, b, e, 3)
        assert list(zip(y.tolist(), z.tolist())) == [(t[1], t[2]) for t in rng]
        (zz,) = O.select(ix, b, e, 1)
        assert zz.tolist() == [t[2] for t in rng]
        (d,) = O.select(ix, b, e, 2, distinct=True)
        assert d.tolist() == [k for k, _ in itertools.groupby(t[1] for t in rng)]
        lo, hi = 5, 20
        y, z = O.select(ix, b, e, 3, clip=(lo, hi))
        assert list(zip(y.tolist(), z.tolist())) == [(t[1], t[2]) for t in rng if lo <= t[1] <= hi]
        skip = [(3, 6), (12, 12), (25, 27)]
        y, z = O.select(ix, b, e, 3, clip=(2, 28), skip=skip)
        keep = [(t[1], t[2]) for t in rng if 2 <= t[1] <= 28
                and not any(a <= t[1] <= c for a, c in skip)]
        assert list(zip(y.tolist(), z.tolist())) == keep
        y, z = O.select(ix, b, e, 3, clip=(9, 4))
        assert len(y) == 0


def nested_loop_join(L, R, nk):
    """Left-major bag join by nested loops; R rows in their order."""
    out = []
    nl, nr = len(L[0]), len(R[0])
    for i in range(nl):
        kl = tuple(int(L[c][i]) for c in range(nk))
        for j in range(nr):
            if tuple(int(R[c][j]) for c in range(nk)) == kl:
                out.append(tuple(int(L[c][i]) for c in range(len(L)))
                           + tuple(int(R[c][j]) for c in range(nk, len(R))))
    return out


@pytest.mark.parametrize("nk", [1, 2])
@pytest.mark.parametrize("seed", [0, 5, 9])
def test_join_matches_nested_loop(nk, seed):
    rng = np.random.default_rng(seed)
    def rel(n, ncols):
        cols = [rng.integers(0, 12, n).astype(np.uint32) for _ in range(ncols)]
        return O.sort_lex(cols) if nk == 2 else O.sort_col0(cols)
    L, R = rel(80, 3), rel(120, 2 + nk - 1)
    got = O.join(L, R, nk)
    want = nested_loop_join(L, R, nk)
    assert list(zip(*[c.tolist() for c in got])) == want
    assert O.join([np.zeros(0, np.uint32)] * 2, R, nk)[0].size == 0


def test_sorts_and_unique():
    rng = np.random.default_rng(11)
    cols = [rng.integers(0, 6, 300).astype(np.uint32) for _ in range(3)]
    got = O.sort_col0(cols)
    rows = list(zip(*[c.tolist() for c in cols]))
    assert list(zip(*[c.tolist() for c in got])) == sorted(rows, key=lambda r: r[0])
    got = O.sort_lex(cols)
    assert list(zip(*[c.tolist() for c in got])) == sorted(rows)
    srt = O.sort_lex(cols)
    got = O.unique_adjacent(srt)
    assert list(zip(*[c.tolist() for c in got])) == [k for k, _ in itertools.groupby(sorted(rows))]
    got = O.unique_adjacent(cols)  # unsorted: adjacent only
    assert list(zip(*[c.tolist() for c in got])) == [k for k, _ in itertools.groupby(rows)]


def small_store(seed=4):
    s, p, o = rand_triples(300, seed, ns=25, npred=4, no=25)
    # distinct triples, as the parity contract asks
    uniq = sorted(set(zip(s.tolist(), p.tolist(), o.tolist())))
    s, p, o = (np.array(c, np.uint32) for c in zip(*uniq))
    return s, p, o


@pytest.mark.parametrize("shape", ["star2", "star3", "chain3", "snow"])
def test_plan_equals_naive_bgp(shape):
    s, p, o = small_store()
    if shape == "star2":
        pats, node = plans.star_plan([1, 2])
    elif shape == "star3":
        pats, node = plans.star_plan([1, 2, 3])
    elif shape == "chain3":
        pats, node = plans.chain_plan([1, 2, 3])
    else:
        pats, node = plans.snowflake_plan([1, 2], [3, 4])
    ex = plans.OracleExecutor(s, p, o)
    rel = ex.run(node)
    names, rows = O.canonical(rel.headers, ex.columns(rel))
    vars_, want = O.naive_bgp(s, p, o, pats)
    assert names == vars_
    assert rows == want
    assert len(rows) > 0


@pytest.mark.parametrize("shape", ["star3", "chain3", "snow", "bound"])
def test_run_plan_equals_naive_bgp(shape):
    """The oracle plan walk used by the benchmark's CPU baseline."""
    s, p, o = small_store()
    if shape == "star3":
        pats, node = plans.star_plan([1, 2, 3])
    elif shape == "chain3":
        pats, node = plans.chain_plan([1, 2, 3])
    elif shape == "snow":
        pats, node = plans.snowflake_plan([1, 2], [3, 4])
    else:
        pats = [("?s", int(p[0]), int(o[0])), ("?s", 2, "?y"), (int(s[3]), 3, "?s")]
        node = ("JOIN", ("JOIN", ("UPLOAD", 0, "?s"), ("UPLOAD", 1, "?s"), "?s"),
                ("UPLOAD", 2, "?s"), "?s")
    ix = {perm: O.build_perm(s, p, o, perm) for perm in ("PSO", "POS")}
    st = {}
    hdr, cols = O.run_plan(ix, pats, node, st)
    assert O.canonical(hdr, cols) == O.naive_bgp(s, p, o, pats)
    assert st["join_rows"] >= len(cols[0])


def test_bound_constant_patterns():
    s, p, o = small_store()
    obj = int(o[0])
    pats = [("?s", int(p[0]), obj), ("?s", 2, "?y")]
    node = ("JOIN", ("UPLOAD", pats[0], "?s"), ("UPLOAD", pats[1], "?s"), "?s")
    ex = plans.OracleExecutor(s, p, o)
    rel = ex.run(node)
    names, rows = O.canonical(rel.headers, ex.columns(rel))
    assert (names, rows) == O.naive_bgp(s, p, o, pats)


def test_golden_fixtures():
    path = os.path.join(GOLDEN, "index_small.npz")
    g = np.load(path, allow_pickle=False)
    for perm in ["SPO", "PSO", "POS", "OSP"]:
        ix = O.build_perm(g["s"], g["p"], g["o"], perm)
        for k in ("l1v", "l1o", "l2v", "l2o", "y", "z"):
            assert np.array_equal(ix[k], g[f"{perm}_{k}"]), (perm, k)
    j = np.load(os.path.join(GOLDEN, "join_small.npz"), allow_pickle=False)
    got = O.join([j["l0"], j["l1"]], [j["r0"], j["r1"]], 1)
    for c in range(3):
        assert np.array_equal(got[c], j[f"out{c}"])


def test_planted_mistake_is_caught():
    """A wrong join (right rows in reverse order) must fail the comparison
    the GPU tests use."""
    rng = np.random.default_rng(2)
    L = O.sort_col0([rng.integers(0, 5, 40).astype(np.uint32), rng.integers(0, 9, 40).astype(np.uint32)])
    R = O.sort_col0([rng.integers(0, 5, 40).astype(np.uint32), rng.integers(0, 9, 40).astype(np.uint32)])
    good = O.join(L, R, 1)
    bad = [c.copy() for c in good]
    bad[2] = bad[2][::-1].copy()
    want = nested_loop_join(L, R, 1)
    assert list(zip(*[c.tolist() for c in good])) == want
    assert list(zip(*[c.tolist() for c in bad])) != want


# ---- checkers used for full-size GPU results --------------------------------
def test_bag_digest_is_the_bag_of_rows():
    """Equal multisets (Counter) <=> equal digests, on small relations with
    duplicate rows; row order does not matter, column names order columns."""
    from collections import Counter
    rng = np.random.default_rng(4)
    base = [rng.integers(0, 4, 300).astype(np.uint32) for _ in range(3)]
    perm = rng.permutation(300)
    shuffled = [c[perm] for c in base]
    d0 = O.bag_digest(["?a", "?b", "?c"], base)
    assert O.bag_digest(["?a", "?b", "?c"], shuffled) == d0
    # the same columns under headers listed in another order
    assert O.bag_digest(["?c", "?a", "?b"], [base[2], base[0], base[1]]) == d0
    rows = Counter(zip(*(c.tolist() for c in base)))
    for trial in range(200):
        other = [c.copy() for c in base]
        i = int(rng.integers(0, 300))
        j = int(rng.integers(0, 3))
        other[j][i] = (other[j][i] + 1 + trial % 3) % 5
        same = Counter(zip(*(c.tolist() for c in other))) == rows
        assert (O.bag_digest(["?a", "?b", "?c"], other) == d0) == same
    # a duplicated row instead of another row (same row count)
    dup = [np.concatenate([c[:-1], c[:1]]) for c in base]
    assert (O.bag_digest(["?a", "?b", "?c"], dup) == d0) == (
        Counter(zip(*(c.tolist() for c in dup))) == rows)
    # two columns swapped under the same headers
    assert O.bag_digest(["?a", "?b", "?c"], [base[1], base[0], base[2]]) != d0
    assert O.bag_digest(["?a"], [np.zeros(0, np.uint32)]) == (("?a",), 0, 0, 0)


@pytest.mark.parametrize("perm", PERM_NAMES)
def test_check_perm_and_perm_digest(perm):
    s, p, o = rand_triples(3000, 9)
    ix = O.build_perm(s, p, o, perm)
    O.check_perm(ix)
    assert O.perm_digest(ix, perm) == O.triples_digest(s, p, o)
    golden = np.load(os.path.join(GOLDEN, "index_small.npz"))
    if f"{perm}_y" in golden:
        gix = {k: golden[f"{perm}_{k}"] for k in ("l1v", "l1o", "l2v", "l2o", "y", "z")}
        O.check_perm(gix)
        assert O.perm_digest(gix, perm) == O.triples_digest(golden["s"], golden["p"], golden["o"])

    def broken(fn):
        b = {k: v.copy() for k, v in ix.items()}
        fn(b)
        return b

    heads = ix["l2o"]
    r = int(heads[len(heads) // 2])  # a row that starts an (X, Y) run
    faults = [
        lambda b: b["z"].__setitem__(slice(r - 1, r + 1), b["z"][r - 1:r + 1][::-1].copy())
        if b["y"][r - 1] == b["y"][r] else b["y"].__setitem__(r, b["y"][r - 1] - 1),
        lambda b: b.__setitem__("l2o", np.delete(b["l2o"], len(heads) // 2)),
        lambda b: b["l2v"].__setitem__(1, b["l2v"][1] + 1),
        lambda b: b["l1v"].__setitem__(slice(0, 2), b["l1v"][1::-1].copy()),
        lambda b: b["l1o"].__setitem__(1, b["l1o"][1] + 1),
        lambda b: b.__setitem__("l2o", np.append(b["l2o"], len(b["y"]) - 1)),
        lambda b: (b.__setitem__("l2o", np.append(b["l2o"], len(b["y"]) - 1)),
                   b.__setitem__("l2v", np.append(b["l2v"], b["y"][-1]))),
    ]
    for k, f in enumerate(faults):
        b = broken(f)
        with pytest.raises(AssertionError):
            O.check_perm(b)
            # a fault that keeps the invariants must change the multiset
            assert O.perm_digest(b, perm) == O.triples_digest(s, p, o), k
    # a changed value with the order kept: invariants hold, multiset differs
    b = broken(lambda b: b["z"].__setitem__(-1, b["z"][-1] + 1))
    O.check_perm(b)
    assert O.perm_digest(b, perm) != O.triples_digest(s, p, o)


# ---- join-variable bounds (pre-scan, clip, after-join update) ---------------
def brute_bounds(s, p, o, patterns):
    """Per join variable, the intersection over the patterns that hold it
    of [min, max] of its values among the triples matching the pattern's
    constants, straight from the triples; [0, 0] for a variable whose
    intersection is empty, and for every variable when a pattern matches
    nothing."""
    cols = (s, p, o)
    count = {}
    for pat in patterns:
        for t in pat:
            if isinstance(t, str):
                count[t] = count.get(t, 0) + 1
    lohi = {v: [0, 0xFFFFFFFF] for v, c in count.items() if c > 1}
    empty = False
    for pat in patterns:
        m = np.ones(len(s), bool)
        for c, t in zip(cols, pat):
            if not isinstance(t, str):
                m &= c == t
        if not m.any():
            empty = True
            continue
        for c, t in zip(cols, pat):
            if isinstance(t, str) and t in lohi:
                lohi[t][0] = max(lohi[t][0], int(c[m].min()))
                lohi[t][1] = min(lohi[t][1], int(c[m].max()))
    if empty:
        return {v: (0, 0) for v in lohi}
    return {v: (lo, hi) if lo <= hi else (0, 0) for v, (lo, hi) in lohi.items()}


def bounded_queries(s, p, o):
    return [plans.star_plan([1, 2, 3]), plans.chain_plan([1, 2, 3]),
            plans.snowflake_plan([1, 2], [3, 4]),
            ([("?s", int(p[0]), int(o[0])), ("?s", 2, "?y"), (int(s[3]), 3, "?s")],
             ("JOIN", ("JOIN", ("UPLOAD", 0, "?s"), ("UPLOAD", 1, "?s"), "?s"),
              ("UPLOAD", 2, "?s"), "?s")),
            # a constant no triple has: every bound collapses to [0, 0]
            ([("?a", 1, "?b"), ("?b", 2, 999_999)],
             ("JOIN", ("UPLOAD", 0, "?b"), ("UPLOAD", 1, "?b"), "?b"))]


def test_prescan_bounds_equal_brute_force():
    s, p, o = small_store()
    # objects of predicate 1 beyond every subject: the chain's ?b narrows
    o = o.copy()
    o[p == 1] += 40
    allp = ("SPO", "PSO", "POS", "OSP", "SOP", "OPS")
    ix = {perm: O.build_perm(s, p, o, perm) for perm in allp}
    narrowed = 0
    for pats, _ in bounded_queries(s, p, o):
        want = brute_bounds(s, p, o, pats)
        assert O.prescan_bounds(ix, pats) == want, pats
        # only the permutations the default store has
        assert O.prescan_bounds({k: ix[k] for k in ("PSO", "POS")}, pats) == want
        narrowed += any(b != O.FULL_BOUND for b in want.values())
    assert narrowed >= 3
    # the reference's updateBoundDict rule on its own
    b = {"?v": (5, 10)}
    O.update_bound(b, "?v", 7, 20)
    assert b["?v"] == (7, 10)
    O.update_bound(b, "?v", 11, 30)  # disjoint
    assert b["?v"] == (0, 0)


def test_run_plan_with_bounds_is_result_neutral_and_smaller():
    s, p, o = small_store()
    o = o.copy()
    o[p == 1] += 40
    ix = {perm: O.build_perm(s, p, o, perm) for perm in ("PSO", "POS")}
    smaller = 0
    for pats, node in bounded_queries(s, p, o):
        st0, st1 = {}, {}
        h0, c0 = O.run_plan(ix, pats, node, st0)
        h1, c1 = O.run_plan(ix, pats, node, st1, bounds=True)
        want = O.naive_bgp(s, p, o, pats)
        assert O.canonical(h0, c0) == want
        assert O.canonical(h1, c1) == want
        assert st1["select_rows"] <= st0["select_rows"]
        assert st1["join_rows"] <= st0["join_rows"]
        smaller += st1["select_rows"] < st0["select_rows"]
    assert smaller >= 2


def test_too_narrow_bound_is_caught():
    """A clip narrower than the true bound changes the result: the
    comparison with the naive evaluator would catch such a mistake."""
    s, p, o = small_store()
    ix = {perm: O.build_perm(s, p, o, perm) for perm in ("PSO", "POS")}
    pats, node = plans.chain_plan([1, 2, 3])
    real = O.prescan_bounds
    try:
        def narrow(ixp, pp):
            b = real(ixp, pp)
            return {v: (lo, max(lo, hi - 3)) for v, (lo, hi) in b.items()}
        O.prescan_bounds = narrow
        h, c = O.run_plan(ix, pats, node, bounds=True)
    finally:
        O.prescan_bounds = real
    assert O.canonical(h, c) != O.naive_bgp(s, p, o, pats)


# ---- equal-width / equal-depth histograms -----------------------------------
def check_hist_width(col, width, limit, hist):
    """Brute force: contiguous buckets of `width` IDs from the first key
    below the limit through the last, each counting its keys."""
    lowers, uppers, freq = (np.asarray(x, np.int64) for x in hist)
    keys = np.asarray(col, np.int64)
    keys = keys[keys < limit]
    if len(keys) == 0:
        assert len(lowers) == 0
        return
    assert lowers[0] == keys.min() and uppers[-1] == keys.max()
    assert (np.diff(lowers) == width).all() and (uppers[:-1] - lowers[:-1] == width - 1).all()
    for lo, hi, f in zip(lowers, uppers, freq):
        assert f == np.count_nonzero((keys >= lo) & (keys <= hi))


def check_hist_depth(col, depth, limit, hist):
    """Buckets partition the sorted rows below the limit in order; a
    bucket from row s ends at the last row of the key held by row
    min(s + depth, m - 1) (found with searchsorted on the key, not by
    walking); lower / upper are the keys at its ends."""
    lowers, uppers, freq = (np.asarray(x, np.int64) for x in hist)
    keys = np.asarray(col, np.int64)
    m = int(np.searchsorted(keys, limit, "left"))
    assert int(freq.sum()) == m
    s = 0
    for lo, hi, f in zip(lowers, uppers, freq):
        e0 = min(s + depth, m - 1)
        end = int(np.searchsorted(keys, keys[e0], "right")) - 1
        end = min(end, m - 1)
        assert s + f - 1 == end and lo == keys[s] and hi == keys[e0]
        s = end + 1
    assert s == m


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_histograms_equal_brute_force(seed):
    rng = np.random.default_rng(seed)
    # Zipf-like keys (long runs at small IDs) plus "literals" above a limit
    keys = np.sort(np.concatenate([rng.zipf(1.3, 20_000).clip(1, 5_000),
                                   rng.integers(60_000, 70_000, 3000)]).astype(np.uint32))
    for limit in (50_000 - 1, 0xFFFFFFFF, 1, 3):
        for w in (1, 7, 100):
            h = O.hist_eq_width(keys, w, limit)
            check_hist_width(keys, w, limit, h)
        for d in (1, 10, 1000, 50_000):
            h = O.hist_eq_depth(keys, d, limit)
            check_hist_depth(keys, d, limit, h)
            if len(h[2]) > 1:
                assert min(h[2][:-1]) >= d + 1  # the reference's depth + 1 rows
    # the span a multiple of the width: the last key still has its bucket
    h = O.hist_eq_width(np.array([1, 50, 101], np.uint32), 100, 1000)
    assert h == ([1, 101], [100, 101], [2, 1])
    assert O.hist_eq_width(np.array([], np.uint32), 10, 99) == ([], [], [])
