"""GPU parity of the batched (multi-input) operators: every input's result
equals the oracle's for that input alone."""
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


def rand_rel(rng, n, ncols, keymax):
    return O.sort_col0([rng.integers(0, keymax, n).astype(np.uint32) for _ in range(ncols)])


@pytest.mark.parametrize("nv", [1024, 2048, 4096])
def test_join_multi(ctx, nv):
    from vedas_amd._native import Rel, join_multi
    ctx.tune("join_nv", nv)
    rng = np.random.default_rng(nv)
    shapes = [(0, 10, 5), (10, 0, 5), (1, 1, 1), (3000, 5000, 40), (20_000, 10, 9000), (5, 70_000, 3),
              (100_000, 90_000, 300_000), (17, 33, 4)] * 3
    Ls, Rs, want = [], [], []
    for nl, nr, km in shapes:
        L = rand_rel(rng, nl, int(rng.integers(1, 4)), km)
        R = rand_rel(rng, nr, int(rng.integers(1, 4)), km)
        Ls.append(Rel.from_columns(ctx, L))
        Rs.append(Rel.from_columns(ctx, R))
        want.append(O.join(L, R, 1))
    outs = join_multi(ctx, Ls, Rs)
    for j, (o, w) in enumerate(zip(outs, want)):
        assert o.nrows == len(w[0]), j
        assert o.ncols == len(w), j
        for g, ww in zip(o.columns(), w):
            assert np.array_equal(g, ww), j
    for r in Ls + Rs + outs:
        r.free()
    ctx.tune("join_nv", 2048)


def test_join_multi_many_tiny_pairs(ctx):
    """More pairs than one chunk of the batched join (4096): 5000 joins of
    0-6 rows each, results equal to the oracle's pair by pair."""
    from vedas_amd._native import Rel, join_multi
    rng = np.random.default_rng(77)
    Ls, Rs, want = [], [], []
    for _ in range(5000):
        L = rand_rel(rng, int(rng.integers(0, 7)), 2, 5)
        R = rand_rel(rng, int(rng.integers(0, 7)), 2, 5)
        Ls.append(Rel.from_columns(ctx, L))
        Rs.append(Rel.from_columns(ctx, R))
        want.append(O.join(L, R, 1))
    outs = join_multi(ctx, Ls, Rs)
    for j, (o, w) in enumerate(zip(outs, want)):
        assert o.nrows == len(w[0]), j
        for g, ww in zip(o.columns(), w):
            assert np.array_equal(g, ww), j
    for r in Ls + Rs + outs:
        r.free()


def test_sort_and_unique_multi(ctx):
    from vedas_amd._native import Rel, sort_col0_multi, unique_adjacent_multi
    rng = np.random.default_rng(1)
    cases = [(0, 2, 9), (1, 1, 9), (2, 3, 2), (5000, 2, 40), (70_000, 4, 2**24), (300, 1, 3),
             (9000, 3, 2**27)] * 4
    rels, cols = [], []
    for n, nc, km in cases:
        c = [rng.integers(0, km, n).astype(np.uint32) for _ in range(nc)]
        cols.append(c)
        rels.append(Rel.from_columns(ctx, c))
    sort_col0_multi(ctx, rels)
    for r, c in zip(rels, cols):
        for g, w in zip(r.columns(), O.sort_col0(c)):
            assert np.array_equal(g, w)
    unique_adjacent_multi(ctx, rels)
    for r, c in zip(rels, cols):
        w = O.unique_adjacent(O.sort_col0(c))
        assert r.nrows == len(w[0])
        for g, ww in zip(r.columns(), w):
            assert np.array_equal(g, ww)
    for r in rels:
        r.free()


def test_batch_equals_single_queries():
    """A seeded mix of stars, chains and bound-constant queries: every
    query's batched result equals its single-query result (canonical
    order), and the oracle's for a sample."""
    from tests import plans
    from vedas_amd._native import gen_triples
    from vedas_amd.engine import Engine, query_text
    from vedas_amd.workload import query_mix
    prm = dict(n=2_000_000, n_entities=200_000, subject_zipf=0.8, n_predicates=64,
               pred_zipf=1.1, literal_frac=0.4, literal_base=50_000_000, n_literals=1_000_000)
    s, p, o = gen_triples(prm["n"], seed=11, **{k: v for k, v in prm.items() if k != "n"},
                          distinct=True)
    mix = query_mix(48, 5, prm, s, p, o, lo=10, hi=3e5)
    texts = [query_text(pats, node) for _, pats, node in mix]
    e = Engine(0)
    try:
        e.load(s, p, o, ("SPO", "PSO", "POS", "OSP"))
        single = []
        for t in texts:
            r = e.query(t)
            single.append(O.canonical(r.headers(), r.columns()))
            r.free()
        res, st = e.batch(texts)
        batched = [O.canonical(r.headers(), r.columns()) for r in res]
        for r in res:
            r.free()
        # the same queries again, mixed in a different order
        order = list(range(len(texts)))[::-1]
        res2, _ = e.batch([texts[i] for i in order])
        again = {i: O.canonical(r.headers(), r.columns()) for i, r in zip(order, res2)}
        for r in res2:
            r.free()
    finally:
        e.close()
    assert st["rounds"] >= 2 and st["calls"] < 10 * st["rounds"]
    nonempty = 0
    for i, (a, b) in enumerate(zip(single, batched)):
        assert a == b, (i, mix[i][0])
        assert again[i] == a, i
        nonempty += len(a[1]) > 0
    assert nonempty > len(texts) // 2
    ex = plans.OracleExecutor(s, p, o)
    for i in range(0, len(mix), 6):
        _, pats, node = mix[i]
        onode = _resolve(node, pats)
        rel = ex.run(onode)
        assert O.canonical(rel.headers, ex.columns(rel)) == single[i], i


def _resolve(node, pats):
    """Plan with pattern indices -> plan with pattern tuples (tests.plans form)."""
    if node[0] == "UPLOAD":
        return ("UPLOAD", pats[node[1]], node[2])
    if node[0] == "JOIN":
        return ("JOIN", _resolve(node[1], pats), _resolve(node[2], pats), node[3])
    return ("INDEXSWAP", _resolve(node[1], pats), node[2])


def test_select_multi(ctx):
    from vedas_amd._native import Index, gen_triples, select_multi
    s, p, o = gen_triples(300_000, seed=5)
    idx = Index(ctx, s, p, o, ["SPO", "PSO", "POS", "OSP"])
    try:
        orc = {k: O.build_perm(s, p, o, k) for k in ["SPO", "PSO", "POS", "OSP"]}
        rng = np.random.default_rng(3)
        jobs = []
        for t in rng.integers(0, len(s), 200):
            jobs.append(("PSO", int(p[t]), None, 3))
            jobs.append(("POS", int(p[t]), int(o[t]), 1))
            jobs.append(("SPO", int(s[t]), None, 3))
        jobs.append(("PSO", 10_000, None, 3))  # absent predicate
        outs = select_multi(ctx, idx, jobs)
        for (perm, k1, k2, cols), r in zip(jobs, outs):
            ix = orc[perm]
            b, e = O.lookup(ix, k1) if k2 is None else O.lookup(ix, k1, k2)
            want = O.select(ix, b, e, cols)
            got = r.columns()
            assert len(got) == len(want)
            for g, w in zip(got, want):
                assert np.array_equal(g, w)
            r.free()
    finally:
        idx.free()


def directory_pair(rng, nl=6000, hot=1_000_000, tail=1_000_000, span=1_000_000):
    """A sparse left side against a dense right side with a hot key: every
    2048-row left tile spans the whole key range, so its right range (2M
    rows) is staged as a directory of every k-th key (k ~ 250-500) and the
    exact bounds come from searches in the gaps between directory entries:
    keys equal to the hot key, keys inside and at the ends of gaps, and keys
    absent from the right side."""
    rk = np.concatenate([np.full(hot, span // 2), rng.integers(0, span, tail)])
    lk = np.concatenate([rng.integers(0, span, nl - 4), [span // 2] * 2, [0, span + 7]])
    R = O.sort_col0([rk.astype(np.uint32), rng.integers(0, 1 << 30, len(rk)).astype(np.uint32)])
    L = O.sort_col0([lk.astype(np.uint32), rng.integers(0, 1 << 30, len(lk)).astype(np.uint32)])
    return L, R


def test_join_multi_directory_gaps(ctx):
    from vedas_amd._native import Rel, join_multi
    rng = np.random.default_rng(5)
    Ls, Rs, want = [], [], []
    for nl in (6000, 300, 2049):
        L, R = directory_pair(rng, nl=nl)
        Ls.append(Rel.from_columns(ctx, L))
        Rs.append(Rel.from_columns(ctx, R))
        want.append(O.join(L, R, 1))
    outs = join_multi(ctx, Ls, Rs)
    for j, (o, w) in enumerate(zip(outs, want)):
        assert o.nrows == len(w[0]) and o.nrows > 2_000_000, j
        for g, ww in zip(o.columns(), w):
            assert np.array_equal(g, ww), j
    for r in Ls + Rs + outs:
        r.free()
