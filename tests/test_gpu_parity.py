"""GPU parity: every C-ABI operator against the oracle on the same seeded
inputs, bit-exact (integer work).  Needs an MI355X."""
import numpy as np
import pytest

from oracle import pyoracle as O
from tests import plans

pytestmark = pytest.mark.gpu

ALL_PERMS = ["SPO", "SOP", "PSO", "POS", "OSP", "OPS"]


@pytest.fixture(scope="module")
def ctx():
    from vedas_amd._native import Ctx
    c = Ctx(0)
    yield c
    c.close()


def cfg0_triples(n=200_000, seed=42, distinct=True):
    from vedas_amd._native import gen_triples
    return gen_triples(n, seed=seed, n_entities=100_000, n_predicates=64, pred_zipf=1.1,
                       literal_frac=0.5, literal_base=50_000_000, n_literals=500_000,
                       distinct=distinct)


def rand_triples(n, seed, ns=300, npred=7, no=400):
    rng = np.random.default_rng(seed)
    return (rng.integers(1, ns, n).astype(np.uint32), rng.integers(1, npred, n).astype(np.uint32),
            rng.integers(1, no, n).astype(np.uint32))


def check_index(ctx, s, p, o, perms=ALL_PERMS):
    from vedas_amd._native import Index
    idx = Index(ctx, s, p, o, perms)
    try:
        for perm in perms:
            got = idx.arrays(perm)
            want = O.build_perm(s, p, o, perm)
            for k in ("l1v", "l1o", "l2v", "l2o", "y", "z"):
                assert np.array_equal(got[k], want[k]), (perm, k, len(got[k]), len(want[k]))
    finally:
        idx.free()


@pytest.mark.parametrize("n", [0, 1, 2, 4095, 4096, 4097, 65537, 300_001])
def test_index_build_sizes(ctx, n):
    print(f"index build n={n}", flush=True)
    s, p, o = rand_triples(n, n + 1)
    check_index(ctx, s, p, o)


def test_index_build_cfg0(ctx):
    s, p, o = cfg0_triples()
    check_index(ctx, s, p, o, ["SPO", "PSO", "POS", "OSP"])


def test_index_build_duplicates_and_runs(ctx):
    n = 50_000
    s = np.full(n, 7, np.uint32)
    p = np.full(n, 3, np.uint32)
    o = (np.arange(n) % 5 + 1).astype(np.uint32)  # heavy duplicates
    check_index(ctx, s, p, o)


def test_index_build_wide_ids(ctx):
    rng = np.random.default_rng(3)
    n = 20_000
    s = rng.integers(1, 1 << 24, n).astype(np.uint32)
    p = rng.integers(1, 1 << 9, n).astype(np.uint32)
    o = rng.integers(1, 1 << 31, n).astype(np.uint32)  # 24 + 9 + 31 = 64 bits
    check_index(ctx, s, p, o, ["PSO", "OSP"])


@pytest.mark.parametrize("n", [1, 3, 5000, 70_001])
def test_index_build_over_64_bits(ctx, n):
    """IDs needing more than 64 bits per packed triple take the two-phase
    path (sort (Y,Z) carrying X, then stable by X)."""
    rng = np.random.default_rng(n)
    s = rng.integers(1, 1 << 26, n).astype(np.uint32)
    p = rng.integers(1, 1 << 11, n).astype(np.uint32)
    o = rng.integers(1, 1 << 29, n).astype(np.uint32)
    s[: n // 2] = s[0]  # long X runs
    check_index(ctx, s, p, o)
    big = np.full(n, 0xFFFFFFF0, np.uint32)
    big[::3] = 0xFFFFFFFF
    check_index(ctx, big, big[::-1].copy(), big, ["SPO", "OSP"])


@pytest.fixture(scope="module")
def store(ctx):
    from vedas_amd._native import Index
    s, p, o = cfg0_triples()
    idx = Index(ctx, s, p, o, ["SPO", "PSO", "POS", "OSP"])
    orc = {perm: O.build_perm(s, p, o, perm) for perm in ["SPO", "PSO", "POS", "OSP"]}
    yield s, p, o, idx, orc
    idx.free()


def test_lookup_single_and_batch(ctx, store):
    s, p, o, idx, orc = store
    rng = np.random.default_rng(5)
    import ctypes as C
    from vedas_amd._native import PERMS, VD_UNBOUND
    for perm in ["SPO", "PSO", "POS", "OSP"]:
        ix = orc[perm]
        xs = ix["l1v"]
        k1 = np.concatenate([rng.choice(xs, 300), np.array([0, 1, xs.max() + 1, 2**32 - 2], np.uint32)]).astype(np.uint32)
        ys = ix["y"]
        k2 = np.where(rng.random(len(k1)) < 0.5, VD_UNBOUND, rng.choice(ys, len(k1))).astype(np.uint32)
        want = [O.lookup(ix, int(a)) if b == VD_UNBOUND else O.lookup(ix, int(a), int(b))
                for a, b in zip(k1, k2)]
        for q in range(0, len(k1), 37):
            kk = None if k2[q] == VD_UNBOUND else int(k2[q])
            gb, ge = idx.lookup(perm, int(k1[q]), kk)
            assert (ge - gb) == (want[q][1] - want[q][0])
            if ge > gb:
                assert (gb, ge) == want[q]
        dk1, dk2 = ctx.upload(k1), ctx.upload(k2)
        db, de = ctx.alloc(8 * len(k1)), ctx.alloc(8 * len(k1))
        ctx.check(ctx.L.vd_index_lookup_batch(ctx.h, idx.h, PERMS[perm], C.c_void_p(dk1),
                                              C.c_void_p(dk2), len(k1), C.c_void_p(db),
                                              C.c_void_p(de)))
        gb = ctx.download(db, len(k1), np.uint64)
        ge = ctx.download(de, len(k1), np.uint64)
        for q in range(len(k1)):
            assert ge[q] - gb[q] == want[q][1] - want[q][0], q
            if want[q][1] > want[q][0]:
                assert (gb[q], ge[q]) == want[q]
        for d in (dk1, dk2, db, de):
            ctx.free(d)
    # every permutation in one batched call
    from vedas_amd._native import lookup_multi
    jobs, want = [], []
    for perm in ["SPO", "PSO", "POS", "OSP"]:
        ix = orc[perm]
        for a in list(rng.choice(ix["l1v"], 20)) + [0, 2**32 - 2]:
            b = None if rng.random() < 0.5 else int(rng.choice(ix["y"]))
            jobs.append((perm, int(a), b))
            want.append(O.lookup(ix, int(a)) if b is None else O.lookup(ix, int(a), b))
    got = lookup_multi(ctx, idx, jobs)
    for g, w in zip(got, want):
        assert g[1] - g[0] == w[1] - w[0]
        if w[1] > w[0]:
            assert g == w


def test_select_modes(ctx, store):
    s, p, o, idx, orc = store
    ix = orc["PSO"]
    for pred in [1, 2, 5, 17, 64]:
        b, e = O.lookup(ix, pred)
        for cols in (1, 2, 3):
            rel = idx.select("PSO", b, e, cols)
            want = O.select(ix, b, e, cols)
            got = rel.columns()
            rel.free()
            assert all(np.array_equal(g, w) for g, w in zip(got, want)), (pred, cols)
        rel = idx.select("PSO", b, e, 2, distinct=True)
        got = rel.columns()
        rel.free()
        assert np.array_equal(got[0], O.select(ix, b, e, 2, distinct=True)[0])
        ys = ix["y"][b:e]
        if len(ys):
            lo, hi = int(np.percentile(ys, 20)), int(np.percentile(ys, 70))
            sk = [(lo + 10, lo + 500), (hi - 300, hi - 100)]
            for clip, skip in [((lo, hi), ()), ((lo, hi), sk), (None, sk), ((hi, lo), ())]:
                rel = idx.select("PSO", b, e, 3, clip=clip, skip=skip)
                got = rel.columns()
                rel.free()
                want = O.select(ix, b, e, 3, clip=clip, skip=skip)
                assert all(np.array_equal(g, w) for g, w in zip(got, want)), (pred, clip, skip)


def test_select_views_and_lifetime(ctx):
    """Unfiltered selects read the permutation in place: operators replace
    view columns instead of writing into them, and a view keeps the index
    arrays alive after the index is freed."""
    from vedas_amd._native import Index
    s, p, o = rand_triples(50_000, 11)
    idx = Index(ctx, s, p, o, ["PSO", "POS"])
    ix = O.build_perm(s, p, o, "PSO")
    b, e = O.lookup(ix, 3)
    assert e - b > 100
    want = O.select(ix, b, e, 3)
    r1 = idx.select("PSO", b, e, 3)
    assert r1.is_view(0) and r1.is_view(1)
    r1.swap_cols(0, 1)
    r1.sort_col0()
    assert not r1.is_view(0) and not r1.is_view(1)
    for g, w in zip(r1.columns(), O.sort_col0([want[1], want[0]])):
        assert np.array_equal(g, w)
    r2 = idx.select("PSO", b, e, 3)
    r2.drop_col(0)
    assert r2.ncols == 1 and r2.is_view(0)
    r3 = idx.select("PSO", b, e, 1)
    r3.materialize()
    assert not r3.is_view(0)
    r4 = idx.select("PSO", b, e, 3)  # the sort above left the index as it was
    for g, w in zip(r4.columns(), want):
        assert np.array_equal(g, w)
    ys = want[0]
    lo, hi = int(np.percentile(ys, 25)), int(np.percentile(ys, 75))
    r6 = idx.select("PSO", b, e, 3, clip=(lo, hi))  # one run of rows: a view
    assert r6.is_view(0) and r6.is_view(1)
    for g, w in zip(r6.columns(), O.select(ix, b, e, 3, clip=(lo, hi))):
        assert np.array_equal(g, w)
    r6.free()
    r5 = idx.select("PSO", b, e, 2)
    r5.unique_adjacent()
    assert np.array_equal(r5.columns()[0], O.unique_adjacent([want[0]])[0])
    idx.free()
    for g, w in zip(r4.columns(), want):  # storage kept alive by the views
        assert np.array_equal(g, w)
    assert np.array_equal(r2.columns()[0], want[1])
    assert np.array_equal(r3.columns()[0], O.select(ix, b, e, 1)[0])
    for r in (r1, r2, r3, r4, r5):
        r.free()


def test_select_key_equals_lookup_then_select(ctx, store):
    s, p, o, idx, orc = store
    for perm in ("PSO", "POS"):
        ix = orc[perm]
        for pred in [1, 2, 5, 17, 64, 10**6]:
            b, e = O.lookup(ix, pred)
            ys = ix["y"][b:e]
            lo, hi = (int(np.percentile(ys, 20)), int(np.percentile(ys, 70))) if len(ys) else (0, 2000)
            sk = [(lo + 10, lo + 500), (max(hi - 300, lo + 501), max(hi - 100, lo + 502))]
            for cols, clip, skip, distinct in [(3, None, (), False), (3, (lo, hi), (), False),
                                               (3, (lo, hi), sk, False), (2, (lo, hi), (), True),
                                               (2, None, (), True)]:
                rel, gb, ge = idx.select_key(perm, pred, None, cols, clip, skip, distinct)
                assert (gb, ge) == (b, e) or (b == e and gb == ge)
                want = O.select(ix, b, e, cols, clip=clip, skip=skip, distinct=distinct)
                got = rel.columns()
                rel.free()
                assert all(np.array_equal(g, w) for g, w in zip(got, want)), (perm, pred, cols)
            if len(ys):  # two bound constants, Z only
                y0 = int(ys[len(ys) // 2])
                rel, gb, ge = idx.select_key(perm, pred, y0, 1)
                bb, ee = O.lookup(ix, pred, y0)
                assert (gb, ge) == (bb, ee)
                assert np.array_equal(rel.columns()[0], O.select(ix, bb, ee, 1)[0])
                rel.free()


def test_select_batch(ctx, store):
    import ctypes as C
    from vedas_amd._native import PERMS, VD_UNBOUND
    s, p, o, idx, orc = store
    rng = np.random.default_rng(9)
    nq = 1000
    pick = rng.integers(0, len(s), nq)
    perm = "SPO"
    ix = orc[perm]
    k1 = s[pick].astype(np.uint32)
    k2 = np.where(rng.random(nq) < 0.5, VD_UNBOUND, p[pick]).astype(np.uint32)
    want_y, want_z, sizes = [], [], []
    for a, bb in zip(k1, k2):
        b, e = O.lookup(ix, int(a)) if bb == VD_UNBOUND else O.lookup(ix, int(a), int(bb))
        want_y.append(ix["y"][b:e])
        want_z.append(ix["z"][b:e])
        sizes.append(e - b)
    total = int(sum(sizes))
    dk1, dk2 = ctx.upload(k1), ctx.upload(k2)
    doff = ctx.alloc(8 * (nq + 1))
    dy, dz = ctx.alloc(4 * total), ctx.alloc(4 * total)
    tot = C.c_uint64()
    ctx.check(ctx.L.vd_select_batch(ctx.h, idx.h, PERMS[perm], C.c_void_p(dk1), C.c_void_p(dk2), nq,
                                    C.c_void_p(doff), C.c_void_p(dy), C.c_void_p(dz), total,
                                    C.byref(tot)))
    assert tot.value == total
    off = ctx.download(doff, nq + 1, np.uint64)
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64))
    assert np.array_equal(ctx.download(dy, total), np.concatenate(want_y))
    assert np.array_equal(ctx.download(dz, total), np.concatenate(want_z))
    # too small an output -> VD_ERANGE and the needed size
    rc = ctx.L.vd_select_batch(ctx.h, idx.h, PERMS[perm], C.c_void_p(dk1), C.c_void_p(dk2), nq,
                               C.c_void_p(doff), C.c_void_p(dy), C.c_void_p(dz), max(1, total - 1),
                               C.byref(tot))
    assert rc == -4 and tot.value == total
    for d in (dk1, dk2, doff, dy, dz):
        ctx.free(d)


def test_select_batch_tiny_and_empty_ranges(ctx, store):
    """Tens of thousands of (S, P) selects of a row or two, a third of them
    absent (empty ranges): output tiles then span more queries than the
    gather stages in LDS, and empty ranges sit between rows."""
    import ctypes as C
    from vedas_amd._native import PERMS
    s, p, o, idx, orc = store
    rng = np.random.default_rng(19)
    nq = 40_000
    pick = rng.integers(0, len(s), nq)
    k1 = s[pick].astype(np.uint32)
    k2 = p[pick].astype(np.uint32)
    k2[rng.random(nq) < 0.33] = 65_000  # no such predicate
    ix = orc["SPO"]
    rng_be = [O.lookup(ix, int(a), int(b)) for a, b in zip(k1, k2)]
    total = int(sum(e - b for b, e in rng_be))
    assert total < 3 * nq
    dk1, dk2 = ctx.upload(k1), ctx.upload(k2)
    doff = ctx.alloc(8 * (nq + 1))
    dy, dz = ctx.alloc(4 * total), ctx.alloc(4 * total)
    tot = C.c_uint64()
    ctx.check(ctx.L.vd_select_batch(ctx.h, idx.h, PERMS["SPO"], C.c_void_p(dk1), C.c_void_p(dk2),
                                    nq, C.c_void_p(doff), C.c_void_p(dy), C.c_void_p(dz), total,
                                    C.byref(tot)))
    assert tot.value == total
    assert np.array_equal(ctx.download(dy, total),
                          np.concatenate([ix["y"][b:e] for b, e in rng_be]))
    assert np.array_equal(ctx.download(dz, total),
                          np.concatenate([ix["z"][b:e] for b, e in rng_be]))
    for d in (dk1, dk2, doff, dy, dz):
        ctx.free(d)


def make_rel(rng, n, ncols, keymax, nk=1):
    cols = [rng.integers(0, keymax, n).astype(np.uint32) for _ in range(ncols)]
    return O.sort_lex(cols) if nk == 2 else O.sort_col0(cols)


@pytest.mark.parametrize("nk", [1, 2])
@pytest.mark.parametrize("shape", [(1000, 1500, 50), (20_000, 30_000, 3000), (5000, 7000, 3),
                                   (100_000, 80_000, 200_000), (1, 1, 1)])
def test_join_parity(ctx, nk, shape):
    from vedas_amd._native import Rel
    nl, nr, keymax = shape
    print(f"join nk={nk} shape={shape}", flush=True)
    rng = np.random.default_rng(nl + nr + nk)
    L = make_rel(rng, nl, 3, keymax, nk)
    R = make_rel(rng, nr, 2 + (nk - 1), keymax, nk)
    want = O.join(L, R, nk)
    gl, gr = Rel.from_columns(ctx, L), Rel.from_columns(ctx, R)
    out = gl.join(gr, nk)
    got = out.columns()
    assert out.nrows == len(want[0])
    for c, (g, w) in enumerate(zip(got, want)):
        assert np.array_equal(g, w), c
    if out.nrows:  # first and last output key, read back with the size
        assert out.key_range() == (int(want[0][0]), int(want[0][-1]))
    for r in (gl, gr, out):
        r.free()


@pytest.mark.parametrize("only", [None, ["join_expand", "radix_onesweep_kv"]])
def test_operators_under_profiling(ctx, only):
    """Per-launch timing on (all launches, or some labels): the size
    read-backs return while timed launches are still pending, and the
    results and the launch counts are unchanged."""
    from vedas_amd._native import Rel
    rng = np.random.default_rng(77)
    L = make_rel(rng, 60_000, 3, 5000)
    R = make_rel(rng, 50_000, 2, 5000)
    want = O.join(L, R, 1)
    ctx.prof(True)
    ctx.prof_only(only)
    ctx.prof_reset()
    try:
        for _ in range(2):
            gl, gr = Rel.from_columns(ctx, L), Rel.from_columns(ctx, R)
            out = gl.join(gr, 1)
            out.swap_cols(0, 1)
            out.sort_col0()
            got = out.columns()
            for g, w in zip(got, O.sort_col0([want[1], want[0]] + want[2:])):
                assert np.array_equal(g, w)
            for r in (gl, gr, out):
                r.free()
        ctx.sync()
        st = ctx.prof_stats()
        assert st["join_expand"]["launches"] == 2 and st["radix_onesweep_kv"]["launches"] >= 2
        if only:
            assert set(st) <= set(only)
        else:
            assert st["join_bounds"]["launches"] == 2 and st["scan_tiles"]["launches"] == 2
    finally:
        ctx.prof(False)
        ctx.prof_only(None)


@pytest.mark.parametrize("nv,rcap", [(1024, 2048), (2048, 4096), (4096, 8192)])
def test_join_tunables_and_skew(ctx, nv, rcap):
    from vedas_amd._native import Rel
    ctx.tune("join_nv", nv)
    ctx.tune("join_rcap", rcap)
    try:
        rng = np.random.default_rng(nv)
        # one hot key on both sides + a uniform tail, plus left keys absent on the right
        lk = np.concatenate([np.full(3000, 5), rng.integers(10, 100_000, 40_000)])
        rk = np.concatenate([np.full(2500, 5), rng.integers(50, 100_000, 60_000)])
        L = O.sort_col0([lk.astype(np.uint32), rng.integers(0, 9, len(lk)).astype(np.uint32)])
        R = O.sort_col0([rk.astype(np.uint32), rng.integers(0, 9, len(rk)).astype(np.uint32)])
        want = O.join(L, R, 1)
        gl, gr = Rel.from_columns(ctx, L), Rel.from_columns(ctx, R)
        out = gl.join(gr, 1)
        got = out.columns()
        assert out.nrows == len(want[0]) and out.nrows > 7_500_000
        for g, w in zip(got, want):
            assert np.array_equal(g, w)
        for r in (gl, gr, out):
            r.free()
    finally:
        ctx.tune("join_nv", 2048)
        ctx.tune("join_rcap", 4096)


@pytest.mark.parametrize("rcap", [2048, 4096, 8192])
def test_join_directory_gaps(ctx, rcap):
    """Left tiles whose right range (2M rows around a hot key) exceeds the
    LDS staging: the bounds are finished by searches in the gaps between
    directory entries (tests/test_gpu_batch.py:directory_pair), bit-equal
    to the oracle for every staging size."""
    from vedas_amd._native import Rel
    from tests.test_gpu_batch import directory_pair
    ctx.tune("join_rcap", rcap)
    try:
        L, R = directory_pair(np.random.default_rng(rcap))
        want = O.join(L, R, 1)
        gl, gr = Rel.from_columns(ctx, L), Rel.from_columns(ctx, R)
        out = gl.join(gr, 1)
        assert out.nrows == len(want[0]) and out.nrows > 2_000_000
        for g, w in zip(out.columns(), want):
            assert np.array_equal(g, w)
        for r in (gl, gr, out):
            r.free()
    finally:
        ctx.tune("join_rcap", 4096)


def test_join_empty_sides(ctx):
    from vedas_amd._native import Rel
    rng = np.random.default_rng(1)
    R = make_rel(rng, 100, 2, 10)
    E = [np.zeros(0, np.uint32), np.zeros(0, np.uint32)]
    gl, gr = Rel.from_columns(ctx, E), Rel.from_columns(ctx, R)
    out = gl.join(gr, 1)
    assert out.nrows == 0 and out.ncols == 3
    out2 = gr.join(gl, 1)
    assert out2.nrows == 0
    # disjoint keys
    A = O.sort_col0([np.arange(0, 100, 2, dtype=np.uint32)])
    B = O.sort_col0([np.arange(1, 100, 2, dtype=np.uint32)])
    ga, gb = Rel.from_columns(ctx, A), Rel.from_columns(ctx, B)
    out3 = ga.join(gb, 1)
    assert out3.nrows == 0
    for r in (gl, gr, out, out2, ga, gb, out3):
        r.free()


@pytest.mark.parametrize("ncols,differ", [(5, 4), (7, 6), (9, 1), (3, 2)])
def test_unique_rows_differing_late(ctx, ncols, differ):
    """Adjacent unique where runs of rows agree on the first `differ`
    columns and differ (or not) only in later ones, across wave and tile
    boundaries (the previous row's value crosses lanes, waves and tiles)."""
    from vedas_amd._native import Rel
    rng = np.random.default_rng(ncols * 10 + differ)
    n = 70_001
    run = rng.integers(1, 300, n).cumsum()
    head = np.searchsorted(run, np.arange(n), side="right").astype(np.uint32)
    cols = [head.copy() for _ in range(differ)]
    cols += [rng.integers(0, 3, n).astype(np.uint32) for _ in range(ncols - differ)]
    for j in (63, 64, 2047, 2048, 4095):  # equal rows straddling boundaries
        for c in cols:
            c[j] = c[j - 1]
    cols = O.sort_lex(cols)
    want = O.unique_adjacent(cols)
    r = Rel.from_columns(ctx, cols)
    try:
        r.unique_adjacent()
        assert r.nrows == len(want[0]) < n
        for g, w in zip(r.columns(), want):
            assert np.array_equal(g, w)
    finally:
        r.free()


@pytest.mark.parametrize("n,ncols,keymax", [(0, 2, 5), (1, 3, 5), (1000, 1, 50), (70_001, 3, 1000),
                                            (300_000, 5, 2**31), (4096, 2, 1),
                                            (100_000, 2, 2**32 - 1), (50_000, 4, 2**20),
                                            (3000, 16, 7)])
def test_sorts_and_unique(ctx, n, ncols, keymax):
    from vedas_amd._native import Rel
    rng = np.random.default_rng(n + ncols)
    cols = [rng.integers(0, keymax, n).astype(np.uint32) for _ in range(ncols)]
    r = Rel.from_columns(ctx, cols)
    r.sort_col0()
    want = O.sort_col0(cols)
    for g, w in zip(r.columns(), want):
        assert np.array_equal(g, w)
    r.free()
    r = Rel.from_columns(ctx, cols)
    r.sort_lex()
    want = O.sort_lex(cols)
    for g, w in zip(r.columns(), want):
        assert np.array_equal(g, w)
    r.unique_adjacent()
    want = O.unique_adjacent(want)
    assert r.nrows == len(want[0])
    for g, w in zip(r.columns(), want):
        assert np.array_equal(g, w)
    r.free()
    r = Rel.from_columns(ctx, cols)
    r.unique_adjacent()
    want = O.unique_adjacent(cols)
    for g, w in zip(r.columns(), want):
        assert np.array_equal(g, w)
    r.free()


@pytest.mark.parametrize("items", [8, 16, 24])
@pytest.mark.parametrize("ncols", [1, 2, 3, 4, 5, 6])
def test_sort_col0_every_tile_shape(ctx, items, ncols):
    """Every radix tile shape (keys only and with 1-3 payload columns, whose
    tiles hold 8192 / 6144 / 4096 rows) at sizes one past a tile boundary,
    with duplicate keys so stability shows."""
    from vedas_amd._native import Rel
    ctx.tune("radix_items", items)
    try:
        # keys below 2^27 with the top byte varying take the 9-bit digit
        # plan (3 passes); the others the 8-bit one
        for n, keymax in ((4097, 2**31), (6145, 300), (8193, 2**20), (24_577, 7),
                          (100_003, 2**32 - 1), (8193, 2**27), (70_001, 2**25 + 3)):
            print(f"sort_col0 items={items} ncols={ncols} n={n}", flush=True)
            rng = np.random.default_rng(n * 31 + ncols)
            cols = [rng.integers(0, keymax, n, dtype=np.uint64).astype(np.uint32)]
            cols += [rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
                     for _ in range(ncols - 1)]
            r = Rel.from_columns(ctx, cols)
            r.sort_col0()
            for g, w in zip(r.columns(), O.sort_col0(cols)):
                assert np.array_equal(g, w)
            r.free()
    finally:
        ctx.tune("radix_items", 16)


@pytest.mark.parametrize("group", [0, 4, 16, 32])
def test_radix_tile_groups(ctx, group):
    """Tiles taken per XCD in groups of consecutive tiles (8 counters, a grid
    rounded up to whole groups per XCD, blocks past the last tile exiting):
    IndexSwap sorts with 0-3 payload columns and an index build, at sizes
    that end inside a group, on a group boundary and past a tile, against
    the oracle.  The default (8) runs in every other test."""
    from vedas_amd._native import Rel
    ctx.tune("radix_group", group)
    try:
        for n, ncols in ((1, 2), (6145, 4), (8192 * 8 * 3 + 1, 2), (300_007, 3),
                         (500_001, 5), (6144 * 37 + 5, 4)):
            print(f"radix_group={group} n={n} ncols={ncols}", flush=True)
            rng = np.random.default_rng(n + group)
            cols = [rng.integers(0, 1 << 27, n, dtype=np.uint64).astype(np.uint32)]
            cols[0][::3] = 77  # duplicate keys: stability shows
            cols += [rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
                     for _ in range(ncols - 1)]
            r = Rel.from_columns(ctx, cols)
            r.sort_col0()
            for g, w in zip(r.columns(), O.sort_col0(cols)):
                assert np.array_equal(g, w)
            r.free()
        s, p, o = rand_triples(250_003, 40 + group)
        check_index(ctx, s, p, o, ["SPO", "PSO", "OSP"])
    finally:
        ctx.tune("radix_group", 8)


def test_sort_col0_views_any_alignment(ctx):
    """IndexSwap sorts of unfiltered selects, which read the permutation in
    place: column 0 starts at every offset modulo 16 bytes (the column
    maximum taken before the pack reads it in 16-byte loads, with a head and
    a tail), over 1 to thousands of rows."""
    from vedas_amd._native import Index
    s, p, o = rand_triples(40_000, 23)
    idx = Index(ctx, s, p, o, ["PSO"])
    ix = O.build_perm(s, p, o, "PSO")
    try:
        for lo in (1000, 1001, 1002, 1003):
            for cnt in (1, 3, 6, 777, 9001):
                print(f"view sort lo={lo} rows={cnt}", flush=True)
                want = O.select(ix, lo, lo + cnt, 3)
                r = idx.select("PSO", lo, lo + cnt, 3)
                assert r.is_view(0)
                r.swap_cols(0, 1)
                r.sort_col0()
                for g, w in zip(r.columns(), O.sort_col0([want[1], want[0]])):
                    assert np.array_equal(g, w)
                r.free()
    finally:
        idx.free()


@pytest.mark.parametrize("ncols", [2, 3, 4, 6])
def test_sort_col0_digit_plans(ctx, ncols):
    """Column-0 values that only vary in some 9-bit or 8-bit digits: the
    passes skip the constant digits, in both digit plans, under every
    radix_bits setting (the rule, always 8-bit, 9-bit up to 27 bits)."""
    from vedas_amd._native import Rel
    rng = np.random.default_rng(ncols)
    n = 50_001
    try:
        for bits in (0, 8, 9):
            ctx.tune("radix_bits", bits)
            for col0 in (rng.integers(0, 64, n).astype(np.uint32) << 18,        # bits 18..23
                         (rng.integers(0, 8, n).astype(np.uint32) << 24) | 5,  # bits 24..26
                         rng.integers(1 << 24, 1 << 27, n).astype(np.uint32),
                         rng.integers(0, 1 << 20, n).astype(np.uint32),
                         rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32),
                         np.full(n, 123456, np.uint32)):
                cols = [col0] + [rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
                                 for _ in range(ncols - 1)]
                r = Rel.from_columns(ctx, cols)
                r.sort_col0()
                for g, w in zip(r.columns(), O.sort_col0(cols)):
                    assert np.array_equal(g, w), bits
                r.free()
    finally:
        ctx.tune("radix_bits", 0)


@pytest.mark.parametrize("perms", [["SOP", "POS", "OPS"], ["OPS", "OSP"], ["POS"], ["SOP", "PSO"]])
def test_index_build_derivation_orders(ctx, perms):
    """Permutation sets whose first fully sorted permutation is not SPO:
    {SOP, POS, OPS} sorts SOP and POS in full and derives OPS from POS;
    {OPS, OSP} derives neither from the other; a single permutation; {SOP,
    PSO} sorts both.  Every set of all six derives five (the default set's
    PSO / OSP from SPO and POS from OSP)."""
    s, p, o = rand_triples(30_011, len(perms) * 7 + 1)
    s[::5] = s[0]  # long runs of X
    check_index(ctx, s, p, o, perms)


@pytest.mark.parametrize("bits", [8, 9])
def test_index_build_digit_widths(ctx, bits):
    """Derived permutations sort on X's bits under the radix_bits setting
    (always bytes / 9-bit digits up to 27 bits), with X of 9, 20 and 27
    bits."""
    ctx.tune("radix_bits", bits)
    try:
        rng = np.random.default_rng(bits)
        n = 40_003
        s = rng.integers(1, 1 << 20, n).astype(np.uint32)
        p = rng.integers(1, 1 << 9, n).astype(np.uint32)
        o = rng.integers(1, 1 << 27, n).astype(np.uint32)
        check_index(ctx, s, p, o)
    finally:
        ctx.tune("radix_bits", 0)


def test_value_bounds_drive_index_swaps(ctx):
    """Selects carry the index's column maxima as value bounds, joins pass
    them on, and an IndexSwap on a bounded column 0 takes its digit plan
    from the bound (no maximum kernel, no read-back): with a bound far
    above the selected values the top digits do not vary, and the sort
    must still equal the oracle's.  A column handed out writable loses its
    bound (the maximum is taken again)."""
    from vedas_amd._native import Index, Rel
    rng = np.random.default_rng(9)
    n = 60_000
    s = rng.integers(1, 400, n).astype(np.uint32)
    p = rng.integers(1, 6, n).astype(np.uint32)
    o = rng.integers(1, 3000, n).astype(np.uint32)
    o[7] = (1 << 27) - 5  # one large object: the O bound needs 27 bits
    s[9] = (1 << 24) - 3  # one large subject: the S bound needs 24 bits
    idx = Index(ctx, s, p, o, ["PSO", "POS"])
    ix = {q: O.build_perm(s, p, o, q) for q in ["PSO", "POS"]}
    ctx.prof(True)
    try:
        for pred in (1, 2, 3):
            b, e = O.lookup(ix["PSO"], pred)
            want = O.select(ix["PSO"], b, e, 3)            # [S, O] of predicate pred
            r = idx.select("PSO", b, e, 3)
            r.swap_cols(0, 1)                               # O first: bound 2^27 - 5
            ctx.prof_reset()
            r.sort_col0()
            st = ctx.prof_stats()
            assert "max_u32" not in st and st["radix_onesweep"]["launches"] == 3
            for g, w in zip(r.columns(), O.sort_col0([want[1], want[0]])):
                assert np.array_equal(g, w)
            r.free()
            # a join of two selects on S, then a swap to the right side's O
            b2, e2 = O.lookup(ix["PSO"], pred + 1)
            want2 = O.select(ix["PSO"], b2, e2, 3)
            jl, jr = idx.select("PSO", b, e, 3), idx.select("PSO", b2, e2, 3)
            out = jl.join(jr, 1)                            # [S, O_l, O_r]
            wj = O.join(want, want2, 1)
            out.swap_cols(0, 2)
            ctx.prof_reset()
            out.sort_col0()
            assert "max_u32" not in ctx.prof_stats()
            got = out.columns()
            for g, w in zip(got, O.sort_col0([wj[2], wj[1], wj[0]])):
                assert np.array_equal(g, w)
            # columns() handed out pointers: the next swap takes the maximum
            out.swap_cols(0, 1)
            ctx.prof_reset()
            out.sort_col0()
            assert ctx.prof_stats()["max_u32"]["launches"] == 1
            for g, w in zip(out.columns(), O.sort_col0([got[1], got[0], got[2]])):
                assert np.array_equal(g, w)
            for x in (jl, jr, out):
                x.free()
    finally:
        ctx.prof(False)
        idx.free()


@pytest.mark.parametrize("items", [8, 24])
def test_index_build_tile_shapes(ctx, items):
    ctx.tune("radix_items", items)
    try:
        s, p, o = rand_triples(70_001, 5 + items)
        check_index(ctx, s, p, o, ["SPO", "PSO", "OSP"])
    finally:
        ctx.tune("radix_items", 16)


@pytest.mark.parametrize("shape", ["star2", "star4", "chain4", "snow"])
def test_plans_cfg0_store(ctx, store, shape):
    s, p, o, idx, orc = store
    if shape == "star2":
        pats, node = plans.star_plan([1, 2])
    elif shape == "star4":
        pats, node = plans.star_plan([3, 4, 5, 6])
    elif shape == "chain4":
        pats, node = plans.chain_plan([1, 1, 1, 2])
    else:
        pats, node = plans.snowflake_plan([1, 2], [1, 3])
    gex = plans.GpuExecutor(ctx, idx)
    rel = gex.run(node)
    gnames, grows = O.canonical(rel.headers, gex.columns(rel))
    gex.free()
    oex = plans.OracleExecutor(s, p, o)
    orel = oex.run(node)
    onames, orows = O.canonical(orel.headers, oex.columns(orel))
    assert gnames == onames
    assert grows == orows
    assert len(grows) > 0


def test_plan_small_store_naive(ctx):
    from vedas_amd._native import Index
    rng = np.random.default_rng(4)
    trip = sorted(set(zip(rng.integers(1, 25, 300).tolist(), rng.integers(1, 5, 300).tolist(),
                          rng.integers(1, 25, 300).tolist())))
    s, p, o = (np.array(c, np.uint32) for c in zip(*trip))
    idx = Index(ctx, s, p, o, ["PSO", "POS"])
    try:
        for pats, node in [plans.star_plan([1, 2, 3]), plans.chain_plan([1, 2, 3]),
                           plans.snowflake_plan([1, 2], [3, 4])]:
            gex = plans.GpuExecutor(ctx, idx)
            rel = gex.run(node)
            got = O.canonical(rel.headers, gex.columns(rel))
            gex.free()
            assert got == O.naive_bgp(s, p, o, pats)
    finally:
        idx.free()


@pytest.mark.parametrize("perm", ["SPO", "OSP", "PSO"])
def test_histograms_equal_oracle(ctx, store, perm):
    """Equal-width and equal-depth histograms of the leading column
    (vd_index_histogram) against the oracle's restatement of
    Histogram::initEqWidth / initEqDepth on the sorted column, with the
    reference's literal limit and with every key."""
    s, p, o, idx, orc = store
    col = {"S": s, "P": p, "O": o}[perm[0]]
    keys = np.sort(col)
    for limit in (50_000_000 - 1, 0xFFFFFFFF, 40_000):
        for w in (1, 100, 7777):
            lo, hi, fr = idx.histogram(perm, False, w, limit)
            want = O.hist_eq_width(keys, w, limit)
            assert (lo.tolist(), hi.tolist(), fr.tolist()) == want, (perm, limit, w)
        for d in (1, 1000, 300_000):
            lo, hi, fr = idx.histogram(perm, True, d, limit)
            want = O.hist_eq_depth(keys, d, limit)
            assert (lo.tolist(), hi.tolist(), fr.tolist()) == want, (perm, limit, d)


@pytest.mark.parametrize("repeat", [False, True])
def test_distinct_rows_skip_the_unique(ctx, repeat):
    """An index of distinct triples gives selects whose rows are known to be
    distinct, joins of them keep the property, and the adjacent unique then
    launches nothing; one repeated triple (counted by the build) leaves
    every select unmarked and the unique runs.  Dropping a column clears
    the mark.  Results equal the oracle's either way."""
    from vedas_amd._native import Index
    s, p, o = rand_triples(50_000, 77)
    t = np.unique(np.stack([s, p, o], 1), axis=0)
    s, p, o = (np.ascontiguousarray(t[:, i]) for i in range(3))
    if repeat:
        s, p, o = (np.append(a, a[123]) for a in (s, p, o))
    idx = Index(ctx, s, p, o, ["PSO", "POS"])
    ix = {q: O.build_perm(s, p, o, q) for q in ("PSO", "POS")}
    ctx.prof(True)
    try:
        b1, e1 = O.lookup(ix["POS"], 2)
        b2, e2 = O.lookup(ix["PSO"], 3)
        L = idx.select("POS", b1, e1, 3)
        R = idx.select("PSO", b2, e2, 3)
        assert L.distinct == R.distinct == (not repeat)
        out = L.join(R, 1)
        assert out.distinct == (not repeat)
        want = O.join(O.select(ix["POS"], b1, e1, 3), O.select(ix["PSO"], b2, e2, 3), 1)
        ctx.prof_reset()
        out.unique_adjacent()
        launched = "unique_count" in ctx.prof_stats()
        assert launched == repeat
        for g, w in zip(out.columns(), O.unique_adjacent(want)):
            assert np.array_equal(g, w)
        out.free()
        out = L.join(R, 1)
        out.drop_col(2)
        assert not out.distinct
        ctx.prof_reset()
        out.unique_adjacent()
        assert "unique_count" in ctx.prof_stats()
        for g, w in zip(out.columns(), O.unique_adjacent(want[:2])):
            assert np.array_equal(g, w)
        for x in (L, R, out):
            x.free()
    finally:
        ctx.prof(False)
        idx.free()
