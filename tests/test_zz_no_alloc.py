"""The no-allocation batched select (vd_select_batch_noalloc) through the C
ABI: the same offsets and rows as vd_select_batch on two calls in a row,
the second smaller, without touching the context's device memory pool;
more queries than its fixed scratch holds are refused.  Needs an MI355X."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run(ctx, idx, name, perm, k1, k2, max_rows):
    from vedas_amd import _native
    n = len(k1)
    d1, d2 = ctx.upload(k1), ctx.upload(k2)
    doff = ctx.alloc(8 * (n + 1))
    dy, dz = ctx.alloc(4 * max(max_rows, 1)), ctx.alloc(4 * max(max_rows, 1))
    tot = C.c_uint64()
    before = ctx.mem()
    ctx.check(getattr(ctx.L, name)(ctx.h, idx.h, _native.PERMS[perm], C.c_void_p(d1),
                                   C.c_void_p(d2), n, C.c_void_p(doff), C.c_void_p(dy),
                                   C.c_void_p(dz), max_rows, C.byref(tot)))
    ctx.sync()
    after = ctx.mem()
    rows = tot.value
    got = rows if max_rows else 0  # max_rows == 0 only counts
    out = (ctx.download(doff, n + 1, np.uint64), ctx.download(dy, got), ctx.download(dz, got),
           rows, before == after)
    for d in (d1, d2, doff, dy, dz):
        ctx.free(d)
    return out


def test_noalloc_select_equals_select_batch():
    from vedas_amd._native import VD_ERANGE, Ctx, Index, VedasError, gen_triples
    print("no-alloc select: 300k triples, two calls (PSO 6000 queries, POS 900)", flush=True)
    ctx = Ctx(0)
    try:
        s, p, o = gen_triples(300_000, seed=5, n_predicates=64, n_entities=50_000)
        idx = Index(ctx, s, p, o, ("SPO", "PSO", "POS", "OSP"))
        rng = np.random.default_rng(11)
        t = rng.integers(0, len(s), 6000)
        calls = [("PSO", p[t].astype(np.uint32), np.full(len(t), 0xFFFFFFFF, np.uint32)),
                 ("POS", p[t[:900]].astype(np.uint32), o[t[:900]].astype(np.uint32))]
        for perm, k1, k2 in calls:
            k1 = k1.copy()
            k1[::97] = 4_000_000_000  # absent constants: empty ranges
            off, y, z, rows, _ = run(ctx, idx, "vd_select_batch", perm, k1, k2, 0)
            cap = int(off[-1])
            off, y, z, rows, _ = run(ctx, idx, "vd_select_batch", perm, k1, k2, cap)
            off2, y2, z2, rows2, same_pool = run(ctx, idx, "vd_select_batch_noalloc", perm, k1,
                                                 k2, cap)
            assert rows > 0 and rows2 == rows
            assert np.array_equal(off, off2)
            assert np.array_equal(y, y2) and np.array_equal(z, z2)
            assert same_pool, "the no-allocation select changed the context's pool"
        # more queries than the scratch holds: refused before any device work
        from vedas_amd._native import lib
        n = (1 << 20) + 1
        d1 = ctx.alloc(4 * n)
        doff = ctx.alloc(8 * (n + 1))
        tot = C.c_uint64()
        rc = lib().vd_select_batch_noalloc(ctx.h, idx.h, 2, C.c_void_p(d1), None, n,
                                           C.c_void_p(doff), None, None, 0, C.byref(tot))
        assert rc == VD_ERANGE
        ctx.free(d1)
        ctx.free(doff)
        idx.free()
    finally:
        ctx.close()
