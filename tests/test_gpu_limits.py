"""The documented hard limits of the C ABI, driven on the GPU: each call
returns VD_ERANGE with a message before it launches work or allocates its
output, and the context stays usable afterwards.  Needs an MI355X."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

ROWS_2_30 = 1 << 30


@pytest.fixture(scope="module")
def ctx():
    from vedas_amd._native import Ctx
    c = Ctx(0)
    yield c
    c.close()


def still_works(ctx):
    """A small sort right after the error, checked against the oracle."""
    from vedas_amd._native import Rel
    cols = [np.array([5, 1, 4, 1, 3], np.uint32), np.array([0, 1, 2, 3, 4], np.uint32)]
    r = Rel.from_columns(ctx, cols)
    try:
        r.sort_col0()
        for g, w in zip(r.columns(), O.sort_col0(cols)):
            assert np.array_equal(g, w)
    finally:
        r.free()


def test_sorts_refuse_2_30_rows(ctx):
    """IndexSwap and the lexicographic sort take fewer than 2^30 rows
    (their 30-bit look-back values): a 2^30-row relation (4 GiB, contents
    never read) is refused by both."""
    from vedas_amd._native import VD_ERANGE, Rel, VedasError
    h = C.c_void_p()
    ctx.check(ctx.L.vd_rel_create(ctx.h, 1, ROWS_2_30, C.byref(h)))
    r = Rel(ctx, h)
    try:
        for op in (r.sort_col0, r.sort_lex):
            with pytest.raises(VedasError) as e:
                op()
            assert e.value.code == VD_ERANGE
            assert "2^30" in str(e.value)
        assert r.nrows == ROWS_2_30
    finally:
        r.free()
    still_works(ctx)


def test_index_build_refuses_2_30_triples(ctx):
    """vd_index_build checks the triple count before it reads a column."""
    from vedas_amd._native import VD_ERANGE
    d = ctx.alloc(16)
    try:
        h = C.c_void_p()
        rc = ctx.L.vd_index_build(ctx.h, C.c_void_p(d), C.c_void_p(d), C.c_void_p(d),
                                  ROWS_2_30, 1, C.byref(h))
        assert rc == VD_ERANGE and not h.value
        assert b"2^30" in ctx.L.vd_ctx_last_error(ctx.h)
    finally:
        ctx.free(d)
    still_works(ctx)


def test_join_tile_over_2_32_matches(ctx):
    """2048 consecutive left rows whose matches exceed 2^32 - 1 right rows
    (the tile-local offsets are 32-bit): 2048 left rows and 2^21 + 1 right
    rows of one key make 2048 * (2^21 + 1) > 2^32 - 1 output rows.  The
    join refuses after its count read-back, before the output exists."""
    from vedas_amd._native import VD_ERANGE, Rel, VedasError
    left = Rel.from_columns(ctx, [np.full(2048, 5, np.uint32), np.arange(2048, dtype=np.uint32)])
    right = Rel.from_columns(ctx, [np.full((1 << 21) + 1, 5, np.uint32)])
    try:
        before = ctx.mem()[0]
        with pytest.raises(VedasError) as e:
            left.join(right, 1)
        assert e.value.code == VD_ERANGE
        assert "2^32" in str(e.value)
        ctx.sync()
        assert ctx.mem()[0] == before  # the scratch went back, no output kept
        # the same left tile against a short right side joins as usual
        small = Rel.from_columns(ctx, [np.full(3, 5, np.uint32)])
        out = left.join(small, 1)
        assert out.nrows == 2048 * 3
        want = O.join([np.full(2048, 5, np.uint32), np.arange(2048, dtype=np.uint32)],
                      [np.full(3, 5, np.uint32)])
        for g, w in zip(out.columns(), want):
            assert np.array_equal(g, w)
        out.free()
        small.free()
    finally:
        left.free()
        right.free()
    still_works(ctx)
