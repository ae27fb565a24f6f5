"""ctypes front of oracle/libcpuport.so, the multi-threaded host port of
the query operators (select with the bound clip, join, stable re-sort,
adjacent unique) that the benchmark times as its CPU baseline.

TEST AND BASELINE INFRASTRUCTURE ONLY: imported by tests/ and bench.py's
cpu_baseline leg, never by the vedas_amd package.  Its results are bit-equal
to the oracle's (pyoracle) for any thread count.
"""
from __future__ import annotations

import ctypes as C
import os
import types

import numpy as np

from . import pyoracle as O

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcpuport.so")
U32P = C.POINTER(C.c_uint32)
_lib = None


def threads() -> int:
    """Worker threads: OMP_NUM_THREADS when set (the host's share of
    cores), else every core the process may use."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built: run python -m vedas_amd.build oracle")
        l = C.CDLL(LIB_PATH)
        vp, u64, i = C.c_void_p, C.c_uint64, C.c_int
        l.cp_select.restype = u64
        l.cp_select.argtypes = [vp, vp, u64, u64, i, i, C.c_uint32, C.c_uint32, vp, vp,
                                C.POINTER(u64), i]
        l.cp_join_bounds.restype = u64
        l.cp_join_bounds.argtypes = [vp, u64, vp, u64, vp, vp, i]
        l.cp_join_expand.restype = None
        l.cp_join_expand.argtypes = [C.POINTER(U32P), u64, i, C.POINTER(U32P), i, vp, vp,
                                     C.POINTER(U32P), i]
        l.cp_sort_col0.restype = i
        l.cp_sort_col0.argtypes = [C.POINTER(U32P), i, u64, i]
        l.cp_unique_adjacent.restype = u64
        l.cp_unique_adjacent.argtypes = [C.POINTER(U32P), i, u64, i]
        _lib = l
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data)


def _colptrs(cols):
    arr = (U32P * max(1, len(cols)))()
    for k, c in enumerate(cols):
        arr[k] = c.ctypes.data_as(U32P)
    return arr


def make_ops(nthreads: int | None = None) -> types.SimpleNamespace:
    """select / join / sort_col0 / unique_adjacent with pyoracle's
    signatures, run on nthreads threads."""
    nt = nthreads or threads()

    def select(ix, b, e, cols=3, clip=None, skip=(), distinct=False):
        assert not skip and not distinct, "the port covers the clip only"
        lo, hi = clip if clip is not None else (0, 0)
        y, z = ix["y"], ix["z"]
        first = C.c_uint64()
        m = lib().cp_select(_p(y), _p(z), b, e, cols, int(clip is not None), lo, hi, None, None,
                            C.byref(first), nt)
        o0 = np.empty(max(m, 1), np.uint32)
        o1 = np.empty(max(m, 1), np.uint32) if cols == 3 else o0
        lib().cp_select(_p(y), _p(z), b, e, cols, int(clip is not None), lo, hi, _p(o0), _p(o1),
                        C.byref(first), nt)
        return [o0[:m], o1[:m]] if cols == 3 else [o0[:m]]

    def join(lcols, rcols, nk=1):
        assert nk == 1, "the port covers 1-key joins"
        lcols = [np.ascontiguousarray(c, np.uint32) for c in lcols]
        rcols = [np.ascontiguousarray(c, np.uint32) for c in rcols]
        nl, nr = len(lcols[0]), len(rcols[0])
        lower = np.empty(max(nl, 1), np.uint64)
        off = np.empty(nl + 1, np.uint64)
        n = lib().cp_join_bounds(_p(lcols[0]), nl, _p(rcols[0]), nr, _p(lower), _p(off), nt)
        out = [np.empty(max(n, 1), np.uint32) for _ in range(len(lcols) + len(rcols) - 1)]
        if n:
            lib().cp_join_expand(_colptrs(lcols), nl, len(lcols), _colptrs(rcols), len(rcols),
                                 _p(lower), _p(off), _colptrs(out), nt)
        return [c[:n] for c in out]

    def sort_col0(cols):
        cols = [np.array(c, np.uint32) for c in cols]
        assert lib().cp_sort_col0(_colptrs(cols), len(cols), len(cols[0]), nt) == 0
        return cols

    def unique_adjacent(cols):
        cols = [np.array(c, np.uint32) for c in cols]
        n = lib().cp_unique_adjacent(_colptrs(cols), len(cols), len(cols[0]), nt)
        return [c[:n] for c in cols]

    return types.SimpleNamespace(select=select, join=join, sort_col0=sort_col0,
                                 unique_adjacent=unique_adjacent, threads=nt)


def run_query(ix_by_perm: dict, patterns, node, ops, stats: dict | None = None,
              bounds: bool = True):
    """One query as the drop-in path runs it: the plan walk with the
    reference's bounds, then the finalise step's adjacent unique
    (QueryPlan.cu:188-204; every column is selected)."""
    h, cols = O.run_plan(ix_by_perm, patterns, node, stats, bounds=bounds, ops=ops)
    if len(cols) and len(cols[0]):
        cols = ops.unique_adjacent(cols)
    return h, cols
