"""ctypes binding of the C ABI in include/vedas_hip.h (libvedas_hip.so).

Thin and strict: a missing library raises at import of the binding, every
non-zero status raises VedasError with the context's last-error message.
There is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvedas_hip.so")

VD_OK, VD_EINVAL, VD_ENOMEM, VD_EHIP, VD_ERANGE = 0, -1, -2, -3, -4
VD_MAX_COLS = 16
VD_UNBOUND = 0xFFFFFFFF
PERMS = {"SPO": 0, "SOP": 1, "PSO": 2, "POS": 3, "OSP": 4, "OPS": 5}
PERM_NAMES = {v: k for k, v in PERMS.items()}


class VedasError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"vedas error {code}: {msg}")
        self.code = code


class PermView(C.Structure):
    _fields_ = [("n", C.c_uint64), ("n_l1", C.c_uint64), ("n_l2", C.c_uint64),
                ("l1_values", C.c_void_p), ("l1_offsets", C.c_void_p),
                ("l2_values", C.c_void_p), ("l2_offsets", C.c_void_p),
                ("col_y", C.c_void_p), ("col_z", C.c_void_p)]


class Interval(C.Structure):
    _fields_ = [("lo", C.c_uint32), ("hi", C.c_uint32)]


class SelectJob(C.Structure):
    _fields_ = [("perm", C.c_int), ("k1", C.c_uint32), ("k2", C.c_uint32), ("cols", C.c_int)]


class GenParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("n", C.c_uint64),
                ("n_entities", C.c_uint32), ("subject_zipf", C.c_double),
                ("n_predicates", C.c_uint32), ("pred_zipf", C.c_double),
                ("literal_frac", C.c_double), ("literal_base", C.c_uint32),
                ("n_literals", C.c_uint32), ("distinct", C.c_int)]


EXPORTS = {
    # name: (restype, argtypes)
    "vd_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "vd_ctx_destroy": (C.c_int, [C.c_void_p]),
    "vd_ctx_last_error": (C.c_char_p, [C.c_void_p]),
    "vd_ctx_sync": (C.c_int, [C.c_void_p]),
    "vd_ctx_stream": (C.c_void_p, [C.c_void_p]),
    "vd_ctx_trim": (C.c_int, [C.c_void_p]),
    "vd_ctx_mem": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "vd_ctx_mem_peak": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.c_int]),
    "vd_ctx_tune": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "vd_dev_alloc": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "vd_dev_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "vd_copy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "vd_copy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "vd_copy_d2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "vd_index_build": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_uint64, C.c_uint32, C.POINTER(C.c_void_p)]),
    "vd_index_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "vd_runs_expand": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64,
                                 C.c_void_p]),
    "vd_index_from_arrays": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(PermView),
                                       C.POINTER(C.c_void_p)]),
    "vd_index_view": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(PermView)]),
    "vd_index_perms": (C.c_uint32, [C.c_void_p]),
    "vd_index_lookup": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_uint32,
                                  C.c_int, C.c_uint32, C.POINTER(C.c_uint64),
                                  C.POINTER(C.c_uint64)]),
    "vd_index_stats": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "vd_index_histogram": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_uint32,
                                     C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_uint64, C.POINTER(C.c_uint64)]),
    "vd_index_lookup_batch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                        C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]),
    "vd_rel_create": (C.c_int, [C.c_void_p, C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]),
    "vd_rel_destroy": (C.c_int, [C.c_void_p, C.c_void_p]),
    "vd_rel_ncols": (C.c_int, [C.c_void_p]),
    "vd_rel_nrows": (C.c_uint64, [C.c_void_p]),
    "vd_rel_col": (C.c_void_p, [C.c_void_p, C.c_int]),
    "vd_rel_swap_cols": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "vd_rel_drop_col": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "vd_rel_materialize": (C.c_int, [C.c_void_p, C.c_void_p]),
    "vd_rel_is_view": (C.c_int, [C.c_void_p, C.c_int]),
    "vd_rel_key_range": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "vd_select_key": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_uint32, C.c_int,
                                C.c_uint32, C.c_int, C.c_int, C.c_uint32, C.c_uint32,
                                C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_void_p),
                                C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "vd_select": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_uint64, C.c_uint64,
                            C.c_int, C.c_int, C.c_uint32, C.c_uint32,
                            C.POINTER(Interval), C.c_int, C.c_int,
                            C.POINTER(C.c_void_p)]),
    "vd_select_batch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                  C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "vd_select_batch_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                          C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "vd_join": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                          C.POINTER(C.c_void_p)]),
    "vd_rel_distinct": (C.c_int, [C.c_void_p]),
    "vd_ctx_record": (C.c_int, [C.c_void_p, C.c_int]),
    "vd_ctx_capture_begin": (C.c_int, [C.c_void_p]),
    "vd_ctx_capture_end": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "vd_graph_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "vd_graph_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "vd_join_driven_applies": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                         C.c_uint32]),
    "vd_join_capped": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_uint64,
                                 C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]),
    "vd_semi_filter": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                 C.c_uint32, C.c_uint64, C.POINTER(C.c_void_p),
                                 C.POINTER(C.c_uint64)]),
    "vd_join_driven": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_uint32,
                                 C.c_uint32, C.c_int, C.POINTER(C.c_int), C.c_int,
                                 C.POINTER(C.c_void_p), C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint32)]),
    "vd_select_multi": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                  C.POINTER(C.c_void_p)]),
    "vd_index_lookup_multi": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                        C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "vd_index_bounds_multi": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                        C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "vd_join_multi": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                C.c_int, C.POINTER(C.c_void_p)]),
    "vd_sort_col0_multi": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int]),
    "vd_unique_adjacent_multi": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int]),
    "vd_sort_col0": (C.c_int, [C.c_void_p, C.c_void_p]),
    "vd_sort_lex": (C.c_int, [C.c_void_p, C.c_void_p]),
    "vd_unique_adjacent": (C.c_int, [C.c_void_p, C.c_void_p]),
    "vd_gen_triples": (C.c_int, [C.POINTER(GenParams), C.c_void_p, C.c_void_p,
                                 C.c_void_p]),
    "vd_prof_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "vd_prof_reset": (C.c_int, [C.c_void_p]),
    "vd_prof_only": (C.c_int, [C.c_void_p, C.c_char_p]),
    "vd_prof_count": (C.c_int, [C.c_void_p]),
    "vd_prof_get": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_char_p),
                              C.POINTER(C.c_uint64), C.POINTER(C.c_double),
                              C.POINTER(C.c_double)]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libvedas_hip.so (built in-tree); raise if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built: run python -m vedas_amd.build")
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def _ptr(a: np.ndarray) -> C.c_void_p:
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


def gen_triples(n: int, seed: int = 42, n_entities: int = 100_000,
                subject_zipf: float = 0.0, n_predicates: int = 64,
                pred_zipf: float = 1.1, literal_frac: float = 0.5,
                literal_base: int = 50_000_000, n_literals: int = 500_000,
                distinct: bool = True):
    """Seeded synthetic triples (host arrays)."""
    prm = GenParams(seed, n, n_entities, subject_zipf, n_predicates, pred_zipf,
                    literal_frac, literal_base, n_literals, int(distinct))
    s = np.empty(n, np.uint32)
    p = np.empty(n, np.uint32)
    o = np.empty(n, np.uint32)
    rc = lib().vd_gen_triples(C.byref(prm), _ptr(s), _ptr(p), _ptr(o))
    if rc != VD_OK:
        raise VedasError(rc, "vd_gen_triples failed")
    return s, p, o


class Ctx:
    """A vd_ctx: one device, one stream."""

    def __init__(self, device: int = 0):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.vd_ctx_create(device, C.byref(h))
        if rc != VD_OK:
            raise VedasError(rc, "vd_ctx_create failed (no GPU?)")
        self.h = h
        self.owned = True

    @classmethod
    def wrap(cls, handle) -> "Ctx":
        """A non-owning view of a context created elsewhere (libvedas.so)."""
        self = cls.__new__(cls)
        self.L = lib()
        self.h = C.c_void_p(handle)
        self.owned = False
        return self

    def check(self, rc: int) -> None:
        if rc != VD_OK:
            msg = self.L.vd_ctx_last_error(self.h)
            raise VedasError(rc, msg.decode() if msg else "")

    def close(self) -> None:
        if self.h and self.owned:
            self.L.vd_ctx_destroy(self.h)
        self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self) -> None:
        self.check(self.L.vd_ctx_sync(self.h))

    def record(self, on: bool) -> None:
        """Log every read-back word of the following calls (vd_ctx_record)."""
        self.check(self.L.vd_ctx_record(self.h, int(on)))

    def capture_begin(self) -> None:
        self.check(self.L.vd_ctx_capture_begin(self.h))

    def capture_end(self, keep: bool = True):
        """The captured calls as a graph handle (or None with keep=False)."""
        g = C.c_void_p()
        self.check(self.L.vd_ctx_capture_end(self.h, C.byref(g) if keep else None))
        return g if keep else None

    def graph_launch(self, g) -> None:
        self.check(self.L.vd_graph_launch(self.h, g))

    def graph_free(self, g) -> None:
        self.check(self.L.vd_graph_free(self.h, g))

    def tune(self, key: str, value: int) -> None:
        self.check(self.L.vd_ctx_tune(self.h, key.encode(), value))

    def mem(self):
        a, b = C.c_uint64(), C.c_uint64()
        self.check(self.L.vd_ctx_mem(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def mem_peak(self, reset: bool = False) -> int:
        """Highest in-use bytes since the last reset (reset: start again)."""
        v = C.c_uint64()
        self.check(self.L.vd_ctx_mem_peak(self.h, C.byref(v), int(reset)))
        return v.value

    # -- device buffers
    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        self.check(self.L.vd_dev_alloc(self.h, nbytes, C.byref(p)))
        return p.value

    def free(self, dptr: int) -> None:
        self.check(self.L.vd_dev_free(self.h, C.c_void_p(dptr)))

    def upload(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        d = self.alloc(max(a.nbytes, 4))
        self.check(self.L.vd_copy_h2d(self.h, C.c_void_p(d), _ptr(a), a.nbytes))
        return d

    def download(self, dptr: int, n: int, dtype=np.uint32) -> np.ndarray:
        out = np.empty(n, dtype)
        if n:
            self.check(self.L.vd_copy_d2h(self.h, _ptr(out), C.c_void_p(dptr), out.nbytes))
        return out

    # -- profiling
    def prof(self, on: bool = True) -> None:
        self.check(self.L.vd_prof_enable(self.h, int(on)))

    def prof_only(self, labels=None) -> None:
        """Time only the launches with these labels (None: all)."""
        self.check(self.L.vd_prof_only(self.h, ",".join(labels).encode() if labels else None))

    def prof_reset(self) -> None:
        self.check(self.L.vd_prof_reset(self.h))

    def prof_stats(self) -> dict:
        n = self.L.vd_prof_count(self.h)
        if n < 0:
            self.check(n)
        out = {}
        for i in range(n):
            nm, la, ms, by = C.c_char_p(), C.c_uint64(), C.c_double(), C.c_double()
            self.check(self.L.vd_prof_get(self.h, i, C.byref(nm), C.byref(la),
                                          C.byref(ms), C.byref(by)))
            out[nm.value.decode()] = {"launches": la.value, "ms": ms.value,
                                      "bytes": by.value}
        return out


class Rel:
    """A vd_rel handle owned by Python."""

    def __init__(self, ctx: Ctx, handle: C.c_void_p):
        self.ctx, self.h = ctx, handle

    @classmethod
    def from_columns(cls, ctx: Ctx, cols) -> "Rel":
        cols = [np.ascontiguousarray(c, dtype=np.uint32) for c in cols]
        n = len(cols[0])
        h = C.c_void_p()
        ctx.check(ctx.L.vd_rel_create(ctx.h, len(cols), n, C.byref(h)))
        r = cls(ctx, h)
        for i, c in enumerate(cols):
            assert len(c) == n
            if n:
                ctx.check(ctx.L.vd_copy_h2d(ctx.h, C.c_void_p(r.col_ptr(i)), _ptr(c), c.nbytes))
        return r

    @property
    def ncols(self) -> int:
        return self.ctx.L.vd_rel_ncols(self.h)

    @property
    def nrows(self) -> int:
        return self.ctx.L.vd_rel_nrows(self.h)

    def col_ptr(self, i: int) -> int:
        p = self.ctx.L.vd_rel_col(self.h, i)
        if not p:
            raise VedasError(VD_EINVAL, f"no column {i}")
        return p

    def columns(self) -> list:
        return [self.ctx.download(self.col_ptr(i), self.nrows) for i in range(self.ncols)]

    def key_range(self):
        """(first, last) of column 0 when the producing join knew them, else None."""
        lo, hi = C.c_uint32(), C.c_uint32()
        if self.ctx.L.vd_rel_key_range(self.h, C.byref(lo), C.byref(hi)):
            return lo.value, hi.value
        return None

    @property
    def distinct(self) -> bool:
        """The rows are known to be pairwise distinct."""
        return bool(self.ctx.L.vd_rel_distinct(self.h))

    def is_view(self, i: int) -> bool:
        """Column i reads index storage in place (read-only)."""
        return bool(self.ctx.L.vd_rel_is_view(self.h, i))

    def materialize(self) -> None:
        """Give every view column a private copy (before writing into it)."""
        self.ctx.check(self.ctx.L.vd_rel_materialize(self.ctx.h, self.h))

    def swap_cols(self, i: int, j: int) -> None:
        self.ctx.check(self.ctx.L.vd_rel_swap_cols(self.h, i, j))

    def drop_col(self, i: int) -> None:
        self.ctx.check(self.ctx.L.vd_rel_drop_col(self.ctx.h, self.h, i))

    def sort_col0(self) -> None:
        self.ctx.check(self.ctx.L.vd_sort_col0(self.ctx.h, self.h))

    def sort_lex(self) -> None:
        self.ctx.check(self.ctx.L.vd_sort_lex(self.ctx.h, self.h))

    def unique_adjacent(self) -> None:
        self.ctx.check(self.ctx.L.vd_unique_adjacent(self.ctx.h, self.h))

    def join(self, other: "Rel", nkeys: int = 1) -> "Rel":
        h = C.c_void_p()
        self.ctx.check(self.ctx.L.vd_join(self.ctx.h, self.h, other.h, nkeys, C.byref(h)))
        return Rel(self.ctx, h)

    def join_capped(self, other: "Rel", max_rows: int, nkeys: int = 1):
        """(join, output rows), or (None, output rows) when the output would
        have more than max_rows rows (nothing is built then)."""
        h, n = C.c_void_p(), C.c_uint64()
        self.ctx.check(self.ctx.L.vd_join_capped(self.ctx.h, self.h, other.h, nkeys, max_rows,
                                                 C.byref(h), C.byref(n)))
        return (Rel(self.ctx, h) if h.value else None), n.value

    def semi_filter(self, col: int, other: "Rel", ocol: int = 0, key_hi: int = 0xFFFFFFFF,
                    max_rows: int = (1 << 64) - 1):
        """(the rows of self whose column `col` value occurs in other's column
        `ocol`, in self's order, kept rows), or (None, kept rows) when more
        than max_rows are kept, or (None, None) when the key bound is beyond
        the filter's bitmap."""
        h, n = C.c_void_p(), C.c_uint64()
        self.ctx.check(self.ctx.L.vd_semi_filter(self.ctx.h, self.h, col, other.h, ocol, key_hi,
                                                 max_rows, C.byref(h), C.byref(n)))
        if not h.value and n.value == (1 << 64) - 1:
            return None, None
        return (Rel(self.ctx, h) if h.value else None), n.value

    def join_driven(self, dkey: int, other: "Rel", src, key_lo: int = 0,
                    key_hi: int = 0xFFFFFFFF, check_dups: bool = False):
        """For every row of self (the driver), in order, the rows of `other`
        (sorted on column 0) whose key equals self's column dkey.  Output
        column c is other's column src[c] (>= 0) or self's column -1 - src[c].
        Returns (Rel, smallest matched key, largest matched key), or None
        when the driven form does not apply."""
        h = C.c_void_p()
        arr = (C.c_int * len(src))(*src)
        lo, hi = C.c_uint32(), C.c_uint32()
        self.ctx.check(self.ctx.L.vd_join_driven(self.ctx.h, self.h, dkey, other.h, key_lo,
                                                 key_hi, len(src), arr, int(check_dups),
                                                 C.byref(h), C.byref(lo), C.byref(hi)))
        if not h.value:
            return None
        return Rel(self.ctx, h), lo.value, hi.value

    def free(self) -> None:
        if self.h:
            self.ctx.L.vd_rel_destroy(self.ctx.h, self.h)
            self.h = None


class Index:
    """HBM-resident permutations built from device triples."""

    def __init__(self, ctx: Ctx, s, p, o, perms=("SPO", "PSO", "POS", "OSP")):
        self.ctx = ctx
        n = len(s)
        self.n = n
        ds, dp, do = ctx.upload(s), ctx.upload(p), ctx.upload(o)
        mask = 0
        for name in perms:
            mask |= 1 << PERMS[name]
        h = C.c_void_p()
        try:
            ctx.check(ctx.L.vd_index_build(ctx.h, C.c_void_p(ds), C.c_void_p(dp),
                                           C.c_void_p(do), n, mask, C.byref(h)))
        finally:
            for d in (ds, dp, do):
                ctx.free(d)
        self.h = h

    @classmethod
    def wrap(cls, ctx: Ctx, handle, n: int) -> "Index":
        """A non-owning view of an index built elsewhere (libvedas.so)."""
        self = cls.__new__(cls)
        self.ctx, self.n, self.h = ctx, n, C.c_void_p(handle)
        self.owned = False
        return self

    @classmethod
    def from_device(cls, ctx: Ctx, ds: int, dp: int, do: int, n: int, perms):
        self = cls.__new__(cls)
        self.ctx, self.n = ctx, n
        mask = 0
        for name in perms:
            mask |= 1 << PERMS[name]
        h = C.c_void_p()
        ctx.check(ctx.L.vd_index_build(ctx.h, C.c_void_p(ds), C.c_void_p(dp),
                                       C.c_void_p(do), n, mask, C.byref(h)))
        self.h = h
        return self

    def view(self, perm: str) -> PermView:
        v = PermView()
        self.ctx.check(self.ctx.L.vd_index_view(self.h, PERMS[perm], C.byref(v)))
        return v

    def arrays(self, perm: str) -> dict:
        v = self.view(perm)
        d = self.ctx.download
        return {"l1v": d(v.l1_values, v.n_l1), "l1o": d(v.l1_offsets, v.n_l1),
                "l2v": d(v.l2_values, v.n_l2), "l2o": d(v.l2_offsets, v.n_l2),
                "y": d(v.col_y, v.n), "z": d(v.col_z, v.n)}

    def histogram(self, perm: str, depth: bool, param: int, id_limit: int = 0xFFFFFFFF):
        """Equal-width (depth=False) or equal-depth histogram of the leading
        column: (lowers, uppers, freqs) arrays."""
        cap = 1 << 16
        while True:
            lo, hi = np.empty(cap, np.uint32), np.empty(cap, np.uint32)
            fr = np.empty(cap, np.uint64)
            nb = C.c_uint64()
            rc = self.ctx.L.vd_index_histogram(self.ctx.h, self.h, PERMS[perm], int(depth), param,
                                               id_limit, _ptr(lo), _ptr(hi), _ptr(fr), cap,
                                               C.byref(nb))
            if rc == VD_ERANGE and nb.value > cap:
                cap = nb.value
                continue
            self.ctx.check(rc)
            k = nb.value
            return lo[:k].copy(), hi[:k].copy(), fr[:k].copy()

    def lookup(self, perm: str, k1: int, k2: int | None = None):
        b, e = C.c_uint64(), C.c_uint64()
        self.ctx.check(self.ctx.L.vd_index_lookup(
            self.ctx.h, self.h, PERMS[perm], k1, int(k2 is not None),
            0 if k2 is None else k2, C.byref(b), C.byref(e)))
        return b.value, e.value

    def select(self, perm: str, begin: int, end: int, cols: int = 3, clip=None,
               skip=(), distinct: bool = False) -> Rel:
        sk = (Interval * max(1, len(skip)))(*[Interval(a, b) for a, b in skip])
        lo, hi = clip if clip is not None else (0, 0)
        h = C.c_void_p()
        self.ctx.check(self.ctx.L.vd_select(
            self.ctx.h, self.h, PERMS[perm], begin, end, cols,
            int(clip is not None), lo, hi, sk, len(skip), int(distinct), C.byref(h)))
        return Rel(self.ctx, h)

    def select_key(self, perm: str, k1: int, k2=None, cols: int = 3, clip=None, skip=(),
                   distinct: bool = False):
        """Lookup + select in one device step; returns (Rel, begin, end)."""
        sk = (Interval * max(1, len(skip)))(*[Interval(a, b) for a, b in skip])
        lo, hi = clip if clip is not None else (0, 0)
        h, b, e = C.c_void_p(), C.c_uint64(), C.c_uint64()
        self.ctx.check(self.ctx.L.vd_select_key(
            self.ctx.h, self.h, PERMS[perm], k1, int(k2 is not None), 0 if k2 is None else k2,
            cols, int(clip is not None), lo, hi, C.cast(sk, C.c_void_p), len(skip),
            int(distinct), C.byref(h), C.byref(b), C.byref(e)))
        return Rel(self.ctx, h), b.value, e.value

    def free(self) -> None:
        if self.h and getattr(self, "owned", True):
            self.ctx.L.vd_index_free(self.ctx.h, self.h)
        self.h = None


# ---- batched operators ---------------------------------------------------------
def _handles(rels):
    arr = (C.c_void_p * max(1, len(rels)))()
    for i, r in enumerate(rels):
        arr[i] = r.h.value if isinstance(r.h, C.c_void_p) else r.h
    return arr


def select_multi(ctx: Ctx, index: "Index", jobs) -> list:
    """jobs: [(perm name, k1, k2 or None, cols)] -> [Rel]"""
    arr = (SelectJob * max(1, len(jobs)))(*[
        SelectJob(PERMS[p], k1, VD_UNBOUND if k2 is None else k2, cols) for p, k1, k2, cols in jobs])
    outs = (C.c_void_p * max(1, len(jobs)))()
    ctx.check(ctx.L.vd_select_multi(ctx.h, index.h, C.cast(arr, C.c_void_p), len(jobs), outs))
    return [Rel(ctx, C.c_void_p(outs[i])) for i in range(len(jobs))]


def lookup_multi(ctx: Ctx, index: "Index", jobs) -> list:
    """jobs: [(perm name, k1, k2 or None)] -> [(begin, end)], one launch."""
    arr = (SelectJob * max(1, len(jobs)))(*[
        SelectJob(PERMS[p], k1, VD_UNBOUND if k2 is None else k2, 1) for p, k1, k2 in jobs])
    b = (C.c_uint64 * max(1, len(jobs)))()
    e = (C.c_uint64 * max(1, len(jobs)))()
    ctx.check(ctx.L.vd_index_lookup_multi(ctx.h, index.h, C.cast(arr, C.c_void_p), len(jobs),
                                          b, e))
    return [(int(b[i]), int(e[i])) for i in range(len(jobs))]


def bounds_multi(ctx: Ctx, index: "Index", jobs) -> list:
    """jobs: [(perm name, k1, k2 or None)] -> [(begin, end, first, last)]:
    the range and the first / last value of the column after the constants
    (Y for one constant, Z for two), one launch."""
    arr = (SelectJob * max(1, len(jobs)))(*[
        SelectJob(PERMS[p], k1, VD_UNBOUND if k2 is None else k2, 1) for p, k1, k2 in jobs])
    m = max(1, len(jobs))
    b, e = (C.c_uint64 * m)(), (C.c_uint64 * m)()
    f, l = (C.c_uint32 * m)(), (C.c_uint32 * m)()
    ctx.check(ctx.L.vd_index_bounds_multi(ctx.h, index.h, C.cast(arr, C.c_void_p), len(jobs),
                                          b, e, f, l))
    return [(int(b[i]), int(e[i]), int(f[i]), int(l[i])) for i in range(len(jobs))]


def join_multi(ctx: Ctx, lefts, rights) -> list:
    outs = (C.c_void_p * max(1, len(lefts)))()
    ctx.check(ctx.L.vd_join_multi(ctx.h, _handles(lefts), _handles(rights), len(lefts), outs))
    return [Rel(ctx, C.c_void_p(outs[i])) for i in range(len(lefts))]


def sort_col0_multi(ctx: Ctx, rels) -> None:
    ctx.check(ctx.L.vd_sort_col0_multi(ctx.h, _handles(rels), len(rels)))


def unique_adjacent_multi(ctx: Ctx, rels) -> None:
    ctx.check(ctx.L.vd_unique_adjacent_multi(ctx.h, _handles(rels), len(rels)))
