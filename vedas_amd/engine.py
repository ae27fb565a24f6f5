"""ctypes binding of include/vedas_host.h (libvedas.so): queries through the
C++ operator API (QueryExecutor -> SelectQueryJob / JoinQueryJob /
IndexSwapJob -> QueryPlan), the drop-in path of the reference's vdQuery."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _native

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvedas.so")


class Stats(C.Structure):
    _fields_ = [("rows", C.c_uint64), ("join_rows", C.c_uint64),
                ("select_rows", C.c_uint64), ("swap_rows", C.c_uint64),
                ("elapsed_ms", C.c_double), ("join_ms", C.c_double),
                ("select_ms", C.c_double), ("swap_ms", C.c_double), ("jobs", C.c_int),
                ("driven_joins", C.c_int), ("filtered_rows", C.c_uint64)]


class BatchStats(C.Structure):
    _fields_ = [("rounds", C.c_int), ("calls", C.c_int), ("join_rows", C.c_uint64),
                ("select_rows", C.c_uint64), ("swap_rows", C.c_uint64),
                ("elapsed_ms", C.c_double)]


class ServeCfg(C.Structure):
    _fields_ = [("slots", C.c_int), ("max_batch", C.c_int), ("mem_budget", C.c_uint64),
                ("plain", C.c_int), ("min_fill", C.c_int), ("max_wait_ms", C.c_double)]


class ServeRec(C.Structure):
    _fields_ = [("arrival_ms", C.c_double), ("start_ms", C.c_double), ("done_ms", C.c_double),
                ("est_bytes", C.c_double), ("rows", C.c_uint64), ("batch", C.c_int),
                ("slot", C.c_int), ("mem_held", C.c_int)]


class TuneResult(C.Structure):
    _fields_ = [("radix_items", C.c_int), ("join_nv", C.c_int), ("join_rcap", C.c_int),
                ("radix_bits", C.c_int), ("base_ms", C.c_double), ("tuned_ms", C.c_double),
                ("evaluations", C.c_int), ("radix_group", C.c_int)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built: run python -m vedas_amd.build")
        _native.lib()  # libvedas_hip.so first (same directory, rpath $ORIGIN)
        l = C.CDLL(LIB_PATH)
        vp, u64, i = C.c_void_p, C.c_uint64, C.c_int
        sig = {
            "vdq_engine_create": (i, [i, C.POINTER(vp)]),
            "vdq_engine_free": (i, [vp]),
            "vdq_engine_load": (i, [vp, vp, vp, vp, u64, C.c_uint32]),
            "vdq_engine_load_device": (i, [vp, vp, vp, vp, u64, C.c_uint32]),
            "vdq_engine_ctx": (vp, [vp]),
            "vdq_dict_create": (i, [i, C.POINTER(vp)]),
            "vdq_dict_free": (i, [vp]),
            "vdq_dict_load_ntriples": (i, [vp, C.c_char_p, C.POINTER(u64)]),
            "vdq_dict_load_ntriples_text": (i, [vp, C.c_char_p, C.POINTER(u64)]),
            "vdq_dict_triples": (i, [vp, vp, vp, vp]),
            "vdq_dict_term": (i, [vp, i, C.c_uint32, C.c_char_p, C.c_size_t]),
            "vdq_dict_translate": (i, [vp, C.c_char_p, C.c_char_p, C.c_size_t]),
            "vdq_engine_load_dict": (i, [vp, vp, C.c_uint32]),
            "vdq_dict_save": (i, [vp, C.c_char_p]),
            "vdq_dict_open": (i, [vp, C.c_char_p]),
            "vdq_engine_set_dict": (i, [vp, vp]),
            "vdq_sparql": (i, [vp, C.c_char_p, C.POINTER(vp)]),
            "vdq_result_is_predicate": (i, [vp, i]),
            "vdq_engine_save": (i, [vp, C.c_char_p]),
            "vdq_engine_open": (i, [vp, C.c_char_p]),
            "vdq_engine_index": (vp, [vp]),
            "vdq_query": (i, [vp, C.c_char_p, C.POINTER(vp)]),
            "vdq_set_option": (i, [C.c_char_p, i]),
            "vdq_get_option": (i, [C.c_char_p, C.POINTER(i)]),
            "vdq_result_shape": (i, [vp, C.POINTER(i), C.POINTER(u64)]),
            "vdq_result_header": (C.c_char_p, [vp, i]),
            "vdq_result_column": (i, [vp, i, vp]),
            "vdq_result_stats": (i, [vp, C.POINTER(Stats)]),
            "vdq_result_free": (i, [vp]),
            "vdq_last_error": (C.c_char_p, []),
            "vdq_batch": (i, [vp, C.POINTER(C.c_char_p), i, C.POINTER(vp),
                              C.POINTER(BatchStats)]),
            "vdq_serve": (i, [vp, C.POINTER(C.c_char_p), C.POINTER(C.c_double), i,
                              C.POINTER(ServeCfg), C.POINTER(ServeRec), C.POINTER(vp)]),
            "vdq_estimate": (i, [vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "vdq_autotune": (i, [vp, C.POINTER(C.c_char_p), i, i, C.POINTER(TuneResult)]),
            "vdq_plan": (i, [C.c_char_p, C.POINTER(u64), C.c_char_p, C.c_size_t]),
            "vdq_graph_create": (i, [vp, C.POINTER(C.c_char_p), i, C.POINTER(vp)]),
            "vdq_graph_replay": (i, [vp]),
            "vdq_graph_result": (vp, [vp, i]),
            "vdq_graph_free": (i, [vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(l, name)
            f.restype, f.argtypes = res, args
        _lib = l
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise _native.VedasError(rc, (lib().vdq_last_error() or b"").decode())


class Result:
    def __init__(self, h):
        self.h = h

    def shape(self):
        nc, nr = C.c_int(), C.c_uint64()
        _check(lib().vdq_result_shape(self.h, C.byref(nc), C.byref(nr)))
        return nc.value, nr.value

    def headers(self):
        nc, _ = self.shape()
        return [lib().vdq_result_header(self.h, i).decode() for i in range(nc)]

    def columns(self):
        nc, nr = self.shape()
        out = []
        for i in range(nc):
            a = np.empty(nr, np.uint32)
            _check(lib().vdq_result_column(self.h, i, C.c_void_p(a.ctypes.data)))
            out.append(a)
        return out

    def is_predicate(self, col: int) -> bool:
        return bool(lib().vdq_result_is_predicate(self.h, col))

    def terms(self, d: "Dict"):
        """Result rows as term strings through the dictionary."""
        cols = self.columns()
        preds = [self.is_predicate(c) for c in range(len(cols))]
        n = len(cols[0]) if cols else 0
        return [tuple(d.term(int(cols[c][r]), preds[c]) for c in range(len(cols)))
                for r in range(n)]

    def stats(self) -> dict:
        s = Stats()
        _check(lib().vdq_result_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def free(self):
        if self.h:
            lib().vdq_result_free(self.h)
            self.h = None


class GraphResult(Result):
    """A result owned by a Graph: rewritten by every replay, freed with it."""

    def free(self):
        self.h = None


class Graph:
    """A batch captured as one HIP graph (vdq_graph_create): replay() reruns
    every kernel of the batch over the same store as one launch and
    rewrites the results in place."""

    def __init__(self, engine: "Engine", texts):
        n = len(texts)
        arr = (C.c_char_p * max(1, n))(*[t.encode() for t in texts])
        h = C.c_void_p()
        _check(lib().vdq_graph_create(engine.h, arr, n, C.byref(h)))
        self.h, self.n = h, n

    def replay(self) -> None:
        _check(lib().vdq_graph_replay(self.h))

    def result(self, i: int) -> GraphResult:
        p = lib().vdq_graph_result(self.h, i)
        if not p:
            raise IndexError(i)
        return GraphResult(C.c_void_p(p))

    def free(self) -> None:
        if self.h:
            _check(lib().vdq_graph_free(self.h))
            self.h = None


class Engine:
    """A device context plus an HBM-resident triple store."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().vdq_engine_create(device, C.byref(h)))
        self.h = h

    def load(self, s, p, o, perms=("PSO", "POS")):
        mask = 0
        for name in perms:
            mask |= 1 << _native.PERMS[name]
        s, p, o = (np.ascontiguousarray(x, np.uint32) for x in (s, p, o))
        _check(lib().vdq_engine_load(self.h, C.c_void_p(s.ctypes.data), C.c_void_p(p.ctypes.data),
                                     C.c_void_p(o.ctypes.data), len(s), mask))

    def load_device(self, ds: int, dp: int, do: int, n: int, perms):
        mask = 0
        for name in perms:
            mask |= 1 << _native.PERMS[name]
        _check(lib().vdq_engine_load_device(self.h, C.c_void_p(ds), C.c_void_p(dp),
                                            C.c_void_p(do), n, mask))

    def load_dict(self, d: "Dict", perms=("PSO", "POS")) -> None:
        """Build the store from a dictionary's triples; SPARQL then works."""
        mask = 0
        for name in perms:
            mask |= 1 << _native.PERMS[name]
        _check(lib().vdq_engine_load_dict(self.h, d.h, mask))
        self._dict = d

    def set_dict(self, d: "Dict") -> None:
        """The dictionary of a store opened from a .vds file (for sparql)."""
        _check(lib().vdq_engine_set_dict(self.h, d.h))
        self._dict = d

    def sparql(self, text: str) -> "Result":
        h = C.c_void_p()
        _check(lib().vdq_sparql(self.h, text.encode(), C.byref(h)))
        return Result(h)

    def save(self, path: str) -> None:
        """Write the store as a .vds file (the reference's default index)."""
        _check(lib().vdq_engine_save(self.h, path.encode()))

    def open(self, path: str) -> None:
        """Replace the store with a .vds file's permutations."""
        _check(lib().vdq_engine_open(self.h, path.encode()))

    def ctx_handle(self):
        return lib().vdq_engine_ctx(self.h)

    def index_handle(self):
        return lib().vdq_engine_index(self.h)

    def query(self, text: str) -> Result:
        r = C.c_void_p()
        _check(lib().vdq_query(self.h, text.encode(), C.byref(r)))
        return Result(r)

    def batch(self, texts):
        """Run many queries together; returns ([Result], stats dict)."""
        n = len(texts)
        arr = (C.c_char_p * max(1, n))(*[t.encode() for t in texts])
        outs = (C.c_void_p * max(1, n))()
        st = BatchStats()
        _check(lib().vdq_batch(self.h, arr, n, outs, C.byref(st)))
        return ([Result(C.c_void_p(outs[i])) for i in range(n)],
                {k: getattr(st, k) for k, _ in BatchStats._fields_})

    def graph(self, texts) -> Graph:
        """The batch of `texts` captured as one replayable HIP graph."""
        return Graph(self, texts)

    def serve(self, texts, arrival_ms, slots=4, max_batch=64, mem_budget=64 << 30,
              plain=False, keep_results=False, min_fill=0, max_wait_ms=0.0):
        """Serve an open-loop arrival stream; returns (records, [Result] or None).
        min_fill / max_wait_ms: the fill policy (see vdq_serve_cfg_t)."""
        n = len(texts)
        arr = (C.c_char_p * max(1, n))(*[t.encode() for t in texts])
        arrival = (C.c_double * max(1, n))(*arrival_ms)
        recs = (ServeRec * max(1, n))()
        outs = (C.c_void_p * max(1, n))() if keep_results else None
        cfg = ServeCfg(slots, max_batch, mem_budget, int(plain), min_fill, max_wait_ms)
        _check(lib().vdq_serve(self.h, arr, arrival, n, C.byref(cfg), recs, outs))
        out = [{k: getattr(recs[i], k) for k, _ in ServeRec._fields_} for i in range(n)]
        res = [Result(C.c_void_p(outs[i])) for i in range(n)] if keep_results else None
        return out, res

    def autotune(self, texts, reps: int = 2) -> dict:
        """Pick the kernel tunables by timing `texts` (coordinate descent);
        returns the chosen values and the before/after times."""
        arr = (C.c_char_p * max(1, len(texts)))(*[t.encode() for t in texts])
        out = TuneResult()
        _check(lib().vdq_autotune(self.h, arr, len(texts), reps, C.byref(out)))
        return {k: getattr(out, k) for k, _ in TuneResult._fields_}

    def estimate(self, text: str):
        r, b = C.c_double(), C.c_double()
        _check(lib().vdq_estimate(self.h, text.encode(), C.byref(r), C.byref(b)))
        return r.value, b.value

    def close(self):
        if self.h:
            lib().vdq_engine_free(self.h)
            self.h = None


class Dict:
    """Term dictionary of an N-Triples store (host only)."""

    def __init__(self, literal_dict: bool = False):
        h = C.c_void_p()
        _check(lib().vdq_dict_create(int(literal_dict), C.byref(h)))
        self.h = h
        self.n = 0

    def load(self, path: str) -> int:
        n = C.c_uint64()
        _check(lib().vdq_dict_load_ntriples(self.h, path.encode(), C.byref(n)))
        self.n = n.value
        return self.n

    def load_text(self, text: str) -> int:
        n = C.c_uint64()
        _check(lib().vdq_dict_load_ntriples_text(self.h, text.encode(), C.byref(n)))
        self.n = n.value
        return self.n

    def save(self, path: str) -> None:
        """Write the .vdd dictionary file."""
        _check(lib().vdq_dict_save(self.h, path.encode()))

    def open(self, path: str) -> None:
        """Replace the terms with a .vdd file's (no triples afterwards)."""
        _check(lib().vdq_dict_open(self.h, path.encode()))
        self.n = 0

    def triples(self):
        import numpy as np
        s, p, o = (np.empty(self.n, np.uint32) for _ in range(3))
        _check(lib().vdq_dict_triples(self.h, s.ctypes.data, p.ctypes.data, o.ctypes.data))
        return s, p, o

    def term(self, ident: int, is_predicate: bool = False) -> str:
        buf = C.create_string_buffer(1 << 16)
        _check(lib().vdq_dict_term(self.h, int(is_predicate), ident, buf, len(buf)))
        return buf.value.decode()

    def translate(self, sparql: str) -> str:
        buf = C.create_string_buffer(1 << 16)
        _check(lib().vdq_dict_translate(self.h, sparql.encode(), buf, len(buf)))
        return buf.value.decode()

    def close(self):
        if self.h:
            lib().vdq_dict_free(self.h)
            self.h = None


def set_option(key: str, value) -> None:
    """Executor switch ("bound_prescan", "bound_after_join"), process-wide."""
    _check(lib().vdq_set_option(key.encode(), int(value)))


def get_option(key: str) -> int:
    v = C.c_int()
    _check(lib().vdq_get_option(key.encode(), C.byref(v)))
    return v.value


def plan(text: str, est_rows=None) -> str:
    """PLAN clause of a query text: its own, or the query graph planner's
    from est_rows (rows of each pattern's index range).  Host only."""
    est = None
    if est_rows is not None:
        est = (C.c_uint64 * max(1, len(est_rows)))(*[int(x) for x in est_rows])
    buf = C.create_string_buffer(1 << 16)
    _check(lib().vdq_plan(text.encode(), est, buf, len(buf)))
    return buf.value.decode()


def parse_plan(clause: str):
    """PLAN clause -> plan tree of ("UPLOAD", index, var) / ("JOIN", l, r,
    var) / ("INDEXSWAP", child, var) nodes."""
    toks = clause.replace("(", " ( ").replace(")", " ) ").split()
    pos = 0

    def node():
        nonlocal pos
        assert toks[pos] == "("
        op = toks[pos + 1]
        pos += 2
        if op == "U":
            out = ("UPLOAD", int(toks[pos]), toks[pos + 1])
            pos += 2
        elif op == "J":
            var = toks[pos]
            pos += 1
            a = node()
            b = node()
            out = ("JOIN", a, b, var)
        elif op == "S":
            var = toks[pos]
            pos += 1
            out = ("INDEXSWAP", node(), var)
        else:
            raise ValueError(op)
        assert toks[pos] == ")"
        pos += 1
        return out

    return node()


# ---- query text helpers ---------------------------------------------------------
def plan_text(node) -> str:
    """Plan tree (tests/plans.py form) -> the PLAN s-expression."""
    op = node[0]
    if op == "UPLOAD":
        return f"(U {node[1]} {node[2]})"
    if op == "JOIN":
        return f"(J {node[3]} {plan_text(node[1])} {plan_text(node[2])})"
    if op == "INDEXSWAP":
        return f"(S {node[2]} {plan_text(node[1])})"
    raise ValueError(op)


def query_text(patterns, node, select="*") -> str:
    """patterns: list of (s, p, o) terms; node: plan tree whose UPLOAD
    leaves hold the pattern tuple itself or its index (None: no PLAN)."""
    idx = {tuple(p): i for i, p in enumerate(patterns)}

    def fix(n):
        if n[0] == "UPLOAD":
            pat = n[1]
            i = pat if isinstance(pat, int) else idx[tuple(pat)]
            return ("UPLOAD", i, n[2])
        if n[0] == "JOIN":
            return ("JOIN", fix(n[1]), fix(n[2]), n[3])
        return ("INDEXSWAP", fix(n[1]), n[2])

    body = " . ".join(" ".join(str(t) for t in pat) for pat in patterns)
    sel = select if isinstance(select, str) else " ".join(select)
    if node is None:  # planned by the engine's query graph planner
        return f"SELECT {sel} WHERE {{ {body} }}"
    return f"SELECT {sel} WHERE {{ {body} }} PLAN {plan_text(fix(node))}"
