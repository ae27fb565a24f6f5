"""ctypes front of oracle/liboracle.so (the C restatement) plus a naive
nested-loop BGP evaluator.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the vedas_amd package.
"""
from __future__ import annotations

import ctypes as C
import itertools
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
PERMS = {"SPO": 0, "SOP": 1, "PSO": 2, "POS": 3, "OSP": 4, "OPS": 5}
PERM_COLS = {"SPO": (0, 1, 2), "SOP": (0, 2, 1), "PSO": (1, 0, 2),
             "POS": (1, 2, 0), "OSP": (2, 0, 1), "OPS": (2, 1, 0)}

_lib = None
U32P = C.POINTER(C.c_uint32)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built: run python -m vedas_amd.build oracle")
        l = C.CDLL(LIB_PATH)
        vp, u64, i = C.c_void_p, C.c_uint64, C.c_int
        l.or_build_perm.restype = i
        l.or_build_perm.argtypes = [vp, vp, vp, u64, i, vp, vp, C.POINTER(u64), vp, vp,
                                    C.POINTER(u64), vp, vp]
        l.or_lookup.restype = None
        l.or_lookup.argtypes = [vp, vp, u64, vp, vp, u64, u64, C.c_uint32, i,
                                C.c_uint32, C.POINTER(u64), C.POINTER(u64)]
        l.or_select.restype = u64
        l.or_select.argtypes = [vp, vp, u64, u64, i, i, C.c_uint32, C.c_uint32, vp, i,
                                i, vp, vp]
        l.or_join.restype = u64
        l.or_join.argtypes = [C.POINTER(U32P), u64, i, C.POINTER(U32P), u64, i, i,
                              C.POINTER(U32P)]
        l.or_sort_col0.restype = i
        l.or_sort_col0.argtypes = [C.POINTER(U32P), i, u64]
        l.or_sort_lex.restype = i
        l.or_sort_lex.argtypes = [C.POINTER(U32P), i, u64]
        l.or_unique_adjacent.restype = u64
        l.or_unique_adjacent.argtypes = [C.POINTER(U32P), i, u64]
        l.or_bag_digest.restype = None
        l.or_bag_digest.argtypes = [C.POINTER(U32P), i, u64, vp]
        l.or_perm_digest.restype = None
        l.or_perm_digest.argtypes = [vp, vp, u64, vp, vp, u64, i, vp]
        l.or_check_perm.restype = i
        l.or_check_perm.argtypes = [vp, vp, u64, vp, vp, u64, vp, vp, u64, C.POINTER(u64)]
        _lib = l
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data)


def _colptrs(cols):
    arr = (U32P * max(1, len(cols)))()
    for i, c in enumerate(cols):
        arr[i] = c.ctypes.data_as(U32P)
    return arr


def build_perm(s, p, o, perm: str) -> dict:
    n = len(s)
    m = max(n, 1)
    out = {k: np.zeros(m, np.uint32) for k in ("l1v", "l1o", "l2v", "l2o", "y", "z")}
    n1, n2 = C.c_uint64(), C.c_uint64()
    s, p, o = (np.ascontiguousarray(x, np.uint32) for x in (s, p, o))
    rc = lib().or_build_perm(_p(s), _p(p), _p(o), n, PERMS[perm], _p(out["l1v"]),
                             _p(out["l1o"]), C.byref(n1), _p(out["l2v"]), _p(out["l2o"]),
                             C.byref(n2), _p(out["y"]), _p(out["z"]))
    assert rc == 0
    for k in ("l1v", "l1o"):
        out[k] = out[k][: n1.value].copy()
    for k in ("l2v", "l2o"):
        out[k] = out[k][: n2.value].copy()
    out["y"] = out["y"][:n].copy()
    out["z"] = out["z"][:n].copy()
    return out


def lookup(ix: dict, k1: int, k2: int | None = None):
    b, e = C.c_uint64(), C.c_uint64()
    lib().or_lookup(_p(ix["l1v"]), _p(ix["l1o"]), len(ix["l1v"]), _p(ix["l2v"]),
                    _p(ix["l2o"]), len(ix["l2v"]), len(ix["y"]), k1,
                    int(k2 is not None), 0 if k2 is None else k2, C.byref(b), C.byref(e))
    return b.value, e.value


def select(ix: dict, b: int, e: int, cols: int = 3, clip=None, skip=(),
           distinct: bool = False):
    m = max(e - b, 1)
    o0 = np.zeros(m, np.uint32)
    o1 = np.zeros(m, np.uint32)
    sk = np.array([v for iv in sorted(skip) for v in iv] or [0], np.uint32)
    lo, hi = clip if clip is not None else (0, 0)
    r = lib().or_select(_p(ix["y"]), _p(ix["z"]), b, e, cols, int(clip is not None),
                        lo, hi, _p(sk), len(skip), int(distinct), _p(o0), _p(o1))
    if cols == 3:
        return [o0[:r].copy(), o1[:r].copy()]
    return [o0[:r].copy()]


def join(lcols, rcols, nk: int = 1):
    lcols = [np.ascontiguousarray(c, np.uint32) for c in lcols]
    rcols = [np.ascontiguousarray(c, np.uint32) for c in rcols]
    nl, nr = len(lcols[0]), len(rcols[0])
    lp, rp = _colptrs(lcols), _colptrs(rcols)
    n = lib().or_join(lp, nl, len(lcols), rp, nr, len(rcols), nk, None)
    ncols = len(lcols) + len(rcols) - nk
    out = [np.zeros(max(n, 1), np.uint32) for _ in range(ncols)]
    lib().or_join(lp, nl, len(lcols), rp, nr, len(rcols), nk, _colptrs(out))
    return [c[:n].copy() for c in out]


def sort_col0(cols):
    cols = [np.array(c, np.uint32) for c in cols]
    assert lib().or_sort_col0(_colptrs(cols), len(cols), len(cols[0])) == 0
    return cols


def sort_lex(cols):
    cols = [np.array(c, np.uint32) for c in cols]
    assert lib().or_sort_lex(_colptrs(cols), len(cols), len(cols[0])) == 0
    return cols


def unique_adjacent(cols):
    cols = [np.array(c, np.uint32) for c in cols]
    n = lib().or_unique_adjacent(_colptrs(cols), len(cols), len(cols[0]))
    return [c[:n].copy() for c in cols]


# ---- plan walk ----------------------------------------------------------------
FULL_BOUND = (0, 0xFFFFFFFF)


def update_bound(bnd: dict, var: str, lo: int, hi: int) -> None:
    """QueryExecutor::updateBoundDict (QueryExecutor.cu:583-594): intersect
    var's bound with [lo, hi]; a disjoint interval gives [0, 0], and a lower
    end of 0 keeps 0."""
    cur_lo, cur_hi = bnd[var]
    if hi < cur_lo or lo > cur_hi:
        lo = hi = 0
    bnd[var] = (0 if lo == 0 else max(cur_lo, lo), min(cur_hi, hi))


def prescan_bounds(ix_by_perm: dict, patterns) -> dict:
    """QueryExecutor::createVariableBound (QueryExecutor.cu:441-479) with
    squeezeQueryBound1Var / 2Var, updateDataBound and updateL2Bound
    (:527-581, 648-675, 680-765): variables occurring more than once start
    at the full range; each pattern, in order, narrows them to the first /
    last value of the variable's column over the pattern's index range (the
    column after the constants, sorted there); an empty pattern sets every
    bound to [0, 0].  For one variable the permutation with both constants
    first is used (the first of two that exist); for two variables, the
    object / predicate first as the reference orders them.  Permutations
    missing from ix_by_perm give no bound."""
    count = {}
    for pat in patterns:
        for t in pat:
            if isinstance(t, str):
                count[t] = count.get(t, 0) + 1
    bnd = {v: FULL_BOUND for v, c in count.items() if c > 1}
    zero = False
    for sv, pv, ov in patterns:
        isv = [isinstance(t, str) for t in (sv, pv, ov)]
        if sum(isv) == 1:
            if isv[0]:
                alts = [(sv, "POS", pv, ov), (sv, "OPS", ov, pv)]
            elif isv[1]:
                alts = [(pv, "OSP", ov, sv), (pv, "SOP", sv, ov)]
            else:
                alts = [(ov, "PSO", pv, sv), (ov, "SPO", sv, pv)]
            items = [a for a in alts if a[1] in ix_by_perm][:1]
        elif sum(isv) == 2:
            if not isv[1]:
                items = [(ov, "POS", pv, None), (sv, "PSO", pv, None)]
            elif not isv[0]:
                items = [(ov, "SOP", sv, None), (pv, "SPO", sv, None)]
            else:
                items = [(pv, "OPS", ov, None), (sv, "OSP", ov, None)]
            items = [a for a in items if a[1] in ix_by_perm]
        else:
            items = []
        for var, perm, k1, k2 in items:
            ix = ix_by_perm[perm]
            b, e = lookup(ix, k1, k2)
            if b == e:
                zero = True
                continue
            col = ix["y"] if k2 is None else ix["z"]
            if var in bnd:
                update_bound(bnd, var, int(col[b]), int(col[e - 1]))
    if zero:
        bnd = {v: (0, 0) for v in bnd}
    return bnd


def run_plan(ix_by_perm: dict, patterns, node, stats: dict | None = None,
             bounds: bool = False, ops=None):
    """Execute a plan tree over the oracle: QueryExecutor::postorderTraversal
    (QueryExecutor.cu:106-135) with the index choice of
    create1VarSelectQueryJob / create2VarSelectQueryJob
    (QueryExecutor.cu:781-971: 2-variable pattern -> PSO [?s, ?o] when the
    node's variable is the subject, else POS [?o, ?s]; s p ?o -> PSO (p, s);
    ?s p o -> POS (p, o)) and the join rules of JoinQueryJob::join
    (JoinQueryJob.cu:125-180: smaller relation left, key moved to column 0
    with a re-sort, output [key, left rest, right rest]).  Nodes:
    ("UPLOAD", pattern or index into patterns, var), ("JOIN", l, r, var),
    ("INDEXSWAP", child, var).  ix_by_perm: {"PSO": arrays, "POS": arrays}
    as build_perm returns (other permutations are used by the pre-scan when
    present).  With bounds, the join-variable bounds of the reference are
    applied: the pre-scan (prescan_bounds), the clip of 2-variable selects
    to their first column's bound (SelectQueryJob.cu:125-141), the empty
    1-variable select's bound (:113-119) and the update after each join
    (JoinQueryJob.cu:388-397, 434-437); the empty-interval filter is not
    restated (it acts only after joins of fewer than EIF_ALPHA = 100 rows and
    only shrinks later selects).  stats (optional) accumulates "join_rows",
    "select_rows" and "swap_rows".  ops (optional) supplies other
    implementations of select / join / sort_col0 with these functions'
    signatures (the multi-threaded CPU port).  Returns (headers, columns)."""
    _select = ops.select if ops is not None else select
    _join = ops.join if ops is not None else join
    _sort_col0 = ops.sort_col0 if ops is not None else sort_col0
    st = stats if stats is not None else {}
    st.setdefault("join_rows", 0)
    st.setdefault("select_rows", 0)
    st.setdefault("swap_rows", 0)
    bnd = prescan_bounds(ix_by_perm, patterns) if bounds else None

    def key_first(h, cols, var):
        i = h.index(var)
        if i == 0:
            return h, cols
        h, cols = list(h), list(cols)
        h[0], h[i] = h[i], h[0]
        cols[0], cols[i] = cols[i], cols[0]
        st["swap_rows"] += len(cols[0])
        return h, _sort_col0(cols)

    def upload(nd):
        pat = patterns[nd[1]] if isinstance(nd[1], int) else nd[1]
        sv, pv, ov = pat
        var = nd[2]
        if isinstance(sv, str) and isinstance(ov, str):
            perm, hdr = ("PSO", [sv, ov]) if sv == var else ("POS", [ov, sv])
            ix = ix_by_perm[perm]
            b, e = lookup(ix, pv)
            clip = bnd[hdr[0]] if bnd is not None and hdr[0] in bnd else None
            return hdr, _select(ix, b, e, 3, clip=clip)
        if isinstance(sv, str):
            ix = ix_by_perm["POS"]
            b, e = lookup(ix, pv, ov)
            hdr = [sv]
        else:
            ix = ix_by_perm["PSO"]
            b, e = lookup(ix, pv, sv)
            hdr = [ov]
        if bnd is not None and hdr[0] in bnd and b == e:
            update_bound(bnd, hdr[0], 0, 0)
        return hdr, _select(ix, b, e, 1)

    def walk(nd):
        op = nd[0]
        if op == "UPLOAD":
            h, cols = upload(nd)
            st["select_rows"] += len(cols[0])
            return h, cols
        if op == "JOIN":
            a, b = walk(nd[1]), walk(nd[2])
            if len(a[1][0]) > len(b[1][0]):
                a, b = b, a
            ha, ca = key_first(a[0], a[1], nd[3])
            hb, cb = key_first(b[0], b[1], nd[3])
            out = _join(ca, cb, 1)
            st["join_rows"] += len(out[0])
            if bnd is not None and nd[3] in bnd:
                if len(out[0]):
                    update_bound(bnd, nd[3], int(out[0][0]), int(out[0][-1]))
                else:
                    update_bound(bnd, nd[3], 0, 0)
            return [nd[3]] + ha[1:] + hb[1:], out
        if op == "INDEXSWAP":
            h, cols = walk(nd[1])
            return key_first(h, cols, nd[2])
        raise ValueError(op)

    out = walk(node)
    if bnd is not None:
        st["bounds"] = dict(bnd)
    return out


# ---- naive evaluator -----------------------------------------------------
def naive_bgp(s, p, o, patterns):
    """Bag of bindings of a basic graph pattern by nested loops over the
    triples (a different algorithm from the plan of selects and joins).
    patterns: list of (S, P, O) with ints for constants and '?x' strings for
    variables.  Returns (sorted variable names, sorted list of row tuples)."""
    triples = list(zip(s.tolist(), p.tolist(), o.tolist()))
    vars_ = sorted({t for pat in patterns for t in pat if isinstance(t, str)})
    rows = []

    def rec(k, env):
        if k == len(patterns):
            rows.append(tuple(env[v] for v in vars_))
            return
        pat = patterns[k]
        for tr in triples:
            e2 = dict(env)
            ok = True
            for term, val in zip(pat, tr):
                if isinstance(term, str):
                    if term in e2 and e2[term] != val:
                        ok = False
                        break
                    e2[term] = val
                elif term != val:
                    ok = False
                    break
            if ok:
                rec(k + 1, e2)

    rec(0, {})
    rows.sort()
    return vars_, rows


def canonical(headers, cols):
    """Rows of a relation with columns ordered by variable name, sorted."""
    order = sorted(range(len(headers)), key=lambda i: headers[i])
    names = [headers[i] for i in order]
    if len(cols) == 0 or len(cols[0]) == 0:
        return names, []
    mat = np.stack([np.asarray(cols[i], np.uint32) for i in order], axis=1)
    rows = sorted(map(tuple, mat.tolist()))
    return names, rows


# ---- histograms -------------------------------------------------------------------
def hist_eq_width(col, width: int, id_limit: int):
    """Histogram::initEqWidth (Histogram.cu:170-199) over a sorted key
    column: only keys below id_limit; buckets of `width` IDs from the first
    key, the last one cut at the last key; freq = keys in the bucket.  The
    reference stops when a bucket would start at the last key itself (its
    loop runs while id < endId), which leaves that key with no bucket; here
    the last key always has one.  Returns (lowers, uppers, freqs) lists."""
    col = np.asarray(col)
    m = int(np.searchsorted(col, id_limit, "left"))
    if m == 0:
        return [], [], []
    start, end = int(col[0]), int(col[m - 1])
    lowers, uppers = [], []
    lo = start
    while lo <= end:
        lowers.append(lo)
        uppers.append(min(lo + width - 1, end))
        lo += width
    freq = [0] * len(lowers)
    k = 0
    for v in col[:m].tolist():
        while v > uppers[k]:
            k += 1
        freq[k] += 1
    return lowers, uppers, freq


def hist_eq_depth(col, depth: int, id_limit: int):
    """Histogram::initEqDepth (Histogram.cu:201-231) over a sorted key
    column: only keys below id_limit; a bucket starts at a row, its end is
    the row `depth` rows later (or the last row), moved on to the end of
    that key's run; lower / upper are the keys at its start and end, freq
    its rows (depth + 1 or more).  Returns (lowers, uppers, freqs)."""
    col = np.asarray(col)
    m = int(np.searchsorted(col, id_limit, "left"))
    c = col[:m].tolist()
    lowers, uppers, freq = [], [], []
    d = 0
    while d < m:
        s = d
        d += depth
        if d >= m:
            d = m - 1
        last = c[d]
        while d < m - 1 and c[d] == c[d + 1]:
            d += 1
        lowers.append(c[s])
        uppers.append(last)
        freq.append(d - s + 1)
        d += 1
    return lowers, uppers, freq


# ---- checkers for full-size results -------------------------------------------
def bag_digest(headers, cols):
    """Order-independent digest of a relation's rows with columns ordered by
    variable name: (names, rows, sum of one 64-bit row hash, sum of another).
    Equal bags of rows give equal digests; compare two relations with ==."""
    order = sorted(range(len(headers)), key=lambda i: headers[i])
    names = tuple(headers[i] for i in order)
    return (names,) + _raw_digest([cols[i] for i in order])


def _raw_digest(cols):
    cs = [np.ascontiguousarray(c, np.uint32) for c in cols]
    n = len(cs[0]) if cs else 0
    out = np.zeros(2, np.uint64)
    lib().or_bag_digest(_colptrs(cs), len(cs), n, _p(out))
    return n, int(out[0]), int(out[1])


def triples_digest(s, p, o):
    """Bag digest of (s, p, o) triples, columns in that order: (rows, sum, sum)."""
    return _raw_digest([s, p, o])


def perm_digest(ix: dict, perm: str):
    """Bag digest of a permutation's triples in (s, p, o) order; equals
    triples_digest of the input exactly when the multisets are equal."""
    out = np.zeros(2, np.uint64)
    n = len(ix["y"])
    l1v = np.ascontiguousarray(ix["l1v"], np.uint32)
    l1o = np.ascontiguousarray(ix["l1o"], np.uint32)
    lib().or_perm_digest(_p(l1v), _p(l1o), len(l1v), _p(ix["y"]), _p(ix["z"]), n,
                         PERMS[perm], _p(out))
    return n, int(out[0]), int(out[1])


CHECK_PERM_FAILURES = {
    1: "L1 empty / not empty against the row count",
    2: "L1 offsets not strictly increasing from 0 below n",
    3: "L1 values not strictly increasing",
    4: "rows not in (X, Y, Z) order",
    5: "(X, Y) run head without L2 entry, or L2 entry off a head",
    6: "L2 value differs from the row's Y",
    7: "L2 entries left after the last row",
}


def check_perm(ix: dict):
    """Raise AssertionError unless the arrays of one permutation are sorted
    by (X, Y, Z) with L1/L2 at exactly the X and (X, Y) run heads."""
    a = {k: np.ascontiguousarray(ix[k], np.uint32) for k in ("l1v", "l1o", "l2v", "l2o", "y", "z")}
    assert len(a["l1v"]) == len(a["l1o"]), "L1 values and offsets differ in length"
    assert len(a["l2v"]) == len(a["l2o"]), "L2 values and offsets differ in length"
    assert len(a["y"]) == len(a["z"]), "Y and Z columns differ in length"
    where = C.c_uint64()
    rc = lib().or_check_perm(_p(a["l1v"]), _p(a["l1o"]), len(a["l1v"]), _p(a["l2v"]),
                             _p(a["l2o"]), len(a["l2v"]), _p(a["y"]), _p(a["z"]), len(a["y"]),
                             C.byref(where))
    assert rc == 0, f"permutation check {rc}: {CHECK_PERM_FAILURES.get(rc)} at {where.value}"


def groupby_starts(keys):
    """(value, first index) of each run of equal consecutive keys."""
    out = []
    i = 0
    for k, g in itertools.groupby(keys):
        out.append((k, i))
        i += sum(1 for _ in g)
    return out
