"""Plan trees for the benchmark query shapes and two executors of them:
one over the HIP C ABI (vedas_amd._native) and one over the oracle.

A plan node mirrors the reference's PlanTreeNode (PlanTreeNode.h:9-20):
("UPLOAD", pattern, var) | ("JOIN", left, right, var) |
("INDEXSWAP", child, var).  Execution follows
QueryExecutor::postorderTraversal (QueryExecutor.cu:106-135) with the
index choice of create1VarSelectQueryJob / create2VarSelectQueryJob
(QueryExecutor.cu:781-971) and the join rules of JoinQueryJob::join
(JoinQueryJob.cu:125-180): smaller relation left, key moved to column 0
with a re-sort when it is not there, output [key, left rest, right rest].
Every variable is selected (SELECT *), so no column is pruned.
"""
from __future__ import annotations

import numpy as np


def is_var(t) -> bool:
    return isinstance(t, str)


# ---- plan shapes -------------------------------------------------------------
def star_plan(preds, subj="?s", objs=None):
    objs = objs or [f"?o{i}" for i in range(len(preds))]
    pats = [(subj, p, o) for p, o in zip(preds, objs)]
    node = ("UPLOAD", pats[0], subj)
    for pat in pats[1:]:
        node = ("JOIN", node, ("UPLOAD", pat, subj), subj)
    return pats, node


def chain_plan(preds, names=None):
    names = names or [f"?v{i}" for i in range(len(preds) + 1)]
    pats = [(names[i], p, names[i + 1]) for i, p in enumerate(preds)]
    # first join on names[1]: pattern 0 joined through its object
    node = ("JOIN", ("UPLOAD", pats[0], names[1]), ("UPLOAD", pats[1], names[1]), names[1])
    for i in range(2, len(preds)):
        v = names[i]
        node = ("JOIN", ("INDEXSWAP", node, v), ("UPLOAD", pats[i], v), v)
    return pats, node


def snowflake_plan(preds_a, preds_b, x="?x", y="?y"):
    """Star on ?x whose first pattern's object ?y is the centre of a second
    star; the two stars are joined on ?y."""
    objs_a = [y] + [f"?a{i}" for i in range(1, len(preds_a))]
    pats_a, node_a = star_plan(preds_a, x, objs_a)
    pats_b, node_b = star_plan(preds_b, y, [f"?b{i}" for i in range(len(preds_b))])
    node = ("JOIN", ("INDEXSWAP", node_a, y), node_b, y)
    return pats_a + pats_b, node


# ---- executors -------------------------------------------------------------------
class Relation:
    def __init__(self, headers, data):
        self.headers = list(headers)
        self.data = data  # backend specific


class Executor:
    """Backend-neutral plan walker; subclasses implement the primitives."""

    def run(self, node) -> Relation:
        op = node[0]
        if op == "UPLOAD":
            return self.upload(node[1], node[2])
        if op == "JOIN":
            a = self.run(node[1])
            b = self.run(node[2])
            return self.join(a, b, node[3])
        if op == "INDEXSWAP":
            r = self.run(node[1])
            return self.swap(r, node[2])
        raise ValueError(op)

    def upload(self, pat, var) -> Relation:
        s, p, o = pat
        assert not is_var(p), "predicate variables need the full index"
        if is_var(s) and is_var(o):
            if s == var:  # PSO: [?s, ?o]
                return self.select2("PSO", p, [s, o])
            return self.select2("POS", p, [o, s])  # POS: [?o, ?s]
        if is_var(s):  # ?s p o -> POS (p, o) -> [?s]
            return self.select1("POS", p, o, s)
        return self.select1("PSO", p, s, o)  # s p ?o -> PSO (p, s) -> [?o]

    def join(self, a: Relation, b: Relation, var) -> Relation:
        if self.size(a) > self.size(b):
            a, b = b, a
        a = self.key_first(a, var)
        b = self.key_first(b, var)
        headers = [var] + a.headers[1:] + b.headers[1:]
        return Relation(headers, self.join_data(a, b))

    def key_first(self, r: Relation, var) -> Relation:
        i = r.headers.index(var)
        if i == 0:
            return r
        h = list(r.headers)
        h[0], h[i] = h[i], h[0]
        return Relation(h, self.swap_sort(r.data, i))

    def swap(self, r: Relation, var) -> Relation:
        i = r.headers.index(var)
        h = list(r.headers)
        h[0], h[i] = h[i], h[0]
        return Relation(h, self.swap_sort(r.data, i))


class OracleExecutor(Executor):
    def __init__(self, s, p, o):
        from oracle import pyoracle as O
        self.O = O
        self.ix = {name: O.build_perm(s, p, o, name) for name in ("PSO", "POS")}

    def size(self, r):
        return len(r.data[0])

    def select2(self, perm, p, headers):
        ix = self.ix[perm]
        b, e = self.O.lookup(ix, p)
        return Relation(headers, self.O.select(ix, b, e, 3))

    def select1(self, perm, k1, k2, header):
        ix = self.ix[perm]
        b, e = self.O.lookup(ix, k1, k2)
        return Relation([header], self.O.select(ix, b, e, 1))

    def join_data(self, a, b):
        return self.O.join(a.data, b.data, 1)

    def swap_sort(self, cols, i):
        cols = list(cols)
        cols[0], cols[i] = cols[i], cols[0]
        return self.O.sort_col0(cols)

    def columns(self, r):
        return r.data


class GpuExecutor(Executor):
    """Plan execution over the HIP C ABI."""

    def __init__(self, ctx, index):
        self.ctx = ctx
        self.index = index
        self.live = []

    def size(self, r):
        return r.data.nrows

    def select2(self, perm, p, headers):
        b, e = self.index.lookup(perm, p)
        rel = self.index.select(perm, b, e, 3)
        self.live.append(rel)
        return Relation(headers, rel)

    def select1(self, perm, k1, k2, header):
        b, e = self.index.lookup(perm, k1, k2)
        rel = self.index.select(perm, b, e, 1)
        self.live.append(rel)
        return Relation([header], rel)

    def join_data(self, a, b):
        rel = a.data.join(b.data, 1)
        self.live.append(rel)
        return rel

    def swap_sort(self, rel, i):
        rel.swap_cols(0, i)
        rel.sort_col0()
        return rel

    def columns(self, r):
        return r.data.columns()

    def free(self):
        for r in self.live:
            r.free()
        self.live = []
