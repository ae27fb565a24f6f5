"""Synthetic workloads: generator parameters of the benchmark stores, the
single-query join set and the seeded mixed query batch.

Sizes are steered with the expected row counts of the generator's model
(triples drawn independently: subject Zipf(s) over E entities, predicate
Zipf over P, object an entity uniform over E with probability 1 - lit or a
literal): for a star around a variable v, E[rows] = sum_v prod_i lambda_i(v)
with lambda = N w_p w_v for a pattern whose subject is v and
lambda = N w_p (1 - lit) / E for a pattern whose object is v.
"""
from __future__ import annotations

import functools

import numpy as np

# 100M-triple store (single-query joins, batches)
STORE_100M = dict(n=100_000_000, n_entities=10_000_000, subject_zipf=0.8, n_predicates=256,
                  pred_zipf=1.1, literal_frac=0.4, literal_base=50_000_000,
                  n_literals=50_000_000)
# 500M-triple store of the served form: 50M entities, 1024 predicates.
# Literal IDs start above every entity ID (> 5e7) and, 5:1 to entities as
# in the other stores, reach 3.1e8: a packed (P, S, O) key then needs
# 11 + 26 + 29 = 66 bits, so this store exercises the two-phase build.
STORE_500M = dict(n=500_000_000, n_entities=50_000_000, subject_zipf=0.8, n_predicates=1024,
                  pred_zipf=1.1, literal_frac=0.4, literal_base=60_000_000,
                  n_literals=250_000_000)
# Seeds of the benchmark workloads.  Every store is drawn with distinct
# triples, so a query's result is its set of bindings whatever the plan.
SEED_100M = 42            # the 100M store (index build, selects, join set, mix)
SEED_SELECTS = 7          # the 10k range selects on it
SEED_MIX = 3              # the 256-query batched mix on it
SEED_500M = 142           # the 500M served store
SEED_SERVE_MIX = 143      # its 1024-request mix
N_SELECTS = 10_000
N_MIX = 256
N_SERVE = 1024
# 1M-triple store of the CPU-runnable case
STORE_1M = dict(n=1_000_000, n_entities=100_000, subject_zipf=0.0, n_predicates=64,
                pred_zipf=1.1, literal_frac=0.5, literal_base=50_000_000, n_literals=500_000)

# Single-query join set on the 100M store.  Four subject patterns around one
# variable cannot stay below 1e10 rows under Zipf(0.8) subjects for any four
# of the 256 predicates (the top subject alone gives >= ~1e10), so the
# 4-pattern star mixes subject and object positions around ?s.
JOIN_SET = {
    "star4": ([("?s", 30, "?a"), ("?s", 31, "?b"), ("?x", 1, "?s"), ("?y", 2, "?s")],
              ("JOIN", ("JOIN", ("JOIN", ("UPLOAD", 0, "?s"), ("UPLOAD", 1, "?s"), "?s"),
                        ("UPLOAD", 2, "?s"), "?s"), ("UPLOAD", 3, "?s"), "?s")),
    "chain4": ([("?a", 1, "?b"), ("?b", 1, "?c"), ("?c", 1, "?d"), ("?d", 1, "?e")],
               ("JOIN", ("INDEXSWAP", ("JOIN", ("INDEXSWAP", ("JOIN", ("UPLOAD", 0, "?b"),
                                                              ("UPLOAD", 1, "?b"), "?b"), "?c"),
                                                 ("UPLOAD", 2, "?c"), "?c"), "?d"),
                ("UPLOAD", 3, "?d"), "?d")),
    "snowflake6": ([("?x", 50, "?y"), ("?x", 51, "?u"), ("?w", 1, "?x"),
                    ("?y", 30, "?v"), ("?y", 31, "?t"), ("?z", 1, "?y")],
                   ("JOIN",
                    ("INDEXSWAP", ("JOIN", ("JOIN", ("UPLOAD", 0, "?x"), ("UPLOAD", 1, "?x"), "?x"),
                                   ("UPLOAD", 2, "?x"), "?x"), "?y"),
                    ("JOIN", ("JOIN", ("UPLOAD", 3, "?y"), ("UPLOAD", 4, "?y"), "?y"),
                     ("UPLOAD", 5, "?y"), "?y"),
                    "?y")),
}


@functools.lru_cache(maxsize=8)
def _weights(m: int, s: float) -> np.ndarray:
    if s == 0.0:
        return np.full(m, 1.0 / m)
    w = np.arange(1, m + 1, dtype=np.float64) ** -s
    return w / w.sum()


@functools.lru_cache(maxsize=8)
def _power_sums(m: int, s: float) -> tuple:
    w = _weights(m, s)
    return tuple(float((w ** k).sum()) for k in range(0, 7))


class Model:
    """Expected sizes under the generator's independence model."""

    def __init__(self, prm: dict):
        self.N = float(prm["n"])
        self.E = float(prm["n_entities"])
        self.wp = _weights(prm["n_predicates"], prm["pred_zipf"])
        self.S = _power_sums(prm["n_entities"], prm["subject_zipf"])
        self.ent = 1.0 - prm["literal_frac"]

    def pred_rows(self, p: int) -> float:
        return self.N * self.wp[p - 1]

    def star(self, subj_preds, obj_preds) -> float:
        """E[rows] of a star with the centre in subject position for
        subj_preds and in object position for obj_preds."""
        a = len(subj_preds)
        r = self.S[a] * (self.E if a == 0 else 1.0)
        for p in subj_preds:
            r *= self.N * self.wp[p - 1]
        for p in obj_preds:
            r *= self.N * self.wp[p - 1] * self.ent / self.E
        return r

    def chain(self, preds) -> list:
        """E[rows] after each join of ?v0 p1 ?v1 . ?v1 p2 ?v2 ..."""
        out = [self.pred_rows(preds[0])]
        r = self.N * self.N * self.wp[preds[0] - 1] * self.wp[preds[1] - 1] * self.ent / self.E
        out.append(r)
        for p in preds[2:]:
            r *= self.ent * self.N * self.wp[p - 1] / self.E
            out.append(r)
        return out


def _star_query(preds_subj, preds_obj, bound=None):
    pats = [("?s", p, f"?o{i}") for i, p in enumerate(preds_subj)]
    pats += [(f"?i{i}", p, "?s") for i, p in enumerate(preds_obj)]
    if bound is not None:
        p, o = bound
        pats = [("?s", p, o)] + pats
    node = ("UPLOAD", 0, "?s")
    for i in range(1, len(pats)):
        node = ("JOIN", node, ("UPLOAD", i, "?s"), "?s")
    return pats, node


def _chain_query(preds, bound_obj=None):
    n = len(preds)
    names = [f"?v{i}" for i in range(n + 1)]
    pats = [(names[i], p, names[i + 1]) for i, p in enumerate(preds)]
    if bound_obj is not None:
        pats[-1] = (names[n - 1], preds[-1], bound_obj)
    node = ("JOIN", ("UPLOAD", 0, names[1]), ("UPLOAD", 1, names[1]), names[1])
    for i in range(2, n):
        v = names[i]
        node = ("JOIN", ("INDEXSWAP", node, v), ("UPLOAD", i, v), v)
    return pats, node


def query_mix(n_queries: int, seed: int, prm: dict, s=None, p=None, o=None,
              lo: float = 1e2, hi: float = 1e8):
    """Seeded mix: stars of 2-5 patterns, chains of 2-4, about a quarter with
    one bound-object constant taken from an existing triple (so the query is
    not empty).  Predicates are redrawn until every expected intermediate is
    at most `hi` (and, for queries without a constant, at least `lo`).
    Returns [(name, patterns, plan)]."""
    rng = np.random.default_rng(seed)
    m = Model(prm)
    P = prm["n_predicates"]
    ws = _weights(prm["n_entities"], prm["subject_zipf"])
    out = []
    want, tries = None, 0  # (bound, is_star, star size, chain length)

    def accept(item):
        nonlocal want, tries
        out.append(item)
        want, tries = None, 0

    while len(out) < n_queries:
        if want is None or tries > 5000:  # infeasible shape: draw another
            want, tries = (s is not None and rng.random() < 0.25, rng.random() < 0.5,
                           int(rng.integers(2, 6)), int(rng.integers(2, 5))), 0
        tries += 1
        bound, is_star, k, L = want
        if is_star:
            if bound:
                t = int(rng.integers(0, len(s)))
                sv, pb, ob = int(s[t]), int(p[t]), int(o[t])
                subj = rng.random(k - 1) < 0.5
                preds = [int(x) for x in rng.integers(1, P + 1, k - 1)]
                r, sizes = 1.0, []
                for q, sj in zip(preds, subj):
                    r *= (m.N * m.wp[q - 1] * ws[sv - 1]) if sj else (m.N * m.wp[q - 1] * m.ent / m.E)
                    sizes.append(r)
                if max(sizes) > hi:
                    continue
                ps = [q for q, sj in zip(preds, subj) if sj]
                po = [q for q, sj in zip(preds, subj) if not sj]
                pats, node = _star_query(ps, po, (pb, ob))
                accept((f"star{k}b", pats, node))
                continue
            a = int(rng.integers(0, min(k, 2) + 1))
            ps = [int(x) for x in rng.integers(1, P + 1, a)]
            po = [int(x) for x in rng.integers(1, P + 1, k - a)]
            sizes = [m.star(ps[:i], []) if i <= a else m.star(ps, po[:i - a])
                     for i in range(1, k + 1)]
            sizes += [m.pred_rows(q) for q in ps + po]
            if not all(lo <= x <= hi for x in sizes):
                continue
            pats, node = _star_query(ps, po)
            accept((f"star{k}", pats, node))
        else:
            preds = [int(x) for x in rng.integers(1, P + 1, L)]
            sizes = m.chain(preds)
            if bound:
                if max(sizes[:-1]) > hi:
                    continue
                # an existing object of the last predicate (random probes;
                # a rare predicate is redrawn)
                t = rng.integers(0, len(p), 4096)
                hit = t[p[t] == preds[-1]]
                if len(hit) == 0:
                    continue
                oc = int(o[hit[0]])
                pats, node = _chain_query(preds, oc)
                accept((f"chain{L}b", pats, node))
                continue
            if not all(lo <= x <= hi for x in sizes):
                continue
            pats, node = _chain_query(preds)
            accept((f"chain{L}", pats, node))
    return out


# Permutations and constant positions of the range selects: one bound
# position (S -> SPO, P -> PSO, O -> OSP) or two ((S, P) -> SPO,
# (P, O) -> POS, (O, S) -> OSP).
SELECT_PERMS = ("SPO", "PSO", "POS", "OSP")


def select_mix(nq: int, seed: int, s, p, o):
    """Range selects with one or two constants taken from uniformly drawn
    existing triples (50/50, positions uniform), so none is empty.  Returns
    {perm: (k1 array, k2 array)} with 0xFFFFFFFF for "no second constant",
    each permutation's selects in draw order."""
    rng = np.random.default_rng(seed)
    pick = rng.integers(0, len(s), nq)
    two = rng.random(nq) < 0.5
    pos = rng.integers(0, 3, nq)
    U = 0xFFFFFFFF
    groups = {k: ([], []) for k in SELECT_PERMS}
    for i, t in enumerate(pick):
        S, P, O = int(s[t]), int(p[t]), int(o[t])
        if not two[i]:
            perm, k1, k2 = [("SPO", S, U), ("PSO", P, U), ("OSP", O, U)][pos[i]]
        else:
            perm, k1, k2 = [("SPO", S, P), ("POS", P, O), ("OSP", O, S)][pos[i]]
        groups[perm][0].append(k1)
        groups[perm][1].append(k2)
    return {k: (np.array(a, np.uint32), np.array(b, np.uint32)) for k, (a, b) in groups.items()
            if a}


def chunk_by_memory(eng, texts, budget):
    """Consecutive groups of queries whose estimated peak HBM bytes (store
    statistics) sum to at most `budget`."""
    groups, cur, acc = [], [], 0.0
    for t in texts:
        need = eng.estimate(t)[1]
        if cur and acc + need > budget:
            groups.append(cur)
            cur, acc = [], 0.0
        cur.append(t)
        acc += need
    if cur:
        groups.append(cur)
    return groups
