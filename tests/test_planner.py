"""The query graph planner (libvedas.so, host only): plans for queries
without PLAN, executed by the oracle plan walk, must give the naive
evaluator's bindings; every pattern is a leaf exactly once."""
import numpy as np
import pytest

from oracle import pyoracle as O
from tests import plans


def small_store(seed=9):
    rng = np.random.default_rng(seed)
    trip = sorted(set(zip(rng.integers(1, 40, 900).tolist(), rng.integers(1, 6, 900).tolist(),
                          rng.integers(1, 40, 900).tolist())))
    return (np.array(c, np.uint32) for c in zip(*trip))


def leaves(node, out):
    if node[0] == "UPLOAD":
        out.append(node[1])
    else:
        leaves(node[1], out)
        if node[0] == "JOIN":
            leaves(node[2], out)
    return out


def est_rows(ix, pats):
    rows = []
    for s, p, o in pats:
        if isinstance(s, str) and isinstance(o, str):
            b, e = O.lookup(ix["PSO"], p)
        elif isinstance(s, str):
            b, e = O.lookup(ix["POS"], p, o)
        else:
            b, e = O.lookup(ix["PSO"], p, s)
        rows.append(e - b)
    return rows


SHAPES = {
    "star3": lambda s, p, o: plans.star_plan([1, 2, 3])[0],
    "star5": lambda s, p, o: plans.star_plan([1, 2, 3, 4, 5])[0],
    "chain3": lambda s, p, o: plans.chain_plan([1, 2, 3])[0],
    "chain4": lambda s, p, o: plans.chain_plan([2, 1, 3, 4])[0],
    "snowflake": lambda s, p, o: plans.snowflake_plan([1, 2], [3, 4])[0],
    "star_in_out": lambda s, p, o: [("?s", 1, "?a"), ("?x", 2, "?s"), ("?s", 3, "?b"), ("?y", 4, "?s")],
    "bound": lambda s, p, o: [("?s", 1, int(o[p == 1][0])), ("?s", 2, "?y"), ("?y", 3, "?z")],
    "bound_subject": lambda s, p, o: [(int(s[p == 2][0]), 2, "?y"), ("?y", 3, "?z"), ("?z", 4, "?w")],
}


@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_planned_queries_equal_naive(shape):
    from vedas_amd.engine import parse_plan, plan, query_text
    s, p, o = small_store()
    ix = {perm: O.build_perm(s, p, o, perm) for perm in ("PSO", "POS")}
    pats = SHAPES[shape](s, p, o)
    clause = plan(query_text(pats, None), est_rows(ix, pats))
    node = parse_plan(clause)
    assert sorted(leaves(node, [])) == list(range(len(pats)))
    hdr, cols = O.run_plan(ix, pats, node)
    assert O.canonical(hdr, cols) == O.naive_bgp(s, p, o, pats)


def test_planner_orders_by_estimates():
    """The smallest pattern of a star is its first join input (the
    reference sorts a vertex's edges by estimate)."""
    from vedas_amd.engine import parse_plan, plan, query_text
    pats = [("?s", 1, "?a"), ("?s", 2, "?b"), ("?s", 3, "?c")]
    node = parse_plan(plan(query_text(pats, None), [500, 20, 900]))
    while node[0] == "JOIN":
        first = node[1]
        node = first
    assert node == ("UPLOAD", 1, "?s")


def test_plan_text_round_trip_and_errors():
    from vedas_amd import _native
    from vedas_amd.engine import plan, query_text
    pats, node = plans.chain_plan([1, 2, 3])
    text = query_text(pats, node)
    assert plan(text) == text.split("PLAN ", 1)[1]
    with pytest.raises(_native.VedasError):
        plan(query_text(pats, None))  # no estimates
    with pytest.raises(_native.VedasError):  # two patterns over the same two terms
        plan(query_text([("?s", 1, "?o"), ("?s", 2, "?o")], None), [5, 5])
