"""GPU tests of the C++ operator API (libvedas.so): whole queries through
QueryExecutor / QueryPlan against the oracle plan and the naive evaluator."""
import numpy as np
import pytest

from oracle import pyoracle as O
from tests import plans

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    from vedas_amd.engine import Engine
    rng = np.random.default_rng(4)
    trip = sorted(set(zip(rng.integers(1, 30, 400).tolist(), rng.integers(1, 5, 400).tolist(),
                          rng.integers(1, 30, 400).tolist())))
    s, p, o = (np.array(c, np.uint32) for c in zip(*trip))
    e = Engine(0)
    e.load(s, p, o, ("PSO", "POS", "SPO", "SOP", "OSP", "OPS"))
    yield s, p, o, e
    e.close()


def run(e, pats, node, select="*"):
    from vedas_amd.engine import query_text
    r = e.query(query_text(pats, node, select))
    try:
        return r.headers(), r.columns(), r.stats()
    finally:
        r.free()


@pytest.mark.parametrize("shape", ["star2", "star3", "chain3", "snow", "single", "bound"])
def test_queries_equal_naive(small, shape):
    s, p, o, e = small
    if shape == "star2":
        pats, node = plans.star_plan([1, 2])
    elif shape == "star3":
        pats, node = plans.star_plan([1, 2, 3])
    elif shape == "chain3":
        pats, node = plans.chain_plan([1, 2, 3])
    elif shape == "snow":
        pats, node = plans.snowflake_plan([1, 2], [3, 4])
    elif shape == "single":
        pats = [("?s", 2, "?o")]
        node = ("UPLOAD", pats[0], "?s")
    else:
        obj = int(o[p == 1][0])
        pats = [("?s", 1, obj), ("?s", 2, "?y"), ("?y", 3, "?z")]
        node = ("JOIN", ("INDEXSWAP", ("JOIN", ("UPLOAD", pats[0], "?s"), ("UPLOAD", pats[1], "?s"), "?s"), "?y"),
                ("UPLOAD", pats[2], "?y"), "?y")
    headers, cols, st = run(e, pats, node)
    got = O.canonical(headers, cols)
    want = O.naive_bgp(s, p, o, pats)
    assert got == want
    assert st["rows"] == len(want[1])


def predicate_variable_queries(s, p, o):
    """Patterns with a predicate variable (routes through SOP, SPO and OSP:
    QueryExecutor.cu:813-860, 871-971), alone and joined."""
    s0, o0 = int(s[5]), int(o[5])
    s1 = int(s[np.flatnonzero(s != s0)[0]])
    o1 = int(o[np.flatnonzero(p == 2)[0]])
    return [
        ([(s0, "?p", o0)], ("UPLOAD", 0, "?p")),                      # s ?p o  -> SOP
        ([(s0, "?p", "?o")], ("UPLOAD", 0, "?p")),                    # s ?p ?o -> SPO
        ([("?s", "?p", o0)], ("UPLOAD", 0, "?s")),                    # ?s ?p o -> OSP
        ([(s0, "?p", "?o"), ("?o", 2, "?z")],                         # join on the object
         ("JOIN", ("UPLOAD", 0, "?o"), ("UPLOAD", 1, "?o"), "?o")),
        ([("?s", "?p", o1), ("?s", 1, "?y")],                         # join on the subject
         ("JOIN", ("UPLOAD", 0, "?s"), ("UPLOAD", 1, "?s"), "?s")),
        ([(s0, "?p", "?o"), (s1, "?p", "?q")],                        # join on ?p
         ("JOIN", ("UPLOAD", 0, "?p"), ("UPLOAD", 1, "?p"), "?p")),
        ([(s0, "?p", o0), (s1, "?p", "?q")],                          # SOP [?p] x SPO [?p, ?q]
         ("JOIN", ("UPLOAD", 0, "?p"), ("UPLOAD", 1, "?p"), "?p")),
        ([("?s", "?p", o1), ("?s", "?r", o0)],                        # OSP x OSP on ?s
         ("JOIN", ("UPLOAD", 0, "?s"), ("UPLOAD", 1, "?s"), "?s")),
    ]


def test_predicate_variable_patterns(small):
    """Every predicate-variable route, plain and batched, against the naive
    evaluator over the triples (all six permutations resident)."""
    from vedas_amd.engine import query_text
    s, p, o, e = small
    qs = predicate_variable_queries(s, p, o)
    texts = [query_text(pats, node) for pats, node in qs]
    nonempty = 0
    for (pats, node), text in zip(qs, texts):
        headers, cols, st = run(e, pats, node)
        want = O.naive_bgp(s, p, o, pats)
        assert O.canonical(headers, cols) == want, pats
        nonempty += len(want[1]) > 0
    assert nonempty >= 6
    res, _ = e.batch(texts)
    for (pats, node), r in zip(qs, res):
        assert O.canonical(r.headers(), r.columns()) == O.naive_bgp(s, p, o, pats), pats
        r.free()


def test_projection_and_finalise(small):
    s, p, o, e = small
    pats, node = plans.star_plan([1, 2, 3])
    headers, cols, _ = run(e, pats, node, select=["?s", "?o1"])
    # column 0 is never dropped; unused columns are; adjacent duplicates go
    assert headers[0] == "?s" and set(headers) == {"?s", "?o1"}
    rows = list(zip(*[c.tolist() for c in cols]))
    assert all(rows[i] != rows[i - 1] for i in range(1, len(rows)))
    names, full = O.naive_bgp(s, p, o, pats)
    keep = [names.index(h) for h in headers]
    assert set(rows) == {tuple(r[k] for k in keep) for r in full}


def test_errors_are_reported(small):
    from vedas_amd._native import VedasError
    s, p, o, e = small
    with pytest.raises(VedasError):  # two patterns over the same two terms: needs a PLAN
        e.query("SELECT * WHERE { ?s 1 ?x . ?s 2 ?x }")
    with pytest.raises(VedasError):  # disconnected patterns: no plan joins them
        e.query("SELECT * WHERE { ?s 1 ?x . ?a 2 ?b }")
    with pytest.raises(VedasError):
        e.query("SELECT * WHERE { ?s 1 ?x } PLAN (U 3 ?s)")
    with pytest.raises(VedasError):
        e.query("SELECT * WHERE { ?s ?p ?x } PLAN (U 0 ?s)")


def test_cfg0_star_query():
    """The cfg0 query (1M triples, 2-pattern subject star on the two most
    frequent predicates) through the operator API vs the oracle plan."""
    from vedas_amd._native import gen_triples
    from vedas_amd.engine import Engine, query_text
    s, p, o = gen_triples(1_000_000, seed=42)
    pats, node = plans.star_plan([1, 2], "?s", ["?x", "?y"])
    e = Engine(0)
    try:
        e.load(s, p, o, ("PSO", "POS"))
        r = e.query(query_text(pats, node))
        got = O.canonical(r.headers(), r.columns())
        st = r.stats()
        r.free()
    finally:
        e.close()
    ex = plans.OracleExecutor(s, p, o)
    rel = ex.run(node)
    want = O.canonical(rel.headers, ex.columns(rel))
    assert got == want
    assert st["join_rows"] == len(want[1]) and 1e5 < st["join_rows"] < 2e6


def test_bound_prescan_on_cfg0_store():
    """The join-variable bound pre-scan (createVariableBound) on the 1M
    store: selects shrink (objects of predicate 1 include literals, which no
    subject of predicate 2 has), results and every intermediate size equal
    the oracle's plan walk with the reference's bounds, and results equal
    the walk without bounds."""
    from vedas_amd._native import gen_triples
    from vedas_amd.engine import Engine, get_option, query_text, set_option
    s, p, o = gen_triples(1_000_000, seed=42)
    lit = int(o[(p == 1) & (o >= 50_000_000)][0])
    queries = [plans.chain_plan([1, 2, 3]), plans.chain_plan([2, 1]),
               plans.snowflake_plan([1, 2], [3, 4]),
               ([("?s", 1, lit), ("?s", 2, "?y"), ("?y", 3, "?z")],
                ("JOIN", ("INDEXSWAP", ("JOIN", ("UPLOAD", 0, "?s"), ("UPLOAD", 1, "?s"), "?s"),
                          "?y"), ("UPLOAD", 2, "?y"), "?y")),
               ([("?a", 1, "?b"), ("?b", 2, 4_000_000_000)],
                ("JOIN", ("UPLOAD", 0, "?b"), ("UPLOAD", 1, "?b"), "?b"))]
    ix = {perm: O.build_perm(s, p, o, perm) for perm in ("PSO", "POS")}
    e = Engine(0)
    assert get_option("bound_prescan") == 1
    shrunk = 0
    try:
        e.load(s, p, o, ("PSO", "POS"))
        for pats, node in queries:
            got = {}
            for on in (0, 1):
                set_option("bound_prescan", on)
                r = e.query(query_text(pats, node))
                got[on] = (O.bag_digest(r.headers(), r.columns()), r.stats())
                r.free()
            st0, st1 = {}, {}
            h0, c0 = O.run_plan(ix, pats, node, st0)
            h1, c1 = O.run_plan(ix, pats, node, st1, bounds=True)
            want = O.bag_digest(h0, c0)
            assert O.bag_digest(h1, c1) == want
            assert got[0][0] == want and got[1][0] == want, pats
            for k in ("select_rows", "join_rows"):
                assert got[0][1][k] == st0[k], (pats, k)
                assert got[1][1][k] == st1[k], (pats, k)
            assert got[1][1]["select_rows"] <= got[0][1]["select_rows"]
            shrunk += got[1][1]["select_rows"] < got[0][1]["select_rows"]
    finally:
        set_option("bound_prescan", 1)
        e.close()
    assert shrunk >= 3


def test_autotune_keeps_results(small):
    """The autotuner only picks kernel shapes: chosen values are valid
    candidates and results stay the same after tuning (and in a serving
    slot created afterwards)."""
    from vedas_amd.engine import query_text
    s, p, o, e = small
    shapes = [plans.star_plan([1, 2, 3]), plans.chain_plan([1, 2, 3]),
              plans.snowflake_plan([1, 2], [3, 4])]
    texts = [query_text(pats, node) for pats, node in shapes]
    before = []
    for t in texts:
        r = e.query(t)
        before.append(O.canonical(r.headers(), r.columns()))
        r.free()
    tr = e.autotune(texts, reps=1)
    assert tr["radix_items"] in (8, 16, 24)
    assert tr["join_nv"] in (1024, 2048, 4096)
    assert tr["join_rcap"] in (2048, 4096, 8192)
    assert tr["radix_bits"] in (0, 8, 9)
    assert tr["radix_group"] in (0, 8, 16)
    assert tr["evaluations"] >= 9 and tr["tuned_ms"] <= tr["base_ms"] + 1e-9
    for t, want in zip(texts, before):
        r = e.query(t)
        assert O.canonical(r.headers(), r.columns()) == want
        r.free()
    recs, outs = e.serve(texts, [0.0] * len(texts), slots=2, max_batch=2, keep_results=True)
    for res, want in zip(outs, before):
        assert O.canonical(res.headers(), res.columns()) == want
        res.free()
    for (pats, _), want in zip(shapes, before):
        assert want == O.naive_bgp(s, p, o, pats)


@pytest.mark.parametrize("shape", ["star3", "chain3", "snow", "bound"])
def test_unplanned_queries_use_the_planner(small, shape):
    """Queries without PLAN are planned by the query graph planner from the
    device lookups' range sizes; single, batched and served runs all equal
    the naive evaluator."""
    from vedas_amd.engine import query_text
    s, p, o, e = small
    if shape == "star3":
        pats = plans.star_plan([1, 2, 3])[0]
    elif shape == "chain3":
        pats = plans.chain_plan([1, 2, 3])[0]
    elif shape == "snow":
        pats = plans.snowflake_plan([1, 2], [3, 4])[0]
    else:
        obj = int(o[p == 1][0])
        pats = [("?s", 1, obj), ("?s", 2, "?y"), ("?y", 3, "?z")]
    text = query_text(pats, None)
    want = O.naive_bgp(s, p, o, pats)
    r = e.query(text)
    assert O.canonical(r.headers(), r.columns()) == want
    r.free()
    res, _ = e.batch([text, text])
    for r in res:
        assert O.canonical(r.headers(), r.columns()) == want
        r.free()
    _, outs = e.serve([text], [0.0], slots=1, max_batch=1, keep_results=True)
    assert O.canonical(outs[0].headers(), outs[0].columns()) == want
    outs[0].free()
