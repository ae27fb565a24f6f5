"""SPARQL over an N-Triples store end to end on the device: the result rows,
turned back into terms through the dictionary, equal a naive evaluation of
the basic graph pattern over the string triples."""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_store(seed=3, n=1500):
    rng = np.random.default_rng(seed)
    trip = set()
    while len(trip) < n:
        s = f"<http://ex.org/e{rng.integers(0, 60)}>"
        p = f"<http://ex.org/p{rng.integers(0, 4)}>"
        o = (f'"lit{rng.integers(0, 20)}"' if rng.random() < 0.3
             else f"<http://ex.org/e{rng.integers(0, 60)}>")
        trip.add((s, p, o))
    return sorted(trip)


def naive(triples, patterns):
    """Bindings of the patterns (terms as strings, variables '?x')."""
    binds = [{}]
    for pat in patterns:
        nxt = []
        for b in binds:
            for t in triples:
                nb, ok = dict(b), True
                for term, val in zip(pat, t):
                    if term.startswith("?"):
                        if nb.setdefault(term, val) != val:
                            ok = False
                            break
                    elif term != val:
                        ok = False
                        break
                if ok:
                    nxt.append(nb)
        binds = nxt
    names = sorted({v for pat in patterns for v in pat if v.startswith("?")})
    return names, sorted(tuple(b[v] for v in names) for b in binds)


def canonical(headers, rows):
    order = sorted(range(len(headers)), key=lambda i: headers[i])
    return [headers[i] for i in order], sorted(tuple(r[i] for i in order) for r in rows)


QUERIES = {
    "star": ("SELECT * WHERE { ?s <http://ex.org/p0> ?a . ?s <http://ex.org/p1> ?b . "
             "?s <http://ex.org/p2> ?c }",
             [("?s", "<http://ex.org/p0>", "?a"), ("?s", "<http://ex.org/p1>", "?b"),
              ("?s", "<http://ex.org/p2>", "?c")]),
    "chain": ("PREFIX ex: <http://ex.org/> SELECT * WHERE { ?a ex:p0 ?b . ?b ex:p1 ?c . ?c ex:p3 ?d }",
              [("?a", "<http://ex.org/p0>", "?b"), ("?b", "<http://ex.org/p1>", "?c"),
               ("?c", "<http://ex.org/p3>", "?d")]),
    "literal": ('SELECT * WHERE { ?s <http://ex.org/p1> "lit3" . ?s <http://ex.org/p0> ?o }',
                [("?s", "<http://ex.org/p1>", '"lit3"'), ("?s", "<http://ex.org/p0>", "?o")]),
    "iri_const": ("PREFIX ex: <http://ex.org/> SELECT * WHERE { ex:e5 ex:p2 ?x . ?x ex:p0 ?y ; ex:p1 ?z }",
                  [("<http://ex.org/e5>", "<http://ex.org/p2>", "?x"),
                   ("?x", "<http://ex.org/p0>", "?y"), ("?x", "<http://ex.org/p1>", "?z")]),
}


@pytest.mark.parametrize("ldict", [False, True])
@pytest.mark.parametrize("name", sorted(QUERIES))
def test_sparql_equals_naive(name, ldict):
    from vedas_amd.engine import Dict, Engine
    triples = make_store()
    d = Dict(ldict)
    d.load_text("".join(f"{s} {p} {o} .\n" for s, p, o in triples))
    e = Engine(0)
    e.load_dict(d, ("PSO", "POS"))
    text, pats = QUERIES[name]
    r = e.sparql(text)
    got = canonical(r.headers(), r.terms(d))
    r.free()
    want = naive(triples, pats)
    assert got == want
    e.close()
    d.close()
