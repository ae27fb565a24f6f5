"""N-Triples loader and SPARQL front-end (host only): the dictionary numbers
terms in first-appearance order as the reference's loader does
(loader.cu:18-53), restated here independently with a regular expression;
SPARQL basic graph patterns translate to the engine's query text."""
import os
import re

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
W3C = os.path.join(HERE, "golden", "w3c_ntriples_test.nt")  # the reference's data/test.nt

TERM = r'(<[^>]*>|_:[A-Za-z0-9_\-.]*[A-Za-z0-9_\-]|"(?:[^"\\]|\\.)*"(?:@[A-Za-z0-9\-]+|\^\^<[^>]*>)?)'
LINE = re.compile(r"^\s*" + TERM + r"\s*" + TERM + r"\s*" + TERM + r"\s*\.\s*(#.*)?$")


def restate(text, ldict):
    so, p, lit, triples = {}, {}, {}, []
    for line in text.splitlines():
        if not line.strip() or line.strip().startswith("#"):
            continue
        m = LINE.match(line)
        assert m, line
        s, pr, o = m.group(1), m.group(2), m.group(3)
        si = so.setdefault(s, len(so) + 1)
        if o.startswith('"') and ldict:
            oi = lit.setdefault(o, 50_000_000 + len(lit))
        else:
            oi = so.setdefault(o, len(so) + 1)
        pi = p.setdefault(pr, len(p) + 1)
        triples.append((si, pi, oi))
    return so, p, lit, triples


@pytest.mark.parametrize("ldict", [False, True])
def test_w3c_file_numbering(ldict):
    from vedas_amd.engine import Dict
    text = open(W3C).read()
    so, p, lit, want = restate(text, ldict)
    d = Dict(ldict)
    n = d.load(W3C)
    s_, p_, o_ = d.triples()
    assert n == len(want) == 33  # statements in the file
    assert list(zip(s_.tolist(), p_.tolist(), o_.tolist())) == want
    for t, i in so.items():
        assert d.term(i) == t
    for t, i in p.items():
        assert d.term(i, True) == t
    for t, i in lit.items():
        assert d.term(i) == t
    assert d.translate('SELECT ?s WHERE { ?s <http://example.org/property> "simple literal" }') == \
        f'SELECT ?s WHERE {{ ?s 1 {lit.get(chr(34) + "simple literal" + chr(34), so.get(chr(34) + "simple literal" + chr(34)))} }}'
    d.close()


def test_sparql_forms():
    from vedas_amd.engine import Dict
    d = Dict(False)
    d.load_text('<http://x/a> <http://x/p> <http://x/b> .\n'
                '<http://x/b> <http://www.w3.org/1999/02/22-rdf-syntax-ns#type> <http://x/C> .\n'
                '<http://x/b> <http://x/q> "v"@en .\n')
    # ids: a=1 b=2 p=1; C=3 type=2; "v"@en=4 q=3
    q = ("PREFIX x: <http://x/>\nSELECT ?s ?o WHERE { ?s x:p ?o . ?o a x:C ; x:q \"v\"@en , ?z }")
    assert d.translate(q) == "SELECT ?s ?o WHERE { ?s 1 ?o . ?o 2 3 . ?o 3 4 . ?o 3 ?z }"
    assert d.translate("select * { ?a <http://x/p> ?b . }") == "SELECT * WHERE { ?a 1 ?b }"
    assert d.translate("SELECT ?a WHERE { ?a <http://x/unknown> <http://x/none> }") == \
        "SELECT ?a WHERE { ?a 0 0 }"
    from vedas_amd import _native
    for bad in ["SELECT ?a WHERE { ?a y:p ?b }", "SELECT WHERE { ?a <p> ?b }",
                "SELECT ?a WHERE { ?a <p> ?b } LIMIT 3", "SELECT ?a WHERE { }"]:
        with pytest.raises(_native.VedasError):
            d.translate(bad)
    with pytest.raises(_native.VedasError):
        d.load_text('"lit" <http://x/p> <http://x/b> .\n')  # literal subject
    with pytest.raises(_native.VedasError):
        d.load_text('<http://x/a> <http://x/p> <http://x/b>\n')  # no final dot
    d.close()
