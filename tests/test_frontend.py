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


def parse_vdd(path):
    """The .vdd layout read independently: per section a count line, then
    'id term' lines (the term is everything after the first space)."""
    lines = open(path).read().split("\n")
    pos, out = 0, []
    while pos < len(lines) and lines[pos] != "":
        n = int(lines[pos])
        sec = {}
        for ln in lines[pos + 1:pos + 1 + n]:
            i, term = ln.split(" ", 1)
            sec[term] = int(i)
        out.append(sec)
        pos += 1 + n
    return out


@pytest.mark.parametrize("ldict", [False, True])
def test_vdd_round_trip(tmp_path, ldict):
    """save writes the reference's .vdd layout with the loader's numbering;
    open reads it back: every term has the same ID and SPARQL translates the
    same way."""
    from vedas_amd.engine import Dict
    text = open(W3C).read()
    so, p, lit, _ = restate(text, ldict)
    d = Dict(ldict)
    d.load(W3C)
    path = str(tmp_path / "w3c.vdd")
    d.save(path)
    secs = parse_vdd(path)
    assert secs[0] == so and secs[1] == p and secs[2] == lit
    d2 = Dict(False)
    d2.open(path)
    for term, i in so.items():
        assert d2.term(i) == term
    for term, i in p.items():
        assert d2.term(i, True) == term
    for term, i in lit.items():
        assert d2.term(i) == term
    q = ("SELECT * WHERE { ?s <http://example.org/property> ?o . "
         "?o <http://example.org/property> ?z }")
    assert d2.translate(q) == d.translate(q)


def test_vdd_reference_style_file(tmp_path):
    """A file as the reference's vdBuild writes it after an ID reassignment:
    IDs in any order and not contiguous, no literal section."""
    from vedas_amd.engine import Dict
    path = tmp_path / "ref.vdd"
    path.write_text("3\n7 <http://x/b>\n2 <http://x/a>\n5 _:n1\n"
                    "2\n2 <http://x/q>\n1 <http://x/p>\n")
    d = Dict(False)
    d.open(str(path))
    assert d.term(2) == "<http://x/a>" and d.term(7) == "<http://x/b>" and d.term(5) == "_:n1"
    assert d.term(1, True) == "<http://x/p>"
    with pytest.raises(Exception):
        d.term(3)  # a gap in the IDs
    assert "7" in d.translate("SELECT * WHERE { <http://x/b> <http://x/q> ?o }")
    for bad in ("x\n", "2\n1 <a>\n", "1\nfoo <a>\n0\n"):
        path.write_text(bad)
        with pytest.raises(Exception):
            d.open(str(path))
