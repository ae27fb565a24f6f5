"""The .vds store file (VedasStorage::write / open, VedasStorage.cu:
452-516): arrays, each a size_t count followed by uint32 values.  The
default index is ten arrays: p_idx_values, p_idx_offsets, ps_s_idx_values,
ps_idx_offsets, po_o_idx_values, po_idx_offsets, ps_data, po_data,
pso_data, pos_data; a full index appends fifteen: s_idx_values,
s_idx_offsets, sp_p_idx_values, sp_idx_offsets, so_o_idx_values,
so_idx_offsets, o_idx_values, o_idx_offsets, os_data, op_p_idx_values,
op_idx_offsets, spo_data, sop_data, ops_data, osp_data.  Files written by
the engine are parsed here by that layout and compared with the oracle's
arrays; files written here from the oracle's arrays (including the
reference's own full-index form, whose subject L1 is empty) are opened by
the engine, checked permutation by permutation and queried."""
import numpy as np
import pytest

from oracle import pyoracle as O
from tests import plans

pytestmark = pytest.mark.gpu

ORDER = [("PSO", "l1v"), ("PSO", "l1o"), ("PSO", "l2v"), ("PSO", "l2o"), ("POS", "l2v"),
         ("POS", "l2o"), ("PSO", "y"), ("POS", "y"), ("PSO", "z"), ("POS", "z")]
FULL_ORDER = ORDER + [("SPO", "l1v"), ("SPO", "l1o"), ("SPO", "l2v"), ("SPO", "l2o"),
                      ("SOP", "l2v"), ("SOP", "l2o"), ("OSP", "l1v"), ("OSP", "l1o"),
                      ("OSP", "y"), ("OPS", "l2v"), ("OPS", "l2o"), ("SPO", "z"), ("SOP", "z"),
                      ("OPS", "z"), ("OSP", "z")]
ALL = ("SPO", "SOP", "PSO", "POS", "OSP", "OPS")


def store(seed=12):
    rng = np.random.default_rng(seed)
    trip = sorted(set(zip(rng.integers(1, 60, 3000).tolist(), rng.integers(1, 7, 3000).tolist(),
                          rng.integers(1, 60, 3000).tolist())))
    return (np.array(c, np.uint32) for c in zip(*trip))


def parse_vds(path, narrays=10):
    raw = open(path, "rb").read()
    pos, out = 0, []
    for _ in range(narrays):
        n = int(np.frombuffer(raw, np.uint64, 1, pos)[0])
        pos += 8
        out.append(np.frombuffer(raw, np.uint32, n, pos).copy())
        pos += 4 * n
    assert pos == len(raw)
    return out


def write_vds(path, ix, order=ORDER, empty=()):
    """The arrays of `order` from the oracle's permutations; those listed in
    `empty` written with no elements (as the reference writes its subject
    L1)."""
    with open(path, "wb") as f:
        for perm, part in order:
            a = np.ascontiguousarray(ix[perm][part], np.uint32)
            if (perm, part) in empty:
                a = a[:0]
            f.write(np.uint64(len(a)).tobytes())
            f.write(a.tobytes())


def test_saved_file_matches_oracle_arrays(tmp_path):
    from vedas_amd.engine import Engine
    s, p, o = store()
    e = Engine(0)
    e.load(s, p, o, ("PSO", "POS"))
    path = str(tmp_path / "store.vds")
    e.save(path)
    e.close()
    arrays = parse_vds(path)
    ix = {perm: O.build_perm(s, p, o, perm) for perm in ("PSO", "POS")}
    for (perm, part), got in zip(ORDER, arrays):
        assert np.array_equal(got, ix[perm][part]), (perm, part)
    # the predicate L1 is shared: POS's equals PSO's
    assert np.array_equal(ix["POS"]["l1v"], ix["PSO"]["l1v"])
    assert np.array_equal(ix["POS"]["l1o"], ix["PSO"]["l1o"])


@pytest.mark.parametrize("shape", ["star3", "chain3", "snow"])
def test_opened_file_answers_queries(tmp_path, shape):
    from vedas_amd.engine import Engine, query_text
    s, p, o = store()
    ix = {perm: O.build_perm(s, p, o, perm) for perm in ("PSO", "POS")}
    path = str(tmp_path / "oracle.vds")
    write_vds(path, ix)
    pats, node = {"star3": plans.star_plan([1, 2, 3]), "chain3": plans.chain_plan([1, 2, 3]),
                  "snow": plans.snowflake_plan([1, 2], [3, 4])}[shape]
    e = Engine(0)
    e.open(path)
    for text in (query_text(pats, node), query_text(pats, None)):
        r = e.query(text)
        assert O.canonical(r.headers(), r.columns()) == O.naive_bgp(s, p, o, pats)
        r.free()
    # round trip through the engine's own writer
    path2 = str(tmp_path / "again.vds")
    e.save(path2)
    e.close()
    assert open(path, "rb").read() == open(path2, "rb").read()


def test_bad_files_are_rejected(tmp_path):
    from vedas_amd import _native
    from vedas_amd.engine import Engine
    e = Engine(0)
    bad = tmp_path / "short.vds"
    bad.write_bytes(np.uint64(5).tobytes() + np.arange(3, dtype=np.uint32).tobytes())
    with pytest.raises(_native.VedasError):
        e.open(str(bad))
    with pytest.raises(_native.VedasError):
        e.open(str(tmp_path / "missing.vds"))
    e.close()


def test_full_index_file_written_and_read(tmp_path):
    """A store of all six permutations saves the full index; the file's 25
    arrays equal the oracle's (the subject L1 filled), and the file opens
    into six permutations equal to the oracle's (Y columns of SPO, SOP, OPS
    and the OSP L2 rebuilt on the device); saving again gives the same
    bytes."""
    from vedas_amd._native import Ctx, Index
    from vedas_amd.engine import Engine
    s, p, o = store(5)
    ix = {perm: O.build_perm(s, p, o, perm) for perm in ALL}
    e = Engine(0)
    e.load(s, p, o, ALL)
    path = str(tmp_path / "full.vds")
    e.save(path)
    e.close()
    arrays = parse_vds(path, 25)
    for (perm, part), got in zip(FULL_ORDER, arrays):
        assert np.array_equal(got, ix[perm][part]), (perm, part)
    e = Engine(0)
    e.open(path)
    view = Index.wrap(Ctx.wrap(e.ctx_handle()), e.index_handle(), len(s))
    for perm in ALL:
        got = view.arrays(perm)
        for k in ("l1v", "l1o", "l2v", "l2o", "y", "z"):
            assert np.array_equal(got[k], ix[perm][k]), (perm, k)
    path2 = str(tmp_path / "again.vds")
    e.save(path2)
    e.close()
    assert open(path, "rb").read() == open(path2, "rb").read()


def test_reference_full_index_file_opens(tmp_path):
    """The reference's full-index file leaves the subject L1 empty
    (VedasStorage.cu:319-339); opening it builds SPO and SOP from the
    triples of PSO.  Every permutation equals the oracle's, and queries over
    every permutation (predicate variables included) equal the naive
    evaluator."""
    from vedas_amd._native import Ctx, Index
    from vedas_amd.engine import Engine, query_text
    s, p, o = store(7)
    ix = {perm: O.build_perm(s, p, o, perm) for perm in ALL}
    path = str(tmp_path / "reference.vds")
    write_vds(path, ix, FULL_ORDER, empty={("SPO", "l1v"), ("SPO", "l1o")})
    e = Engine(0)
    e.open(path)
    view = Index.wrap(Ctx.wrap(e.ctx_handle()), e.index_handle(), len(s))
    for perm in ALL:
        got = view.arrays(perm)
        for k in ("l1v", "l1o", "l2v", "l2o", "y", "z"):
            assert np.array_equal(got[k], ix[perm][k]), (perm, k)
    s0, o0 = int(s[3]), int(o[3])
    queries = [plans.star_plan([1, 2, 3]), plans.chain_plan([1, 2]),
               ([(s0, "?p", "?o"), ("?o", 2, "?z")],
                ("JOIN", ("UPLOAD", 0, "?o"), ("UPLOAD", 1, "?o"), "?o")),
               ([("?s", "?p", o0)], ("UPLOAD", 0, "?s")),
               ([(s0, "?p", o0)], ("UPLOAD", 0, "?p"))]
    for pats, node in queries:
        r = e.query(query_text(pats, node))
        assert O.canonical(r.headers(), r.columns()) == O.naive_bgp(s, p, o, pats), pats
        r.free()
    # written back, the subject L1 is filled in
    path2 = str(tmp_path / "filled.vds")
    e.save(path2)
    e.close()
    path3 = str(tmp_path / "oracle_full.vds")
    write_vds(path3, ix, FULL_ORDER)
    assert open(path2, "rb").read() == open(path3, "rb").read()


def test_bad_run_arrays_are_rejected(tmp_path):
    """Offsets out of order, not from row 0 or past the rows are refused on
    the host before anything reaches the device."""
    from vedas_amd import _native
    from vedas_amd.engine import Engine
    s, p, o = store(9)
    ix = {perm: O.build_perm(s, p, o, perm) for perm in ALL}
    e = Engine(0)
    for perm, part, fix in [("PSO", "l1o", lambda a: a[::-1].copy()),
                            ("PSO", "l2o", lambda a: a + 1),
                            ("OSP", "l1o", lambda a: np.append(a[:-1], 10 ** 6).astype(np.uint32)),
                            ("OPS", "l2o", lambda a: a[::-1].copy())]:
        bad = {k: dict(v) for k, v in ix.items()}
        bad[perm][part] = fix(ix[perm][part])
        path = str(tmp_path / "bad.vds")
        write_vds(path, bad, FULL_ORDER)
        with pytest.raises(_native.VedasError):
            e.open(path)
    e.close()


def test_store_and_dictionary_files_answer_sparql(tmp_path):
    """N-Triples -> store + dictionary -> .vds + .vdd -> a new engine and
    dictionary opened from the files answers SPARQL with the same terms."""
    import os
    from vedas_amd.engine import Dict, Engine
    w3c = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                       "w3c_ntriples_test.nt")
    d = Dict(True)
    d.load(w3c)
    e = Engine(0)
    e.load_dict(d, ALL)
    q = ("SELECT * WHERE { ?s <http://example.org/property> ?o . "
         "?s ?p <http://example.org/resource2> }")
    r = e.sparql(q)
    want = sorted(r.terms(d))
    r.free()
    assert want
    e.save(str(tmp_path / "w3c.vds"))
    d.save(str(tmp_path / "w3c.vdd"))
    e.close()
    d2 = Dict(False)
    d2.open(str(tmp_path / "w3c.vdd"))
    e2 = Engine(0)
    e2.open(str(tmp_path / "w3c.vds"))
    e2.set_dict(d2)
    r = e2.sparql(q)
    assert sorted(r.terms(d2)) == want
    r.free()
    e2.close()
