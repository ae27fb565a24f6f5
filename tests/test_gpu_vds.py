"""The .vds store file (the reference's default index, VedasStorage::write /
open, VedasStorage.cu:452-515): ten arrays, each a size_t count followed by
uint32 values, in the order p_idx_values, p_idx_offsets, ps_s_idx_values,
ps_idx_offsets, po_o_idx_values, po_idx_offsets, ps_data, po_data,
pso_data, pos_data.  The file written by the engine is parsed here by that
layout and compared with the oracle's arrays; a file written here from the
oracle's arrays is opened by the engine and queried."""
import numpy as np
import pytest

from oracle import pyoracle as O
from tests import plans

pytestmark = pytest.mark.gpu

ORDER = [("PSO", "l1v"), ("PSO", "l1o"), ("PSO", "l2v"), ("PSO", "l2o"), ("POS", "l2v"),
         ("POS", "l2o"), ("PSO", "y"), ("POS", "y"), ("PSO", "z"), ("POS", "z")]


def store(seed=12):
    rng = np.random.default_rng(seed)
    trip = sorted(set(zip(rng.integers(1, 60, 3000).tolist(), rng.integers(1, 7, 3000).tolist(),
                          rng.integers(1, 60, 3000).tolist())))
    return (np.array(c, np.uint32) for c in zip(*trip))


def parse_vds(path):
    raw = open(path, "rb").read()
    pos, out = 0, []
    for _ in range(10):
        n = int(np.frombuffer(raw, np.uint64, 1, pos)[0])
        pos += 8
        out.append(np.frombuffer(raw, np.uint32, n, pos).copy())
        pos += 4 * n
    assert pos == len(raw)
    return out


def write_vds(path, ix):
    with open(path, "wb") as f:
        for perm, part in ORDER:
            a = np.ascontiguousarray(ix[perm][part], np.uint32)
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
