"""Full-size parity of the served form on the 500M-triple store (50M
entities with Zipf(0.8) subjects, 1024 predicates, 66-bit packed keys,
distinct triples): the two-phase index build of SPO / PSO / POS / OSP
checked at full size and against the oracle's own build, then the
1024-request open-loop mix served with continuous batching under a memory
budget: every result equals the oracle's.  Needs an MI355X."""
import numpy as np
import pytest

from oracle import pyoracle as O
from tests import scale
from vedas_amd import workload as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def served():
    from vedas_amd.engine import query_text
    st = scale.build_store(W.STORE_500M, W.SEED_500M)
    gold = scale.golden("store_500M")
    arrays = scale.checked_arrays(st, gold, keep=("PSO", "POS"))
    mix = W.query_mix(W.N_SERVE, W.SEED_SERVE_MIX, W.STORE_500M, st["s"], st["p"], st["o"])
    st["texts"] = [query_text(pats, node) for _, pats, node in mix]
    st["mix"] = mix
    for k in ("s", "p", "o"):
        del st[k]
    yield st, gold, arrays
    st["eng"].close()


@pytest.mark.timeout(900)
def test_cfg4_index_build_500m(served):
    st, gold, arrays = served
    assert arrays["_same_store"], "the generator did not reproduce the oracle's store"
    assert st["idx"].n == W.STORE_500M["n"]


@pytest.mark.timeout(900)
def test_cfg4_served_mix_500m(served):
    """All 1024 requests served at 50k q/s offered on four slots with a
    16 GiB intermediate-memory budget: each result equals the oracle's
    digest of that query (and a live plan walk on the checked arrays for a
    sample), and a sample run alone and plain gives the same."""
    st, gold, arrays = served
    eng, texts = st["eng"], st["texts"]
    rng = np.random.default_rng(7)
    arr = list(np.cumsum(rng.exponential(1000.0 / 50_000.0, len(texts))))
    recs, res = eng.serve(texts, arr, slots=4, max_batch=128, mem_budget=16 << 30,
                          keep_results=True, min_fill=48, max_wait_ms=1.0)
    got = []
    for r in res:
        got.append(O.bag_digest(r.headers(), r.columns()))
        r.free()
    assert all(rc["done_ms"] >= rc["start_ms"] for rc in recs)
    ix = {k: arrays[k] for k in ("PSO", "POS")}
    for i in range(0, len(texts), 64):
        _, pats, node = st["mix"][i]
        h, cols = O.run_plan(ix, pats, node, bounds=True)
        assert O.bag_digest(h, cols) == got[i], i
        r = eng.query(texts[i])
        assert O.bag_digest(r.headers(), r.columns()) == got[i], i
        r.free()
    if arrays["_same_store"]:
        g = gold["serve_mix"]
        assert [x["text"] for x in g] == texts
        bad = [i for i, (a, b) in enumerate(zip(got, g)) if a != scale.digest_of(b)]
        assert not bad, bad[:10]
    scale.note(f"served {len(texts)} requests, {sum(d[1] for d in got)} result rows, "
               f"{len({rc['batch'] for rc in recs})} batches")
