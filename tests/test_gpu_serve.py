"""GPU tests of the serving loop: every request's result is the same served
in the mix (continuous batches on several streams) as run alone."""
import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def served():
    from vedas_amd._native import gen_triples
    from vedas_amd.engine import Engine, query_text
    from vedas_amd.workload import query_mix
    prm = dict(n=1_000_000, n_entities=100_000, subject_zipf=0.8, n_predicates=64,
               pred_zipf=1.1, literal_frac=0.4, literal_base=50_000_000, n_literals=500_000)
    s, p, o = gen_triples(prm["n"], seed=21, **{k: v for k, v in prm.items() if k != "n"},
                          distinct=True)
    mix = query_mix(40, 9, prm, s, p, o, lo=10, hi=2e5)
    texts = [query_text(pats, node) for _, pats, node in mix]
    e = Engine(0)
    e.load(s, p, o, ("SPO", "PSO", "POS", "OSP"))
    alone = []
    for t in texts:
        r = e.query(t)
        alone.append(O.canonical(r.headers(), r.columns()))
        r.free()
    yield e, texts, alone
    e.close()


def arrivals(n, qps, seed):
    rng = np.random.default_rng(seed)
    return list(np.cumsum(rng.exponential(1000.0 / qps, n)))


@pytest.mark.parametrize("slots,max_batch,budget,fill", [(3, 8, 64 << 30, 0), (2, 32, 64 << 30, 0),
                                                         (4, 4, 1 << 20, 0), (4, 16, 64 << 30, 12)])
def test_served_results_equal_alone(served, slots, max_batch, budget, fill):
    """fill > 0: the fill policy holds small batches back (up to 1 ms)
    while two batches run; every request still completes, with the same
    result."""
    e, texts, alone = served
    recs, res = e.serve(texts, arrivals(len(texts), 20000.0 if fill else 2000.0, slots),
                        slots=slots, max_batch=max_batch, mem_budget=budget, keep_results=True,
                        min_fill=fill, max_wait_ms=1.0 if fill else 0.0)
    got = [O.canonical(r.headers(), r.columns()) for r in res]
    for r in res:
        r.free()
    for i, (a, b) in enumerate(zip(alone, got)):
        assert a == b, i
    for rc in recs:
        assert rc["done_ms"] >= rc["start_ms"] >= 0
        assert rc["start_ms"] + 1e-3 >= rc["arrival_ms"] - 1.0  # started after arriving
        assert rc["slot"] >= 0 and rc["batch"] >= 0
    assert len({rc["batch"] for rc in recs}) < len(texts) or max_batch == 1
    assert max(rc["slot"] for rc in recs) < slots
    if budget < (1 << 30):  # a budget that binds: admission holds requests back
        assert sum(rc["mem_held"] for rc in recs) > 0


def test_plain_serving_and_estimates(served):
    e, texts, alone = served
    recs, res = e.serve(texts, arrivals(len(texts), 500.0, 1), plain=True, keep_results=True)
    got = [O.canonical(r.headers(), r.columns()) for r in res]
    for r in res:
        r.free()
    assert got == alone
    assert len({rc["batch"] for rc in recs}) == len(texts)  # one request per batch
    rows, peak = e.estimate(texts[0])
    assert rows >= 0 and peak > 0


def test_estimates_against_measured_peaks(served):
    """The admission estimate (vdq_estimate: histogram moments, views free)
    against the measured peak of device memory of each query of the mix run
    alone (allocator peak above the resident store).  Reported per query;
    every estimate must be within a factor of 8 of the measurement, both
    taken with a 16 MiB floor (below that the pool's power-of-two size
    classes and a cold bound constant's tiny result make the ratio
    meaningless, and admission budgets are GiBs)."""
    from vedas_amd._native import Ctx
    e, texts, alone = served
    ctx = Ctx.wrap(e.ctx_handle())
    ratios = []
    for t in texts:
        _, peak = e.estimate(t)
        ctx.sync()
        base = ctx.mem()[0]
        ctx.mem_peak(reset=True)
        r = e.query(t)
        ctx.sync()
        meas = ctx.mem_peak() - base
        r.free()
        ratios.append((peak + (16 << 20)) / (meas + (16 << 20)))
        if not 0.5 <= ratios[-1] <= 2:
            print(f"estimate {peak:.3g} B, measured {meas} B: {t}")
    print("estimate / measured peak per query:", " ".join(f"{x:.2f}" for x in ratios))
    assert all(1 / 8 <= x <= 8 for x in ratios), ratios
