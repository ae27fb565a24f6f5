"""CPU checks of the native libraries: they load, export every symbol the
public header declares, and the host-only entry points behave."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vd[a-z]*_[a-z0-9_]+)\s*\(", text)))


def test_hip_lib_exports_header():
    from vedas_amd import _native
    lib = C.CDLL(_native.LIB_PATH)
    names = declared("vedas_hip.h")
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes table binds exactly the declared C ABI
    assert set(_native.EXPORTS) == set(names)


def test_host_lib_loads_and_exports():
    path = os.path.join(ROOT, "vedas_amd", "libvedas.so")
    lib = C.CDLL(path)
    names = declared("vedas_host.h")
    assert names, "no declarations parsed"
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ctx_create_without_gpu_fails_cleanly():
    from vedas_amd import _native
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    h = C.c_void_p()
    rc = _native.lib().vd_ctx_create(0, C.byref(h))
    assert rc != 0 and not h.value


def test_generator_deterministic_and_distinct():
    from vedas_amd._native import gen_triples
    a = gen_triples(300_000, seed=42)
    b = gen_triples(300_000, seed=42)
    c = gen_triples(300_000, seed=43)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[0], c[0])
    s, p, o = a
    assert len(set(zip(s.tolist(), p.tolist(), o.tolist()))) == len(s)
    assert s.min() >= 1 and s.max() <= 100_000
    assert p.min() >= 1 and p.max() <= 64
    lit = o >= 50_000_000
    assert 0.45 < lit.mean() < 0.55
    assert o[~lit].max() <= 100_000 and o[lit].max() < 50_000_000 + 500_000
    # Zipf(1.1) over 64: share of predicate 1 = 1 / H(64, 1.1)
    h = sum(k ** -1.1 for k in range(1, 65))
    assert abs((p == 1).mean() - 1 / h) < 0.01


def test_generator_chunks_are_independent():
    """Triples are drawn in 2^20-triple chunks: later chunks must not
    repeat earlier ones (a shifted stream would)."""
    from vedas_amd._native import gen_triples
    n = 3 << 20
    s, p, o = gen_triples(n, seed=3, n_entities=10_000_000, n_predicates=256,
                          literal_frac=0.4, n_literals=50_000_000, distinct=False)
    key = (s.astype(np.uint64) << np.uint64(32)) | o
    assert len(np.unique(key)) > 0.99 * n
    c = 1 << 20
    assert not np.array_equal(s[1:c], s[c:2 * c - 1])


def test_generator_zipf_subjects():
    from vedas_amd._native import gen_triples
    s, p, o = gen_triples(200_000, seed=1, n_entities=10_000, subject_zipf=0.8,
                          n_predicates=256, literal_frac=0.4, n_literals=50_000,
                          distinct=False)
    h = sum(k ** -0.8 for k in range(1, 10_001))
    assert abs((s == 1).mean() - 1 / h) < 0.005
    assert s.max() <= 10_000 and p.max() <= 256


def test_generator_rejects_bad_params():
    from vedas_amd._native import GenParams, lib
    prm = GenParams(1, 10, 0, 0.0, 4, 1.1, 0.5, 100, 10, 0)
    s = np.empty(10, np.uint32)
    rc = lib().vd_gen_triples(C.byref(prm), s.ctypes.data, s.ctypes.data, s.ctypes.data)
    assert rc == -1


def test_executor_options_round_trip():
    """The process-wide executor switches (no device work): each option
    reads back what was set, filtered_swap takes a row count and rejects a
    negative one, unknown names fail."""
    from vedas_amd.engine import get_option, set_option
    saved = {k: get_option(k) for k in ("bound_prescan", "bound_after_join", "driven_swap",
                                        "filtered_swap")}
    try:
        for k in ("bound_prescan", "bound_after_join", "driven_swap"):
            set_option(k, 0)
            assert get_option(k) == 0
            set_option(k, 1)
            assert get_option(k) == 1
        assert saved["filtered_swap"] == 1 << 20  # the default: relations of 2^20 rows
        set_option("filtered_swap", 12345)
        assert get_option("filtered_swap") == 12345
        with pytest.raises(Exception):
            set_option("filtered_swap", -1)
        assert get_option("filtered_swap") == 12345
        with pytest.raises(Exception):
            set_option("no_such_option", 1)
    finally:
        for k, v in saved.items():
            set_option(k, v)
