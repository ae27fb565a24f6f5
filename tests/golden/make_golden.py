"""Writes the small golden fixtures under tests/golden/.

Expected outputs come from plain Python (sorted tuples, run starts, nested
loops), not from the oracle or the GPU path, so the fixtures pin both.
The reference ships no expected outputs for this path.
Run: python tests/golden/make_golden.py
"""
import itertools
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PERM_COLS = {"SPO": (0, 1, 2), "PSO": (1, 0, 2), "POS": (1, 2, 0), "OSP": (2, 0, 1)}


def starts(keys):
    out, i = [], 0
    for k, g in itertools.groupby(keys):
        out.append((k, i))
        i += sum(1 for _ in g)
    return out


def index_fixture():
    rng = np.random.default_rng(20240101)
    n = 2000
    s = rng.integers(1, 60, n).astype(np.uint32)
    p = rng.integers(1, 9, n).astype(np.uint32)
    o = rng.integers(1, 90, n).astype(np.uint32)
    d = {"s": s, "p": p, "o": o}
    cols = (s.tolist(), p.tolist(), o.tolist())
    for perm, (xi, yi, zi) in PERM_COLS.items():
        tup = sorted(zip(cols[xi], cols[yi], cols[zi]))
        l1 = starts([t[0] for t in tup])
        l2 = starts([(t[0], t[1]) for t in tup])
        d[f"{perm}_y"] = np.array([t[1] for t in tup], np.uint32)
        d[f"{perm}_z"] = np.array([t[2] for t in tup], np.uint32)
        d[f"{perm}_l1v"] = np.array([k for k, _ in l1], np.uint32)
        d[f"{perm}_l1o"] = np.array([i for _, i in l1], np.uint32)
        d[f"{perm}_l2v"] = np.array([k[1] for k, _ in l2], np.uint32)
        d[f"{perm}_l2o"] = np.array([i for _, i in l2], np.uint32)
    np.savez_compressed(os.path.join(HERE, "index_small.npz"), **d)


def join_fixture():
    rng = np.random.default_rng(7)
    def rel(n):
        k = rng.integers(0, 40, n)
        v = rng.integers(0, 1000, n)
        order = sorted(range(n), key=lambda i: k[i])  # stable by key
        return (np.array([k[i] for i in order], np.uint32),
                np.array([v[i] for i in order], np.uint32))
    l0, l1 = rel(300)
    r0, r1 = rel(500)
    out = []
    for i in range(len(l0)):
        for j in range(len(r0)):
            if l0[i] == r0[j]:
                out.append((l0[i], l1[i], r1[j]))
    o = np.array(out, np.uint32).reshape(-1, 3)
    np.savez_compressed(os.path.join(HERE, "join_small.npz"), l0=l0, l1=l1, r0=r0,
                        r1=r1, out0=o[:, 0].copy(), out1=o[:, 1].copy(),
                        out2=o[:, 2].copy())


if __name__ == "__main__":
    index_fixture()
    join_fixture()
    print("written")
