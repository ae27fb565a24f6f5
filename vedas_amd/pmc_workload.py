"""Workload for hardware-counter runs (rocprofv3 --pmc): a device-to-device
copy of a known size (calibrates FETCH_SIZE / WRITE_SIZE against known
bytes) followed by the dominant kernels at benchmark shape -- one
permutation build of the 100M-triple store (pack, 8 radix passes, unpack)
and one IndexSwap re-sort of a 2-column relation.

Usage: rocprofv3 --pmc FETCH_SIZE -- python3 -m vedas_amd.pmc_workload
"""
import sys
import time

import numpy as np


def main():
    from vedas_amd import _native as N
    from vedas_amd.workload import STORE_100M
    prm = {k: v for k, v in STORE_100M.items() if k != "n"}
    s, p, o = N.gen_triples(STORE_100M["n"], seed=42, distinct=False, **prm)
    ctx = N.Ctx(0)
    # calibration: 1 GiB copy (the runtime's blit kernel)
    nbytes = 1 << 30
    a, b = ctx.alloc(nbytes), ctx.alloc(nbytes)
    ctx.check(ctx.L.vd_copy_d2d(ctx.h, N.C.c_void_p(b), N.C.c_void_p(a), nbytes))
    ctx.sync()
    print(f"copy {nbytes} bytes", flush=True)
    ds, dp, do = ctx.upload(s), ctx.upload(p), ctx.upload(o)
    t0 = time.perf_counter()
    idx = N.Index.from_device(ctx, ds, dp, do, len(s), ["PSO"])
    ctx.sync()
    print(f"PSO build {len(s)} triples {1e3 * (time.perf_counter() - t0):.1f} ms", flush=True)
    # IndexSwap shape: the PSO select of predicate 1, re-sorted by its object
    bb, ee = idx.lookup("PSO", 1)
    rel = idx.select("PSO", bb, ee, 3)
    rel.swap_cols(0, 1)
    rel.sort_col0()
    ctx.sync()
    print(f"swap sort of {rel.nrows} rows", flush=True)
    rel.free()
    idx.free()
    return 0


if __name__ == "__main__":
    sys.exit(main())
