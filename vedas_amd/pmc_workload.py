"""Workloads for hardware-counter runs (rocprofv3 --pmc).

Every mode first runs a device-to-device copy of 1 GiB through the runtime's
copy kernel: FETCH_SIZE / WRITE_SIZE of that dispatch calibrate the counters
against a known byte count.  Then:

  joinset  the single-query join set (star4, chain4, snowflake6) on the
           100M-triple store, twice, exactly as the benchmark's timed step
           runs it (index build of SPO/PSO/POS/OSP first);
  build    one PSO permutation build of the 100M store plus one IndexSwap
           re-sort of a 2-column select.

Usage (one counter per run; see vedas_amd.pmc_summary for the table):
  rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT -o f -- \
      python3 -m vedas_amd.pmc_workload joinset
"""
import sys
import time


def _calibrate(N, ctx):
    nbytes = 1 << 30
    a, b = ctx.alloc(nbytes), ctx.alloc(nbytes)
    ctx.check(ctx.L.vd_copy_d2d(ctx.h, N.C.c_void_p(b), N.C.c_void_p(a), nbytes))
    ctx.sync()
    ctx.free(a)
    ctx.free(b)
    print(f"copy {nbytes} bytes", flush=True)


def joinset():
    from vedas_amd import _native as N
    from vedas_amd.engine import Engine, query_text
    from vedas_amd.workload import JOIN_SET, STORE_100M
    prm = {k: v for k, v in STORE_100M.items() if k != "n"}
    s, p, o = N.gen_triples(STORE_100M["n"], seed=42, distinct=True, **prm)
    eng = Engine(0)
    ctx = N.Ctx.wrap(eng.ctx_handle())
    _calibrate(N, ctx)
    ds, dp, do = ctx.upload(s), ctx.upload(p), ctx.upload(o)
    eng.load_device(ds, dp, do, len(s), ("SPO", "PSO", "POS", "OSP"))
    ctx.sync()
    for d in (ds, dp, do):
        ctx.free(d)
    print("index built", flush=True)
    for rep in range(2):
        for q in ("star4", "chain4", "snowflake6"):
            t0 = time.perf_counter()
            r = eng.query(query_text(JOIN_SET[q][0], JOIN_SET[q][1]))
            st = r.stats()
            r.free()
            print(f"run {rep} {q}: {st['join_rows']} join rows, "
                  f"{1e3 * (time.perf_counter() - t0):.1f} ms", flush=True)
    eng.close()


def build():
    from vedas_amd import _native as N
    from vedas_amd.workload import STORE_100M
    prm = {k: v for k, v in STORE_100M.items() if k != "n"}
    s, p, o = N.gen_triples(STORE_100M["n"], seed=42, distinct=True, **prm)
    ctx = N.Ctx(0)
    _calibrate(N, ctx)
    ds, dp, do = ctx.upload(s), ctx.upload(p), ctx.upload(o)
    t0 = time.perf_counter()
    idx = N.Index.from_device(ctx, ds, dp, do, len(s), ["PSO"])
    ctx.sync()
    print(f"PSO build {len(s)} triples {1e3 * (time.perf_counter() - t0):.1f} ms", flush=True)
    bb, ee = idx.lookup("PSO", 1)
    rel = idx.select("PSO", bb, ee, 3)
    rel.swap_cols(0, 1)
    rel.sort_col0()
    ctx.sync()
    print(f"swap sort of {rel.nrows} rows", flush=True)
    rel.free()
    idx.free()


def main(argv):
    mode = argv[1] if len(argv) > 1 else "joinset"
    {"joinset": joinset, "build": build}[mode]()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
