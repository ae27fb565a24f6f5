"""Turn rocprofv3 CSV output into the short text summaries kept under
profiles/.

  trace    --kernel-trace --stats: per-kernel calls and average duration,
           with the launch label the library's own per-launch timer uses;
  traffic  two --pmc runs (FETCH_SIZE, then WRITE_SIZE) of the same
           workload: HBM bytes per launch of each kernel, each counter
           scaled by what it reads for the calibration copy of known size
           (the first __amd_rocclr_copyBuffer dispatch, 1 GiB);
  counters one --pmc run of several counters: their mean per launch for
           each kernel whose name matches a pattern, with the LDS
           bank-conflict rate and the share of wave cycles spent waiting.

  python3 -m vedas_amd.prof_summary trace STATS_CSV OUT.md "COMMAND"
  python3 -m vedas_amd.prof_summary traffic FETCH_CSV WRITE_CSV OUT.md "COMMAND"
  python3 -m vedas_amd.prof_summary counters PMC_CSV PATTERN OUT.md "COMMAND"

The traffic table is also read by bench.py (label -> bytes per launch).
"""
from __future__ import annotations

import csv
import re
import sys

# Kernel symbol (as rocprofv3 prints it) -> launch label of the library's
# per-launch timer (vd_prof_*), so both views name the same kernel.
LABELS = [
    (r"k_rs_onesweep<\d+, \d+, [1-9][,>]", "radix_onesweep_kv"),
    (r"k_rs_onesweep<\d+, \d+, 0[,>]", "radix_onesweep"),
    (r"k_rs_hist\b", "radix_hist"),
    (r"k_rs_scan_hist\b|k_swap_plan_scan\b", "radix_scan_hist"),
    (r"k_join_expand_multi\b", "join_expand"),
    (r"k_join_expand<", "join_expand"),
    (r"k_join_bounds_multi\b", "join_bounds"),
    (r"k_join_bounds<", "join_bounds"),
    (r"k_join_partition", "join_partition"),
    (r"k_scan_tile_totals\b", "scan_tiles"),
    (r"k_join_scan_total\b", "scan_tiles"),
    (r"k_rb_copy\b", "read_back"),
    (r"k_interleave<", "interleave"),
    (r"k_semi_mark\b", "semi_mark"),
    (r"k_semi_count\b", "semi_count"),
    (r"k_semi_scatter\b", "semi_scatter"),
    (r"k_join_ranges", "join_ranges"),
    (r"k_pair_totals\b", "join_totals"),
    (r"k_pack_xyz\b|k_pack_split\b|k_pack_xpos\b", "index_pack"),
    (r"k_unpack_count<", "index_unpack"),
    (r"k_compact_heads<", "index_heads"),
    (r"k_split_gather\b", "index_gather"),
    (r"k_index_stats\b", "index_stats"),
    (r"k_lookup_batch\b", "index_lookup"),
    (r"k_copy_pieces\b", "select_copy"),
    (r"k_select_pieces\b", "select_pieces"),
    (r"k_gather_ranges\b", "select_gather"),
    (r"k_sel_lookup_multi\b", "select_lookup"),
    (r"k_sel_bounds_multi\b", "select_bounds"),
    (r"k_hist_", "histogram"),
    (r"k_runs_expand\b", "runs_expand"),
    (r"k_run_table\b", "run_table"),
    (r"k_drive_bounds\b", "drive_bounds"),
    (r"k_mark_l1\b|k_mark_y\b|k_scatter_heads\b", "l2_heads"),
    (r"k_pack_rowid\b|k_pack_multi\b", "sort_pack"),
    (r"k_swap_hist\b", "sort_hist"),
    (r"k_repack\b", "index_repack"),
    (r"k_gather_perm\b|k_gather_sort_multi\b", "sort_gather"),
    (r"k_compose_perm\b", "sort_compose"),
    (r"k_max_col0_multi\b", "sort_max"),
    (r"k_unique_count", "unique_count"),
    (r"k_unique_scatter", "unique_scatter"),
    (r"k_unique_totals\b", "unique_totals"),
    (r"k_tile_scan<", "scan_tile"),
    (r"k_tile_reduce<", "scan_reduce"),
    (r"k_fill<", "fill"),
    (r"k_max_u32\b", "max_u32"),
    (r"__amd_rocclr_copyBuffer", "runtime_copy"),
]


def label_of(name: str) -> str:
    for pat, lab in LABELS:
        if re.search(pat, name):
            return lab
    return "other"


def short(name: str) -> str:
    name = re.sub(r"vd::\(anonymous namespace\)::", "", name)
    return name.split("(")[0].replace("void ", "")[:60]


def trace(stats_csv: str, out: str, command: str) -> None:
    rows = list(csv.DictReader(open(stats_csv)))
    lines = [f"# Kernel time: {command.split(' -- ')[-1]}", "",
             f"Command: `{command}`", "",
             "| kernel | launch label | calls | avg us | total ms | % |",
             "|---|---|---:|---:|---:|---:|"]
    for r in rows[:25]:
        lines.append(f"| {short(r['Name'])} | {label_of(r['Name'])} | {r['Calls']} | "
                     f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['TotalDurationNs']) / 1e6:.2f} | "
                     f"{float(r['Percentage']):.1f} |")
    open(out, "w").write("\n".join(lines) + "\n")


def _per_dispatch(path: str):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Dispatch_Id"]))
    return [(r["Kernel_Name"], float(r["Counter_Value"])) for r in rows]


def read_traffic(path: str) -> dict:
    """label -> HBM bytes per launch, from a traffic summary written below."""
    out = {}
    for line in open(path):
        if not line.startswith("| ") or line.startswith("| label") or line.startswith("|---"):
            continue
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 5:
            out[cells[0]] = float(cells[4])
    return out


def traffic(fetch_csv: str, write_csv: str, out: str, command: str) -> None:
    fetch, write = _per_dispatch(fetch_csv), _per_dispatch(write_csv)
    copy_bytes = float(1 << 30)
    f_copy = next(v for k, v in fetch if "copyBuffer" in k) * 1024.0
    w_copy = next(v for k, v in write if "copyBuffer" in k) * 1024.0
    f_scale, w_scale = copy_bytes / f_copy, copy_bytes / w_copy
    agg = {}
    for (name, fv), (name2, wv) in zip(fetch, write):
        if name != name2:
            raise SystemExit("FETCH and WRITE runs dispatched different kernels")
        lab = label_of(name)
        if lab == "runtime_copy":
            continue
        a = agg.setdefault(lab, [0, 0.0, 0.0])
        a[0] += 1
        a[1] += fv * 1024.0
        a[2] += wv * 1024.0
    lines = [f"# HBM traffic per launch: {command.split(' -- ')[-1]}", "",
             f"Commands: `{command.replace('COUNTER', 'FETCH_SIZE')}` and the same with "
             "`WRITE_SIZE` (one counter per run).", "",
             f"Calibration: the 1 GiB runtime copy reads FETCH_SIZE = {f_copy / 1024:.0f} KiB and "
             f"WRITE_SIZE = {w_copy / 1024:.0f} KiB, so FETCH_SIZE is scaled by {f_scale:.3f} and "
             f"WRITE_SIZE by {w_scale:.3f}.", "",
             "| label | launches | FETCH_SIZE KiB/launch | WRITE_SIZE KiB/launch | "
             "HBM bytes/launch (scaled) |",
             "|---|---:|---:|---:|---:|"]
    for lab, (cnt, fb, wb) in sorted(agg.items(), key=lambda kv: -(kv[1][1] + kv[1][2])):
        tot = (fb * f_scale + wb * w_scale) / cnt
        lines.append(f"| {lab} | {cnt} | {fb / cnt / 1024:.0f} | {wb / cnt / 1024:.0f} | {tot:.0f} |")
    open(out, "w").write("\n".join(lines) + "\n")


def counters(pmc_csv: str, pattern: str, out: str, command: str) -> None:
    rows = list(csv.DictReader(open(pmc_csv)))
    per = {}  # kernel -> dispatch -> counter -> value
    for r in rows:
        if not re.search(pattern, r["Kernel_Name"]):
            continue
        k = short(r["Kernel_Name"])
        d = per.setdefault(k, {}).setdefault(r["Dispatch_Id"], {})
        d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    names = sorted({c for k in per.values() for d in k.values() for c in d})
    lines = [f"# Hardware counters per launch: {command.split(' -- ')[-1]}", "",
             f"Command: `{command}`", "",
             "Means over the launches of each kernel.  LDS conflict rate = "
             "SQ_LDS_BANK_CONFLICT / (SQ_LDS_IDX_ACTIVE - SQ_LDS_BANK_CONFLICT); "
             "wait share = SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES (cycles a wave waits for an "
             "issued instruction's result, over its resident cycles); cycles per wave = "
             "SQ_WAVE_CYCLES / SQ_WAVES.", "",
             "| kernel | launches | " + " | ".join(names) +
             " | LDS conflict rate | wait share | cycles per wave |",
             "|---|---:|" + "---:|" * (len(names) + 3)]
    for k, ds in sorted(per.items()):
        n = len(ds)
        mean = {c: sum(d.get(c, 0.0) for d in ds.values()) / n for c in names}
        conf = mean.get("SQ_LDS_BANK_CONFLICT", 0.0)
        act = mean.get("SQ_LDS_IDX_ACTIVE", 0.0)
        rate = conf / (act - conf) if act > conf else float("nan")
        wc = mean.get("SQ_WAVE_CYCLES", 0.0)
        wait = mean.get("SQ_WAIT_INST_ANY", 0.0) / wc if wc else float("nan")
        cpw = wc / mean["SQ_WAVES"] if mean.get("SQ_WAVES") else float("nan")
        lines.append(f"| {k} | {n} | " + " | ".join(f"{mean[c]:.4g}" for c in names) +
                     f" | {rate:.3f} | {wait:.3f} | {cpw:.0f} |")
    open(out, "w").write("\n".join(lines) + "\n")


def main(argv):
    if argv[1] == "trace":
        trace(argv[2], argv[3], argv[4])
    elif argv[1] == "traffic":
        traffic(argv[2], argv[3], argv[4], argv[5])
    elif argv[1] == "counters":
        counters(argv[2], argv[3], argv[4], argv[5])
    else:
        raise SystemExit(__doc__)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
