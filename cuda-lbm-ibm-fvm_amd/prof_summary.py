"""Summaries of rocprofv3 CSV output for the kernel-time, HBM-traffic and
counter tables kept as text under profiles/.

  python prof_summary.py stats DIR
      per-kernel calls / average / min / max / share from DIR/*kernel_stats.csv
  python prof_summary.py traffic FETCH_DIR WRITE_DIR --copy-bytes N [--tune TUNE.json] [--json OUT]
      HBM bytes per launch from one FETCH_SIZE and one WRITE_SIZE pass
      (KiB per dispatch), each scaled by what the counter reads for the first
      k_copy dispatch of the run, a plain copy that reads and writes N bytes
  python prof_summary.py counters DIR [DIR ...]
      mean of every counter per kernel over the dispatches of DIR/*counter_collection.csv
  python prof_summary.py valu SQ_DIR STATS_DIR [--cells N]
      SQ_INSTS_VALU per kernel, per cell, and the time its issue takes on the
      CUs' SIMDs beside the kernel-trace average

Markdown goes to stdout; --json writes the per-kernel traffic that bench.py
reads for its roofline line.
"""
import argparse
import collections
import csv
import glob
import json
import os
import sys

# bench.py stage name of each step kernel (by the kernel's name up to its template arguments)
STAGE_OF = {
    "vof::k_predict": "predictor",
    "vof::k_predict_band": "predictor",
    "vof::k_jacobiT": "jacobi",
    "vof::k_rhs_blk": "rhs",
    "vof::k_rhs": "rhs",
    "vof::k_jacobi2_dma": "jacobi",
    "vof::k_jacobi2_blk": "jacobi",
    "vof::k_jacobi2": "jacobi",
    "vof::k_jacobi_blk": "jacobi",
    "vof::k_jacobi": "jacobi",
    "vof::k_correct_blk": "correction",
    "vof::k_correct": "correction",
    "vof::k_fct_zx": "fct_zx",
    "vof::k_fct_xy": "fct_xy",
    "vof::k_fct_xb": "fct_x",
}
# algorithmic bytes per interior cell of one launch at FP32 (read + write of
# each field the stage must touch once; a fused FCT pair moves 16 B/cell)
ALGO_B_PER_CELL = {"predictor": 28, "rhs": 20, "correction": 32, "fct_x": 12, "fct_y": 12, "fct_z": 12,
                   "fct_zx": 16, "fct_xy": 16}


def short_name(name):
    name = name.replace("void ", "")
    return name.split("(")[0]


def stage_of(kname):
    base = kname.split("<")[0]
    if base == "vof::k_fct_line":
        axis = kname.split("<")[1].split(",")[1].strip()
        return {"1": "fct_y", "2": "fct_z"}.get(axis)
    return STAGE_OF.get(base)


def find(d, pattern):
    hits = sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True))
    if not hits:
        sys.exit(f"no {pattern} under {d}")
    return hits[0]


def cmd_stats(a):
    rows = list(csv.DictReader(open(find(a.dir, "*kernel_stats.csv"))))
    print("| Kernel | Calls | Average (us) | Min (us) | Max (us) | Share |")
    print("|---|---|---|---|---|---|")
    for r in rows:
        print(f"| `{short_name(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
              f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | {float(r['Percentage']):.1f}% |")


def per_dispatch(d, counter):
    """(kernel, value) per dispatch, in dispatch order; a counter summed over
    its instances within a dispatch."""
    acc = collections.OrderedDict()
    for r in csv.DictReader(open(find(d, "*counter_collection.csv"))):
        if r["Counter_Name"] != counter:
            continue
        key = int(r["Dispatch_Id"])
        k, v = acc.get(key, (short_name(r["Kernel_Name"]), 0.0))
        acc[key] = (k, v + float(r["Counter_Value"]))
    return [acc[k] for k in sorted(acc)]


def cmd_traffic(a):
    fetch = per_dispatch(a.fetch_dir, "FETCH_SIZE")
    write = per_dispatch(a.write_dir, "WRITE_SIZE")
    cal_f = next(v for k, v in fetch if k == "vof::k_copy")
    cal_w = next(v for k, v in write if k == "vof::k_copy")
    ff = a.copy_bytes / 1024.0 / cal_f
    fw = a.copy_bytes / 1024.0 / cal_w
    mean = collections.defaultdict(lambda: [0.0, 0, 0.0, 0])
    for k, v in fetch:
        mean[k][0] += v
        mean[k][1] += 1
    for k, v in write:
        mean[k][2] += v
        mean[k][3] += 1
    tuning = {}
    if a.tune:
        with open(a.tune) as fh:
            t = json.load(fh)
        tuning = next(iter(t.values())) if t else {}
    out = {"workload": a.workload, "cells": a.cells, "fetch_factor": ff, "write_factor": fw, "tuning": tuning,
           "kernels": {}}
    print(f"Calibration: the first k_copy dispatch reads and writes {a.copy_bytes} bytes; it reads FETCH_SIZE = "
          f"{cal_f:.0f} KiB and WRITE_SIZE = {cal_w:.0f} KiB, so bytes = FETCH_SIZE x 1024 x {ff:.4f} and "
          f"WRITE_SIZE x 1024 x {fw:.4f}.  Values are means over the profiled launches of each kernel.\n")
    print("| Kernel | Launches | FETCH_SIZE KiB | WRITE_SIZE KiB | MB read | MB written | total MB | algorithmic MB "
          "| ratio |")
    print("|---|---|---|---|---|---|---|---|---|")
    for k, (fs, nf, ws, nw) in mean.items():
        st = stage_of(k)
        if st is None or nf == 0 or nw == 0:
            continue
        fk, wk = fs / nf, ws / nw
        rb, wb = fk * 1024 * ff, wk * 1024 * fw
        algo = None
        if st == "jacobi":
            sweeps = 2 if "jacobi2" in k else 1
            algo = 12 * sweeps * a.cells
        elif st in ALGO_B_PER_CELL:
            algo = ALGO_B_PER_CELL[st] * a.cells
        ratio = (rb + wb) / algo if algo else None
        out["kernels"].setdefault(st, {})
        out["kernels"][st] = {"kernel": k, "fetch_kib": fk, "write_kib": wk, "read_bytes": rb, "write_bytes": wb,
                              "bytes_per_launch": rb + wb, "algorithmic_bytes": algo, "launches_profiled": nf}
        print(f"| `{k}` | {nf} | {fk:.0f} | {wk:.0f} | {rb / 1e6:.1f} | {wb / 1e6:.1f} | {(rb + wb) / 1e6:.1f} | "
              f"{algo / 1e6 if algo else float('nan'):.1f} | {ratio if ratio else float('nan'):.2f} |")
    if a.json:
        with open(a.json, "w") as fh:
            json.dump(out, fh, indent=1)


def cmd_counters(a):
    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    for d in a.dirs:
        for r in csv.DictReader(open(find(d, "*counter_collection.csv"))):
            k = short_name(r["Kernel_Name"])
            if not k.startswith("vof::"):
                continue
            vals[k][(r["Counter_Name"], r["Dispatch_Id"])].append(float(r["Counter_Value"]))
    for k, c in sorted(vals.items()):
        per = collections.defaultdict(list)
        for (name, _), v in c.items():
            per[name].append(sum(v))
        print(f"### `{k}`\n")
        print("| Counter | mean per dispatch |")
        print("|---|---|")
        for name in sorted(per):
            v = per[name]
            print(f"| {name} | {sum(v) / len(v):.4g} |")
        print()


def cmd_valu(a):
    """VALU issue per kernel: SQ_INSTS_VALU per dispatch, per interior cell
    (lane-instructions: x 64 / cells) and the time the CU's four 16-lane SIMDs
    need to issue them (4 cycles per wave64 instruction, a.cus x 4 SIMDs, at
    a.ghz), against the kernel-trace average."""
    avg = {short_name(r["Name"]): float(r["AverageNs"]) / 1e3
           for r in csv.DictReader(open(find(a.stats_dir, "*kernel_stats.csv")))}
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    for r in csv.DictReader(open(find(a.sq_dir, "*counter_collection.csv"))):
        k = short_name(r["Kernel_Name"])
        if k.startswith("vof::") and k != "vof::k_copy":
            per[k][(r["Counter_Name"], r["Dispatch_Id"])] += float(r["Counter_Value"])
    print("| Kernel | VALU / dispatch | SALU / dispatch | VALU / cell | VALU issue (us) | average (us) | VALU share |")
    print("|---|---|---|---|---|---|---|")
    for k, c in sorted(per.items()):
        def mean(name):
            v = [x for (n, _), x in c.items() if n == name]
            return sum(v) / len(v) if v else float("nan")
        valu, salu = mean("SQ_INSTS_VALU"), mean("SQ_INSTS_SALU")
        t_issue = valu * 4 / (a.cus * 4) / (a.ghz * 1e3)
        t = avg.get(k, float("nan"))
        print(f"| `{k}` | {valu:.3e} | {salu:.3e} | {valu * 64 / a.cells:.0f} | {t_issue:.1f} | {t:.2f} | "
              f"{t_issue / t:.0%} |")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    p = sub.add_parser("stats")
    p.add_argument("dir")
    p = sub.add_parser("traffic")
    p.add_argument("fetch_dir")
    p.add_argument("write_dir")
    p.add_argument("--copy-bytes", type=int, required=True)
    p.add_argument("--cells", type=int, default=200 ** 3)
    p.add_argument("--workload", default="200^3 FP32 dam break")
    p.add_argument("--tune", default=None)
    p.add_argument("--json", default=None)
    p = sub.add_parser("counters")
    p.add_argument("dirs", nargs="+")
    p = sub.add_parser("valu")
    p.add_argument("sq_dir")
    p.add_argument("stats_dir")
    p.add_argument("--cells", type=int, default=200 ** 3)
    p.add_argument("--cus", type=int, default=256)
    p.add_argument("--ghz", type=float, default=2.4)
    a = ap.parse_args()
    {"stats": cmd_stats, "traffic": cmd_traffic, "counters": cmd_counters, "valu": cmd_valu}[a.cmd](a)


if __name__ == "__main__":
    main()
