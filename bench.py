"""Benchmark of the VOF timestep on one MI355X.

Workload (one "step" = one full time step of the hot path): the 200^3 FP32
dam break in a 0.1 m cube, liquid block x<0.04, y<0.04, z<0.06 m, rho 1000/50,
sigma 0.007 (unused by the reference physics), g = (0,0,-9.81), dt = 4e-6 s,
10 Jacobi sweeps per step.  Fields are resident in HBM before the timed
region; the timed region replays one hipGraph per step.

Prints ONE JSON line.  `value` is interior cell-updates per second over the
timed steps (wall clock between two stream synchronisations).  `roofline`
describes the kernel with the largest share of step time, from HIP events
recorded live on the solver's stream.  `cpu_baseline` times the CPU oracle
(the reference's kernel bodies as OpenMP loops) on a bounded sample of the
same workload.
"""
from __future__ import annotations

import argparse
import subprocess
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cuda-lbm-ibm-fvm_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "cell-updates/s per timestep at 200³ + % HBM peak; batched & adjoint vs plain"

# Algorithmic HBM bytes per interior cell for each stage (FP32; x2 for FP64):
# every array the stage must read or write, touched once.
STAGE_BYTES = {
    "predictor": 28,   # read u,v,w,F; write u*,v*,w*
    "rhs": 20,         # read u*,v*,w*,F; write rhs
    "jacobi": 12,      # per sweep: read p, rhs; write p'
    "correction": 32,  # read u*,v*,w*,p,F; write u,v,w
    "fct_x": 12,       # read F, u; write F
    "fct_y": 12,
    "fct_z": 12,
}
LAUNCHES_PER_STEP = {"predictor": 1, "rhs": 1, "jacobi": 10, "correction": 1, "fct_x": 1, "fct_y": 1, "fct_z": 1}
# Two FCT sweeps in one launch (z then x in the sweep orders y-z-x and z-x-y,
# x then y in x-y-z): credited with both sweeps' 24 B/cell, the "effective"
# figure as for two Jacobi sweeps per launch; what such a launch must move
# (read F and the two velocities; write F) is 16.
FCT_PAIR_BYTES = 24
FCT_PAIR_COMPULSORY_BYTES = 16
# Fields each launch must read and write once (the compulsory stream mix):
# a plain stream kernel moving the same mix (vof_stream_bandwidth, inputs from
# HBM) gives the attainable time that launch is compared with.
STREAM_MIX = {
    "predictor": (4, 3), "rhs": (4, 1), "jacobi": (2, 1), "correction": (5, 3),
    "fct_x": (2, 1), "fct_y": (2, 1), "fct_z": (2, 1), "fct_zx": (3, 1), "fct_xy": (3, 1),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--n", type=int, default=200)
    ap.add_argument("--precision", type=int, default=4)
    ap.add_argument("--eager", action="store_true", help="launch kernels per step instead of graph replay")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-instance ensemble leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the reverse-mode leg")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--tune-cache", default=None,
                    help="JSON file of autotuned launch shapes: read if it holds this shape, written otherwise "
                         "(lets profiler runs reuse the benchmark's choice without re-running trials)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--stage-steps", type=int, default=30)
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="run one plain copy of this many bytes first (byte-counter calibration)")
    ap.add_argument("--jacobi-tblock", type=int, default=0,
                    help="Jacobi sweeps per launch (1 .. 5) before the autotuner (use with --no-autotune to keep it)")
    ap.add_argument("--kz-jacobi", type=int, default=0, help="Jacobi z-chunk length, as --jacobi-tblock")
    ap.add_argument("--kz-predict", type=int, default=0,
                    help="predictor z-chunk length (4 or 8: bands of 4 rows per lane in FP32), as --jacobi-tblock")
    ap.add_argument("--case", choices=["dam", "ripple"], default="dam",
                    help="initial state: the dam break of the metric (default), or the liquid layer with a rippled "
                         "surface (mostly liquid: every FCT wave does the limiter arithmetic), for profiling")
    ap.add_argument("--study", choices=["precision", "long", "adjoint-sizes"], default=None,
                    help="instead of the benchmark line, print the FP32-vs-FP64 study of the 200^3 dam break "
                         "(1000 steps) and the 384^3 ripple layer (500 steps); 'long': the 200^3 dam break "
                         "over 20000 steps; 'adjoint-sizes': reverse-mode cost over the plain step at 64^3 .. "
                         "160^3")
    return ap.parse_args()


# Launch-shape fields that change what each kernel moves (the others only
# change other kernels' launches).
TUNE_FIELDS = {
    "predictor": ("kz_predict",), "rhs": ("kz_rhs",), "jacobi": ("kz_jacobi", "jacobi_tblock"),
    "correction": ("kz_correct",), "fct_x": ("kz_fct_x",), "fct_y": ("fct_chunk",), "fct_z": ("fct_chunk",),
    "fct_zx": ("kz_fct_x", "fct_zx_fused"), "fct_xy": ("fct_chunk", "fct_xy_fused"),
}


def load_traffic(kernel, tuning):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (profiles/traffic_200.json, written from rocprofv3 FETCH_SIZE / WRITE_SIZE
    runs scaled by a copy kernel of known size) and a label naming that file
    and the launch shapes it was profiled with; (None, reason) when the file
    is absent or the kernel was profiled with other launch shapes than this
    run's (only the fields that change that kernel's launches are compared)."""
    path = os.path.join(ROOT, "profiles", "traffic_200.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
        byts = t["kernels"][kernel]["bytes_per_launch"]
        prof = t.get("tuning", {})
    except (OSError, KeyError, ValueError):
        return None, "no counter summary for this kernel"
    fields = TUNE_FIELDS.get(kernel, tuple(prof))
    if not prof or any(tuning.get(k) != prof.get(k) for k in fields):
        return None, f"profiles/traffic_200.json was profiled with tuning {prof}, this run uses {tuning}"
    same = {k: prof[k] for k in fields}
    return byts, f"rocprofv3 FETCH_SIZE + WRITE_SIZE from profiles/traffic_200.json, {kernel} launch shapes {same}"


from vof_cases import SEED, ensemble_config, field_errors, ripple_F  # noqa: E402


def adjoint_sizes(vof, sizes=(64, 96, 128, 160), steps=100, every=20):
    """Reverse-mode cost over the plain step across grid sizes: the dam break
    from rest (FP32 forward, FP64 adjoint accumulation), `steps` steps with a
    checkpoint every `every`; forward with checkpoints, backward, and the same
    steps as plain graph replays, each timed on the host around a sync."""
    out = {"workload": f"dam break from rest, FP32, {steps} steps, checkpoint every {every}", "sizes": {}}
    for n in sizes:
        s = vof.Solver(vof.default_params(n=n, precision=4, g=(0.0, 0.0, -9.81), dt=4e-6))
        s.init_box(0.04, 0.04, 0.06)
        s.adj_forward(1, steps, every)
        s.adj_backward(1.0)  # warm-up
        s.init_box(0.04, 0.04, 0.06)
        s.sync()
        t0 = time.perf_counter()
        s.adj_forward(1, steps, every)
        s.sync()
        t1 = time.perf_counter()
        s.adj_backward(1.0)
        s.sync()
        t2 = time.perf_counter()
        s.init_box(0.04, 0.04, 0.06)
        s.step(1, 10, graph=True)
        s.sync()
        t3 = time.perf_counter()
        s.step(11, steps, graph=True)
        s.sync()
        t4 = time.perf_counter()
        s.close()
        out["sizes"][f"{n}^3"] = {"forward_ms": 1e3 * (t1 - t0), "backward_ms": 1e3 * (t2 - t1),
                                  "plain_ms": 1e3 * (t4 - t3), "backward_over_plain": (t2 - t1) / (t4 - t3)}
        print(f"adjoint sizes: {n}^3 backward / plain {(t2 - t1) / (t4 - t3):.2f}", file=sys.stderr, flush=True)
    return out


def make_solver(vof, n, prec, case="dam"):
    prm = vof.default_params(n=n, precision=prec, g=(0.0, 0.0, -9.81))
    s = vof.Solver(prm)
    if case == "ripple":
        s.upload("F", ripple_F(n, dtype=s.dtype))
    else:
        s.init_box(0.04, 0.04, 0.06)
    return s


def batched_leg(vof, steps=100, warmup=10):
    """Aggregate cell-updates/s of the ensemble advanced as one batch (one launch
    per stage, hipGraph per step) and as 64 plain solvers run one after the other."""
    prms, boxes = ensemble_config(vof)
    cells = sum(p.nx * p.ny * p.nz for p in prms)
    b = vof.Batch(prms)
    b.init_box(*zip(*boxes))
    b.autotune(1, 3, None)
    b.step(1, warmup, graph=True)
    b.sync()
    t0 = time.perf_counter()
    b.step(1 + warmup, steps, graph=True)
    b.sync()
    tb = time.perf_counter() - t0
    b.close()
    tp = 0.0
    for p, bx in zip(prms, boxes):
        s = vof.Solver(p)
        s.init_box(*bx)
        s.step(1, warmup, graph=True)
        s.sync()
        t0 = time.perf_counter()
        s.step(1 + warmup, steps, graph=True)
        s.sync()
        tp += time.perf_counter() - t0
        s.close()
    batched = cells * steps / tb
    plain = cells * steps / tp
    return {"value": batched, "unit": "cell-updates/s", "plain_value": plain, "ratio_vs_plain": batched / plain,
            "instances": len(prms), "interior_cells": cells, "steps": steps,
            "ms_per_step_batched": 1e3 * tb / steps, "ms_per_step_plain_sum": 1e3 * tp / steps}


def adjoint_leg(vof, n=96, steps=400, every=20, seed=SEED):
    """Reverse mode on the differentiation workload: 96^3 dam break (liquid
    block x<0.04, y<0.04, z<0.06 m), initial velocity = seeded Gaussian noise of
    std 1e-3 m/s smoothed over 4 cells, 400 steps, checkpoint every 20;
    objective J = liquid-centroid height after the last step.  Cost: device time
    of the backward sweep (interval recomputation included) over that of the
    checkpointed forward, and over the plain graph-replayed step.  Accuracy:
    the FP64 gradient against central differences (relative step 1e-4) for
    rho_l and one smooth velocity direction, and FP32 against FP64."""
    import numpy as np
    import vof_autograd as va

    def solver(prec, **kw):
        p = vof.default_params(n=n, precision=prec, g=(0.0, 0.0, -9.81), dt=4e-6, **kw)
        s = vof.Solver(p)
        s.init_box(0.04, 0.04, 0.06)
        return s

    s = solver(4)
    vel = {k: va.smoothed_noise(s.shape, 1e-3, 4.0, seed + i) for i, k in enumerate("UVW")}
    state = {"F": s.download("F").astype(np.float64), "P": np.zeros(s.shape), **vel}
    s.close()

    def load(s, st):
        for k, a in st.items():
            s.upload(k, a.astype(s.dtype))

    out = {"workload": f"dam-break {n}^3, u0 noise 1e-3 m/s (4-cell smoothing), {steps} steps, "
                       f"checkpoint every {every}, J = liquid-centroid height"}
    grads = {}
    for prec in (4, 8):
        s = solver(prec)
        s.autotune(1, 4, None)  # launch shapes for this grid (state restored)
        times = []
        for rep in range(4):  # the first run allocates and captures graphs; median of the other three
            load(s, state)
            J = s.adj_forward(1, steps, every)
            sc = s.adj_backward(1.0)
            if rep:
                times.append(s.adj_times())
        fwd_ms, bwd_ms = sorted(times, key=lambda fb: fb[1] / fb[0])[1]
        grads[prec] = {k: s.adj_download(k) for k in "UVW"}
        load(s, state)
        plain_ms = s.time_steps(1, 3, steps, graph=True)
        key = "fp32" if prec == 4 else "fp64"
        out[key] = {"J": J, "fwd_ms": fwd_ms, "bwd_ms": bwd_ms, "ratio": bwd_ms / fwd_ms,
                    "ratio_vs_plain_step": bwd_ms / plain_ms, "plain_fwd_ms": plain_ms,
                    "dJ_drho_l": sc["rho_l"], "dJ_drho_g": sc["rho_g"], "dJ_dsigma": sc["sigma"]}
        s.close()
        if prec == 8:
            h = 1e-4

            def J_of(st, **kw):
                t = solver(8, **kw)
                load(t, st)
                v = t.adj_forward(1, steps, steps)
                t.close()
                return v

            d = {k: va.smoothed_noise(state["U"].shape, 1.0, 4.0, seed + 100 + i) for i, k in enumerate("UVW")}
            scale = np.sqrt(sum((state[k] ** 2).sum() for k in "UVW") / sum((d[k] ** 2).sum() for k in "UVW"))
            d = {k: a * scale for k, a in d.items()}
            fd_u = (J_of(dict(state, **{k: state[k] + h * d[k] for k in "UVW"})) -
                    J_of(dict(state, **{k: state[k] - h * d[k] for k in "UVW"}))) / (2 * h)
            ad_u = float(sum((grads[8][k] * d[k]).sum() for k in "UVW"))
            fd_r = (J_of(state, rho_l=1000.0 * (1 + h)) - J_of(state, rho_l=1000.0 * (1 - h))) / (2 * h * 1000.0)
            out["fd_check_fp64"] = {
                "h_relative": h,
                "u0_direction": {"adjoint": ad_u, "fd": fd_u, "rel_err": abs(ad_u - fd_u) / abs(fd_u)},
                "rho_l": {"adjoint": sc["rho_l"], "fd": fd_r, "rel_err": abs(sc["rho_l"] - fd_r) / abs(fd_r)},
            }
    num = np.sqrt(sum(((grads[4][k] - grads[8][k]) ** 2).sum() for k in "UVW"))
    den = np.sqrt(sum((grads[8][k] ** 2).sum() for k in "UVW"))
    out["fp32_vs_fp64_dJ_du0_rel_l2"] = float(num / den)
    out["ratio"] = out["fp32"]["ratio"]
    out["csf"] = csf_adjoint_leg(vof)
    return out


def csf_adjoint_leg(vof, n=64, steps=200, every=20, sigma=0.07):
    """Reverse mode with surface tension (VOF_FLAG_CSF): a density-matched
    (rho 1/1) ellipsoidal drop flattened against the floor of the 0.1 m box
    (semi-axes 25, 25, 12 mm), sigma 0.07 N/m, dt 2e-5 s, 200 steps on 64^3,
    checkpoint every 20; surface tension rounds the drop and lifts its
    centroid.  Cost ratio in FP32 and FP64; dJ/dsigma against central
    differences (relative step 1e-4) in FP64."""
    from vof_cases import drop_F

    def solver(prec, sg=sigma):
        p = vof.default_params(n=n, precision=prec, g=(0.0, 0.0, -9.81), dt=2e-5, rho_l=1.0, rho_g=1.0,
                               flags=vof.FLAG_CSF, sigma=sg)
        s = vof.Solver(p)
        s.upload("F", drop_F(n, dtype=s.dtype))
        return s

    out = {"workload": f"drop {n}^3 with CSF (sigma {sigma} N/m, rho 1/1), {steps} steps, checkpoint every {every}"}
    for prec in (4, 8):
        s = solver(prec)
        s.autotune(1, 3, None)
        for rep in range(2):
            s.upload("F", drop_F(n, dtype=s.dtype))
            for k in ("U", "V", "W", "P"):
                s.upload(k, np.zeros(s.shape, dtype=s.dtype))
            J = s.adj_forward(1, steps, every)
            sc = s.adj_backward(1.0)
            fwd_ms, bwd_ms = s.adj_times()
        s.close()
        out["fp32" if prec == 4 else "fp64"] = {"J": J, "fwd_ms": fwd_ms, "bwd_ms": bwd_ms, "ratio": bwd_ms / fwd_ms,
                                                 "dJ_dsigma": sc["sigma"]}
        if prec == 8:
            def J_of(sg):
                t = solver(8, sg)
                v = t.adj_forward(1, steps, steps)
                t.close()
                return v

            # J is piecewise smooth in sigma: a relative step of 1e-4 already
            # crosses limiter/clip branch changes over 200 steps, 1e-6 does
            # not (and still sits far above FP64 round-off)
            checks = {}
            for h in (1e-4, 1e-6):
                fd = (J_of(sigma * (1 + h)) - J_of(sigma * (1 - h))) / (2 * h * sigma)
                checks[f"h_relative_{h:g}"] = {"fd": fd, "rel_err": abs(sc["sigma"] - fd) / abs(fd)}
            out["fd_check_fp64_dJ_dsigma"] = {"adjoint": sc["sigma"], **checks}
    return out


def precision_study(vof, long_run=False):
    """FP32 fields (FP64 sums) against an all-FP64 run of the same case, both on
    the GPU: per-field errors after the last step (p with its interior mean
    removed, as p is defined up to a constant), the Jacobi residual, liquid
    volume, F range and non-finite count every `every` steps, and the
    graph-replayed step rate of each precision with the step's algorithmic
    bytes (236 B/cell FP32, x2 FP64) against the 8 TB/s HBM peak.  long_run:
    the 200^3 dam break for 20000 steps (0.08 s of flow) instead."""
    if long_run:
        cases = [("dam_break_200_20000", "dam-break 200^3, box IC x<0.04 y<0.04 z<0.06 m, g=(0,0,-9.81), dt=4e-6 s, "
                  "20000 steps", 200, 20000, 1000, "box")]
    else:
        cases = [
            ("dam_break_200", "dam-break 200^3, box IC x<0.04 y<0.04 z<0.06 m, g=(0,0,-9.81), dt=4e-6 s", 200, 1000,
             50, "box"),
            ("ripple_384", "liquid layer 0.06 m deep at 384^3, free surface = 4 plane waves of up to 2 cells "
             "(wavelengths U[8,32] cells, seed 20261203), g=(0,0,-9.81), dt=4e-6 s", 384, 500, 25, "ripple"),
        ]
    out = {"study": "long run" if long_run else "precision", "cases": {}}
    for key, desc, n, steps, every, ic in cases:
        runs = {}
        for prec in (8, 4):
            s = vof.Solver(vof.default_params(n=n, precision=prec, g=(0.0, 0.0, -9.81)))
            if ic == "box":
                s.init_box(0.04, 0.04, 0.06)
            else:
                s.upload("F", ripple_F(n, dtype=s.dtype))
            s.autotune(1, 3, None)  # trials run on a snapshot; the state is restored
            hist = [dict(step=0, **{k: s.diagnostics()[k] for k in ("liquid_volume", "residual_l2")})]
            for c in range(steps // every):
                s.step(1 + c * every, every, graph=True)
                d = s.diagnostics()
                hist.append(dict(step=(c + 1) * every, liquid_volume=d["liquid_volume"],
                                 residual_l2=d["residual_l2"], nonfinite=d["nonfinite"], f_min=d["f_min"],
                                 f_max=d["f_max"]))
            fields = {k: s.download(k) for k in ("F", "U", "V", "W", "P")}
            ms = s.time_steps(1 + steps, 10, 100, graph=True) / 100
            print(f"precision study: {key} {'FP64' if prec == 8 else 'FP32'} done", file=sys.stderr, flush=True)
            s.close()
            esz = 2 if prec == 8 else 1
            runs[prec] = dict(history=hist, fields=fields, ms_per_step=ms, cell_updates_per_s=n ** 3 / (ms * 1e-3),
                              step_hbm_frac=236 * esz * n ** 3 / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS)
        f64, f32 = runs[8]["fields"], runs[4]["fields"]
        errs = {}
        for k in ("F", "U", "V", "W"):
            errs[k] = field_errors(f32[k], f64[k])
        p32 = f32["P"][1:-1, 1:-1, 1:-1].astype(np.float64)
        p64 = f64["P"][1:-1, 1:-1, 1:-1]
        errs["P_mean_removed"] = field_errors(p32 - p32.mean(), p64 - p64.mean(), interior=False)
        v0 = runs[8]["history"][0]["liquid_volume"]
        out["cases"][key] = {
            "workload": desc, "steps": steps,
            "fp32_vs_fp64": errs,
            "liquid_volume_drift": {str(prec): runs[prec]["history"][-1]["liquid_volume"] / v0 - 1.0 for prec in (4, 8)},
            "history": {str(prec): runs[prec]["history"] for prec in (4, 8)},
            "rate": {str(prec): {k: runs[prec][k] for k in ("ms_per_step", "cell_updates_per_s", "step_hbm_frac")}
                     for prec in (4, 8)},
        }
    return out


def transfer_leg(s, cells, ms_step):
    """Host <-> device cost of handing the state over as host buffers: upload
    and download of F, u, v, w, p through the C ABI (pageable numpy arrays,
    the solver's stream), and the step rate if every step were such a round
    trip.  Not part of `value`, which is measured on HBM-resident fields."""
    names = ("F", "U", "V", "W", "P")
    host = {k: s.download(k) for k in names}
    s.sync()
    t0 = time.perf_counter()
    for k in names:
        s.upload(k, host[k])
    s.sync()
    t_up = time.perf_counter() - t0
    t0 = time.perf_counter()
    for k in names:
        host[k] = s.download(k)
    t_down = time.perf_counter() - t0
    nbytes = sum(a.nbytes for a in host.values())
    return {"bytes_each_way": nbytes, "upload_ms": 1e3 * t_up, "download_ms": 1e3 * t_down,
            "upload_gbs": nbytes / t_up / 1e9, "download_gbs": nbytes / t_down / 1e9,
            "cell_updates_per_s_one_step_per_round_trip": cells / (ms_step * 1e-3 + t_up + t_down)}


def _host_cpu():
    """This host's CPU as lscpu reports it: model, sockets, cores per socket,
    threads per core, logical CPUs; plus the CPUs this process may run on."""
    info = {"model": "unknown", "sockets": None, "cores_per_socket": None, "threads_per_core": None,
            "logical_cpus": os.cpu_count() or 0}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "logical_cpus"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys and v.strip():
                v = v.strip()
                info[keys[k.strip()]] = int(v) if v.isdigit() else v
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except OSError:
        info["affinity_cpus"] = None
    return info


def _cpu_has_avx512():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("flags"):
                    f = set(line.split(":", 1)[1].split())
                    return {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"} <= f
    except OSError:
        pass
    return False


_CPU_TIMING = r"""
import json, os, sys, time
sys.path.insert(0, os.path.join(sys.argv[1], "oracle"))
sys.path.insert(0, os.path.join(sys.argv[1], "cuda-lbm-ibm-fvm_amd"))
import oracle, vof
n, prec, budget = int(sys.argv[2]), int(sys.argv[3]), float(sys.argv[4])
o = oracle.Oracle(vof.default_params(n=n, precision=prec, g=(0.0, 0.0, -9.81)))
o.init_box(0.04, 0.04, 0.06)
t0 = time.perf_counter()
o.step(1, 1)
per = max(time.perf_counter() - t0, 1e-6)
k = max(1, min(1000, int(budget / 3 / per)))
rates, istep = [], 2
for _ in range(3):
    t0 = time.perf_counter()
    o.step(istep, k)
    rates.append(n ** 3 * k / (time.perf_counter() - t0))
    istep += k
print(json.dumps({"rates": rates, "steps": k, "threads": oracle.threads()}))
"""


def cpu_baseline(n, prec, budget_s):
    """Time the oracle (the reference's kernel bodies as OpenMP loops) on the
    same workload, on this host's cores: one warm-up step, then three timed
    runs of as many steps as fit a third of the budget each; the median run is
    reported.  The timing build is oracle/liboracle_fast.so (-O3, x86-64-v4,
    contraction allowed; the parity build keeps one rounding per source
    operation) when the CPU has AVX-512, else the parity build.  It runs in a
    child process whose OpenMP threads are bound one per place with
    OMP_PROC_BIND=close, OMP_PLACES=cores; OMP_NUM_THREADS is the host's."""
    fast = os.path.join(ROOT, "oracle", "liboracle_fast.so")
    env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores")
    build = "parity build (-O3 -march=x86-64-v3 -ffp-contract=off)"
    if os.path.exists(fast) and _cpu_has_avx512():
        env["VOF_ORACLE_LIB"] = fast
        build = "timing build (-O3 -march=x86-64-v4, FMA contraction allowed)"
    r = subprocess.run([sys.executable, "-c", _CPU_TIMING, ROOT, str(n), str(prec), str(budget_s)], env=env,
                       capture_output=True, text=True, timeout=int(10 * budget_s) + 120)
    if r.returncode != 0:
        raise RuntimeError("cpu baseline child failed: " + r.stderr[-2000:])
    res = json.loads(r.stdout.strip().splitlines()[-1])
    rate = sorted(res["rates"])[1]
    host = _host_cpu()
    esz = 2 if prec == 8 else 1
    return {
        "value": rate,
        "unit": "cell-updates/s",
        "cores": res["threads"],
        "kind": "port",
        "gbs_algorithmic": rate * 236 * esz / 1e9,
        "host": host,
        "omp": {"OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS", "unset"), "OMP_PROC_BIND": "close",
                "OMP_PLACES": "cores"},
        "build": build,
        "sample": f"median of 3 runs of {res['steps']} steps (after 1 warm-up) of the same {n}^3 "
                  f"{'FP64' if prec == 8 else 'FP32'} dam break, OpenMP oracle restating voftrial1.cu's kernels "
                  f"on {res['threads']} threads",
    }


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (the hot path is a single-GPU stencil step); got --gpus {a.gpus}",
              file=sys.stderr)
        sys.exit(2)
    import vof

    vof.load_library()
    if a.study in ("precision", "long"):
        print(json.dumps(precision_study(vof, long_run=(a.study == "long"))), flush=True)
        return
    if a.study == "adjoint-sizes":
        print(json.dumps(adjoint_sizes(vof)), flush=True)
        return
    if a.calibrate_copy:
        vof.copy_bandwidth(a.calibrate_copy, 1)
    n, prec = a.n, a.precision
    esz = 8 if prec == 8 else 4
    cells = n ** 3
    s = make_solver(vof, n, prec, a.case)
    if a.jacobi_tblock or a.kz_jacobi or a.kz_predict:
        s.set_tuning(jacobi_tblock=a.jacobi_tblock, kz_jacobi=a.kz_jacobi, kz_predict=a.kz_predict)
    if not a.no_autotune:
        # launch shapes chosen by timing candidates on this GPU (state restored)
        s.autotune(1, 4, a.tune_cache)
    graph = not a.eager

    # warm-up (also captures the three graphs of the sweep rotation)
    s.step(1, a.warmup, graph=graph)
    s.sync()
    istep = 1 + a.warmup
    t0 = time.perf_counter()
    s.step(istep, a.steps, graph=graph)
    s.sync()
    wall = time.perf_counter() - t0
    istep += a.steps
    value = cells * a.steps / wall
    ms_step = 1e3 * wall / a.steps

    # per-stage device times from HIP events on the solver's stream
    st_steps = 3 * max(1, (a.stage_steps + 2) // 3)  # whole rotations of the sweep order
    _, stages = s.time_steps(istep, 0, st_steps, stages=True)
    istep += st_steps
    per_step = {k: v / st_steps for k, v in stages.items()}
    tuning = s.get_tuning()
    # With T sweeps per launch (jacobi_tblock T, temporal blocking: T-sweep
    # launches while T sweeps are left, then pairs, then one) a launch performs
    # on average sweeps / launches sweeps of 12 algorithmic B/cell each: the
    # "effective" figure of the roofline, beside the counter bytes.
    launches = dict(LAUNCHES_PER_STEP)
    tb, sweeps = tuning.get("jacobi_tblock", 1), LAUNCHES_PER_STEP["jacobi"]
    deep = sweeps // tb if tb >= 3 else 0
    left = sweeps - deep * tb
    launches["jacobi"] = deep + ((left // 2 + left % 2) if tb >= 2 else left)
    sweeps_per_launch = sweeps / launches["jacobi"]
    stage_bytes = dict(STAGE_BYTES, fct_zx=FCT_PAIR_BYTES, fct_xy=FCT_PAIR_BYTES)
    # FCT launches of the three sweep orders (:1384-1396), with fused pairs
    zx = tuning.get("fct_zx_fused", 1) == 2 and tuning.get("kz_fct_x") in (4, 8, 16)
    xy = tuning.get("fct_xy_fused", 1) == 2 and tuning.get("fct_chunk") in (4, 8, 16)
    orders = [["fct_xy", "fct_z"] if xy else ["fct_x", "fct_y", "fct_z"],
              ["fct_y", "fct_zx"] if zx else ["fct_y", "fct_z", "fct_x"],
              ["fct_zx", "fct_y"] if zx else ["fct_z", "fct_x", "fct_y"]]
    for k in ("fct_x", "fct_y", "fct_z", "fct_zx", "fct_xy"):
        launches[k] = sum(o.count(k) for o in orders) / 3
    kern = {}
    for k, b in stage_bytes.items():
        if launches[k] == 0:
            continue
        t_launch = per_step[k] / launches[k] * 1e-3
        if t_launch <= 0:
            continue
        byts = b * (sweeps_per_launch if k == "jacobi" else 1) * (esz // 4) * cells
        kern[k] = {"ms_per_launch": t_launch * 1e3, "bytes_per_launch": byts, "gbs": byts / t_launch / 1e9,
                   "frac": byts / t_launch / 1e9 / HBM_PEAK_GBS}
    kern["jacobi"]["sweeps_per_launch"] = sweeps_per_launch
    # what a launch must move whatever its sweeps: read p and rhs, write p
    kern["jacobi"]["compulsory_bytes_per_launch"] = STAGE_BYTES["jacobi"] * (esz // 4) * cells
    kern["jacobi"]["frac_compulsory"] = (kern["jacobi"]["compulsory_bytes_per_launch"] /
                                         (kern["jacobi"]["ms_per_launch"] * 1e-3) / 1e9 / HBM_PEAK_GBS)
    for k in ("fct_zx", "fct_xy"):
        if k in kern:
            kern[k]["sweeps_per_launch"] = 2
            kern[k]["compulsory_bytes_per_launch"] = FCT_PAIR_COMPULSORY_BYTES * (esz // 4) * cells
    # attainable time of each launch's compulsory stream mix at this field size
    field_bytes = (n + 2) ** 3 * esz
    stream_us = {}
    for k in kern:
        mix = STREAM_MIX[k]
        if mix not in stream_us:
            stream_us[mix] = (vof.stream_time(field_bytes, *mix, resident=False),
                              vof.stream_time(field_bytes, *mix, resident=True))
        t_hbm, t_res = stream_us[mix]
        kern[k]["stream_mix"] = f"{mix[0]} read, {mix[1]} written"
        kern[k]["attainable_us"] = t_hbm
        kern[k]["attainable_us_resident"] = t_res
        kern[k]["frac_attainable"] = t_hbm / (kern[k]["ms_per_launch"] * 1e3)
    dom = max(kern, key=lambda k: per_step[k])
    d = kern[dom]
    traffic, traffic_source = (load_traffic(dom, tuning) if (n == 200 and prec == 4 and a.case == "dam")
                               else (None, "counter bytes are profiled for the 200^3 FP32 dam break only"))
    step_bytes = sum(STAGE_BYTES[k] * LAUNCHES_PER_STEP[k] for k in STAGE_BYTES) * (esz // 4) * cells
    diag = s.diagnostics()
    pcie = transfer_leg(s, cells, ms_step)
    s.close()
    copy_hbm = vof.copy_bandwidth(1 << 30, 10)
    copy_mall = vof.copy_bandwidth(1 << 26, 50)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == 8 else "f32",
        "data": "synthetic",
        "config": {
            "workload": ((f"dam-break {n}^3 {'FP64' if prec == 8 else 'FP32'}, 0.1 m cube, liquid x<0.04 y<0.04 "
                          f"z<0.06 m") if a.case == "dam" else
                         (f"ripple layer {n}^3 {'FP64' if prec == 8 else 'FP32'}, 0.1 m cube, liquid below 0.06 m "
                          f"with 4 seeded plane waves")) +
                        f", rho 1000/50, g=(0,0,-9.81), dt=4e-6 s, 10 Jacobi sweeps/step, "
                        f"{'eager launches' if a.eager else 'hipGraph per step'}",
            "grid": [n, n, n],
            "interior_cells": cells,
            "jacobi_sweeps": 10,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": d["gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": d["frac"],
            "traffic": traffic,
            "traffic_source": traffic_source,
            "compulsory_bytes_per_launch": d.get("compulsory_bytes_per_launch"),
            "frac_compulsory": d.get("frac_compulsory"),
            "bytes_per_launch": d["bytes_per_launch"],
            "ms_per_launch": d["ms_per_launch"],
            "sweeps_per_launch": d.get("sweeps_per_launch"),
            "attainable_copy_gbs": {"hbm_1GiB": round(copy_hbm, 1), "mall_64MiB": round(copy_mall, 1)},
            "attainable": {"stream_mix": d["stream_mix"], "us_hbm": d["attainable_us"],
                           "us_resident": d["attainable_us_resident"], "frac": d["frac_attainable"],
                           "source": "vof_stream_bandwidth: plain 16 B/lane stream kernel over the same number of "
                                     "fields of the same size, inputs from HBM (us_resident: one set repeated)"},
        },
        "tuning": tuning,
        "step_hbm_frac": step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "stage_ms_per_step": {k: round(v, 5) for k, v in per_step.items()},
        "kernels": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in kern.items()},
        "final_liquid_volume": diag["liquid_volume"],
        "pcie": pcie,
    }
    if not a.no_batched:
        out["batched"] = batched_leg(vof)
    if not a.no_adjoint:
        out["adjoint"] = adjoint_leg(vof)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, prec, a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
