"""Liquid-volume accounting of the VOF step on the CPU oracle.

Where the change of the liquid volume sum(F vol) over a run comes from:
per directional sweep, stage 1 conserves sum(Ftd dv) and stage 4 sum(F dv)
(telescoping fluxes; dv = vol - dt h^2 du is the sweep's dilated volume), so
sum(F vol) moves by
  clip1  the stage-1 clip of values outside [0, 1],
  clamp4 stage 4's clamp01 (its clipping part clip4, the rest the rounding of
         every F to the grid of 1 + F: 2^-23 in FP32, 2^-52 in FP64),
  dil    sum F (vol - dv): the dilation of a divergent velocity field (first
         order: dt vol sum F div u, 'div') plus the operator-splitting remainder,
and post_process_f's clamp01 adds 'post'.  What the terms do not account for
is the rounding of the sums themselves ('rest').

    python tests/conservation_study.py [n] [steps]

runs the ripple-layer workload (FP32 and FP64, 10 and 100 Jacobi sweeps per
step) and prints the table.  tests/test_conservation.py checks on a small
case that the terms add up to the measured change.
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "cuda-lbm-ibm-fvm_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

from oracle import Oracle  # noqa: E402
from vof import default_params  # noqa: E402
from vof_cases import ripple_F  # noqa: E402


def account(n, steps, prec, sweeps, init="ripple", **kw):
    o = Oracle(default_params(n=n, precision=prec, g=(0.0, 0.0, -9.81), jacobi_sweeps=sweeps, **kw))
    if init == "ripple":
        o.set("F", ripple_F(n))
    else:
        o.init_box(0.04, 0.04, 0.06)
    v0 = o.diagnostics()["liquid_volume"]
    o.stats_enable(True)
    o.step(1, steps)
    v1 = o.diagnostics()["liquid_volume"]
    st = o.stats()
    o.close()
    terms = st["clip1"] + st["clamp4"] + st["dil"] + st["post"]
    row = {k: v / v0 for k, v in st.items()}
    row.update(change=(v1 - v0) / v0, rest=(v1 - v0 - terms) / v0, round4=(st["clamp4"] - st["clip4"]) / v0,
               split=(st["dil"] - st["div"]) / v0)
    return row


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 96
    steps = int(sys.argv[2]) if len(sys.argv) > 2 else 500
    cols = ("change", "dil", "div", "split", "clip1", "clip4", "round4", "post", "rest")
    print(f"ripple layer {n}^3, {steps} steps; every term relative to the initial volume")
    print("| run | " + " | ".join(cols) + " |")
    print("|---" * (len(cols) + 1) + "|")
    for prec in (4, 8):
        for sweeps in (10, 100):
            r = account(n, steps, prec, sweeps)
            print(f"| FP{8 * prec} {sweeps} sweeps | " + " | ".join(f"{r[c]:+.3e}" for c in cols) + " |", flush=True)


if __name__ == "__main__":
    main()
