"""Writes the golden trajectories in this folder from the CPU oracle.

The reference ships no test vectors.  These fixtures freeze the oracle's
output on small grids (state after three steps, every field with its ghost
layer) so that both the oracle and the HIP library are held to the same bits
from one change to the next: tests/test_golden.py (oracle, CPU) and
tests/test_gpu_golden.py (library, GPU).  Rerun only on a deliberate change of
the arithmetic, and say why in the commit.

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "cuda-lbm-ibm-fvm_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

SEED = 20261203
FIELDS = ("F", "U", "V", "W", "P")


def cases():
    """name -> (parameter overrides, initial-state builder, first step, steps)."""
    def box(o):
        o.init_box(0.04, 0.05, 0.06)

    def random_state(o):
        rng = np.random.default_rng(SEED)
        sh = o.shape
        F = rng.random(sh)
        F[F < 0.35] = 0.0
        F[F > 0.8] = 1.0
        st = dict(F=F, U=rng.uniform(-1.0, 1.0, sh), V=rng.uniform(-1.0, 1.0, sh), W=rng.uniform(-1.0, 1.0, sh),
                  P=np.zeros(sh))
        for k, a in st.items():
            o.set(k, a.astype(o.dtype))

    def sphere(o):
        from test_csf import sphere_F

        rng = np.random.default_rng(SEED + 1)
        F = sphere_F(12, (0.045, 0.055, 0.05), 0.027)
        o.set("F", F.astype(o.dtype))
        for k in "UVW":
            o.set(k, rng.normal(0.0, 0.05, F.shape).astype(o.dtype))

    g = (0.0, 0.0, -9.81)
    return {
        "dam_break_fp32": (dict(nx=12, ny=10, nz=14, precision=4, g=g, dt=1e-3), box, 1, 3),
        "dam_break_fp64": (dict(nx=12, ny=10, nz=14, precision=8, g=g, dt=1e-3), box, 1, 3),
        "reference_ic_fp32": (dict(n=11, precision=4, dt=1e-3), lambda o: o.init_ic(1), 1, 3),
        "random_fp32": (dict(nx=13, ny=11, nz=9, precision=4, dt=2e-4), random_state, 2, 3),
        "csf_fp64": (dict(n=12, precision=8, g=g, dt=1e-4, flags=0x2, sigma=0.07, rho_l=1.0, rho_g=1.0), sphere, 1, 3),
    }


def run_oracle(orc, vof, kw, init, first, steps):
    o = orc.Oracle(vof.default_params(**kw))
    init(o)
    o.step(first, steps)
    out = {k: o.get(k).copy() for k in FIELDS}
    o.close()
    return out


def main():
    import oracle as orc
    import vof

    orc.load()
    meta = {}
    for name, (kw, init, first, steps) in cases().items():
        out = run_oracle(orc, vof, kw, init, first, steps)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        meta[name] = {"params": {k: list(v) if isinstance(v, tuple) else v for k, v in kw.items()},
                      "first_istep": first, "steps": steps}
        print(name, {k: a.shape for k, a in out.items()})
    with open(os.path.join(HERE, "cases.json"), "w") as fh:
        json.dump(meta, fh, indent=1)


if __name__ == "__main__":
    main()
