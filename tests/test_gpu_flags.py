"""GPU parity of the two extension flags against the CPU oracle.

* VOF_FLAG_NO_SLIP: the correction's closed-form wall conditions with
  negated tangential ghosts, bit for bit on every stored entry, signs of
  zeros included, then whole steps.
* VOF_FLAG_RHO_PREDICTOR: the density-weighted viscous predictor, stage and
  whole steps, alone, with CSF and with no-slip walls, and in an ensemble of
  uneven instances.
* Reverse mode refuses both (VOF_ERR_UNSUPPORTED).

The oracle's forms are pinned by closed forms in tests/test_oracle_pins.py
(no-slip ghost algebra, the discrete shear eigenmode, momentum conservation,
the steady two-layer shear, the one-fluid limit).
"""
import numpy as np
import pytest

from test_gpu_parity import SEED, assert_same, load_state, make_pair, random_state

pytestmark = pytest.mark.gpu

NO_SLIP, RHO_PRED, CSF = 0x8, 0x4, 0x2


def same_bits(a, b, what):
    assert_same(a, b, what)
    sa, sb = np.signbit(np.asarray(a)), np.signbit(np.asarray(b))
    bad = (sa != sb) & (np.asarray(a) == 0)
    assert not bad.any(), f"{what}: {int(bad.sum())} zeros of another sign, first {np.argwhere(bad)[:3].tolist()}"


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("grid", [(13, 11, 9), (70, 5, 7)])
def test_no_slip_correction_and_ghosts(vofmod, orc, prec, grid):
    nx, ny, nz = grid
    print(f"no-slip correction prec={prec} grid={grid}", flush=True)
    s, o = make_pair(vofmod, orc, nx=nx, ny=ny, nz=nz, precision=prec, dt=1e-3, flags=NO_SLIP)
    rng = np.random.default_rng(SEED + 7 * nx)
    st = random_state(s.shape, rng, s.dtype)
    load_state(s, o, st)
    s.debug_stage(vofmod.STAGE_PREDICT)
    for name in ("nu_rho", "advect_u", "advect_v", "advect_w"):
        o.stage(name)
    s.debug_stage(vofmod.STAGE_RHS)
    for _ in range(2):
        s.debug_stage(vofmod.STAGE_JACOBI)
        o.stage("jacobi")
    s.debug_stage(vofmod.STAGE_CORRECT)
    o.stage("update")
    o.stage("bc_vel")
    for f in ("U", "V", "W"):
        same_bits(s.download(f), o.get(f), f)
    s.close()
    o.close()


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("grid", [(13, 11, 9), (16, 16, 16), (70, 5, 7)])
def test_rho_predictor_stage(vofmod, orc, prec, grid):
    nx, ny, nz = grid
    print(f"rho predictor stage prec={prec} grid={grid}", flush=True)
    s, o = make_pair(vofmod, orc, nx=nx, ny=ny, nz=nz, precision=prec, dt=1e-3, flags=RHO_PRED)
    rng = np.random.default_rng(SEED + 11 * nx)
    st = random_state(s.shape, rng, s.dtype)
    load_state(s, o, st)
    s.debug_stage(vofmod.STAGE_PREDICT)
    for name in ("nu_rho", "advect_u", "advect_v", "advect_w"):
        o.stage(name)
    for f in ("USTAR", "VSTAR", "WSTAR"):
        assert_same(s.download(f), o.get(f), f)
    s.close()
    o.close()


@pytest.mark.parametrize("flags", [NO_SLIP, RHO_PRED, NO_SLIP | RHO_PRED, NO_SLIP | RHO_PRED | CSF])
@pytest.mark.parametrize("prec", [4, 8])
def test_whole_steps_with_flags(vofmod, orc, flags, prec):
    """Dam break, 9 steps (every sweep rotation three times), eager and graph,
    every stored entry of F, u, v, w, p equal to the oracle's."""
    print(f"whole steps flags={flags:#x} prec={prec}", flush=True)
    kw = dict(nx=24, ny=20, nz=22, precision=prec, g=(0.0, -5.0, -9.81), dt=2e-5, flags=flags, sigma=0.07)
    for graph in (False, True):
        s, o = make_pair(vofmod, orc, **kw)
        s.init_box(0.04, 0.05, 0.06)
        o.init_box(0.04, 0.05, 0.06)
        s.step(1, 9, graph=graph)
        o.step(1, 9)
        for f in ("F", "U", "V", "W", "P"):
            same_bits(s.download(f), o.get(f), f"{f} graph={graph}")
        s.close()
        o.close()


def test_flags_in_an_uneven_ensemble(vofmod, orc):
    """Three instances of different sizes with both flags, one launch per
    stage: each equals its own oracle run bit for bit."""
    print("ensemble with no-slip and rho predictor", flush=True)
    grids = [(16, 12, 14), (24, 24, 24), (10, 18, 12)]
    prms = [vofmod.default_params(nx=a, ny=b, nz=c, precision=4, g=(0.0, 0.0, -9.81), dt=2e-5,
                                  flags=NO_SLIP | RHO_PRED, rho_g=25.0 + 20 * n) for n, (a, b, c) in enumerate(grids)]
    bt = vofmod.Batch(prms)
    bt.init_box([0.04, 0.05, 0.03], [0.04, 0.03, 0.06], [0.06, 0.05, 0.04])
    bt.step(1, 6, graph=True)
    for n, prm in enumerate(prms):
        o = orc.Oracle(prm)
        o.init_box([0.04, 0.05, 0.03][n], [0.04, 0.03, 0.06][n], [0.06, 0.05, 0.04][n])
        o.step(1, 6)
        for f in ("F", "U", "V", "W", "P"):
            same_bits(bt.download(n, f), o.get(f), f"instance {n} {f}")
        o.close()
    bt.close()


@pytest.mark.parametrize("flags", [NO_SLIP, RHO_PRED])
def test_reverse_mode_refuses_the_extension_flags(vofmod, flags):
    s = vofmod.Solver(vofmod.default_params(n=12, precision=8, flags=flags))
    s.init_ic(1)
    with pytest.raises(vofmod.VofError) as e:
        s.adj_forward(1, 2, 1)
    assert e.value.status == 4
    s.close()
