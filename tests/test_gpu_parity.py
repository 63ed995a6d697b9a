"""GPU parity: the HIP library against the CPU oracle, through the C ABI.

Bar: exact equality (IEEE ==, so +0 == -0) of every field the reference
exposes, stage by stage and after whole steps, in FP32 and FP64.  Both sides
are built without FMA contraction and keep the reference's operation order.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 20261203
# range of the reference clamp01f in FP32 (it returns the rounding error of 1+v)
F32_LO, F32_HI = -(2.0 ** -25), 1.0 + 2.0 ** -23


def random_state(shape, rng, dtype, vel=1.0):
    """Random volume fraction (many exact 0 and 1, interface values between)
    and random face velocities / pressure; p ghosts zero as in the reference."""
    F = rng.random(shape)
    F[F < 0.35] = 0.0
    F[F > 0.8] = 1.0
    u = rng.uniform(-vel, vel, shape)
    v = rng.uniform(-vel, vel, shape)
    w = rng.uniform(-vel, vel, shape)
    p = rng.normal(0.0, 50.0, shape)
    p[0, :, :] = p[-1, :, :] = 0
    p[:, 0, :] = p[:, -1, :] = 0
    p[:, :, 0] = p[:, :, -1] = 0
    return {k: a.astype(dtype) for k, a in dict(F=F, U=u, V=v, W=w, P=p).items()}


def interior(a):
    return a[1:-1, 1:-1, 1:-1]


def assert_same(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, what
    bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
    if bad.any():
        idx = np.argwhere(bad)[:5]
        diff = np.max(np.abs(a.astype(np.float64) - b.astype(np.float64))[bad])
        raise AssertionError(f"{what}: {bad.sum()} of {a.size} differ (max |d| {diff:.3e}); first {idx.tolist()}"
                             f" gpu {a[tuple(idx[0])]!r} oracle {b[tuple(idx[0])]!r}")


def make_pair(vofmod, orc, **kw):
    prm = vofmod.default_params(**kw)
    s = vofmod.Solver(prm)
    o = orc.Oracle(vofmod.default_params(**kw))
    return s, o


def load_state(s, o, st):
    for k, a in st.items():
        s.upload(k, a)
        o.set(k, a)


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("grid", [(13, 11, 9), (16, 16, 16), (70, 5, 7)])
def test_stages(vofmod, orc, prec, grid):
    nx, ny, nz = grid
    print(f"stage parity prec={prec} grid={grid}", flush=True)
    s, o = make_pair(vofmod, orc, nx=nx, ny=ny, nz=nz, precision=prec, dt=1e-3)
    rng = np.random.default_rng(SEED + nx)
    st = random_state(s.shape, rng, s.dtype)
    load_state(s, o, st)
    # predictor
    s.debug_stage(vofmod.STAGE_PREDICT)
    for name in ("nu_rho", "advect_u", "advect_v", "advect_w"):
        o.stage(name)
    for f in ("USTAR", "VSTAR", "WSTAR"):
        assert_same(s.download(f), o.get(f), f)
    # rhs (formed once; the oracle forms it inside each sweep)
    s.debug_stage(vofmod.STAGE_RHS)
    assert_same(interior(s.download("RHS")), interior(o.get("RHS")), "RHS")
    # Jacobi sweeps: single-sweep launches, then two-sweep (temporally blocked) launches
    for it in range(3):
        s.debug_stage(vofmod.STAGE_JACOBI)
        o.stage("jacobi")
        assert_same(s.download("P"), o.get("P"), f"P sweep {it}")
    for it in range(2):
        s.debug_stage(vofmod.STAGE_JACOBI2)
        o.stage("jacobi")
        o.stage("jacobi")
        assert_same(s.download("P"), o.get("P"), f"P double sweep {it}")
    # correction + velocity BC, all stored entries
    s.debug_stage(vofmod.STAGE_CORRECT)
    o.stage("update")
    o.stage("bc_vel")
    for f in ("U", "V", "W"):
        assert_same(s.download(f), o.get(f), f)
    # FCT sweeps in a rotated order, then clamp + F BC
    for gs, on in ((vofmod.STAGE_FCT_Y, "fct_y"), (vofmod.STAGE_FCT_Z, "fct_z"), (vofmod.STAGE_FCT_X, "fct_x")):
        s.debug_stage(gs)
        o.stage(on)
        assert_same(interior(s.download("F")), interior(o.get("F")), on)
    s.debug_stage(vofmod.STAGE_GHOST_F)
    o.stage("clamp")
    o.stage("bc_F")
    assert_same(s.download("F"), o.get("F"), "F with ghosts")
    s.close()
    o.close()


@pytest.mark.parametrize("kz", [4, 8, 16])
@pytest.mark.parametrize("grid", [(70, 20, 9), (5, 4, 3), (131, 36, 17), (62, 8, 5), (61, 22, 10), (1, 4, 1)])
def test_predictor_row_bands(vofmod, orc, grid, kz):
    """FP32 predictor with bands of 4 rows per lane (chunks of 4 or 8 planes,
    ny a multiple of 4: one band of 4 rows, several, positions wrapping from
    one band row into the next, a one-cell-wide grid) and the one-row mapping
    it falls back to (ny 22, chunks of 16), against the oracle's u*, v*, w*."""
    nx, ny, nz = grid
    print(f"predictor bands grid={grid} kz={kz}", flush=True)
    s, o = make_pair(vofmod, orc, nx=nx, ny=ny, nz=nz, precision=4, dt=1e-3)
    s.set_tuning(kz_predict=kz)
    rng = np.random.default_rng(SEED + 31 + nx)
    st = random_state(s.shape, rng, s.dtype)
    load_state(s, o, st)
    s.debug_stage(vofmod.STAGE_PREDICT)
    for name in ("nu_rho", "advect_u", "advect_v", "advect_w"):
        o.stage(name)
    for f in ("USTAR", "VSTAR", "WSTAR"):
        assert_same(s.download(f), o.get(f), f)
    s.close()
    o.close()


@pytest.mark.parametrize("prec", [4, 8])
def test_fct_each_direction_random(vofmod, orc, prec):
    """Each sweep direction alone on a random interface field with a large CFL."""
    for gs, on in ((vofmod.STAGE_FCT_X, "fct_x"), (vofmod.STAGE_FCT_Y, "fct_y"), (vofmod.STAGE_FCT_Z, "fct_z")):
        s, o = make_pair(vofmod, orc, nx=37, ny=29, nz=23, precision=prec, dt=2e-3)
        rng = np.random.default_rng(SEED + 7)
        st = random_state(s.shape, rng, s.dtype, vel=1.5)
        load_state(s, o, st)
        s.debug_stage(gs)
        o.stage(on)
        assert_same(interior(s.download("F")), interior(o.get("F")), on)
        s.close()
        o.close()


def run_steps_pair(vofmod, orc, steps, first=1, graph=False, ic=1, box=None, **kw):
    s, o = make_pair(vofmod, orc, **kw)
    if box:
        s.init_box(*box)
        o.init_box(*box)
    else:
        s.init_ic(ic)
        o.init_ic(ic)
    assert_same(s.download("F"), o.get("F"), "initial F")
    s.step(first, steps, graph=graph)
    o.step(first, steps)
    return s, o


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("shape", [3, 4, 8, 16, 32])
def test_launch_shapes_match_oracle(vofmod, orc, prec, shape):
    """Every kernel variant behind the launch shapes: the Jacobi z-chunk
    (3: streaming march; 4, 8, 16: operands preloaded) and the y/z FCT chunk
    (3: streaming march; 8, 16, 32: register-blocked lines), on a grid whose
    sizes leave partial chunks and with a large CFL on a random interface."""
    s, o = make_pair(vofmod, orc, nx=37, ny=45, nz=27, precision=prec, dt=2e-3)
    s.set_tuning(kz_jacobi=min(shape, 16), fct_chunk=shape, jacobi_tblock=1)
    rng = np.random.default_rng(SEED + 13 + shape)
    st = random_state(s.shape, rng, s.dtype, vel=1.5)
    load_state(s, o, st)
    s.debug_stage(vofmod.STAGE_PREDICT)
    s.debug_stage(vofmod.STAGE_RHS)
    for name in ("nu_rho", "advect_u", "advect_v", "advect_w"):
        o.stage(name)
    for it in range(2):
        s.debug_stage(vofmod.STAGE_JACOBI)
        o.stage("jacobi")
        assert_same(s.download("P"), o.get("P"), f"P sweep {it}")
    for gs, on in ((vofmod.STAGE_FCT_Y, "fct_y"), (vofmod.STAGE_FCT_Z, "fct_z")):
        load_state(s, o, st)
        s.debug_stage(gs)
        o.stage(on)
        assert_same(interior(s.download("F")), interior(o.get("F")), on)
    s.close()
    o.close()
    # whole steps (the last sweep of each step also writes the F ghosts)
    s, o = make_pair(vofmod, orc, nx=37, ny=45, nz=27, precision=prec, dt=2e-3)
    s.set_tuning(kz_jacobi=min(shape, 16), fct_chunk=shape)
    load_state(s, o, st)
    s.step(1, 3)
    o.step(1, 3)
    for k in ("F", "U", "V", "W"):
        assert_same(s.download(k), o.get(k), f"{k} after 3 steps")
    assert_same(interior(s.download("P")), interior(o.get("P")), "P after 3 steps")
    s.close()
    o.close()


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("kz", [4, 8, 16])
@pytest.mark.parametrize("fused", [1, 2])
def test_fct_pairs_fused_match_oracle(vofmod, orc, prec, kz, fused):
    """FCT sweep pairs in one launch (fused 2) or two (fused 1): z then x in
    the orders y-z-x and z-x-y (chunks of kz planes), x then y in x-y-z
    (chunks of kz rows).  Whole steps bit for bit, F ghosts included, on a grid
    with partial chunks, a large CFL and random F in the ghost layer (which a
    fused launch never reads for its second sweep).  Steps 1..4 cover z+x as
    the last sweeps of a step (order y-z-x, which also clamps twice and writes
    the ghosts), z+x as the first (z-x-y) and x+y as the first (x-y-z); eager
    and graph replay."""
    print(f"fct pairs prec={prec} chunk={kz} fused={fused}", flush=True)
    rng = np.random.default_rng(SEED + 17 + kz)
    for graph in (False, True):
        s, o = make_pair(vofmod, orc, nx=37, ny=45, nz=27, precision=prec, dt=2e-3)
        s.set_tuning(kz_fct_x=kz, fct_chunk=kz, fct_zx_fused=fused, fct_xy_fused=fused)
        t = s.get_tuning()
        assert (t["fct_zx_fused"], t["fct_xy_fused"]) == (fused, fused)
        st = random_state(s.shape, rng, s.dtype, vel=1.5)
        load_state(s, o, st)
        for step in range(1, 5):
            s.step(step, 1, graph=graph)
            o.step(step, 1)
            for k in ("F", "U", "V", "W"):
                assert_same(s.download(k), o.get(k), f"{k} after step {step} (graph {graph})")
        s.close()
        o.close()


def _ghost_handoffs(nx, ny, nz):
    """Where the last sweep's i ghosts change writer at a wave edge: the fused
    z+x launch (waves of 64 lanes over j*(nx+2)+i storing lanes 3..60) and the
    y and z line kernels (waves of 64 lanes over (o-1)*(nx+2)+i)."""
    sx = nx + 2
    return {
        "zx: ghost i=0 on lane 60": any((j * sx) % 58 == 57 for j in range(1, ny + 1)),
        "zx: ghost i=nx+1 on lane 3": any((j * sx + nx + 1) % 58 == 0 for j in range(1, ny + 1)),
        "y line: ghost i=0 on lane 63": any(((k - 1) * sx) % 64 == 63 for k in range(1, nz + 1)),
        "y line: ghost i=nx+1 on lane 0": any(((k - 1) * sx + nx + 1) % 64 == 0 for k in range(1, nz + 1)),
        "z line: ghost i=0 on lane 63": any(((j - 1) * sx) % 64 == 63 for j in range(1, ny + 1)),
        "z line: ghost i=nx+1 on lane 0": any(((j - 1) * sx + nx + 1) % 64 == 0 for j in range(1, ny + 1)),
    }


@pytest.mark.parametrize("prec", [4, 8])
def test_last_sweep_ghosts_at_wave_edges(vofmod, orc, prec):
    """The last FCT sweep of a step writes the i ghosts from the ghost-column
    lanes (a lane shift of the neighbour's value, stored with the row), except
    where the neighbour is in another wave or is a halo lane, where the
    interior lane writes them.  On a 17x54x38 grid every such hand-off occurs
    (checked here); three steps put each last-sweep kernel (z line, fused z+x,
    y line) last once.  Every F value, ghosts included, bit for bit."""
    nx, ny, nz = 17, 54, 38
    cases = _ghost_handoffs(nx, ny, nz)
    assert all(cases.values()), cases
    print(f"last-sweep ghosts at wave edges prec={prec}", flush=True)
    s, o = make_pair(vofmod, orc, nx=nx, ny=ny, nz=nz, precision=prec, dt=2e-3)
    s.set_tuning(kz_fct_x=8, fct_chunk=8, fct_zx_fused=2, fct_xy_fused=2)
    rng = np.random.default_rng(SEED + 23)
    st = random_state(s.shape, rng, s.dtype, vel=1.5)
    load_state(s, o, st)
    for step in range(3, 6):  # sweep orders x-y-z, y-z-x, z-x-y
        s.step(step, 1)
        o.step(step, 1)
        assert_same(s.download("F"), o.get("F"), f"F after step {step}")
    s.close()
    o.close()


@pytest.mark.parametrize("grid", [(1, 7, 5), (3, 1, 6), (2, 2, 2), (5, 4, 1), (1, 1, 1), (63, 2, 3), (64, 3, 2)])
def test_degenerate_grids(vofmod, orc, grid):
    """Grids one to three cells thick in some direction (a lane that is both
    the first and the last interior cell of its row, planes that are both the
    first and the last) and rows of 63/64 cells (ghost columns at wave edges):
    three whole steps, one per sweep order, with the default launch shapes
    (fused FCT pairs, two-sweep Jacobi), every field bit for bit."""
    nx, ny, nz = grid
    print(f"degenerate grid {grid}", flush=True)
    s, o = make_pair(vofmod, orc, nx=nx, ny=ny, nz=nz, precision=4, dt=2e-3)
    rng = np.random.default_rng(SEED + 29)
    st = random_state(s.shape, rng, s.dtype, vel=1.5)
    load_state(s, o, st)
    for step in range(1, 4):
        s.step(step, 1, graph=(step == 2))
        o.step(step, 1)
        for k in ("F", "U", "V", "W"):
            assert_same(s.download(k), o.get(k), f"{k} after step {step}")
        assert_same(interior(s.download("P")), interior(o.get("P")), f"P after step {step}")
    s.close()
    o.close()


def _jacobi2_pair(vofmod, orc, kz, prec, nx=131, ny=33, nz=21):
    print(f"two-sweep jacobi kz={kz} prec={prec} grid={(nx, ny, nz)}", flush=True)
    s, o = make_pair(vofmod, orc, nx=nx, ny=ny, nz=nz, precision=prec, dt=1e-3)
    s.set_tuning(kz_jacobi=kz)
    rng = np.random.default_rng(SEED + 11)
    st = random_state(s.shape, rng, s.dtype)
    load_state(s, o, st)
    s.debug_stage(vofmod.STAGE_PREDICT)
    s.debug_stage(vofmod.STAGE_RHS)
    for name in ("nu_rho", "advect_u", "advect_v", "advect_w"):
        o.stage(name)
    for it in range(3):
        s.debug_stage(vofmod.STAGE_JACOBI2)
        o.stage("jacobi")
        o.stage("jacobi")
        assert_same(s.download("P"), o.get("P"), f"P double sweep {it}")
    s.close()
    o.close()


@pytest.mark.parametrize("nz", [25, 26, 27, 51, 60])
def test_jacobi2_streamed_chunk_kinds(vofmod, orc, nz):
    """The FP32 two-sweep kernel streamed through LDS (25-plane chunks) in each
    of its marches: interior chunks, the first chunk (k0 = 1) and the last one
    (ending at nz) with their boundary planes at fixed steps, and the chunks
    that fit neither (nz 25 and 26; the chunk k0 = 26 of nz 51, which ends at
    nz - 1); two x tiles and three row bands."""
    _jacobi2_pair(vofmod, orc, 25, 4, nx=70, ny=30, nz=nz)


@pytest.mark.parametrize("nz", [20, 21, 22, 41, 50])
def test_jacobi2_streamed_short_chunks(vofmod, orc, nz):
    """The same kernel with 20-plane chunks, each chunk kind, on a grid whose
    flattened band positions wrap inside tiles (nx 61, ny 23)."""
    _jacobi2_pair(vofmod, orc, 20, 4, nx=61, ny=23, nz=nz)


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("kz", [1, 2, 3, 4, 5, 8, 16, 25])
def test_jacobi2_tiles_and_chunks(vofmod, orc, kz, prec):
    """Two-sweep launches on a grid with several x tiles, row bands and z chunks
    (partial last chunks; 25 is one chunk over all 21 planes).  FP32 with 4, 8
    and 16 preloads the chunk, other lengths and FP64 stream the planes through
    registers."""
    _jacobi2_pair(vofmod, orc, kz, prec)


@pytest.mark.parametrize("grid", [(70, 20, 25), (125, 29, 26), (62, 15, 31), (70, 20, 60), (131, 33, 81)])
def test_jacobi2_lds_streamed_chunks(vofmod, orc, grid):
    """FP32 two sweeps streamed through LDS requests (chunks of 25 planes, the
    last one shifted back to end at nz): one chunk that holds the first and
    the last plane (nz 25), two overlapping chunks (26, 31), and columns with
    interior chunks between them (60, 81), with one and several x tiles and
    row bands."""
    _jacobi2_pair(vofmod, orc, 25, 4, *grid)


def _jacobiT_pair(vofmod, orc, T, kz, prec, nx, ny, nz, launches=2):
    print(f"{T}-sweep jacobi kz={kz} prec={prec} grid={(nx, ny, nz)}", flush=True)
    s, o = make_pair(vofmod, orc, nx=nx, ny=ny, nz=nz, precision=prec, dt=1e-3)
    s.set_tuning(kz_jacobi=kz, jacobi_tblock=T)
    rng = np.random.default_rng(SEED + 12)
    st = random_state(s.shape, rng, s.dtype)
    load_state(s, o, st)
    s.debug_stage(vofmod.STAGE_PREDICT)
    s.debug_stage(vofmod.STAGE_RHS)
    for name in ("nu_rho", "advect_u", "advect_v", "advect_w"):
        o.stage(name)
    for it in range(launches):
        s.debug_stage(vofmod.STAGE_JACOBI_T)
        for _ in range(T):
            o.stage("jacobi")
        assert_same(s.download("P"), o.get("P"), f"P after {T}-sweep launch {it}")
    s.close()
    o.close()


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("kz", [1, 3, 8, 25])
@pytest.mark.parametrize("T", [3, 4, 5])
def test_jacobiT_tiles_and_chunks(vofmod, orc, T, kz, prec):
    """T sweeps per launch (T = 3, 4, 5) on a grid with several x tiles, row
    bands and z chunks (partial last chunks; 25 is one chunk over all 21
    planes), two launches in a row: bitwise equal to 2T single sweeps."""
    _jacobiT_pair(vofmod, orc, T, kz, prec, 131, 33, 21)


@pytest.mark.parametrize("grid", [(7, 3, 2), (1, 1, 1), (58, 12, 4), (3, 50, 9), (57, 11, 3)])
@pytest.mark.parametrize("T", [3, 5])
def test_jacobiT_small_grids(vofmod, orc, T, grid):
    """T-sweep launches on grids thinner than the sweep halo: fewer planes than
    sweeps, one cell, one exact tile (58 x 12 for T = 3), one row band of
    many rows."""
    _jacobiT_pair(vofmod, orc, T, 4, 4, *grid)


def test_jacobi_tblock_range(vofmod):
    s = vofmod.Solver(n=12)
    with pytest.raises(vofmod.VofError, match="sweeps per Jacobi launch must be 1 .. 5"):
        s.set_tuning(jacobi_tblock=6)
    assert s.get_tuning()["jacobi_tblock"] != 6
    s.set_tuning(jacobi_tblock=2)
    with pytest.raises(vofmod.VofError, match="stage 10 needs jacobi_tblock 3"):
        s.debug_stage(vofmod.STAGE_JACOBI_T)
    s.close()


@pytest.mark.parametrize("tblock", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("prec", [4, 8])
def test_whole_steps_reference_case(vofmod, orc, prec, tblock):
    print(f"whole steps prec={prec} n=24 tblock={tblock}", flush=True)
    s, o = run_steps_pair(vofmod, orc, 30, n=24, precision=prec, jacobi_tblock=tblock)
    for f in ("F", "U", "V", "W", "P", "USTAR", "VSTAR", "WSTAR"):
        assert_same(s.download(f), o.get(f), f)
    d1, d2 = s.diagnostics(), o.diagnostics()
    for k in ("liquid_volume", "p_mean", "residual_l2", "rhs_l2"):
        assert d1[k] == pytest.approx(d2[k], rel=1e-12, abs=1e-300), k
    assert d1["f_min"] == d2["f_min"] and d1["f_max"] == d2["f_max"]
    assert d1["nonfinite"] == 0


def test_whole_steps_config_box_gravity_noncubic(vofmod, orc):
    """Non-cubic grid, extension box IC and gravity along -z (config style)."""
    s, o = run_steps_pair(vofmod, orc, 25, nx=30, ny=22, nz=26, precision=4, g=(0.0, 0.0, -9.81), dt=2e-5,
                          box=(0.04, 0.04, 0.06))
    for f in ("F", "U", "V", "W", "P"):
        assert_same(s.download(f), o.get(f), f)


def test_whole_steps_random_velocity_large_dt(vofmod, orc):
    """Random initial velocities and a large time step drive every FCT branch
    (both upwind signs, active limiter, clipping) through whole steps."""
    s, o = make_pair(vofmod, orc, nx=28, ny=20, nz=24, precision=4, dt=2e-4)
    rng = np.random.default_rng(SEED + 3)
    st = random_state(s.shape, rng, s.dtype, vel=0.8)
    st["P"][:] = 0
    load_state(s, o, st)
    s.step(1, 12)
    o.step(1, 12)
    for f in ("F", "U", "V", "W", "P"):
        assert_same(s.download(f), o.get(f), f)
    assert float(interior(s.download("F")).min()) >= F32_LO and float(interior(s.download("F")).max()) <= F32_HI


FUSED = dict(jacobi_tblock=2, kz_jacobi=8, kz_fct_x=8, fct_chunk=8, fct_zx_fused=2, fct_xy_fused=2)


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("tuning", [None, FUSED], ids=["default", "fused"])
def test_whole_steps_ripple_layer(vofmod, orc, prec, tuning):
    """Liquid-rich state: the ripple layer (liquid below a surface of four
    seeded plane waves) at 48^3 with a large time step, six steps (each sweep
    order twice), with the default and the fused launch shapes.  Below the
    surface whole waves have no antidiffusion and take the FCT's stage-4
    shortcut; across it they do the limiter arithmetic.  F is compared bit for
    bit (zero signs included), the other fields by value."""
    from vof_cases import ripple_F
    n = 48
    print(f"ripple layer {n}^3 prec={prec} tuning={tuning}", flush=True)
    s, o = make_pair(vofmod, orc, n=n, precision=prec, g=(0.0, 0.0, -9.81), dt=2e-4)
    if tuning:
        s.set_tuning(**tuning)
    F = ripple_F(n, dtype=s.dtype)
    s.upload("F", F)
    o.set("F", F)
    for step in range(1, 7):
        s.step(step, 1, graph=(step > 3))
        o.step(step, 1)
        a, b = s.download("F"), o.get("F")
        ui = np.uint64 if a.dtype == np.float64 else np.uint32
        assert np.array_equal(a.view(ui), b.view(ui)), f"F bits after step {step}"
        for f in ("U", "V", "W"):
            assert_same(s.download(f), o.get(f), f"{f} after step {step}")
        assert_same(interior(s.download("P")), interior(o.get("P")), f"P after step {step}")
    s.close()
    o.close()


def test_graph_equals_eager(vofmod):
    print("graph == eager", flush=True)
    a = vofmod.Solver(n=20)
    b = vofmod.Solver(n=20)
    a.init_ic(1)
    b.init_ic(1)
    a.step(1, 14, graph=False)
    b.step(1, 14, graph=True)
    for f in ("F", "U", "V", "W", "P"):
        assert_same(a.download(f), b.download(f), f)
    # eager work between replays, replays again
    a.step(15, 3)
    b.step(15, 3, graph=True)
    assert_same(a.download("F"), b.download("F"), "F after more replays")


def test_remove_p_mean(vofmod, orc):
    flags = vofmod.FLAG_REMOVE_P_MEAN
    s, o = run_steps_pair(vofmod, orc, 9, n=18, precision=8, flags=flags)
    for f in ("F", "U", "V", "W"):
        np.testing.assert_allclose(s.download(f), o.get(f), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(s.download("P"), o.get("P"), rtol=1e-9, atol=1e-9)
    assert abs(s.diagnostics()["p_mean"]) < 1e-9


def test_watchdog(vofmod):
    """VOF_FLAG_WATCHDOG: a healthy run passes; a NaN planted in u (from the
    host) reaches the checked fields within one step and vof_step returns
    VOF_ERR_NUMERICS with the count in the message."""
    s = vofmod.Solver(vofmod.default_params(n=24, precision=4, g=(0.0, 0.0, -9.81), dt=2e-5,
                                            flags=vofmod.FLAG_WATCHDOG))
    s.init_box(0.04, 0.04, 0.06)
    s.step(1, 20, graph=True)
    u = s.download("U")
    u[12, 12, 12] = np.nan
    s.upload("U", u)
    with pytest.raises(vofmod.VofError, match="watchdog: instance 0 has [1-9][0-9]* non-finite") as e:
        s.step(21, 1)
    assert e.value.status == 5
    s.close()


@pytest.mark.parametrize("tuning", [None, dict(jacobi_tblock=2, kz_jacobi=8, kz_correct=8, kz_rhs=8, kz_fct_x=8),
                                    dict(jacobi_tblock=2, kz_jacobi=3, fct_chunk=16, kz_fct_x=16, kz_predict=4),
                                    dict(jacobi_tblock=4, kz_jacobi=5), dict(jacobi_tblock=3, kz_jacobi=8)])
def test_batch_equals_plain(vofmod, tuning):
    """Instances of a ragged batch reproduce plain runs bit for bit, also when
    the batch runs other launch shapes (two-sweep Jacobi launches, preloaded or
    streamed, other chunk lengths) than the plain runs' defaults."""
    print(f"batch vs plain, batch tuning {tuning}", flush=True)
    sizes = [(12, 12, 12), (20, 16, 18), (9, 14, 11)]
    prms = [vofmod.default_params(nx=a, ny=b, nz=c, g=(0, 0, -9.81), rho_g=25.0 + 10 * i, dt=2e-5)
            for i, (a, b, c) in enumerate(sizes)]
    boxes = [(0.04, 0.05, 0.06), (0.03, 0.04, 0.07), (0.05, 0.02, 0.05)]
    bt = vofmod.Batch(prms)
    if tuning:
        bt.set_tuning(**tuning)
    bt.init_box(*zip(*boxes))
    bt.step(1, 11, graph=True)
    for i, prm in enumerate(prms):
        s = vofmod.Solver(vofmod.default_params(nx=prm.nx, ny=prm.ny, nz=prm.nz, g=(0, 0, -9.81),
                                                rho_g=prm.rho_g, dt=2e-5))
        s.init_box(*boxes[i])
        s.step(1, 11)
        for f in ("F", "U", "V", "W", "P"):
            assert_same(bt.download(i, f), s.download(f), f"inst {i} {f}")
        s.close()


@pytest.mark.parametrize("tblock", [1, 2, 5])
def test_full_size_200_parity_three_steps(vofmod, orc, tblock):
    """Reference default case at full size: bitwise after one step of each rotation."""
    print(f"200^3 parity, 3 steps, tblock={tblock}", flush=True)
    s, o = run_steps_pair(vofmod, orc, 3, n=200, precision=4, jacobi_tblock=tblock)
    for f in ("F", "U", "V", "W", "P"):
        assert_same(s.download(f), o.get(f), f)


def test_autotune_keeps_state_and_caches(vofmod, tmp_path):
    """Tuning trials run on a snapshot: the run continues bit-identically, and a
    second solver of the same shape picks the cached choice."""
    a = vofmod.Solver(n=24)
    b = vofmod.Solver(n=24)
    a.init_ic(1)
    b.init_ic(1)
    a.step(1, 5)
    b.step(1, 5)
    cache = tmp_path / "tune.json"
    t = a.autotune(6, 2, str(cache))
    assert cache.exists() and "24x24x24" in cache.read_text()
    a.step(6, 7, graph=True)
    b.step(6, 7)
    for f in ("F", "U", "V", "W", "P"):
        assert_same(a.download(f), b.download(f), f)
    c = vofmod.Solver(n=24)
    assert c.autotune(1, 2, str(cache)) == t
    bt = vofmod.Batch([vofmod.default_params(n=12), vofmod.default_params(n=17)])
    bt.init_ic(1)
    tb = bt.autotune(1, 2, None)
    assert tb["fct_chunk"] > 0


def test_full_size_200_invariants(vofmod):
    """200^3, 300 steps: F stays in [0,1], no NaN, volume drift bounded."""
    s = vofmod.Solver(n=200)
    s.init_ic(1)
    v0 = s.diagnostics()["liquid_volume"]
    s.step(1, 300, graph=True)
    d = s.diagnostics()
    assert d["nonfinite"] == 0
    assert F32_LO <= d["f_min"] and d["f_max"] <= F32_HI
    assert abs(d["liquid_volume"] - v0) / v0 < 1e-3
    print("200^3 volume drift", (d["liquid_volume"] - v0) / v0, flush=True)


def test_vtk_from_gpu_matches_oracle_writer(vofmod, orc, tmp_path):
    s = vofmod.Solver(n=10)
    s.init_ic(1)
    s.step(1, 5)
    a = tmp_path / "gpu.vtk"
    b = tmp_path / "orc.vtk"
    s.write_vtk(str(a))
    orc.write_vtk_float(str(b), s.download("F"))
    assert a.read_bytes() == b.read_bytes()
