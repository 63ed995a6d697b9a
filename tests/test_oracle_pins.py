"""Pins of the CPU oracle that fix each stage against the mathematics — CPU only.

The reference ships no tests or vectors, so every restated function must be
fixed by something the mathematics decides without the oracle's own code:

* hand-derived known answers of the FCT sweep on dyadic inputs (every value
  exact in binary, so FP32 and FP64 must give these bits);
* axis-rotation equivariance: on a cubic grid the y and z operators of the
  predictor, the Jacobi right-hand side and sweep, the correction and the FCT
  are the x operator with the axes renamed, so their outputs on a rotated
  state are the rotated outputs (to the round-off of the different summation
  order);
* closed forms on non-cubic grids (linear fields), where an axis mix-up
  (dxi for dyi, the wrong face index) changes the value;
* the interface-sharpening property of the limited downwind anti-diffusion;
* the exact liquid volume of the initial condition.

Each check is chosen so that a named plausible mistake fails it;
tests/test_oracle_mutants.py plants those mistakes in copies of the oracle and
asserts that these pins catch every one.
"""
import numpy as np
import pytest

from oracle import Oracle
from vof import default_params

f32 = np.float32


def rot(a):
    """Rename the axes x -> y -> z -> x of a (z, y, x) array: the result B has
    B[k', j', i'] = A[k = i', j = k', i = j'] (new y index = old x index, new z
    index = old y index, new x index = old z index)."""
    return np.ascontiguousarray(a.transpose(1, 2, 0))


def rot_state(st):
    """The same flow with its axes renamed: the old x velocity becomes the new
    y velocity, old y the new z, old z the new x (MAC faces stay faces)."""
    return {"F": rot(st["F"]), "U": rot(st["W"]), "V": rot(st["U"]), "W": rot(st["V"])}


def close(a, b, rel=1e-13):
    scale = max(np.max(np.abs(a)), np.max(np.abs(b)), 1e-300)
    return np.max(np.abs(a - b)) <= rel * scale


def random_state(n, seed, smooth_F=True):
    rng = np.random.default_rng(seed)
    shape = (n + 2,) * 3
    F = rng.uniform(0.05, 0.95, shape) if smooth_F else rng.random(shape)
    return {"F": F, "U": rng.normal(size=shape), "V": rng.normal(size=shape), "W": rng.normal(size=shape)}


# ----------------------------------------------------------------- predictor
def _predict(p, st):
    o = Oracle(p)
    for k in ("F", "U", "V", "W"):
        o.set(k, st[k])
    for s in ("nu_rho", "advect_u", "advect_v", "advect_w"):
        o.stage(s)
    out = {k: o.get(k) for k in ("USTAR", "VSTAR", "WSTAR")}
    o.close()
    return out


def test_predictor_v_and_w_faces_are_the_u_face_rotated():
    """advect_v (voftrial1.cu:287-342) and advect_w (:344-399) on the rotated
    state equal advect_u (:230-285) rotated: this pins their stencils, upwind
    sides, 4-point averages and gravity components to the u face, which the
    closed forms in test_oracle.py pin absolutely.  Velocities of either sign,
    F in (0,1) so that nu varies from cell to cell."""
    n = 7
    g = (1.5, -5.0, -9.81)
    st = random_state(n, 31)
    p = default_params(n=n, g=g, precision=8, dt=1e-3)
    a = _predict(p, st)
    # one rotation: v* of the rotated flow is u* of the original, w* is v*, u* is w*
    st1 = rot_state(st)
    p1 = default_params(n=n, g=(g[2], g[0], g[1]), precision=8, dt=1e-3)
    b = _predict(p1, st1)
    assert close(b["VSTAR"], rot(a["USTAR"]))
    assert close(b["WSTAR"], rot(a["VSTAR"]))
    assert close(b["USTAR"], rot(a["WSTAR"]))
    # and the values are not trivially equal (the flow is genuinely 3-D)
    assert not close(a["USTAR"], a["VSTAR"], rel=1e-3)


@pytest.mark.parametrize("comp", [0, 1, 2])
def test_predictor_quadratic_profiles_on_a_non_cubic_grid(comp):
    """One face component phi (u, v or w) = (m + 3)^2 along one axis d at a
    time, the other two components zero, F = 0 (nu = nu_g), no gravity, on a
    grid with dx, dy, dz all different.  Transverse axes (d not the face's
    own): only viscosity acts, phi* = phi + dt nu 2/h_d^2.  Own axis: phi > 0
    is advected upwind (backward difference) as well:
    phi* = phi + dt (nu 2/h_d^2 - phi (2m + 5)/h_d)  (m = index along d).
    An h_d of the wrong axis, a wrong upwind side or a dropped viscous term
    changes the value."""
    nx, ny, nz = 6, 5, 7
    L = (0.1, 0.1, 0.2)
    p = default_params(nx=nx, ny=ny, nz=nz, Lx=L[0], Ly=L[1], Lz=L[2], g=(0, 0, 0), precision=8)
    h = (L[0] / nx, L[1] / ny, L[2] / nz)
    shape = (nz + 2, ny + 2, nx + 2)
    idx = np.meshgrid(np.arange(nz + 2), np.arange(ny + 2), np.arange(nx + 2), indexing="ij")
    m_of = {0: idx[2], 1: idx[1], 2: idx[0]}
    names = ("U", "V", "W")
    star = ("USTAR", "VSTAR", "WSTAR")
    # faces each component owns: interior faces with index >= 2 along its axis
    own = [np.s_[1:-1, 1:-1, 2:-1], np.s_[1:-1, 2:-1, 1:-1], np.s_[2:-1, 1:-1, 1:-1]][comp]
    nu = p.nu_g
    for d in range(3):
        st = {k: np.zeros(shape) for k in ("F", "U", "V", "W")}
        m = m_of[d].astype(float)
        st[names[comp]] = (m + 3.0) ** 2
        got = _predict(p, st)[star[comp]][own]
        phi = st[names[comp]][own]
        mm = m[own]
        if d == comp:
            expect = phi + p.dt * (nu * 2.0 / h[d] ** 2 - phi * (2.0 * mm + 5.0) / h[d])
        else:
            expect = phi + p.dt * nu * 2.0 / h[d] ** 2
        np.testing.assert_allclose(got, expect, rtol=1e-13, err_msg=f"axis {d}")


# ----------------------------------------------------------------- rhs, Jacobi
def _rhs_and_sweep(p, st, P):
    o = Oracle(p)
    o.set("F", st["F"])
    o.stage("nu_rho")
    o.set("USTAR", st["U"]), o.set("VSTAR", st["V"]), o.set("WSTAR", st["W"])
    o.set("P", P)
    rhs = o.get("RHS")
    o.stage("jacobi")
    out = o.get("P")
    o.close()
    return rhs, out


def test_jacobi_rhs_and_sweep_rotate_with_the_axes():
    """The right-hand side rho/dt div(u*) (:418-422) and one sweep (:424-438)
    on the rotated state are the rotated results: the v* and w* terms and the
    an/as/af/ab coefficients are the u* term and ae/aw with the axes renamed."""
    n = 6
    st = random_state(n, 32)
    rng = np.random.default_rng(33)
    P = np.zeros((n + 2,) * 3)
    P[1:-1, 1:-1, 1:-1] = rng.normal(size=(n, n, n)) * 1e3
    p = default_params(n=n, precision=8)
    r0, s0 = _rhs_and_sweep(p, st, P)
    r1, s1 = _rhs_and_sweep(p, rot_state(st), rot(P))
    assert close(r1, rot(r0))
    assert close(s1, rot(s0))


def test_jacobi_rhs_closed_form_through_every_face_velocity():
    """Linear face velocities u* = a i, v* = b j, w* = c k on a non-cubic grid:
    div u* = a/dx + b/dy + c/dz in every interior cell, so rhs is
    rho/dt (a dxi + b dyi + c dzi) with rho from F (:418-422).  An axis mix-up
    (dxi for dyi, a j-1 face for j+1) changes the value."""
    nx, ny, nz = 5, 4, 6
    p = default_params(nx=nx, ny=ny, nz=nz, Lx=0.1, Ly=0.1, Lz=0.18, precision=8)  # dx .02, dy .025, dz .03
    o = Oracle(p)
    k, j, i = np.meshgrid(np.arange(nz + 2), np.arange(ny + 2), np.arange(nx + 2), indexing="ij")
    a, b, c = 0.5, -2.0, 3.25
    o.set("F", np.full(o.shape, 0.25))
    o.stage("nu_rho")
    o.set("USTAR", a * i), o.set("VSTAR", b * j), o.set("WSTAR", c * k)
    rho = 50 * 0.75 + 1000 * 0.25
    dxi, dyi, dzi = nx / 0.1, ny / 0.1, nz / 0.18
    expect = rho / p.dt * (a * dxi + b * dyi + c * dzi)
    got = o.get("RHS")[1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(got, expect, rtol=1e-12)


def test_jacobi_sweep_with_rhs_is_the_matrix_form():
    """One sweep on a non-cubic grid with a right-hand side driven through all
    of u*, v*, w*: p' = D^-1 (b - N p), with A = D + N the assembled Neumann
    operator and b = rho/dt * (discrete MAC divergence of u*)."""
    from test_oracle import neumann_matrix

    n = (5, 4, 6)
    L = (0.1, 0.1, 0.18)  # dx .02, dy .025, dz .03
    p = default_params(nx=n[0], ny=n[1], nz=n[2], Lx=L[0], Ly=L[1], Lz=L[2], precision=8)
    o = Oracle(p)
    rng = np.random.default_rng(34)
    shape = o.shape
    us, vs, ws = (rng.normal(size=shape) for _ in range(3))
    P = np.zeros(shape)
    P[1:-1, 1:-1, 1:-1] = rng.normal(size=(n[2], n[1], n[0]))
    o.set("F", np.ones(shape))
    o.stage("nu_rho")
    o.set("USTAR", us), o.set("VSTAR", vs), o.set("WSTAR", ws)
    o.set("P", P)
    o.stage("jacobi")
    h = [L[d] / n[d] for d in range(3)]
    div = ((us[1:-1, 1:-1, 2:] - us[1:-1, 1:-1, 1:-1]) / h[0] + (vs[1:-1, 2:, 1:-1] - vs[1:-1, 1:-1, 1:-1]) / h[1]
           + (ws[2:, 1:-1, 1:-1] - ws[1:-1, 1:-1, 1:-1]) / h[2])
    b = (1000.0 / p.dt * div).reshape(-1)
    A = neumann_matrix(*n, *h)
    D = np.diag(A)
    x = P[1:-1, 1:-1, 1:-1].reshape(-1)
    expect = (b - (A - np.diag(D)) @ x) / D
    np.testing.assert_allclose(o.get("P")[1:-1, 1:-1, 1:-1].reshape(-1), expect, rtol=1e-12,
                               atol=1e-12 * np.abs(expect).max())


# ----------------------------------------------------------------- correction
def test_correction_closed_form_on_every_face_non_cubic():
    """update_u/v/w (:441-493) with p = 7 i + 3 j - 2 k on a non-cubic grid and
    constant rho: u = u* - dt/rho * 7/dx, v = v* - dt/rho * 3/dy,
    w = w* - dt/rho * (-2)/dz on every interior face; wall faces untouched."""
    nx, ny, nz = 6, 5, 7
    p = default_params(nx=nx, ny=ny, nz=nz, Lx=0.1, Ly=0.07, Lz=0.13, precision=8)
    o = Oracle(p)
    k, j, i = np.meshgrid(np.arange(nz + 2), np.arange(ny + 2), np.arange(nx + 2), indexing="ij")
    P = 7.0 * i + 3.0 * j - 2.0 * k
    o.set("F", np.full(o.shape, 0.25))
    o.stage("nu_rho")
    rho = 50 * 0.75 + 1000 * 0.25
    o.set("USTAR", np.full(o.shape, 0.5)), o.set("VSTAR", np.full(o.shape, -0.25))
    o.set("WSTAR", np.full(o.shape, 0.125))
    o.set("P", P)
    o.stage("update")
    dxi, dyi, dzi = nx / 0.1, ny / 0.07, nz / 0.13
    np.testing.assert_allclose(o.get("U")[1:-1, 1:-1, 2:-1], 0.5 - p.dt / rho * 7.0 * dxi, rtol=1e-14)
    np.testing.assert_allclose(o.get("V")[1:-1, 2:-1, 1:-1], -0.25 - p.dt / rho * 3.0 * dyi, rtol=1e-14)
    np.testing.assert_allclose(o.get("W")[2:-1, 1:-1, 1:-1], 0.125 + p.dt / rho * 2.0 * dzi, rtol=1e-14)
    # faces the correction does not own are left as they were (zero here)
    assert np.all(o.get("U")[:, :, :2] == 0) and np.all(o.get("V")[:, :2, :] == 0)
    assert np.all(o.get("W")[:2, :, :] == 0)


def test_correction_rotates_with_the_axes():
    """Random rho (from F), u*, p: the v and w corrections on the rotated state
    are the u correction rotated, including the face density average."""
    n = 6
    st = random_state(n, 35)
    rng = np.random.default_rng(36)
    P = rng.normal(size=(n + 2,) * 3) * 10.0

    def run(s, Pp):
        o = Oracle(default_params(n=n, precision=8))
        o.set("F", s["F"])
        o.stage("nu_rho")
        o.set("USTAR", s["U"]), o.set("VSTAR", s["V"]), o.set("WSTAR", s["W"])
        o.set("P", Pp)
        o.stage("update")
        out = {k: o.get(k) for k in ("U", "V", "W")}
        o.close()
        return out

    a = run(st, P)
    b = run(rot_state(st), rot(P))
    for got, want in ((b["V"], a["U"]), (b["W"], a["V"]), (b["U"], a["W"])):
        m = np.zeros(got.shape, bool)
        m[1:-1, 1:-1, 1:-1] = True  # interior faces (walls and ghosts hold u* there)
        assert close(got[m], rot(want)[m])


# ----------------------------------------------------------------- FCT known answers
# Hand derivation (dyadic: dx = 2^-4 from L = 0.5 and n = 8, dt = 2^-6, so
# dt*dy*dz/vol = 2^-2 per unit velocity and every value below is exact in
# binary).  Faces f = 1..9 (face f is the low face of cell f; 1 and 9 are
# walls).  Donor-cell flux L_f = u_f dt F_upwind, downwind flux
# H_f = u_f dt F_downwind, a_f = H_f - L_f (:566-573); stage 1 (:533)
#   Ftd_c = ((F_c + (L_c - L_{c+1}) dy dz/vol) vol) / dv_c,
#   dv_c = vol - dt dy dz (u_{c+1} - u_c);
# Zalesak ratios with Q = (Fmax - Ftd) dx, (Ftd - Fmin) dx (:574-585), the
# limiter C_f by the sign of a_f (:602-606) and stage 4 (:638-645)
#   F_c = clamp01(Ftd_c - ((a_{c+1} C_{c+1} - a_c C_c) / dy) vol / dv_c).
#
# Case 1: F = [1, 1, .75, .25, 0, 0, 0, 0], u = 1 on faces 2..8:
#   Ftd = [1, 1, .8125, .375, .0625, 0, 0, 0]   (cell 1: ((1-.25) vol)/(.75 vol))
#   a_3 = -2^-8, a_4 = -2^-7, a_5 = -2^-8, others 0
#   (rp, rm): c2 (0, 0), c3 (1, 1), c4 (1, 1), c5 (0, 1), others (0, 0)
#   C_3 = min(rp2, rm3) = 0, C_4 = min(rp3, rm4) = 1, C_5 = min(rp4, rm5) = 1
#   F = [1, 1, .9375, .3125, 0, 0, 0, 0]   (donor cell alone: Ftd)
# Case 2: same F, u = 1 on faces 2..4 and 3 on faces 5..8, so cell 4
#   dilates: dv_4 = vol/2 (u_5 - u_4 = 2) and vol/dv_4 = 2 exactly:
#   Ftd = [1, 1, .8125, .5, .1875, 0, 0, 0]
#   a_3 = -2^-8, a_4 = -2^-7, a_5 = -3 2^-8
#   (rp, rm): c3 (1, 1), c4 (1, 1) [Q+/P+ = 5/3 -> 1], c5 (0, 1)
#   C_4 = C_5 = 1, C_3 = 0
#   F = [1, 1, .9375, .625, 0, 0, 0, 0]   (cell 4: .5 - ((-2^-4 vol)/(vol/2)))
# Both conserve sum F dv (cell 1 keeps F = 1 on dv = .75 vol).
KAT_F = [1.0, 1.0, 0.75, 0.25, 0.0, 0.0, 0.0, 0.0]
# Case 3 (stage-1 clip, :534): F = [1, 1, .5, 0, 0, 0, 0, 0], u = 2, 4, 6, 6, ...
#   on faces 2..8 (an outflow Courant number of 1.5 from cell 3, so stage 1
#   overshoots): cell 1 ((1 - .5) vol)/(vol/2) = 1, cell 2 ((1 + .5 - 1) vol)/(vol/2) = 1,
#   cell 3 ((.5 + 1 - .75) vol)/(vol/2) = 1.5 -> clamp01 -> 1, cell 4 .75;
#   a_3 = -2^-5, a_4 = -3 2^-6; rm_3 = min(1, 2^-6 / 2^-5) = .5, every C = 0:
#   F = Ftd = [1, 1, 1, .75, 0, 0, 0, 0].
KAT_CASES = {
    "clipping": ([2.0, 4.0, 6.0, 6.0, 6.0, 6.0, 6.0], [1.0, 1.0, 1.0, 0.75, 0.0, 0.0, 0.0, 0.0],
                 [1.0, 1.0, 1.0, 0.75, 0, 0, 0, 0], [1.0, 1.0, 0.5, 0.0, 0.0, 0.0, 0.0, 0.0]),
    "uniform": ([1.0] * 7, [1.0, 1.0, 0.8125, 0.375, 0.0625, 0.0, 0.0, 0.0], [1.0, 1.0, 0.9375, 0.3125, 0, 0, 0, 0]),
    "dilating": ([1.0] * 3 + [3.0] * 4, [1.0, 1.0, 0.8125, 0.5, 0.1875, 0.0, 0.0, 0.0],
                 [1.0, 1.0, 0.9375, 0.625, 0, 0, 0, 0]),
}
SWEEP = {0: "fct_x", 1: "fct_y", 2: "fct_z"}
VEL = {0: "U", 1: "V", 2: "W"}


def _sweep_line(axis, F_cells, u_faces, prec, n=8):
    """Run one sweep along `axis` on a line profile broadcast over the other
    two axes; return F and Ftd along the line (interior cells)."""
    p = default_params(n=n, L=0.5, dt=2.0 ** -6, precision=prec)
    o = Oracle(p)
    Fl = np.zeros(n + 2)
    Fl[1:-1] = F_cells
    Fl[0], Fl[-1] = Fl[1], Fl[-2]  # zero-gradient ghosts, as apply_bc_F leaves them
    Vl = np.zeros(n + 2)
    Vl[2:n + 1] = u_faces  # faces 2..n; walls 1 and n+1 stay 0
    shape = o.shape
    sl = [1, 1, 1]
    sl[2 - axis] = -1
    o.set("F", np.broadcast_to(Fl.reshape(sl), shape).copy())
    o.set(VEL[axis], np.broadcast_to(Vl.reshape(sl), shape).copy())
    o.stage(SWEEP[axis])
    pick = lambda a: [a[2, 2, :], a[2, :, 2], a[:, 2, 2]][axis][1:-1]
    F, Ftd = pick(o.get("F")), pick(o.get("FTD"))
    o.close()
    return F, Ftd


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("case", sorted(KAT_CASES))
def test_fct_hand_derived_known_answers(case, axis, prec):
    """The limited downwind anti-diffusion is active and exactly as derived
    above: a donor-cell-only sweep (a = 0) returns Ftd, a central high-order
    flux halves a, a wrong limiter side or a missing vol/dv gives other bits."""
    u, ftd_want, f_want, *f0 = KAT_CASES[case]
    F, Ftd = _sweep_line(axis, f0[0] if f0 else KAT_F, u, prec)
    assert np.array_equal(Ftd, np.array(ftd_want)), Ftd
    assert np.array_equal(F, np.array(f_want, dtype=float)), F


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("case", sorted(KAT_CASES))
def test_fct_hand_derived_known_answers_mirrored(case, axis):
    """The mirror image of each case (profile reversed, face velocities
    reversed and negated) gives the mirrored answer exactly: this pins the
    u < 0 branches of both fluxes and of the limiter choice absolutely."""
    u, ftd_want, f_want, *f0 = KAT_CASES[case]
    u_m = [-x for x in u[::-1]]
    F, Ftd = _sweep_line(axis, (f0[0] if f0 else KAT_F)[::-1], u_m, 8)
    assert np.array_equal(Ftd, np.array(ftd_want[::-1])), Ftd
    assert np.array_equal(F, np.array(f_want[::-1], dtype=float)), F


def _mixed_after_sweeps(nsweeps, cfl=0.25, n=96):
    p = default_params(nx=n, ny=2, nz=2, Lx=0.1 * n / 8, Ly=0.025, Lz=0.025, precision=8, dt=1e-3)
    o = Oracle(p)
    dx = p.Lx / n
    shape = o.shape
    F = np.zeros(shape)
    F[:, :, : n // 4 + 1] = 1.0
    U = np.zeros(shape)
    U[:, :, 2:n + 1] = cfl * dx / p.dt
    o.set("F", F), o.set("U", U)
    for _ in range(nsweeps):
        o.stage("fct_x")
        o.stage("bc_F")
    line = o.get("F")[2, 2, 1:-1]
    o.close()
    return int(np.sum((line > 1e-3) & (line < 1 - 1e-3))), line


def test_fct_keeps_a_step_sharp():
    """A step advected over 160 sweeps at CFL 1/4 (40 cells) keeps at most 3
    mixed cells (0.001 < F < 0.999) with the limited downwind anti-diffusion,
    as after 20 sweeps; donor-cell upwinding alone smears it like sqrt(N)
    (a 1-D donor-cell march of the same case leaves 34 mixed cells)."""
    m20, _ = _mixed_after_sweeps(20)
    m160, line = _mixed_after_sweeps(160)
    assert m160 <= 3 and m20 <= 3, (m20, m160)
    # the front moved by the Courant number per sweep: liquid now fills ~24+40 cells
    assert abs(line.sum() - (24 + 40)) < 1.0


# ----------------------------------------------------------------- FCT rotation
def test_fct_y_and_z_are_the_x_sweep_rotated():
    """fct_y (:648-785) and fct_z (:787-905) on the rotated state equal fct_x
    (:509-646) rotated, on random F in (0,1) and random face velocities of
    either sign (cubic grid: dx = dy = dz bitwise, so the reference's
    differently associated scale factors agree to round-off)."""
    n = 8
    rng = np.random.default_rng(37)
    shape = (n + 2,) * 3
    F = rng.uniform(0.05, 0.95, shape)
    U = rng.uniform(-1.0, 1.0, shape)
    U[:, :, 1] = U[:, :, n + 1] = 0.0  # wall faces
    p = default_params(n=n, precision=8, dt=2e-3)

    def sweep(stage, field, Fi, Vi):
        o = Oracle(p)
        o.set("F", Fi), o.set(field, Vi)
        o.stage(stage)
        out = o.get("F")
        o.close()
        return out

    fx = sweep("fct_x", "U", F, U)
    fy = sweep("fct_y", "V", rot(F), rot(U))
    fz = sweep("fct_z", "W", rot(rot(F)), rot(rot(U)))
    inner = (slice(1, -1),) * 3
    assert close(fy[inner], rot(fx)[inner], rel=1e-13)
    assert close(fz[inner], rot(rot(fx))[inner], rel=1e-13)
    assert not np.allclose(fx[inner], F[inner])


# ----------------------------------------------------------------- diagnostics
@pytest.mark.parametrize("prec", [4, 8])
def test_diagnostics_ic_volume_is_exact(prec):
    """The FP64 liquid volume of the initial condition is the exact cell count
    times vol = (dx dy) dz in field precision: 453,389 cells for the
    reference IC at 200^3 (FP32 predicate, N13) and 26 x 26 x 39 for the
    config box at 64^3 (lower-face strict rule); F min/max 0 and 1."""
    o = Oracle(default_params(n=200, precision=prec))
    o.init_ic(1)
    d = o.diagnostics()
    dtype = np.float32 if prec == 4 else np.float64
    h = dtype(0.1) / dtype(200)
    vol = float(h * h * h)
    assert d["liquid_volume"] == 453389 * vol
    assert d["f_min"] == 0.0 and d["f_max"] == 1.0 and d["nonfinite"] == 0
    o.close()
    o = Oracle(default_params(n=64, precision=prec))
    o.init_box(0.04, 0.04, 0.06)
    h = dtype(0.1) / dtype(64)
    assert o.diagnostics()["liquid_volume"] == 26 * 26 * 39 * float(h * h * h)
    o.close()


def test_diagnostics_sums_min_max_and_residual():
    """On a random state: the FP64 volume is vol * sum(F) and the p-mean
    sum(p)/n over interior cells only (ghosts hold other values), F min/max
    are the interior extremes, the residual is ||b - A p|| with the assembled
    Neumann operator A and the MAC divergence b (both zero-coefficient
    boundaries), and one planted NaN in u counts once."""
    import math

    from test_oracle import neumann_matrix

    n = (5, 4, 6)
    L = (0.1, 0.1, 0.18)
    p = default_params(nx=n[0], ny=n[1], nz=n[2], Lx=L[0], Ly=L[1], Lz=L[2], precision=8)
    o = Oracle(p)
    rng = np.random.default_rng(38)
    shape = o.shape
    F = rng.random(shape)
    F[0], F[-1] = 5.0, -3.0  # ghost planes: must not enter any sum or extreme
    P = rng.normal(size=shape) * 100.0
    us, vs, ws = (rng.normal(size=shape) for _ in range(3))
    U = rng.normal(size=shape)
    U[3, 2, 4] = np.nan
    o.set("F", F), o.set("P", P), o.set("U", U)
    o.stage("nu_rho")
    o.set("USTAR", us), o.set("VSTAR", vs), o.set("WSTAR", ws)
    d = o.diagnostics()
    inner = (slice(1, -1),) * 3
    h = [L[k] / n[k] for k in range(3)]
    vol = (h[0] * h[1]) * h[2]
    assert math.isclose(d["liquid_volume"], vol * math.fsum(F[inner].ravel()), rel_tol=1e-14)
    assert math.isclose(d["p_mean"], math.fsum(P[inner].ravel()) / P[inner].size, rel_tol=1e-12)
    assert d["f_min"] == F[inner].min() and d["f_max"] == F[inner].max()
    assert d["nonfinite"] == 1
    rho = 50.0 * (1 - F[inner]) + 1000.0 * F[inner]
    div = ((us[1:-1, 1:-1, 2:] - us[inner]) / h[0] + (vs[1:-1, 2:, 1:-1] - vs[inner]) / h[1]
           + (ws[2:, 1:-1, 1:-1] - ws[inner]) / h[2])
    b = (rho / p.dt * div).ravel()
    A = neumann_matrix(*n, *h)
    r = b - A @ P[inner].ravel()
    assert math.isclose(d["residual_l2"], float(np.sqrt(np.sum(r * r))), rel_tol=1e-10)
    assert math.isclose(d["rhs_l2"], float(np.sqrt(np.sum(b * b))), rel_tol=1e-12)
    o.close()


# ----------------------------------------------------------------- no-slip walls (extension)
FLAG_RHO_PREDICTOR, FLAG_NO_SLIP = 0x4, 0x8


def test_no_slip_velocity_bc_closed_form():
    """VOF_FLAG_NO_SLIP: the j, i, k passes of apply_bc_velocity (:959-1051)
    with every tangential copy negated.  In closed form (s = -1 on a ghost
    row/plane/column of that axis, c* the index clamped to the interior):
      u(i,j,k) = s_k * 0 on the wall faces i = 1, nx+1 (zeroed by the i pass,
                 after the j pass), else s_j s_k u(i, cj, ck);
      v(i,j,k) = s_i s_k * 0 on j = 1, ny+1, else s_i s_k v(ci, j, ck);
      w(i,j,k) = +0 on k = 1, nz+1 (the k pass is last), else s_i s_j w(ci, cj, k).
    Checked bit for bit, signs of zeros included."""
    nx, ny, nz = 5, 4, 6
    p = default_params(nx=nx, ny=ny, nz=nz, precision=8, flags=FLAG_NO_SLIP)
    o = Oracle(p)
    rng = np.random.default_rng(40)
    pre = {k: rng.normal(size=o.shape) for k in ("U", "V", "W")}
    for k, a in pre.items():
        o.set(k, a)
    o.stage("bc_vel")
    sgn = lambda n: np.where((np.arange(n + 2) >= 1) & (np.arange(n + 2) <= n), 1.0, -1.0)
    si, sj, sk = sgn(nx), sgn(ny), sgn(nz)
    ci, cj, ck = (np.clip(np.arange(n + 2), 1, n) for n in (nx, ny, nz))
    S = lambda a, b: a[:, None, None] if b == "k" else (a[None, :, None] if b == "j" else a[None, None, :])
    u = pre["U"][ck][:, cj] * S(sj, "j") * S(sk, "k")
    u[:, :, [1, nx + 1]] = (np.zeros((nz + 2, ny + 2, 2)) * S(sk, "k"))
    v = pre["V"][ck][:, :, ci] * S(si, "i") * S(sk, "k")
    v[:, [1, ny + 1], :] = np.zeros((nz + 2, 2, nx + 2)) * S(si, "i") * S(sk, "k")
    w = pre["W"][:, cj][:, :, ci] * S(si, "i") * S(sj, "j")
    w[[1, nz + 1], :, :] = 0.0
    for name, want in (("U", u), ("V", v), ("W", w)):
        got = o.get(name)
        assert np.array_equal(got, want), name
        assert np.array_equal(np.signbit(got), np.signbit(want)), name + " signed zeros"
    o.close()


@pytest.mark.parametrize("comp", [0, 1, 2])
def test_no_slip_shear_mode_decays_at_the_discrete_eigenvalue(comp):
    """phi = sin(pi y_a / L_a) sin(pi y_b / L_b) across the two axes a, b
    transverse to the face component phi (cell centres y = (m - 1/2) h), uniform
    along its own axis: with the no-slip ghosts (phi(0) = -phi(1),
    phi(n+1) = -phi(n)) this is an eigenvector of the discrete Laplacian with
    eigenvalue sum_t (2 cos(pi h_t / L_t) - 2) / h_t^2, so one predictor step
    (F = 0, no gravity, the other components zero) gives
    phi* = phi (1 + dt nu_g lambda) on faces away from the normal walls.  The
    reference's free-slip copies (phi(0) = phi(1)) give another value next to
    every wall."""
    n = (8, 6, 7)
    L = (0.08, 0.06, 0.1)
    p = default_params(nx=n[0], ny=n[1], nz=n[2], Lx=L[0], Ly=L[1], Lz=L[2], g=(0, 0, 0), precision=8,
                       flags=FLAG_NO_SLIP, dt=1e-3)
    o = Oracle(p)
    shape = o.shape
    axes = [t for t in range(3) if t != comp]
    idx = np.meshgrid(np.arange(n[2] + 2), np.arange(n[1] + 2), np.arange(n[0] + 2), indexing="ij")
    m_of = {0: idx[2], 1: idx[1], 2: idx[0]}
    h = [L[t] / n[t] for t in range(3)]
    phi = np.ones(shape)
    for t in axes:
        phi = phi * np.sin(np.pi * (m_of[t] - 0.5) * h[t] / L[t])
    names = ("U", "V", "W")
    st = {k: np.zeros(shape) for k in names}
    st[names[comp]] = phi
    o.set("F", np.zeros(shape))
    for k in names:
        o.set(k, st[k])
    o.stage("bc_vel")  # normal walls to 0, no-slip ghosts
    for s_ in ("nu_rho", "advect_u", "advect_v", "advect_w"):
        o.stage(s_)
    got = o.get(("USTAR", "VSTAR", "WSTAR")[comp])
    lam = sum((2 * np.cos(np.pi * h[t] / L[t]) - 2) / h[t] ** 2 for t in axes)
    want = phi * (1 + p.dt * p.nu_g * lam)
    sel = [slice(1, -1)] * 3
    sel[2 - comp] = slice(3, n[comp])  # own-axis faces 3 .. n-1 (the walls' neighbours see u = 0)
    sel = tuple(sel)
    np.testing.assert_allclose(got[sel], want[sel], rtol=1e-13, atol=1e-16)
    o.close()


# ----------------------------------------------------------------- density-weighted viscosity (extension)
def _predict_flags(p, st):
    o = Oracle(p)
    for k in ("F", "U", "V", "W"):
        o.set(k, st[k])
    for s_ in ("nu_rho", "advect_u", "advect_v", "advect_w"):
        o.stage(s_)
    out = {k: o.get(k) for k in ("USTAR", "VSTAR", "WSTAR", "RHO")}
    o.close()
    return out


def test_rho_predictor_reduces_to_the_reference_form_for_one_fluid():
    """With F uniform, mu = rho nu is one constant, every edge average of it
    is exact (sums of equal values times 1/4) and rho_f = rho, so the flux form
    is nu times the reference's Laplacian up to the rounding of the different
    association: the two predictors agree to 1e-13 (random velocities)."""
    n = 7
    st = random_state(n, 41)
    st["F"] = np.full((n + 2,) * 3, 0.375)
    # viscosities large enough that the viscous increment is comparable to the
    # velocity itself (nu dt / h^2 ~ 0.25)
    kw = dict(n=n, precision=8, dt=1e-3, g=(1.0, -2.0, 0.5), nu_l=0.05, nu_g=0.04)
    a = _predict_flags(default_params(**kw), st)
    b = _predict_flags(default_params(flags=FLAG_RHO_PREDICTOR, **kw), st)
    for k, v in zip(("USTAR", "VSTAR", "WSTAR"), ("U", "V", "W")):
        inc_a, inc_b = a[k] - st[v], b[k] - st[v]
        m = a[k] != 0  # faces the predictor writes
        assert close(inc_b[m], inc_a[m], rel=1e-12), k
        assert np.max(np.abs(inc_a[m])) > 0.1 * np.max(np.abs(st[v]))


@pytest.mark.parametrize("comp", [0, 1, 2])
def test_rho_predictor_conserves_momentum_across_a_density_field(comp):
    """phi varying only across its own axis (so advection vanishes: its own
    difference is zero and the other velocities are zero), random F in (0,1),
    free-slip ghosts (zero-gradient copies): the viscous fluxes through the
    edges between faces telescope, so in every slice normal to the face's axis
    sum(rho_f (phi* - phi)/dt) = 0 up to round-off, whatever the density
    field.  A face density of one cell, an edge viscosity averaged over the
    wrong cells or an unbalanced m+/m- breaks the telescoping."""
    n = (6, 5, 7)
    p = default_params(nx=n[0], ny=n[1], nz=n[2], precision=8, dt=1.0, g=(0, 0, 0), flags=FLAG_RHO_PREDICTOR)
    rng = np.random.default_rng(42 + comp)
    shape = (n[2] + 2, n[1] + 2, n[0] + 2)
    F = rng.uniform(0.05, 0.95, shape)
    own = 2 - comp  # numpy axis of the component's own direction
    prof = rng.normal(size=[s if ax != own else 1 for ax, s in enumerate(shape)])
    prof = np.pad(prof[tuple(slice(1, -1) if ax != own else slice(None) for ax in range(3))],
                  [(1, 1) if ax != own else (0, 0) for ax in range(3)], mode="edge")  # zero-gradient ghosts
    phi = np.broadcast_to(prof, shape).copy()
    names = ("U", "V", "W")
    st = {"F": F, "U": np.zeros(shape), "V": np.zeros(shape), "W": np.zeros(shape)}
    st[names[comp]] = phi
    out = _predict_flags(p, st)
    star = out[("USTAR", "VSTAR", "WSTAR")[comp]]
    rho = out["RHO"]
    lo = [slice(1, -1)] * 3
    lo[own] = slice(1, -2)
    hi = [slice(1, -1)] * 3
    hi[own] = slice(2, -1)
    rf = 0.5 * (rho[tuple(lo)] + rho[tuple(hi)])  # faces 2..n along the own axis
    acc = rf * (star[tuple(hi)] - phi[tuple(hi)]) / p.dt
    other = tuple(ax for ax in range(3) if ax != own)
    tot = acc.sum(axis=other)
    scale = np.abs(acc).sum(axis=other)
    assert np.all(np.abs(tot) <= 1e-12 * scale), (tot, scale)
    assert np.all(scale > 0)


def test_rho_predictor_two_layer_shear_is_steady():
    """Two fluids in layers (liquid below the cell boundary z = 3 h, gas above),
    u = u(z) with u(k+1) - u(k) = tau / mu_edge(k + 1/2), mu_edge the edge
    average of the flux form (here (mu(k) + mu(k+1))/2): the shear stress is
    continuous across the density jump, so the viscous term vanishes and
    u* = u on every face whose two z-edges are interior.  The reference's
    single-cell nu (flag off) makes the interface faces accelerate."""
    n = 8
    m = 3
    p_on = default_params(n=n, precision=8, dt=1e-3, g=(0, 0, 0), flags=FLAG_RHO_PREDICTOR)
    shape = (n + 2,) * 3
    F = np.zeros(shape)
    F[: m + 1] = 1.0  # planes 0..m liquid (plane 0 is the ghost)
    rho = 50.0 * (1 - F[:, 0, 0]) + 1000.0 * F[:, 0, 0]
    nu = 1.5e-5 * (1 - F[:, 0, 0]) + 1e-6 * F[:, 0, 0]
    mu = rho * nu
    tau = 2e-4
    uz = np.zeros(n + 2)
    for k in range(1, n + 1):
        uz[k + 1 if k + 1 <= n + 1 else k] = uz[k] + tau / (0.5 * (mu[k] + mu[k + 1])) if k + 1 <= n + 1 else uz[k]
    uz[0] = uz[1]
    U = np.broadcast_to(uz[:, None, None], shape).copy()
    st = {"F": F, "U": U, "V": np.zeros(shape), "W": np.zeros(shape)}
    on = _predict_flags(p_on, st)["USTAR"]
    off = _predict_flags(default_params(n=n, precision=8, dt=1e-3, g=(0, 0, 0)), st)["USTAR"]
    sel = (slice(2, n), slice(1, -1), slice(2, n + 1))  # faces k = 2..n-1, all j, i = 2..nx
    np.testing.assert_allclose(on[sel], U[sel], rtol=1e-12, atol=0)
    assert np.max(np.abs(off[sel] - U[sel])) > 1e-6 * np.max(np.abs(U))


@pytest.mark.parametrize("comp", [0, 1, 2])
def test_rho_predictor_own_axis_stress_uses_each_cells_viscosity(comp):
    """Along a face component's own axis the stress on the cell between faces
    m and m+1 is carried by that cell's viscosity.  phi = 0 on every even face
    and distinct values A, B on the odd faces around it, F (so mu and rho)
    distinct per cell along the axis, everything uniform across: on an even
    face advection vanishes (phi = 0 there, the other components zero), so
    phi* = dt (mu_m B + mu_(m-1) A) / h^2 / (0.5 (rho_(m-1) + rho_m)).
    A swap of the two cells' viscosities, or one cell's density for the face,
    changes the value."""
    n = (8, 9, 10)
    L = (0.08, 0.09, 0.12)
    p = default_params(nx=n[0], ny=n[1], nz=n[2], Lx=L[0], Ly=L[1], Lz=L[2], precision=8, dt=1e-3, g=(0, 0, 0),
                       flags=FLAG_RHO_PREDICTOR)
    own = 2 - comp
    N = n[comp]
    shape = (n[2] + 2, n[1] + 2, n[0] + 2)
    rng = np.random.default_rng(43 + comp)
    line_F = rng.uniform(0.05, 0.95, N + 2)
    line_phi = np.zeros(N + 2)
    line_phi[1::2] = rng.uniform(0.5, 2.0, len(line_phi[1::2])) * rng.choice([-1, 1], len(line_phi[1::2]))
    bshape = [1, 1, 1]
    bshape[own] = -1
    F = np.broadcast_to(line_F.reshape(bshape), shape).copy()
    phi = np.broadcast_to(line_phi.reshape(bshape), shape).copy()
    names = ("U", "V", "W")
    st = {"F": F, "U": np.zeros(shape), "V": np.zeros(shape), "W": np.zeros(shape)}
    st[names[comp]] = phi
    out = _predict_flags(p, st)
    star = out[("USTAR", "VSTAR", "WSTAR")[comp]]
    rho = 50.0 * (1 - line_F) + 1000.0 * line_F
    mu = rho * (1e-6 * line_F + 1.5e-5 * (1 - line_F))
    hi = N / L[comp]
    pick = [2, 2, 2]
    for m in range(2, N + 1, 2):  # even faces
        A, B = line_phi[m - 1], line_phi[m + 1]
        want = p.dt * ((mu[m] * B + mu[m - 1] * A) * hi * hi / (0.5 * (rho[m - 1] + rho[m])))
        pick[own] = m
        got = star[tuple(pick)]
        assert got == pytest.approx(want, rel=1e-12, abs=0), (m, got, want)
