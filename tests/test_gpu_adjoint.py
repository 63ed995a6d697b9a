"""GPU reverse mode: the hand-written HIP adjoint through the C ABI.

Pins, strongest first:
* the vector-Jacobian product of whole checkpointed step sequences against
  PyTorch autograd through the FP64 restatement of the step (tests/adjoint_ref.py,
  itself checked against the oracle's forward), with random seeds on every
  final-state field, in FP64: agreement to 1e-9 relative (the two sides add
  in different orders);
* central finite differences (relative step 1e-4) of the GPU's own FP64
  forward on a dam break, for the initial velocity, rho_l and rho_g;
* the FP32 adjoint against the FP64 one, the torch.autograd.Function contract,
  and that the checkpointed forward is the plain step, bit for bit.
"""
import numpy as np
import pytest
import torch

import adjoint_ref as ar

pytestmark = pytest.mark.gpu

T = torch.float64


def random_state(shape, seed, vel):
    rng = np.random.default_rng(seed)
    F = rng.random(shape)
    F[F < 0.35] = 0.0
    F[F > 0.8] = 1.0
    u, v, w = (rng.uniform(-vel, vel, shape) for _ in range(3))
    p = np.zeros(shape)
    p[1:-1, 1:-1, 1:-1] = rng.normal(0, 30.0, (shape[0] - 2, shape[1] - 2, shape[2] - 2))
    return dict(F=F, U=u, V=v, W=w, P=p)


def dam_state(vofmod, prm, box, vel_std, seed, width=2.0):
    """Dam-break F from the solver's own IC, plus smoothed Gaussian velocity noise."""
    import vof_autograd

    s = vofmod.Solver(prm)
    s.init_box(*box)
    F = s.download("F").astype(np.float64)
    s.close()
    shape = F.shape
    st = {"F": F, "P": np.zeros(shape)}
    for i, k in enumerate("UVW"):
        st[k] = vof_autograd.smoothed_noise(shape, vel_std, width, seed + i)
    return st


def rel_err(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def gpu_vjp(vofmod, prm, st, first, nsteps, every, gJ, seeds):
    s = vofmod.Solver(prm)
    for k, a in st.items():
        s.upload(k, a)
    J = s.adj_forward(first, nsteps, every)
    for k, a in seeds.items():
        s.adj_seed(k, a)
    sc = s.adj_backward(gJ)
    g = {k: s.adj_download(k) for k in "FUVWP"}
    s.close()
    return J, sc, g


@pytest.mark.parametrize("first,nsteps,every", [(1, 1, 1), (2, 3, 2), (3, 4, 3), (1, 5, 8)])
def test_vjp_matches_taped_reference_fp64(vofmod, first, nsteps, every):
    nx, ny, nz = 11, 9, 8
    dt, g = 2e-4, (0.5, -5.0, -9.81)
    prm = vofmod.default_params(nx=nx, ny=ny, nz=nz, precision=8, dt=dt, g=g)
    st = random_state((nz + 2, ny + 2, nx + 2), 40 + first, 0.8)
    rng = np.random.default_rng(7 + first)
    seeds = {k: rng.normal(0, 1, st["F"].shape) for k in "FUVWP"}
    gJ = 0.7
    print(f"adjoint vs taped reference: first {first} nsteps {nsteps} every {every}", flush=True)
    J, sc, gg = gpu_vjp(vofmod, prm, st, first, nsteps, every, gJ, seeds)

    leaf = lambda x: torch.tensor(float(x), dtype=T, requires_grad=True)
    P = ar.Prm(nx, ny, nz, dt=dt, g=tuple(leaf(x) for x in g))
    P.rho_l, P.rho_g, P.nu_l, P.nu_g = leaf(1000.0), leaf(50.0), leaf(1e-6), leaf(1.5e-5)
    x0 = [torch.tensor(st[k], dtype=T, requires_grad=True) for k in "FUVWP"]
    s = tuple(x0)
    coeffs = ar.jacobi_coeffs(P)
    for m in range(nsteps):
        s = ar.step(P, s, first + m, coeffs)
    Jt = ar.centroid_height(P, s[0])
    L = gJ * Jt + sum((torch.tensor(seeds[k]) * t).sum() for k, t in zip("FUVWP", s))
    L.backward()
    assert J == pytest.approx(Jt.item(), rel=1e-10)
    for k, t in zip("FUVWP", x0):
        e = rel_err(gg[k], t.grad.numpy())
        assert e < 1e-9, f"d/d{k}0: relative error {e:.2e}"
    ref = {"rho_l": P.rho_l, "rho_g": P.rho_g, "nu_l": P.nu_l, "nu_g": P.nu_g,
           "gx": P.g[0], "gy": P.g[1], "gz": P.g[2]}
    for k, t in ref.items():
        assert sc[k] == pytest.approx(t.grad.item(), rel=1e-9, abs=1e-30), k
    assert sc["sigma"] == 0.0


@pytest.mark.parametrize("sweeps", [3, 4, 7])
def test_reverse_jacobi_pairs_on_several_tiles(vofmod, sweeps):
    """The reverse Jacobi runs two sweeps per launch (a last single sweep when
    the count is odd) on 64 x 16 tiles marching through z chunks: a grid of
    two x tiles, three row bands and several chunks, odd and even sweep counts,
    against the taped FP64 reference."""
    nx, ny, nz = 70, 30, 21
    dt, g = 2e-4, (0.5, -5.0, -9.81)
    prm = vofmod.default_params(nx=nx, ny=ny, nz=nz, precision=8, dt=dt, g=g, jacobi_sweeps=sweeps)
    st = random_state((nz + 2, ny + 2, nx + 2), 60 + sweeps, 0.8)
    rng = np.random.default_rng(17 + sweeps)
    seeds = {k: rng.normal(0, 1, st["F"].shape) for k in "FUVWP"}
    print(f"reverse Jacobi pairs: {sweeps} sweeps on {nx}x{ny}x{nz}", flush=True)
    J, sc, gg = gpu_vjp(vofmod, prm, st, 1, 2, 2, 0.7, seeds)
    leaf = lambda x: torch.tensor(float(x), dtype=T, requires_grad=True)
    P = ar.Prm(nx, ny, nz, dt=dt, g=tuple(leaf(x) for x in g), sweeps=sweeps)
    P.rho_l, P.rho_g, P.nu_l, P.nu_g = leaf(1000.0), leaf(50.0), leaf(1e-6), leaf(1.5e-5)
    x0 = [torch.tensor(st[k], dtype=T, requires_grad=True) for k in "FUVWP"]
    s = tuple(x0)
    coeffs = ar.jacobi_coeffs(P)
    for m in range(2):
        s = ar.step(P, s, 1 + m, coeffs)
    L = 0.7 * ar.centroid_height(P, s[0]) + sum((torch.tensor(seeds[k]) * t).sum() for k, t in zip("FUVWP", s))
    L.backward()
    for k, t in zip("FUVWP", x0):
        e = rel_err(gg[k], t.grad.numpy())
        assert e < 1e-9, f"d/d{k}0: relative error {e:.2e}"
    assert sc["rho_l"] == pytest.approx(P.rho_l.grad.item(), rel=1e-9, abs=1e-30)


def centroid(vofmod, prm, st, first, nsteps):
    s = vofmod.Solver(prm)
    for k, a in st.items():
        s.upload(k, a)
    J = s.adj_forward(first, nsteps, nsteps)
    s.close()
    return J


def test_gradient_matches_central_differences_fp64(vofmod):
    """Dam break with smoothed initial-velocity noise (1e-3 m/s, Gaussian width
    4 cells); smooth perturbation directions of the size of u0 and relative
    steps 1e-4, as in the 96^3 study, here at 32^3 over 200 steps."""
    import vof_autograd as va

    n, nsteps = 32, 200
    kw = dict(n=n, precision=8, g=(0.0, 0.0, -9.81), dt=4e-6)
    prm = vofmod.default_params(**kw)
    st = dam_state(vofmod, prm, (0.04, 0.04, 0.06), 1e-3, 1000, width=4.0)
    J, sc, g = gpu_vjp(vofmod, prm, st, 1, nsteps, 20, 1.0, {})
    h = 1e-4
    nu = np.sqrt(sum((st[k] ** 2).sum() for k in "UVW"))
    for q in range(3):
        d = {k: va.smoothed_noise(st["U"].shape, 1.0, 4.0, 700 + 3 * q + i) for i, k in enumerate("UVW")}
        nd = np.sqrt(sum((d[k] ** 2).sum() for k in "UVW"))
        d = {k: a * (nu / nd) for k, a in d.items()}
        plus = dict(st, **{k: st[k] + h * d[k] for k in "UVW"})
        minus = dict(st, **{k: st[k] - h * d[k] for k in "UVW"})
        fd = (centroid(vofmod, prm, plus, 1, nsteps) - centroid(vofmod, prm, minus, 1, nsteps)) / (2 * h)
        ad = sum((g[k] * d[k]).sum() for k in "UVW")
        print(f"direction {q}: adjoint {ad:.10e} fd {fd:.10e}", flush=True)
        assert ad == pytest.approx(fd, rel=1e-4)
    for name in ("rho_l", "rho_g"):
        x = getattr(prm, name)
        pp = vofmod.default_params(**dict(kw, **{name: x * (1 + h)}))
        pm = vofmod.default_params(**dict(kw, **{name: x * (1 - h)}))
        fd = (centroid(vofmod, pp, st, 1, nsteps) - centroid(vofmod, pm, st, 1, nsteps)) / (2 * h * x)
        print(f"{name}: adjoint {sc[name]:.10e} fd {fd:.10e}", flush=True)
        assert sc[name] == pytest.approx(fd, rel=1e-4)


def test_fp32_adjoint_close_to_fp64(vofmod):
    n, nsteps = 20, 30
    kw = dict(n=n, g=(0.0, 0.0, -9.81), dt=2e-5)
    st = dam_state(vofmod, vofmod.default_params(precision=8, **kw), (0.04, 0.04, 0.06), 0.05, 21)
    _, sc8, g8 = gpu_vjp(vofmod, vofmod.default_params(precision=8, **kw), st, 1, nsteps, 7, 1.0, {})
    st32 = {k: a.astype(np.float32) for k, a in st.items()}
    _, sc4, g4 = gpu_vjp(vofmod, vofmod.default_params(precision=4, **kw), st32, 1, nsteps, 7, 1.0, {})
    for k in "UVW":
        e = np.linalg.norm(g4[k] - g8[k]) / np.linalg.norm(g8[k])
        print(f"d/d{k}0 FP32 vs FP64: relative L2 {e:.2e}", flush=True)
    assert sc4["rho_l"] == pytest.approx(sc8["rho_l"], rel=5e-2)
    assert sc4["rho_g"] == pytest.approx(sc8["rho_g"], rel=5e-2)


def test_checkpointed_forward_is_the_plain_step_and_state_is_restored(vofmod):
    prm = vofmod.default_params(n=18, precision=4, g=(0.0, 0.0, -9.81))
    a = vofmod.Solver(prm)
    a.init_box(0.04, 0.05, 0.06)
    a.step(1, 13)
    ref = {k: a.download(k) for k in "FUVWP"}
    a.close()
    b = vofmod.Solver(prm)
    b.init_box(0.04, 0.05, 0.06)
    b.adj_forward(1, 13, 4)
    for k in "FUVWP":
        assert np.array_equal(b.download(k), ref[k]), k
    b.adj_backward(1.0)
    for k in "FUVWP":
        assert np.array_equal(b.download(k), ref[k]), f"{k} after backward"
    f_ms, b_ms = b.adj_times()
    assert f_ms > 0 and b_ms > 0
    b.adj_release()
    b.close()


def test_autograd_function_contract(vofmod):
    import vof_autograd as va

    n, nsteps = 12, 6
    prm = vofmod.default_params(n=n, precision=8, g=(0.0, 0.0, -9.81), dt=2e-5)
    st = dam_state(vofmod, prm, (0.04, 0.04, 0.06), 0.05, 31)
    spec = va.AdjointSpec(prm, first_istep=1, nsteps=nsteps, checkpoint_every=4)
    xs = [torch.tensor(st[k], requires_grad=True) for k in "FUVWP"]
    rl = torch.tensor(1000.0, dtype=T, requires_grad=True)
    rg = torch.tensor(50.0, dtype=T, requires_grad=True)
    sg = torch.tensor(0.007, dtype=T, requires_grad=True)
    J = va.centroid_height(*xs, rl, rg, sg, spec)
    (2.0 * J).backward()
    Jd, sc, g = gpu_vjp(vofmod, prm, st, 1, nsteps, 4, 2.0, {})
    assert J.item() == Jd
    for k, x in zip("FUVWP", xs):
        assert np.array_equal(x.grad.numpy(), g[k]) or rel_err(x.grad.numpy(), g[k]) < 1e-12, k
    assert rl.grad.item() == pytest.approx(sc["rho_l"], rel=1e-12)
    assert rg.grad.item() == pytest.approx(sc["rho_g"], rel=1e-12)
    assert sg.grad.item() == 0.0


def test_adjoint_rejects_unsupported_flags(vofmod):
    prm = vofmod.default_params(n=8, precision=4, flags=vofmod.FLAG_REMOVE_P_MEAN)
    s = vofmod.Solver(prm)
    with pytest.raises(vofmod.VofError):
        s.adj_forward(1, 2, 1)
    s.close()


def fold_ghosts(g):
    """Add each ghost entry to the interior cell it copies (the F ghost fill),
    leaving zero ghosts."""
    out = np.zeros_like(g)
    nz, ny, nx = (s - 2 for s in g.shape)
    ck = np.clip(np.arange(nz + 2), 1, nz)
    cj = np.clip(np.arange(ny + 2), 1, ny)
    ci = np.clip(np.arange(nx + 2), 1, nx)
    np.add.at(out, (ck[:, None, None], cj[None, :, None], ci[None, None, :]), g)
    return out


@pytest.mark.parametrize("first,nsteps,every", [(1, 1, 1), (2, 3, 2)])
def test_csf_vjp_matches_taped_reference_fp64(vofmod, first, nsteps, every):
    from test_csf import sphere_F

    n, sigma, dt = 12, 0.07, 1e-4
    prm = vofmod.default_params(n=n, precision=8, dt=dt, g=(0.0, 0.0, -9.81), flags=vofmod.FLAG_CSF,
                                sigma=sigma, rho_l=1.0, rho_g=1.0)
    rng = np.random.default_rng(50 + first)
    F = sphere_F(n, (0.045, 0.055, 0.05), 0.027)
    st = {"F": F, "P": np.zeros(F.shape)}
    for k in "UVW":
        st[k] = rng.normal(0, 0.05, F.shape)
    seeds = {k: rng.normal(0, 1, F.shape) for k in "FUVWP"}
    gJ = 0.7
    J, sc, gg = gpu_vjp(vofmod, prm, st, first, nsteps, every, gJ, seeds)

    leaf = lambda x: torch.tensor(float(x), dtype=T, requires_grad=True)
    P = ar.Prm(n, n, n, dt=dt, g=(0.0, 0.0, -9.81), rho_l=leaf(1.0), rho_g=leaf(1.0), csf=True, sigma=leaf(sigma))
    x0 = [torch.tensor(st[k], dtype=T, requires_grad=True) for k in "FUVWP"]
    s = tuple(x0)
    coeffs = ar.jacobi_coeffs(P)
    for m in range(nsteps):
        s = ar.step(P, s, first + m, coeffs)
    L = gJ * ar.centroid_height(P, s[0]) + sum((torch.tensor(seeds[k]) * t).sum() for k, t in zip("FUVWP", s))
    L.backward()
    ref = {k: t.grad.numpy() for k, t in zip("FUVWP", x0)}
    ref["F"] = fold_ghosts(ref["F"])
    for k in "FUVW":
        e = rel_err(gg[k], ref[k])
        assert e < 1e-9, f"d/d{k}0: relative error {e:.2e}"
    pin = (slice(1, -1),) * 3
    assert rel_err(gg["P"][pin], ref["P"][pin]) < 1e-9
    for name, t in (("sigma", P.sigma), ("rho_l", P.rho_l), ("rho_g", P.rho_g)):
        assert sc[name] == pytest.approx(t.grad.item(), rel=1e-9, abs=1e-30), name


def test_csf_sigma_gradient_matches_central_differences(vofmod):
    from test_csf import sphere_F

    n, nsteps, sigma = 24, 30, 0.07
    kw = dict(n=n, precision=8, dt=1e-4, g=(0.0, 0.0, -9.81), flags=vofmod.FLAG_CSF, rho_l=1.0, rho_g=1.0)
    F = sphere_F(n, (0.045, 0.055, 0.05), 0.027)
    st = {"F": F, "U": np.zeros(F.shape), "V": np.zeros(F.shape), "W": np.zeros(F.shape), "P": np.zeros(F.shape)}
    J, sc, _ = gpu_vjp(vofmod, vofmod.default_params(sigma=sigma, **kw), st, 1, nsteps, 10, 1.0, {})
    h = 1e-4
    fd = (centroid(vofmod, vofmod.default_params(sigma=sigma * (1 + h), **kw), st, 1, nsteps) -
          centroid(vofmod, vofmod.default_params(sigma=sigma * (1 - h), **kw), st, 1, nsteps)) / (2 * h * sigma)
    print(f"dJ/dsigma: adjoint {sc['sigma']:.10e} fd {fd:.10e}", flush=True)
    assert sc["sigma"] != 0.0
    assert sc["sigma"] == pytest.approx(fd, rel=1e-4)


def test_gradient_study_96_400_steps_8_directions(vofmod):
    """The differentiation workload at its size: 96^3 dam break, initial
    velocity = seeded Gaussian noise (std 1e-3 m/s, 4-cell smoothing), 400
    steps, checkpoint every 20, J = liquid-centroid height.  FP64 forward and
    adjoint against central differences of the FP64 forward (relative step
    1e-4) along 8 seeded smooth directions (seed + 1000 + 3q + component),
    scaled to the size of u0, and for rho_l and rho_g; dJ/dsigma is 0 without
    surface tension on both sides.

    J is piecewise smooth (upwind, limiter and clip branches), so a
    difference quotient can straddle a branch change that the adjoint, the
    derivative of the branches taken, does not see.  The bar is therefore
    set on the gradient's scale: |adjoint - fd| <= 1e-4 |g| |d| on every
    direction (measured <= 2.7e-6), and 1e-3 relative on every direction
    whose derivative is not nearly orthogonal to the gradient (measured
    <= 5.4e-4; half of the directions are within 1e-4).  The FP32-forward
    adjoint's scalar gradients stay within 5e-3 of the FP64 ones (measured
    <= 1.8e-3, for nu_g)."""
    import vof_autograd as va

    SEED = 20261203
    n, steps, every = 96, 400, 20

    def solver(prec, **kw):
        p = vofmod.default_params(n=n, precision=prec, g=(0.0, 0.0, -9.81), dt=4e-6, **kw)
        s = vofmod.Solver(p)
        s.init_box(0.04, 0.04, 0.06)
        return s

    s = solver(8)
    vel = {k: va.smoothed_noise(s.shape, 1e-3, 4.0, SEED + i) for i, k in enumerate("UVW")}
    state = {"F": s.download("F").astype(np.float64), "P": np.zeros(s.shape), **vel}
    s.close()

    def load(t, st):
        for k, a in st.items():
            t.upload(k, a.astype(t.dtype))

    grads, scal = {}, {}
    for prec in (8, 4):
        print(f"96^3 adjoint, 400 steps, prec {prec}", flush=True)
        t = solver(prec)
        load(t, state)
        t.adj_forward(1, steps, every)
        scal[prec] = t.adj_backward(1.0)
        grads[prec] = {k: t.adj_download(k) for k in "UVW"}
        t.close()

    def J_of(st, **kw):
        t = solver(8, **kw)
        load(t, st)
        v = t.adj_forward(1, steps, steps)
        t.close()
        return v

    h = 1e-4
    nu = float(np.sqrt(sum((state[k] ** 2).sum() for k in "UVW")))
    gnorm = float(np.sqrt(sum((grads[8][k] ** 2).sum() for k in "UVW")))
    for q in range(8):
        d = {k: va.smoothed_noise(state["U"].shape, 1.0, 4.0, SEED + 1000 + 3 * q + i) for i, k in enumerate("UVW")}
        nd = float(np.sqrt(sum((d[k] ** 2).sum() for k in "UVW")))
        d = {k: a * (nu / nd) for k, a in d.items()}
        fd = (J_of({**state, **{k: state[k] + h * d[k] for k in "UVW"}}) -
              J_of({**state, **{k: state[k] - h * d[k] for k in "UVW"}})) / (2 * h)
        ad = float(sum((grads[8][k] * d[k]).sum() for k in "UVW"))
        print(f"direction {q}: adjoint {ad:.10e} fd {fd:.10e} rel {abs(ad - fd) / abs(fd):.2e} "
              f"scaled {abs(ad - fd) / (gnorm * nu):.2e}", flush=True)
        assert abs(ad - fd) <= 1e-4 * gnorm * nu
        if abs(fd) >= 1e-3 * gnorm * nu:
            assert abs(ad - fd) <= 1e-3 * abs(fd)
    for name, x in (("rho_l", 1000.0), ("rho_g", 50.0)):
        fd = (J_of(state, **{name: x * (1 + h)}) - J_of(state, **{name: x * (1 - h)})) / (2 * h * x)
        print(f"{name}: adjoint {scal[8][name]:.10e} fd {fd:.10e}", flush=True)
        assert scal[8][name] == pytest.approx(fd, rel=1e-4)
    assert scal[8]["sigma"] == 0.0 and scal[4]["sigma"] == 0.0
    assert J_of(state, sigma=0.007 * (1 + h)) == J_of(state, sigma=0.007 * (1 - h))
    for k in ("rho_l", "rho_g", "nu_l", "nu_g", "gz"):
        assert scal[4][k] == pytest.approx(scal[8][k], rel=5e-3), k


def test_fp32_gradient_gap_is_not_input_rounding(vofmod):
    """The FP32-forward gradient dJ/du0 of the differentiation workload (96^3
    dam break, seeded smoothed u0 noise of 1e-3 m/s, 400 steps, K = 20) differs
    from the FP64-forward one by ~0.12 in relative L2 (measured 0.120), while
    FP64 gradients from u0 rounded to FP32 (a 6e-8 relative change) or scaled
    by 1 + 1e-12 noise differ from it by ~4e-8 and ~2e-8: the gap comes from
    FP32 arithmetic along the trajectory (limiter branches taken differently),
    not from the gradient's sensitivity to its input."""
    import vof_autograd as va

    SEED = 20261203
    n, steps, every = 96, 400, 20
    s = vofmod.Solver(vofmod.default_params(n=n, precision=8, g=(0.0, 0.0, -9.81), dt=4e-6))
    s.init_box(0.04, 0.04, 0.06)
    st = {"F": s.download("F").astype(np.float64), "P": np.zeros(s.shape)}
    s.close()
    st.update({k: va.smoothed_noise(st["F"].shape, 1e-3, 4.0, SEED + i) for i, k in enumerate("UVW")})

    def grad(prec, state):
        t = vofmod.Solver(vofmod.default_params(n=n, precision=prec, g=(0.0, 0.0, -9.81), dt=4e-6))
        for k, a in state.items():
            t.upload(k, a.astype(t.dtype))
        t.adj_forward(1, steps, every)
        t.adj_backward(1.0)
        g = np.concatenate([t.adj_download(k).ravel() for k in "UVW"])
        t.close()
        return g

    rng = np.random.default_rng(SEED + 77)
    g8 = grad(8, st)
    cases = {
        "u0 rounded to FP32": (8, {**st, **{k: st[k].astype(np.float32).astype(np.float64) for k in "UVW"}}),
        "u0 x (1 + 1e-12 noise)": (8, {**st, **{k: st[k] * (1 + 1e-12 * rng.standard_normal(st[k].shape))
                                               for k in "UVW"}}),
        "FP32 forward": (4, st),
    }
    rel = {}
    for name, (prec, state) in cases.items():
        rel[name] = float(np.linalg.norm(grad(prec, state) - g8) / np.linalg.norm(g8))
        print(f"{name}: relative L2 against the FP64 gradient {rel[name]:.3e}", flush=True)
    assert rel["u0 rounded to FP32"] < 1e-6
    assert rel["u0 x (1 + 1e-12 noise)"] < 1e-6
    assert 1e-2 < rel["FP32 forward"] < 0.5
