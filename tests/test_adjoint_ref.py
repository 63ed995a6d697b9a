"""The FP64 PyTorch restatement used as the taped gradient reference must
reproduce the oracle's FP64 step before its gradients are trusted (CPU)."""
import numpy as np
import pytest
import torch

import adjoint_ref as ar
from oracle import Oracle
from vof import default_params


def _state(shape, seed, vel):
    rng = np.random.default_rng(seed)
    F = rng.random(shape)
    F[F < 0.35] = 0.0
    F[F > 0.8] = 1.0
    u, v, w = (rng.uniform(-vel, vel, shape) for _ in range(3))
    p = np.zeros(shape)
    p[1:-1, 1:-1, 1:-1] = rng.normal(0, 30.0, (shape[0] - 2, shape[1] - 2, shape[2] - 2))
    return dict(F=F, U=u, V=v, W=w, P=p)


@pytest.mark.parametrize("first", [1, 2, 3])
def test_torch_step_matches_oracle_fp64(first):
    nx, ny, nz = 11, 9, 8
    kw = dict(nx=nx, ny=ny, nz=nz, precision=8, dt=2e-4, g=(0.5, -5.0, -9.81))
    o = Oracle(default_params(**kw))
    st = _state(o.shape, 5 + first, 0.8)
    for k, a in st.items():
        o.set(k, a)
    P = ar.Prm(nx, ny, nz, dt=2e-4, g=(0.5, -5.0, -9.81))
    ts = tuple(torch.tensor(st[k], dtype=torch.float64) for k in ("F", "U", "V", "W", "P"))
    coeffs = ar.jacobi_coeffs(P)
    for s in range(3):
        ts = ar.step(P, ts, first + s, coeffs)
    o.step(first, 3)
    for name, t in zip(("F", "U", "V", "W", "P"), ts):
        ref = o.get(name)
        # bit for bit (IEEE ==): same operations, same association, FP64 scalars
        np.testing.assert_array_equal(t.numpy(), ref, err_msg=name)


def test_torch_gradient_matches_central_differences():
    """Autograd through three steps vs central differences of the same FP64
    function (pins the taped reference's differentiation, not the physics)."""
    nx = ny = nz = 8
    P = ar.Prm(nx, ny, nz, dt=5e-5, g=(0.0, 0.0, -9.81))
    F0 = torch.zeros(nz + 2, ny + 2, nx + 2, dtype=torch.float64)
    F0[1:5, 1:5, 1:4] = 1.0
    F0 = ar.ghost_F(P, ar.interior(F0))
    rng = np.random.default_rng(3)
    u0 = torch.tensor(rng.normal(0, 0.05, F0.shape))
    v0 = torch.tensor(rng.normal(0, 0.05, F0.shape))
    w0 = torch.tensor(rng.normal(0, 0.05, F0.shape))
    p0 = torch.zeros_like(F0)
    coeffs = ar.jacobi_coeffs(P)

    def J(rho_l, u):
        P.rho_l = rho_l
        st = (F0, u, v0, w0, p0)
        for s in range(3):
            st = ar.step(P, st, 1 + s, coeffs)
        return ar.centroid_height(P, st[0])

    rl = torch.tensor(1000.0, dtype=torch.float64, requires_grad=True)
    uu = u0.clone().requires_grad_(True)
    j = J(rl, uu)
    j.backward()
    d = torch.tensor(rng.normal(0, 1.0, F0.shape))
    d = d * (u0.norm() / d.norm())          # direction of the size of u0
    h = 1e-4                                # relative step
    with torch.no_grad():
        fd_u = (J(torch.tensor(1000.0), u0 + h * d) - J(torch.tensor(1000.0), u0 - h * d)) / (2 * h)
        fd_r = (J(torch.tensor(1000.0 * (1 + h)), u0) - J(torch.tensor(1000.0 * (1 - h)), u0)) / (2 * h * 1000.0)
    assert float((uu.grad * d).sum()) == pytest.approx(float(fd_u), rel=1e-4, abs=1e-12)
    assert float(rl.grad) == pytest.approx(float(fd_r), rel=1e-4, abs=1e-14)


def _csf_state(n, seed):
    from test_csf import sphere_F

    rng = np.random.default_rng(seed)
    F = sphere_F(n, (0.045, 0.055, 0.05), 0.027)
    u, v, w = (rng.normal(0, 0.05, F.shape) for _ in range(3))
    return dict(F=F, U=u, V=v, W=w, P=np.zeros(F.shape))


@pytest.mark.parametrize("first,rho,seed", [(1, 1000.0, 1), (2, 1000.0, 2), (1, 1.0, 51)])
def test_torch_csf_step_matches_oracle_fp64(first, rho, seed):
    """Three steps; the light-fluid case (rho 1) moves the interface enough
    that some vertex normals hit inputs where CPU torch.sqrt is one ulp off."""
    n, sigma = 12, 0.07
    dens = dict(rho_l=rho, rho_g=rho if rho == 1.0 else 50.0)
    kw = dict(n=n, precision=8, dt=1e-4, g=(0.0, 0.0, -9.81), flags=0x2, sigma=sigma, **dens)
    o = Oracle(default_params(**kw))
    st = _csf_state(n, seed)
    for k, a in st.items():
        o.set(k, a)
    P = ar.Prm(n, n, n, dt=1e-4, g=(0.0, 0.0, -9.81), csf=True, sigma=sigma, **dens)
    ts = tuple(torch.tensor(st[k], dtype=torch.float64) for k in ("F", "U", "V", "W", "P"))
    coeffs = ar.jacobi_coeffs(P)
    for s in range(3):
        ts = ar.step(P, ts, first + s, coeffs)
    o.step(first, 3)
    for name, t in zip(("F", "U", "V", "W", "P"), ts):
        np.testing.assert_array_equal(t.numpy(), o.get(name), err_msg=name)


def test_sqrt_rn_is_correctly_rounded():
    """The restatement's square root against Python's math.sqrt (IEEE,
    correctly rounded) on inputs spanning the vertex-normal range."""
    import math

    x = np.random.default_rng(3).uniform(0.0, 1.0, 4096) * 10.0 ** np.random.default_rng(4).integers(-20, 4, 4096)
    got = ar.sqrt_rn(torch.tensor(x, dtype=torch.float64)).numpy()
    np.testing.assert_array_equal(got, np.array([math.sqrt(v) for v in x]))


def test_torch_csf_sigma_gradient_matches_central_differences():
    """A light density-matched fluid (rho 1) so that surface tension moves the
    interface within two steps and the difference quotient resolves it."""
    n = 10
    st = _csf_state(n, 7)
    P = ar.Prm(n, n, n, rho_l=1.0, rho_g=1.0, dt=1e-4, g=(0.0, 0.0, -9.81), csf=True)
    coeffs = ar.jacobi_coeffs(P)
    x0 = tuple(torch.tensor(st[k], dtype=torch.float64) for k in ("F", "U", "V", "W", "P"))

    def J(sigma):
        P.sigma = sigma
        s = x0
        for m in range(2):
            s = ar.step(P, s, 1 + m, coeffs)
        return ar.centroid_height(P, s[0])

    sg = torch.tensor(0.07, dtype=torch.float64, requires_grad=True)
    J(sg).backward()
    h = 1e-4
    with torch.no_grad():
        fd = (J(0.07 * (1 + h)) - J(0.07 * (1 - h))) / (2 * h * 0.07)
    assert float(sg.grad) != 0.0
    assert float(sg.grad) == pytest.approx(float(fd), rel=1e-4)
