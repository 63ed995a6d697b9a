"""CSF surface tension (extension; the reference never reads sigma), pinned on
the CPU oracle by closed forms rather than by the reference:
* a planar interface carries no net force (the curvatures on its two sides
  cancel exactly on the interface faces);
* sigma = 0 reproduces the parity step bit for bit;
* the face curvature of a sphere of radius R, weighted by the jump of F, is 2/R;
* a density-matched static drop develops the Laplace pressure jump 2 sigma / R.
"""
import numpy as np
import pytest

from vof import FLAG_CSF, default_params


def sphere_F(n, center, R, L=0.1, sub=4):
    """Volume fraction of a sphere, by sub-sampling each cell sub^3 times,
    with the ghost layer copying the nearest interior cell."""
    h = L / n
    off = (np.arange(sub) + 0.5) / sub
    x = (np.arange(n)[:, None] + off[None, :]).ravel() * h
    dx2 = (x - center[0]) ** 2
    dy2 = (x - center[1]) ** 2
    dz2 = (x - center[2]) ** 2
    inside = (dz2[:, None, None] + dy2[None, :, None] + dx2[None, None, :]) <= R * R
    f = inside.reshape(n, sub, n, sub, n, sub).mean(axis=(1, 3, 5))
    F = np.pad(f, 1, mode="edge")
    return F


def test_planar_interface_has_no_net_force(orc):
    n = 12
    kw = dict(n=n, precision=8, g=(0.0, 0.0, 0.0), dt=1e-4)
    rng = np.random.default_rng(1)
    F = np.zeros((n + 2,) * 3)
    F[: n // 2 + 1] = 1.0  # liquid below z = L/2 (k = 0..n/2), ghosts included
    u = rng.normal(0, 0.01, F.shape)
    outs = []
    for flags, sigma in ((0, 0.07), (FLAG_CSF, 0.07)):
        o = orc.Oracle(default_params(flags=flags, sigma=sigma, **kw))
        o.set("F", F)
        o.set("U", u)
        for st in ("nu_rho", "advect_u", "advect_v", "advect_w"):
            o.stage(st)
        if flags:
            o.stage("curvature")
            k = o.get("KAPPA")
            assert np.abs(k).max() > 0  # the normalised normal jumps: kappa = +-1/dz beside the plane
            o.stage("csf")
        outs.append([o.get(f) for f in ("USTAR", "VSTAR", "WSTAR")])
        o.close()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", [4, 8])
def test_zero_sigma_is_the_parity_step(orc, prec):
    n = 10
    kw = dict(n=n, precision=prec, g=(0.0, 0.0, -9.81), dt=2e-5, sigma=0.0)
    res = []
    for flags in (0, FLAG_CSF):
        o = orc.Oracle(default_params(flags=flags, **kw))
        o.init_box(0.04, 0.05, 0.06)
        o.step(1, 3)
        res.append([o.get(f) for f in ("F", "U", "V", "W", "P")])
        o.close()
    for a, b in zip(*res):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("n,rf", [(24, 0.3), (48, 0.2)])
def test_sphere_curvature_is_two_over_R(orc, n, rf):
    """Face curvature weighted by the jump of F across the face (the weight with
    which it enters the force) for spheres of 7 and 10 cells radius."""
    L = 0.1
    R = rf * L
    o = orc.Oracle(default_params(n=n, precision=8, flags=FLAG_CSF))
    F = sphere_F(n, (0.5 * L, 0.5 * L, 0.5 * L), R)
    o.set("F", F)
    o.stage("curvature")
    k = o.get("KAPPA")
    o.close()
    num = den = 0.0
    for ax in range(3):
        hi = [slice(1, -1)] * 3
        lo = [slice(1, -1)] * 3
        hi[ax], lo[ax] = slice(2, -1), slice(1, -2)
        w = np.abs(F[tuple(hi)] - F[tuple(lo)])
        num += (0.5 * (k[tuple(hi)] + k[tuple(lo)]) * w).sum()
        den += w.sum()
    assert num / den == pytest.approx(2.0 / R, rel=0.02)


def test_static_drop_laplace_pressure(orc):
    """Density-matched drop at rest, no gravity: one step with the Poisson
    iteration run to convergence; the pressure jump between the drop centre and
    the corner of the box is sigma kappa = 2 sigma / R."""
    n, L, sigma = 24, 0.1, 0.07
    R = 0.3 * L
    prm = default_params(n=n, precision=8, flags=FLAG_CSF, sigma=sigma, rho_l=1000.0, rho_g=1000.0,
                         nu_l=1e-6, nu_g=1e-6, g=(0.0, 0.0, 0.0), dt=1e-5, jacobi_sweeps=4000)
    o = orc.Oracle(prm)
    o.set("F", sphere_F(n, (0.5 * L, 0.5 * L, 0.5 * L), R))
    o.step(1, 1)
    p = o.get("P")
    o.close()
    c = n // 2
    dp = float(p[c:c + 2, c:c + 2, c:c + 2].mean() - p[1:3, 1:3, 1:3].mean())
    assert dp == pytest.approx(2 * sigma / R, rel=0.03)
