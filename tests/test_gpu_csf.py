"""GPU parity of the CSF extension (VOF_FLAG_CSF) against the CPU oracle:
the smoothing/curvature/face-force stage and whole steps, bit for bit (IEEE ==)
in FP32 and FP64; the batched and graph-replayed steps equal the plain one."""
import numpy as np
import pytest

from test_csf import sphere_F

pytestmark = pytest.mark.gpu


def csf_pair(vofmod, orc, **kw):
    kw.setdefault("sigma", 0.07)
    prm = dict(flags=vofmod.FLAG_CSF, **kw)
    return vofmod.Solver(vofmod.default_params(**prm)), orc.Oracle(vofmod.default_params(**prm))


def drop_state(shape, L, seed, dtype):
    nz, ny, nx = (s - 2 for s in shape)
    assert nx == ny == nz
    F = sphere_F(nx, (0.45 * L, 0.55 * L, 0.5 * L), 0.27 * L)
    rng = np.random.default_rng(seed)
    st = {"F": F}
    for k in "UVW":
        st[k] = rng.normal(0, 0.05, shape)
    st["P"] = np.zeros(shape)
    return {k: a.astype(dtype) for k, a in st.items()}


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("n", [20, 33])
def test_csf_stage_matches_oracle(vofmod, orc, prec, n):
    s, o = csf_pair(vofmod, orc, n=n, precision=prec, g=(0.0, 0.0, -9.81), dt=1e-4)
    st = drop_state(s.shape, 0.1, n, s.dtype)
    for k, a in st.items():
        s.upload(k, a)
        o.set(k, a)
    s.debug_stage(vofmod.STAGE_PREDICT)
    for name in ("nu_rho", "advect_u", "advect_v", "advect_w"):
        o.stage(name)
    s.debug_stage(vofmod.STAGE_CSF)
    o.stage("curvature")
    o.stage("csf")
    assert np.array_equal(s.download("KAPPA"), o.get("KAPPA"))
    for f in ("USTAR", "VSTAR", "WSTAR"):
        assert np.array_equal(s.download(f), o.get(f)), f
    s.close()
    o.close()


@pytest.mark.parametrize("prec", [4, 8])
def test_csf_steps_match_oracle(vofmod, orc, prec):
    s, o = csf_pair(vofmod, orc, n=24, precision=prec, g=(0.0, 0.0, -9.81), dt=2e-5)
    s.init_box(0.04, 0.05, 0.06)
    o.init_box(0.04, 0.05, 0.06)
    s.step(1, 6)
    o.step(1, 6)
    for f in ("F", "U", "V", "W", "P"):
        assert np.array_equal(s.download(f), o.get(f)), f
    s.close()
    o.close()


def test_csf_graph_and_batch_equal_plain(vofmod):
    prms = [vofmod.default_params(n=n, precision=4, g=(0.0, 0.0, -9.81), flags=vofmod.FLAG_CSF, sigma=sg)
            for n, sg in ((20, 0.07), (28, 0.02), (16, 0.003))]
    boxes = [(0.04, 0.05, 0.06), (0.05, 0.04, 0.05), (0.06, 0.06, 0.04)]
    ref = []
    for p, bx in zip(prms, boxes):
        s = vofmod.Solver(p)
        s.init_box(*bx)
        s.step(1, 5)
        ref.append([s.download(f) for f in ("F", "U", "P")])
        g = vofmod.Solver(p)
        g.init_box(*bx)
        g.step(1, 5, graph=True)
        for a, f in zip(ref[-1], ("F", "U", "P")):
            assert np.array_equal(g.download(f), a), f
        s.close()
        g.close()
    b = vofmod.Batch(prms)
    b.init_box(*zip(*boxes))
    b.step(1, 5, graph=True)
    for i in range(len(prms)):
        for a, f in zip(ref[i], ("F", "U", "P")):
            assert np.array_equal(b.download(i, f), a), (i, f)
    b.close()


def test_static_drop_parasitic_currents(vofmod):
    """Density-matched drop at rest (R = 0.25 L, sigma 0.07 N/m, no gravity,
    48^3 FP32, 10 Jacobi sweeps): the parasitic currents of the CSF term stay
    small (measured on MI355X after 300 steps: max |u| = 3.3e-3 m/s,
    Ca = mu max|u| / sigma = 4.7e-5; the bound is ten times that) and the
    symmetric drop does not move (centroid drift ~1e-10 m)."""
    from test_csf import sphere_F

    n, L, sigma = 48, 0.1, 0.07
    prm = vofmod.default_params(n=n, precision=4, flags=vofmod.FLAG_CSF, sigma=sigma, rho_l=1000.0, rho_g=1000.0,
                                nu_l=1e-6, nu_g=1e-6, g=(0.0, 0.0, 0.0), dt=1e-4)
    s = vofmod.Solver(prm)
    s.upload("F", sphere_F(n, (0.5 * L, 0.5 * L, 0.5 * L), 0.25 * L).astype(s.dtype))
    z = (np.arange(n) + 0.5) * (L / n)

    def centroid():
        F = s.download("F")[1:-1, 1:-1, 1:-1].astype(np.float64)
        return float((F.sum(axis=(1, 2)) * z).sum() / F.sum())

    c0 = centroid()
    s.step(1, 300, graph=True)
    umax = max(float(np.abs(s.download(f)[1:-1, 1:-1, 1:-1]).max()) for f in ("U", "V", "W"))
    ca = 1000.0 * 1e-6 * umax / sigma
    print(f"parasitic max|u| {umax:.3e} m/s, Ca {ca:.2e}", flush=True)
    assert np.isfinite(umax) and ca < 5e-4
    assert abs(centroid() - c0) < 1e-7
    s.close()
