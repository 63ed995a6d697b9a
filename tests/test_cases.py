"""The named synthetic workloads: the ripple layer's volume fraction and the
ensemble draws (CPU)."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cuda-lbm-ibm-fvm_amd"))

from vof_cases import SEED, ensemble_config, field_errors, ripple_F, ripple_height  # noqa: E402


def test_ripple_column_volume_is_the_surface_height():
    # each column holds (h - 0) / dz cells of liquid: sum_k F = h / dz exactly up to rounding
    n = 48
    dz = 0.1 / n
    h = ripple_height(n)
    F = ripple_F(n)
    np.testing.assert_allclose(F[1:-1, 1:-1, 1:-1].sum(axis=0), h / dz, rtol=1e-12)
    assert F.min() == 0.0 and F.max() == 1.0
    # amplitude at most two cells around the 0.06 m depth
    assert np.all(np.abs(h - 0.06) <= 2 * 0.1 / n + 1e-15)


def test_ripple_ghosts_copy_the_nearest_interior_cell():
    F = ripple_F(20)
    i = np.clip(np.arange(22), 1, 20)
    np.testing.assert_array_equal(F, F[i][:, i][:, :, i])


def test_ripple_is_seeded():
    np.testing.assert_array_equal(ripple_height(32), ripple_height(32, seed=SEED))
    assert not np.array_equal(ripple_height(32), ripple_height(32, seed=SEED + 1))


def test_ensemble_draws_are_fixed_by_the_seed(vofmod):
    a, ba = ensemble_config(vofmod)
    b, bb = ensemble_config(vofmod)
    assert len(a) == 64 and ba == bb
    assert [p.nx for p in a] == [p.nx for p in b]
    assert {p.nx for p in a} <= {32, 48, 64, 80, 96}
    assert all(25.0 <= p.rho_g <= 100.0 and 0.003 <= p.sigma <= 0.02 for p in a)


def test_field_errors_closed_form():
    ref = np.ones((4, 4, 4))
    a = ref.copy()
    a[1, 1, 1] = 1.5  # one interior cell of 8 off by 0.5
    e = field_errors(a, ref)
    assert e["max_abs"] == 0.5
    assert abs(e["rel_l2"] - 0.5 / np.sqrt(8)) < 1e-15


def test_drop_volume_and_floor_contact():
    # sub-sampled ellipsoid: volume within 1% of 4/3 pi a b c at 64^3; touches the floor, not the walls
    from vof_cases import drop_F

    n = 64
    F = drop_F(n)
    vol = F[1:-1, 1:-1, 1:-1].sum() * (0.1 / n) ** 3
    assert abs(vol - 4.0 / 3.0 * np.pi * 0.025 * 0.025 * 0.012) / vol < 0.01
    assert F[1].max() > 0.0 and F[1:-1, 1:-1, 1].max() == 0.0
    np.testing.assert_array_equal(F[0], F[1])  # ghost plane copies the first interior plane
