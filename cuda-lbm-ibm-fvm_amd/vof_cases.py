"""Named synthetic workloads of the VOF step, with fixed seeds.

The reference ships one case, its dam break (voftrial1.cu:72-93, defaults
:1258-1283).  These are the inputs the measurements and precision studies
of this library are quoted on:

* dam break: 0.1 m cube, liquid block x < 0.04, y < 0.04, z < 0.06 m,
  rho 1000/50, g = (0, 0, -9.81), dt = 4e-6 s, 10 Jacobi sweeps per step
  (64^3 FP64, 200^3 FP32 and FP64, 96^3 for the gradient);
* ensemble: 64 dam breaks of uneven size and physics (seed + i);
* ripple: a liquid layer 0.06 m deep whose free surface carries a sum of
  plane waves of amplitude up to two cells (384^3);
* drop: a density-matched ellipsoidal drop flattened against the floor, for
  the surface-tension gradient (64^3).

Base seed 20261203; instance i of the ensemble uses seed + i.
"""
from __future__ import annotations

import numpy as np

SEED = 20261203
DAM_BOX = (0.04, 0.04, 0.06)  # liquid block extents (m) in the 0.1 m cube


def dam_break_params(vof, n: int, precision: int = 4, **kw):
    """Dam-break parameters on an n^3 grid (gravity along -z)."""
    return vof.default_params(n=n, precision=precision, g=(0.0, 0.0, -9.81), **kw)


def ensemble_config(vof, ninst: int = 64, seed: int = SEED):
    """64-instance dam-break ensemble: instance i draws, from seed+i, a cubic grid
    N in {32,48,64,80,96}, liquid block fractions of the 0.1 m edge in
    U[0.2,0.6] (x, y) and U[0.3,0.8] (z), gas density U[25,100], sigma
    U[0.003,0.02]; liquid 1000 kg/m^3, g = (0,0,-9.81), dt = 4e-6 s."""
    prms, boxes = [], []
    for i in range(ninst):
        r = np.random.default_rng(seed + i)
        n = int(r.choice([32, 48, 64, 80, 96]))
        fx, fy, fz = r.uniform(0.2, 0.6), r.uniform(0.2, 0.6), r.uniform(0.3, 0.8)
        rho_g = float(r.uniform(25.0, 100.0))
        sigma = float(r.uniform(0.003, 0.02))
        prms.append(vof.default_params(n=n, g=(0.0, 0.0, -9.81), rho_g=rho_g, sigma=sigma))
        boxes.append((0.1 * fx, 0.1 * fy, 0.1 * fz))
    return prms, boxes


def ripple_height(n: int, L: float = 0.1, depth: float = 0.06, seed: int = SEED, modes: int = 4) -> np.ndarray:
    """Free-surface height h(x, y) at the n x n column centres, shape (n, n)
    indexed [j, i]: depth + (2 dx / modes) * sum_m sin(2 pi (x cos a_m +
    y sin a_m) / lam_m + phi_m), with wavelengths lam_m ~ U[8, 32] cells,
    directions a_m and phases phi_m ~ U[0, 2 pi) drawn from `seed`."""
    r = np.random.default_rng(seed)
    dx = L / n
    lam = r.uniform(8.0, 32.0, modes) * dx
    ang = r.uniform(0.0, 2.0 * np.pi, modes)
    phi = r.uniform(0.0, 2.0 * np.pi, modes)
    c = (np.arange(n) + 0.5) * dx
    x, y = c[None, :], c[:, None]
    h = np.zeros((n, n))
    for m in range(modes):
        h += np.sin(2.0 * np.pi * (x * np.cos(ang[m]) + y * np.sin(ang[m])) / lam[m] + phi[m])
    return depth + (2.0 * dx / modes) * h


def ripple_F(n: int, L: float = 0.1, depth: float = 0.06, seed: int = SEED, modes: int = 4,
             dtype=np.float64) -> np.ndarray:
    """Volume fraction of the ripple layer on the stored (n+2)^3 grid: interior
    cell (i, j, k) holds clip((h(i, j) - (k-1) dz) / dz, 0, 1) (the column
    under a flat surface at the column-centre height); every ghost equals the
    nearest interior cell, as the reference's F boundary condition leaves it."""
    dz = L / n
    h = ripple_height(n, L, depth, seed, modes)
    zlo = (np.arange(n) * dz)[:, None, None]
    Fi = np.clip((h[None, :, :] - zlo) / dz, 0.0, 1.0)
    return np.pad(Fi, 1, mode="edge").astype(dtype)


def field_errors(a: np.ndarray, ref: np.ndarray, interior: bool = True) -> dict:
    """max |a - ref| and ||a - ref||_2 / ||ref||_2 in FP64, over interior cells."""
    a = np.asarray(a, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    if interior:
        a, ref = a[1:-1, 1:-1, 1:-1], ref[1:-1, 1:-1, 1:-1]
    d = a - ref
    nr = float(np.sqrt(np.sum(ref * ref)))
    return {"max_abs": float(np.max(np.abs(d))), "rel_l2": float(np.sqrt(np.sum(d * d)) / nr) if nr > 0 else 0.0}


def drop_F(n: int, center=(0.05, 0.05, 0.012), semi_axes=(0.025, 0.025, 0.012), L: float = 0.1, sub: int = 4,
           dtype=np.float64) -> np.ndarray:
    """Volume fraction of an ellipsoidal drop on the stored (n+2)^3 grid, each
    interior cell sampled sub^3 times; ghosts copy the nearest interior cell.
    The default drop is flattened against the floor (its lowest point at
    z = 0), so surface tension pulls it rounder and raises its centroid."""
    h = L / n
    off = (np.arange(sub) + 0.5) / sub
    c = (np.arange(n)[:, None] + off[None, :]).ravel() * h
    qx = ((c - center[0]) / semi_axes[0]) ** 2
    qy = ((c - center[1]) / semi_axes[1]) ** 2
    qz = ((c - center[2]) / semi_axes[2]) ** 2
    inside = (qz[:, None, None] + qy[None, :, None] + qx[None, None, :]) <= 1.0
    f = inside.reshape(n, sub, n, sub, n, sub).mean(axis=(1, 3, 5))
    return np.pad(f, 1, mode="edge").astype(dtype)
