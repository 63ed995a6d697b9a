"""FP64 taped reference of the VOF step (test infrastructure).

A vectorised PyTorch restatement of the reference step (voftrial1.cu:1361-1401),
in the same operation order as the CPU oracle, so that torch.autograd gives a
taped FP64 reference gradient of the objective.  Checked against the oracle's
forward (tests/test_adjoint_ref.py) before being trusted as a gradient oracle.

Layout: tensors of shape (nz+2, ny+2, nx+2), x fastest, one ghost layer.
"""
from __future__ import annotations

import torch

T = torch.float64


class Prm:
    def __init__(self, nx, ny, nz, L=(0.1, 0.1, 0.1), rho_l=1000.0, rho_g=50.0, nu_l=1e-6, nu_g=1.5e-5,
                 g=(0.0, 0.0, -9.81), dt=4e-6, sweeps=10, csf=False, sigma=0.0):
        self.csf, self.sigma = csf, sigma  # surface tension (extension); sigma may be a 0-d tensor
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dx, self.dy, self.dz = L[0] / nx, L[1] / ny, L[2] / nz
        self.dxi, self.dyi, self.dzi = 1.0 / self.dx, 1.0 / self.dy, 1.0 / self.dz
        self.vol = self.dx * self.dy * self.dz
        self.rho_g, self.nu_l, self.nu_g = rho_g, nu_l, nu_g
        self.rho_l = rho_l  # float or 0-d tensor (differentiable)
        self.g = g
        self.dt = dt
        self.sweeps = sweeps


def clamp01(v):
    one = torch.ones((), dtype=v.dtype)
    zero = torch.zeros((), dtype=v.dtype)
    return ((one + v) - torch.maximum(one, v)) - torch.minimum(zero, v)


class _Sqrt(torch.autograd.Function):
    """Correctly rounded FP64 square root (numpy's), derivative 1/(2 sqrt x).

    CPU torch.sqrt in FP64 is not correctly rounded (it differs from IEEE sqrt
    by one ulp on some inputs), while the oracle's std::sqrt and the device
    sqrt are; this keeps the restatement bitwise with both."""

    @staticmethod
    def forward(ctx, x):
        import numpy as np
        out = torch.from_numpy(np.sqrt(x.detach().cpu().numpy())).to(x.dtype)
        ctx.save_for_backward(out)
        return out

    @staticmethod
    def backward(ctx, g):
        (out,) = ctx.saved_tensors
        return g * 0.5 / out


def sqrt_rn(x):
    return _Sqrt.apply(x)


def interior(a):
    return a[1:-1, 1:-1, 1:-1]


def put(base, region, value):
    out = base.clone()
    out[region] = value
    return out


def mom(P, c, xm, xp, ym, yp, zm, zp, ax, ay, az, n0, g):
    ddx = torch.where(ax > 0, (c - xm) * P.dxi, (xp - c) * P.dxi)
    ddy = torch.where(ay > 0, (c - ym) * P.dyi, (yp - c) * P.dyi)
    ddz = torch.where(az > 0, (c - zm) * P.dzi, (zp - c) * P.dzi)
    visc = (n0 * (xm - 2.0 * c + xp) * P.dxi * P.dxi + n0 * (ym - 2.0 * c + yp) * P.dyi * P.dyi +
            n0 * (zm - 2.0 * c + zp) * P.dzi * P.dzi)
    adv = -ax * ddx - ay * ddy - az * ddz
    return c + P.dt * (visc + adv + g)


def predictor(P, u, v, w, nu):
    """advect_{u,v,w}_kernel (:230-399); faces on walls stay zero."""
    K, J = slice(1, -1), slice(1, -1)
    # u* on faces i = 2..nx
    I2 = slice(2, -1)
    uc = u[K, J, I2]
    vh = 0.25 * (v[K, J, 1:-2] + v[K, 2:, 1:-2] + v[K, J, I2] + v[K, 2:, I2])
    wh = 0.25 * (w[K, J, 1:-2] + w[2:, J, 1:-2] + w[K, J, I2] + w[2:, J, I2])
    us_v = mom(P, uc, u[K, J, 1:-2], u[K, J, 3:], u[K, :-2, I2], u[K, 2:, I2], u[:-2, J, I2], u[2:, J, I2],
               uc, vh, wh, nu[K, J, I2], P.g[0])
    us = put(torch.zeros_like(u), (K, J, I2), us_v)
    # v* on faces j = 2..ny
    J2 = slice(2, -1)
    I = slice(1, -1)
    vc = v[K, J2, I]
    uh = 0.25 * (u[K, 1:-2, I] + u[K, J2, I] + u[K, 1:-2, 2:] + u[K, J2, 2:])
    wh = 0.25 * (w[2:, 1:-2, I] + w[K, 1:-2, I] + w[K, J2, I] + w[2:, J2, I])
    vs_v = mom(P, vc, v[K, J2, :-2], v[K, J2, 2:], v[K, 1:-2, I], v[K, 3:, I], v[:-2, J2, I], v[2:, J2, I],
               uh, vc, wh, nu[K, J2, I], P.g[1])
    vs = put(torch.zeros_like(v), (K, J2, I), vs_v)
    # w* on faces k = 2..nz
    K2 = slice(2, -1)
    wc = w[K2, J, I]
    uh = 0.25 * (u[1:-2, J, 2:] + u[1:-2, J, I] + u[K2, J, 2:] + u[K2, J, I])
    vh = 0.25 * (v[1:-2, 2:, I] + v[1:-2, J, I] + v[K2, J, I] + v[K2, 2:, I])
    ws_v = mom(P, wc, w[K2, J, :-2], w[K2, J, 2:], w[K2, :-2, I], w[K2, 2:, I], w[1:-2, J, I], w[3:, J, I],
               uh, vh, wc, nu[K2, J, I], P.g[2])
    ws = put(torch.zeros_like(w), (K2, J, I), ws_v)
    return us, vs, ws


def rhs_of(P, us, vs, ws, rho):
    """rho/dt * div(u*) (:418-422) on interior cells."""
    K = J = I = slice(1, -1)
    div = ((us[K, J, 2:] - us[K, J, I]) * P.dxi + (vs[K, 2:, I] - vs[K, J, I]) * P.dyi +
           (ws[2:, J, I] - ws[K, J, I]) * P.dzi)
    return rho[K, J, I] / P.dt * div


def jacobi_coeffs(P):
    nx, ny, nz = P.nx, P.ny, P.nz
    i = torch.arange(1, nx + 1).view(1, 1, -1)
    j = torch.arange(1, ny + 1).view(1, -1, 1)
    k = torch.arange(1, nz + 1).view(-1, 1, 1)
    # FP64 scalars as tensors: torch.where on two Python floats would build a
    # float32 tensor and round the coefficients.
    ax, ay, az = (torch.tensor(x, dtype=T) for x in (P.dxi * P.dxi, P.dyi * P.dyi, P.dzi * P.dzi))
    z = torch.zeros((), dtype=T)
    sh = (nz, ny, nx)
    ae = torch.where(i != nx, ax, z).expand(sh)
    aw = torch.where(i != 1, ax, z).expand(sh)
    an = torch.where(j != ny, ay, z).expand(sh)
    as_ = torch.where(j != 1, ay, z).expand(sh)
    af = torch.where(k != nz, az, z).expand(sh)
    ab = torch.where(k != 1, az, z).expand(sh)
    ap = -(ae + aw + an + as_ + af + ab)
    return ae, aw, an, as_, af, ab, ap


def jacobi(P, p, rhs, coeffs):
    ae, aw, an, as_, af, ab, ap = coeffs
    K = J = I = slice(1, -1)
    val = (rhs - ae * p[K, J, 2:] - aw * p[K, J, :-2] - an * p[K, 2:, I] - as_ * p[K, :-2, I] -
           af * p[2:, J, I] - ab * p[:-2, J, I]) / ap
    return put(torch.zeros_like(p), (K, J, I), val)


def correction(P, us, vs, ws, p, rho):
    """update_{u,v,w}_kernel (:441-493) then the velocity BC closed forms."""
    K = J = I = slice(1, -1)
    dt = torch.tensor(P.dt, dtype=T)  # python_float / tensor would multiply by a rounded reciprocal
    r = 0.5 * (rho[K, J, 2:-1] + rho[K, J, 1:-2])
    u_in = us[K, J, 2:-1] - dt / r * (p[K, J, 2:-1] - p[K, J, 1:-2]) * P.dxi
    r = 0.5 * (rho[K, 2:-1, I] + rho[K, 1:-2, I])
    v_in = vs[K, 2:-1, I] - dt / r * (p[K, 2:-1, I] - p[K, 1:-2, I]) * P.dyi
    r = 0.5 * (rho[2:-1, J, I] + rho[1:-2, J, I])
    w_in = ws[2:-1, J, I] - dt / r * (p[2:-1, J, I] - p[1:-2, J, I]) * P.dzi
    return u_in, v_in, w_in


def bc_velocity(P, u_old, v_old, w_old, u_in, v_in, w_in):
    """Closed form of set_bc_vel_j/i/k (:959-1051): normal wall faces 0,
    tangential ghosts = nearest interior face, plane 0 of the normal axis kept."""
    nx, ny, nz = P.nx, P.ny, P.nz
    cj = torch.clamp(torch.arange(ny + 2), 1, ny)
    ck = torch.clamp(torch.arange(nz + 2), 1, nz)
    ci = torch.clamp(torch.arange(nx + 2), 1, nx)
    # u: planes i = 2..nx from u_in at (i, cj, ck); i = 1, nx+1 zero; i = 0 from u_old(0, cj, ck)
    body = torch.zeros(nz, ny, nx + 2, dtype=T)
    body = put(body, (slice(None), slice(None), slice(2, nx + 1)), u_in)
    body = put(body, (slice(None), slice(None), slice(0, 1)), u_old[1:-1, 1:-1, 0:1])
    u = body[ck - 1][:, cj - 1]
    # v
    body = torch.zeros(nz, ny + 2, nx, dtype=T)
    body = put(body, (slice(None), slice(2, ny + 1), slice(None)), v_in)
    body = put(body, (slice(None), slice(0, 1), slice(None)), v_old[1:-1, 0:1, 1:-1])
    v = body[ck - 1][:, :, ci - 1]
    # w
    body = torch.zeros(nz + 2, ny, nx, dtype=T)
    body = put(body, (slice(2, nz + 1), slice(None), slice(None)), w_in)
    body = put(body, (slice(0, 1), slice(None), slice(None)), w_old[0:1, 1:-1, 1:-1])
    w = body[:, cj - 1][:, :, ci - 1]
    return u, v, w


def fct_line(F, vel, P, axis):
    """One directional FCT sweep (:509-905) along `axis` (0 x, 1 y, 2 z); returns
    the new interior values.  Lines run along the last dimension after a move."""
    dim = 2 - axis
    Fm = torch.movedim(F, dim, -1)
    Vm = torch.movedim(vel, dim, -1)
    Fm = Fm[tuple(slice(1, -1) for _ in range(2))]  # interior of the two transverse axes
    Vm = Vm[tuple(slice(1, -1) for _ in range(2))]
    dt, vol = P.dt, P.vol
    zero = torch.zeros((), dtype=T)
    one = torch.ones((), dtype=T)
    if axis == 0:
        kdv, qf, s4 = dt * P.dy * P.dz, P.dx, P.dy
        s1 = lambda d: d * P.dy * P.dz / vol
    elif axis == 1:
        kdv, qf, s4 = dt * P.dx * P.dz, P.dx, P.dy
        dxdy = P.dx * P.dy
        s1 = lambda d: d * P.dy / dxdy
    else:
        kdv, qf, s4 = dt * P.dx * P.dy, P.dz, P.dz
        s1 = lambda d: d * P.dy * P.dx / vol
    vf = Vm[..., 1:]            # faces 1..n+1
    Flo = Fm[..., :-1]          # F(f-1)
    Fhi = Fm[..., 1:]           # F(f)
    vdt = vf * dt
    L = torch.where(vf >= 0, vdt * Flo, vdt * Fhi)
    H = torch.where(vf <= 0, vdt * Flo, vdt * Fhi)
    a = H - L                   # index f-1 for face f = 1..n+1
    dv = vol - kdv * (Vm[..., 2:] - Vm[..., 1:-1])  # cells 1..n
    Fc = Fm[..., 1:-1]
    out = (Fc + s1(L[..., :-1] - L[..., 1:])) * vol / dv
    Ftd = torch.where((out > 1) | (out < 0), clamp01(out), out)
    pad = torch.nn.functional.pad
    Fe = pad(Ftd, (1, 1))       # ghost Ftd read as 0
    fmax = torch.maximum(Ftd, torch.maximum(Fe[..., :-2], Fe[..., 2:]))
    fmin = torch.minimum(Ftd, torch.minimum(Fe[..., :-2], Fe[..., 2:]))
    a_lo, a_hi = a[..., :-1], a[..., 1:]   # faces c and c+1 of cell c
    pp = torch.maximum(zero, a_lo) - torch.minimum(zero, a_hi)
    qp = (fmax - Ftd) * qf
    pps = torch.where(pp > 0, pp, one)
    rp = torch.where(pp > 0, torch.minimum(one, qp / pps), zero)
    pm = torch.maximum(zero, a_hi) - torch.minimum(zero, a_lo)
    qm = (Ftd - fmin) * qf
    pms = torch.where(pm > 0, pm, one)
    rm = torch.where(pm > 0, torch.minimum(one, qm / pms), zero)
    rpe, rme = pad(rp, (1, 1)), pad(rm, (1, 1))   # ghost rp, rm = 0
    # limiter on faces 2..n+1 (face f between cells f-1 and f); face 1 is zero
    af = a[..., 1:]
    Cf = torch.where(af >= 0, torch.minimum(rpe[..., 2:], rme[..., 1:-1]),
                     torch.minimum(rpe[..., 1:-1], rme[..., 2:]))
    C = pad(Cf, (1, 0))
    aC = a * C
    Fn = clamp01(Ftd - ((aC[..., 1:] - aC[..., :-1]) / s4) * vol / dv)
    return Fn  # (transverse interior..., n) along the moved axis


def fct_sweep(F, vel, P, axis):
    Fn = fct_line(F, vel, P, axis)
    dim = 2 - axis
    Fm = torch.movedim(F, dim, -1).clone()
    Fm[1:-1, 1:-1, 1:-1] = Fn
    return torch.movedim(Fm, -1, dim)


def ghost_F(P, F_int):
    ci = torch.clamp(torch.arange(P.nx + 2), 1, P.nx) - 1
    cj = torch.clamp(torch.arange(P.ny + 2), 1, P.ny) - 1
    ck = torch.clamp(torch.arange(P.nz + 2), 1, P.nz) - 1
    return F_int[ck][:, cj][:, :, ci]


def csf(P, F, rho, us, vs, ws):
    """CSF surface tension (extension, VOF_FLAG_CSF): F smoothed by the separable
    (1/4,1/2,1/4) filter on interior cells, unit normals at vertices from the
    smoothed field (indices clamped to the interior), kappa = -div of the unit
    normal per cell, face acceleration sigma kappa_f dF/dn / rho_f added to u*."""
    rx = (0.25 * F[:, :, :-2] + 0.5 * F[:, :, 1:-1]) + 0.25 * F[:, :, 2:]
    ry = (0.25 * rx[:, :-2, :] + 0.5 * rx[:, 1:-1, :]) + 0.25 * rx[:, 2:, :]
    fs = (0.25 * ry[:-2] + 0.5 * ry[1:-1]) + 0.25 * ry[2:]
    Fp = torch.nn.functional.pad(fs[None, None], (1, 1, 1, 1, 1, 1), mode="replicate")[0, 0]
    f000, f100, f010, f110 = Fp[:-1, :-1, :-1], Fp[:-1, :-1, 1:], Fp[:-1, 1:, :-1], Fp[:-1, 1:, 1:]
    f001, f101, f011, f111 = Fp[1:, :-1, :-1], Fp[1:, :-1, 1:], Fp[1:, 1:, :-1], Fp[1:, 1:, 1:]
    nx_ = (((f100 + f110) + (f101 + f111)) - ((f000 + f010) + (f001 + f011))) * (0.25 * P.dxi)
    ny_ = (((f010 + f110) + (f011 + f111)) - ((f000 + f100) + (f001 + f101))) * (0.25 * P.dyi)
    nz_ = (((f001 + f101) + (f011 + f111)) - ((f000 + f100) + (f010 + f110))) * (0.25 * P.dzi)
    s = (nx_ * nx_ + ny_ * ny_) + nz_ * nz_
    mag = torch.where(s > 0, sqrt_rn(torch.where(s > 0, s, torch.ones((), dtype=T))), torch.zeros((), dtype=T))
    inv = torch.ones((), dtype=T) / (mag + 1e-8 * P.dxi)
    mx, my, mz = nx_ * inv, ny_ * inv, nz_ * inv
    n3 = (P.nz, P.ny, P.nx)

    def v(m, di, dj, dk):
        return m[dk:dk + n3[0], dj:dj + n3[1], di:di + n3[2]]

    kx = (((v(mx, 1, 0, 0) + v(mx, 1, 1, 0)) + (v(mx, 1, 0, 1) + v(mx, 1, 1, 1))) -
          ((v(mx, 0, 0, 0) + v(mx, 0, 1, 0)) + (v(mx, 0, 0, 1) + v(mx, 0, 1, 1)))) * (0.25 * P.dxi)
    ky = (((v(my, 0, 1, 0) + v(my, 1, 1, 0)) + (v(my, 0, 1, 1) + v(my, 1, 1, 1))) -
          ((v(my, 0, 0, 0) + v(my, 1, 0, 0)) + (v(my, 0, 0, 1) + v(my, 1, 0, 1)))) * (0.25 * P.dyi)
    kz = (((v(mz, 0, 0, 1) + v(mz, 1, 0, 1)) + (v(mz, 0, 1, 1) + v(mz, 1, 1, 1))) -
          ((v(mz, 0, 0, 0) + v(mz, 1, 0, 0)) + (v(mz, 0, 1, 0) + v(mz, 1, 1, 0)))) * (0.25 * P.dzi)
    kap = torch.nn.functional.pad(-((kx + ky) + kz), (1, 1, 1, 1, 1, 1))

    K = J = I = slice(1, -1)

    def acc(c, m, h):
        kf = 0.5 * (kap[m] + kap[c])
        gF = (F[c] - F[m]) * h
        rf = 0.5 * (rho[m] + rho[c])
        return P.sigma * kf * gF / rf

    I2 = slice(2, -1)
    us = put(us, (K, J, I2), us[K, J, I2] + P.dt * acc((K, J, I2), (K, J, slice(1, -2)), P.dxi))
    vs = put(vs, (K, I2, I), vs[K, I2, I] + P.dt * acc((K, I2, I), (K, slice(1, -2), I), P.dyi))
    ws = put(ws, (I2, J, I), ws[I2, J, I] + P.dt * acc((I2, J, I), (slice(1, -2), J, I), P.dzi))
    return us, vs, ws


def step(P, st, istep, coeffs=None):
    F, u, v, w, p = st
    coeffs = coeffs if coeffs is not None else jacobi_coeffs(P)
    f = clamp01(F)
    rho = P.rho_g * (1.0 - f) + P.rho_l * f
    nu = P.nu_l * f + P.nu_g * (1.0 - f)
    us, vs, ws = predictor(P, u, v, w, nu)
    if P.csf:
        us, vs, ws = csf(P, F, rho, us, vs, ws)
    b = rhs_of(P, us, vs, ws, rho)
    for _ in range(P.sweeps):
        p = jacobi(P, p, b, coeffs)
    u_in, v_in, w_in = correction(P, us, vs, ws, p, rho)
    u, v, w = bc_velocity(P, u, v, w, u_in, v_in, w_in)
    order = {0: (0, 1, 2), 1: (1, 2, 0), 2: (2, 0, 1)}[istep % 3]
    vel = (u, v, w)
    for ax in order:
        F = fct_sweep(F, vel[ax], P, ax)
    Fi = clamp01(interior(F))
    F = ghost_F(P, Fi)
    return F, u, v, w, p


def centroid_height(P, F):
    """J = sum F z / sum F over interior cells, z the cell-centre height."""
    z = ((torch.arange(1, P.nz + 1, dtype=T) - 0.5) * P.dz).view(-1, 1, 1)
    Fi = interior(F)
    return (Fi * z).sum() / Fi.sum()
