// vof_oracle.cpp — CPU restatement of the reference VOF timestep (voftrial1.cu).
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it is loaded by
// tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
// nothing in the product (the HIP library, the CLI) links or calls it.
//
// Each function restates one reference kernel body as an OpenMP host loop over
// the same index range, with the same floating-point operation order (built with
// -ffp-contract=off so every source operation is one IEEE rounding), the same
// scratch arrays and the same zero-initialised never-written entries.  It is
// templated on the real type: float reproduces the reference, double is the
// "all-FP64" run.  Citations are voftrial1.cu line numbers.
//
// Parity pins (tests/test_oracle.py, tests/test_oracle_pins.py): clamp01f KATs,
// IC cell counts, zero-velocity, linear- and quadratic-field predictor identities
// on non-cubic grids, axis-rotation equivariance of every y/z operator (predictor,
// rhs, Jacobi, correction, FCT), Jacobi shift commutation, direct-solve fixed point
// and one-sweep matrix form with a right-hand side, hand-derived FCT known answers
// (downwind anti-diffusion, limiter, dilation, stage-1 clip), interface sharpness,
// FCT boundedness / u=0 identity / mirror symmetry / telescoping of sum(Ftd*dv),
// BC closed forms, diagnostics sums, VTK bytes.  tests/test_oracle_mutants.py
// plants plausible mistakes in copies of this file and checks each fails a pin.
#include "../include/vof_capi.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <new>
#include <vector>

#include <omp.h>

namespace {

// voftrial1.cu:52-56 — a + b + c - max - min, evaluated left to right.
template <class R>
inline R median3(R a, R b, R c) {
    const R mx = std::fmax(std::fmax(a, b), c);
    const R mn = std::fmin(std::fmin(a, b), c);
    return a + b + c - mx - mn;
}
// voftrial1.cu:58-60
template <class R>
inline R clamp01(R v) { return median3<R>(R(0), R(1), v); }
// voftrial1.cu:62-64
inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

enum OrcField {
    O_F = 0, O_FTD, O_AX, O_AY, O_AZ, O_CX, O_CY, O_CZ, O_RP, O_RM,
    O_U, O_V, O_W, O_US, O_VS, O_WS, O_P, O_PT, O_RHO, O_NU, O_FSM, O_KAPPA,
    O_S1, O_S4,  // stage-1 value before its clip, stage-4 value before clamp01 (volume accounting only)
    O_NFIELDS
};

struct OrcBase {
    virtual ~OrcBase() {}
    int prec = 4;
};

template <class R>
struct Orc : OrcBase {
    // Params (voftrial1.cu:27-41), derived as in :1262-1291.
    int nx, ny, nz, imin, imax, jmin, jmax, kmin, kmax;
    R Lx, Ly, Lz, rho_l, rho_g, nu_l, nu_g, sigma, gx, gy, gz, dt, dx, dy, dz, dxi, dyi, dzi;
    int NX, NY, NZ;
    size_t ntot;
    int sweeps;
    uint32_t flags;
    double fparams[3];  // Lx, Ly, Lz as given (for the float IC predicate)
    std::vector<R> a[O_NFIELDS];
    int pcur = O_P;  // which buffer holds p after the ping-pong (:1374)
    double last_pmean = 0.0;
    // Liquid-volume accounting (test infrastructure; off unless asked for):
    // where the change of sum(F vol) over a step comes from.
    bool stats_on = false;
    double st_clip1 = 0, st_clamp4 = 0, st_clip4 = 0, st_dil = 0, st_post = 0, st_div = 0;

    inline size_t id(int i, int j, int k) const {  // voftrial1.cu:66-70
        return (static_cast<size_t>(k) * static_cast<size_t>(NY) + static_cast<size_t>(j)) *
                   static_cast<size_t>(NX) + static_cast<size_t>(i);
    }

    explicit Orc(const vof_params& p) {
        prec = sizeof(R);
        nx = p.nx; ny = p.ny; nz = p.nz;
        imin = 1; imax = imin + nx - 1;
        jmin = 1; jmax = jmin + ny - 1;
        kmin = 1; kmax = kmin + nz - 1;
        Lx = R(p.Lx); Ly = R(p.Ly); Lz = R(p.Lz);
        rho_l = R(p.rho_l); rho_g = R(p.rho_g);
        nu_l = R(p.nu_l); nu_g = R(p.nu_g);
        sigma = R(p.sigma);
        gx = R(p.gx); gy = R(p.gy); gz = R(p.gz);
        dt = R(p.dt);
        dx = Lx / static_cast<R>(nx);
        dy = Ly / static_cast<R>(ny);
        dz = Lz / static_cast<R>(nz);
        dxi = R(1) / dx; dyi = R(1) / dy; dzi = R(1) / dz;
        NX = nx + 2; NY = ny + 2; NZ = nz + 2;
        ntot = static_cast<size_t>(NX) * NY * NZ;
        sweeps = p.jacobi_sweeps;
        flags = p.flags;
        fparams[0] = p.Lx; fparams[1] = p.Ly; fparams[2] = p.Lz;
        for (int f = 0; f < O_NFIELDS; ++f) a[f].assign(ntot, R(0));  // alloc_field: zeroed (:1200-1203)
    }

    R* fld(int f) { return a[f].data(); }

    // ---- set_init_F_kernel (:72-93), evaluated in float as the reference does,
    //      whatever R is, so FP32 and FP64 runs start from the same F.
    void init_ic(int ic) {
        std::fill(a[O_F].begin(), a[O_F].end(), R(0));
        if (ic == 1) {
            const float fLx = float(fparams[0]), fLy = float(fparams[1]), fLz = float(fparams[2]);
            const float fdx = fLx / float(nx), fdy = fLy / float(ny), fdz = fLz / float(nz);
            R* F = fld(O_F);
#pragma omp parallel for
            for (int k = 0; k < NZ; ++k)
                for (int j = 0; j < NY; ++j)
                    for (int i = 0; i < NX; ++i) {
                        const float x = float(clampi(i - 1, 0, nx)) * fdx;
                        const float y = float(clampi(j - 1, 0, ny)) * fdy;
                        const float z = float(clampi(k - 1, 0, nz)) * fdz;
                        if ((x >= 0.0f && x <= fLx / 3.0f) && (y >= 0.0f && y <= fLy / 2.0f) &&
                            (z >= 0.0f && z <= fLz / 3.0f))
                            F[id(i, j, k)] = R(1);
                    }
        }
        bc_F();  // :1335
    }

    // Extension IC: lower-face strict '<' box, integer rule (no reference counterpart).
    void init_box(double bx, double by, double bz) {
        auto count = [](double b, int n, double L) {
            const double q = b * double(n) / L;
            int c = int(std::ceil(q - 1e-9));
            return std::max(0, std::min(n, c));
        };
        const int cx_ = count(bx, nx, fparams[0]), cy_ = count(by, ny, fparams[1]),
                  cz_ = count(bz, nz, fparams[2]);
        R* F = fld(O_F);
        std::fill(a[O_F].begin(), a[O_F].end(), R(0));
        for (int k = 1; k <= cz_; ++k)
            for (int j = 1; j <= cy_; ++j)
                for (int i = 1; i <= cx_; ++i) F[id(i, j, k)] = R(1);
        bc_F();
    }

    // ---- cal_nu_rho_kernel (:95-104), all nTot entries.
    void nu_rho() {
        const R* F = fld(O_F);
        R* rho = fld(O_RHO);
        R* nu = fld(O_NU);
#pragma omp parallel for
        for (long long t = 0; t < (long long)ntot; ++t) {
            const R f = clamp01(F[t]);
            rho[t] = rho_g * (R(1) - f) + rho_l * f;
            nu[t] = nu_l * f + nu_g * (R(1) - f);
        }
    }

    // ---- Density-weighted viscous term (VOF_FLAG_RHO_PREDICTOR; extension,
    //      no reference counterpart: the reference's predictor takes rho and
    //      never reads it, :233, :290, :347, and uses nu of one cell,
    //      :277-280).  Face phi at index c between cell a = c - off[d] (low)
    //      and cell b = c along its own axis d:
    //        visc = ((T_x + T_y) + T_z) / (0.5 (rho_a + rho_b)),
    //        T_t  = (m+ (phi(c+e_t) - phi) - m- (phi - phi(c-e_t))) * h_t^-1 * h_t^-1,
    //      with mu = rho nu per cell; along d, m+ = mu_b and m- = mu_a; across,
    //      m+- = 0.25 ((mu_a + mu_b) + (mu_a(+-e_t) + mu_b(+-e_t))) ordered
    //      low edge first: m- = 0.25 ((mu_a(-e_t) + mu_b(-e_t)) + (mu_a + mu_b)).
    //      This is div(mu grad phi) / rho in flux form: the fluxes through the
    //      edges between faces telescope, so sum(rho_f visc) over faces only
    //      sees the walls.  The association is the contract the HIP kernel repeats.
    inline R mu_at(size_t q) const { return a[O_RHO][q] * a[O_NU][q]; }
    R visc_rho(const R* phi, size_t c, int d) const {
        const size_t off[3] = {1, size_t(NX), size_t(NX) * size_t(NY)};
        const R hi[3] = {dxi, dyi, dzi};
        const size_t ca = c - off[d], cb = c;
        const R ma = mu_at(ca), mb = mu_at(cb);
        R T[3];
        for (int t = 0; t < 3; ++t) {
            const size_t o = off[t];
            R mp, mm;
            if (t == d) {
                mp = mb;
                mm = ma;
            } else {
                mp = R(0.25) * ((ma + mb) + (mu_at(ca + o) + mu_at(cb + o)));
                mm = R(0.25) * ((mu_at(ca - o) + mu_at(cb - o)) + (ma + mb));
            }
            T[t] = (mp * (phi[c + o] - phi[c]) - mm * (phi[c] - phi[c - o])) * hi[t] * hi[t];
        }
        const R rf = R(0.5) * (a[O_RHO][ca] + a[O_RHO][cb]);
        return ((T[0] + T[1]) + T[2]) / rf;
    }

    // ---- advect_u_kernel (:230-285): faces i = imin+1..imax.
    void advect_u() {
        const R *u = fld(O_U), *v = fld(O_V), *w = fld(O_W), *nu = fld(O_NU);
        R* us = fld(O_US);
        const bool rho_visc = (flags & VOF_FLAG_RHO_PREDICTOR) != 0;
#pragma omp parallel for collapse(2)
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin + 1; i <= imax; ++i) {
                    const size_t c = id(i, j, k);
                    const R vh = R(0.25) * (v[id(i - 1, j, k)] + v[id(i - 1, j + 1, k)] +
                                            v[id(i, j, k)] + v[id(i, j + 1, k)]);
                    const R wh = R(0.25) * (w[id(i - 1, j, k)] + w[id(i - 1, j, k + 1)] +
                                            w[id(i, j, k)] + w[id(i, j, k + 1)]);
                    const R uc = u[c];
                    const R uxm = u[id(i - 1, j, k)], uxp = u[id(i + 1, j, k)];
                    const R uym = u[id(i, j - 1, k)], uyp = u[id(i, j + 1, k)];
                    const R uzm = u[id(i, j, k - 1)], uzp = u[id(i, j, k + 1)];
                    const R ddx = (uc > R(0)) ? (uc - uxm) * dxi : (uxp - uc) * dxi;
                    const R ddy = (vh > R(0)) ? (uc - uym) * dyi : (uyp - uc) * dyi;
                    const R ddz = (wh > R(0)) ? (uc - uzm) * dzi : (uzp - uc) * dzi;
                    const R n0 = nu[c];
                    const R visc = rho_visc ? visc_rho(u, c, 0)
                                            : n0 * (uxm - R(2) * uc + uxp) * dxi * dxi +
                                                  n0 * (uym - R(2) * uc + uyp) * dyi * dyi +
                                                  n0 * (uzm - R(2) * uc + uzp) * dzi * dzi;
                    const R adv = -uc * ddx - vh * ddy - wh * ddz;
                    us[c] = uc + dt * (visc + adv + gx);
                }
    }

    // ---- advect_v_kernel (:287-342): faces j = jmin+1..jmax.
    void advect_v() {
        const R *u = fld(O_U), *v = fld(O_V), *w = fld(O_W), *nu = fld(O_NU);
        R* vs = fld(O_VS);
        const bool rho_visc = (flags & VOF_FLAG_RHO_PREDICTOR) != 0;
#pragma omp parallel for collapse(2)
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin + 1; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    const size_t c = id(i, j, k);
                    const R uh = R(0.25) * (u[id(i, j - 1, k)] + u[id(i, j, k)] +
                                            u[id(i + 1, j - 1, k)] + u[id(i + 1, j, k)]);
                    const R wh = R(0.25) * (w[id(i, j - 1, k + 1)] + w[id(i, j - 1, k)] +
                                            w[id(i, j, k)] + w[id(i, j, k + 1)]);
                    const R vc = v[c];
                    const R vxm = v[id(i - 1, j, k)], vxp = v[id(i + 1, j, k)];
                    const R vym = v[id(i, j - 1, k)], vyp = v[id(i, j + 1, k)];
                    const R vzm = v[id(i, j, k - 1)], vzp = v[id(i, j, k + 1)];
                    const R ddx = (uh > R(0)) ? (vc - vxm) * dxi : (vxp - vc) * dxi;
                    const R ddy = (vc > R(0)) ? (vc - vym) * dyi : (vyp - vc) * dyi;
                    const R ddz = (wh > R(0)) ? (vc - vzm) * dzi : (vzp - vc) * dzi;
                    const R n0 = nu[c];
                    const R visc = rho_visc ? visc_rho(v, c, 1)
                                            : n0 * (vxm - R(2) * vc + vxp) * dxi * dxi +
                                                  n0 * (vym - R(2) * vc + vyp) * dyi * dyi +
                                                  n0 * (vzm - R(2) * vc + vzp) * dzi * dzi;
                    const R adv = -uh * ddx - vc * ddy - wh * ddz;
                    vs[c] = vc + dt * (visc + adv + gy);
                }
    }

    // ---- advect_w_kernel (:344-399): faces k = kmin+1..kmax.
    void advect_w() {
        const R *u = fld(O_U), *v = fld(O_V), *w = fld(O_W), *nu = fld(O_NU);
        R* ws = fld(O_WS);
        const bool rho_visc = (flags & VOF_FLAG_RHO_PREDICTOR) != 0;
#pragma omp parallel for collapse(2)
        for (int k = kmin + 1; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    const size_t c = id(i, j, k);
                    const R uh = R(0.25) * (u[id(i + 1, j, k - 1)] + u[id(i, j, k - 1)] +
                                            u[id(i + 1, j, k)] + u[id(i, j, k)]);
                    const R vh = R(0.25) * (v[id(i, j + 1, k - 1)] + v[id(i, j, k - 1)] +
                                            v[id(i, j, k)] + v[id(i, j + 1, k)]);
                    const R wc = w[c];
                    const R wxm = w[id(i - 1, j, k)], wxp = w[id(i + 1, j, k)];
                    const R wym = w[id(i, j - 1, k)], wyp = w[id(i, j + 1, k)];
                    const R wzm = w[id(i, j, k - 1)], wzp = w[id(i, j, k + 1)];
                    const R ddx = (uh > R(0)) ? (wc - wxm) * dxi : (wxp - wc) * dxi;
                    const R ddy = (vh > R(0)) ? (wc - wym) * dyi : (wyp - wc) * dyi;
                    const R ddz = (wc > R(0)) ? (wc - wzm) * dzi : (wzp - wc) * dzi;
                    const R n0 = nu[c];
                    const R visc = rho_visc ? visc_rho(w, c, 2)
                                            : n0 * (wxm - R(2) * wc + wxp) * dxi * dxi +
                                                  n0 * (wym - R(2) * wc + wyp) * dyi * dyi +
                                                  n0 * (wzm - R(2) * wc + wzp) * dzi * dzi;
                    const R adv = -uh * ddx - vh * ddy - wc * ddz;
                    ws[c] = wc + dt * (visc + adv + gz);
                }
    }

    // The Jacobi right-hand side exactly as the sweep forms it (:418-422).
    inline R rhs_at(int i, int j, int k) const {
        const R *us = a[O_US].data(), *vs = a[O_VS].data(), *ws = a[O_WS].data(),
                *rho = a[O_RHO].data();
        const size_t c = id(i, j, k);
        return rho[c] / dt *
               ((us[id(i + 1, j, k)] - us[c]) * dxi + (vs[id(i, j + 1, k)] - vs[c]) * dyi +
                (ws[id(i, j, k + 1)] - ws[c]) * dzi);
    }

    // ---- solve_p_jacobi_kernel (:401-439): one sweep p -> pt.
    void jacobi_sweep(const R* p, R* pt) {
#pragma omp parallel for collapse(2)
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    const R rhs = rhs_at(i, j, k);
                    const R ae = (i != imax) ? (dxi * dxi) : R(0);
                    const R aw = (i != imin) ? (dxi * dxi) : R(0);
                    const R an = (j != jmax) ? (dyi * dyi) : R(0);
                    const R as = (j != jmin) ? (dyi * dyi) : R(0);
                    const R af = (k != kmax) ? (dzi * dzi) : R(0);
                    const R ab = (k != kmin) ? (dzi * dzi) : R(0);
                    const R ap = -(ae + aw + an + as + af + ab);
                    pt[id(i, j, k)] = (rhs - ae * p[id(i + 1, j, k)] - aw * p[id(i - 1, j, k)] -
                                       an * p[id(i, j + 1, k)] - as * p[id(i, j - 1, k)] -
                                       af * p[id(i, j, k + 1)] - ab * p[id(i, j, k - 1)]) / ap;
                }
    }

    // :1371-1375 — sweeps with pointer ping-pong; pcur tracks the live buffer.
    void jacobi() {
        for (int it = 0; it < sweeps; ++it) {
            const int src = pcur, dst = (pcur == O_P) ? O_PT : O_P;
            jacobi_sweep(fld(src), fld(dst));
            pcur = dst;
        }
    }

    // FP64 mean of p over interior cells, optionally removed (extension flag).
    void remove_p_mean() {
        R* p = fld(pcur);
        double s = 0.0;
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) s += double(p[id(i, j, k)]);
        const double m = s / (double(nx) * ny * nz);
        last_pmean = m;
        const R mr = R(m);
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) p[id(i, j, k)] = p[id(i, j, k)] - mr;
    }

    // ---- update_{u,v,w}_kernel (:441-493).
    void update_uvw() {
        const R *rho = fld(O_RHO), *us = fld(O_US), *vs = fld(O_VS), *ws = fld(O_WS), *p = fld(pcur);
        R *u = fld(O_U), *v = fld(O_V), *w = fld(O_W);
#pragma omp parallel for collapse(2)
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin + 1; i <= imax; ++i) {
                    const size_t c = id(i, j, k), m = id(i - 1, j, k);
                    const R r = R(0.5) * (rho[c] + rho[m]);
                    u[c] = us[c] - dt / r * (p[c] - p[m]) * dxi;
                }
#pragma omp parallel for collapse(2)
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin + 1; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    const size_t c = id(i, j, k), m = id(i, j - 1, k);
                    const R r = R(0.5) * (rho[c] + rho[m]);
                    v[c] = vs[c] - dt / r * (p[c] - p[m]) * dyi;
                }
#pragma omp parallel for collapse(2)
        for (int k = kmin + 1; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    const size_t c = id(i, j, k), m = id(i, j, k - 1);
                    const R r = R(0.5) * (rho[c] + rho[m]);
                    w[c] = ws[c] - dt / r * (p[c] - p[m]) * dzi;
                }
    }

    // ---- apply_bc_velocity (:1107-1120): set_bc_vel_j, then _i, then _k
    //      (:959-1051), each over full (n+2)^2 planes.  The reference's
    //      tangential ghosts copy the interior (free slip, t = 1).  With
    //      VOF_FLAG_NO_SLIP (extension) they are the negated interior value
    //      (t = -1): the wall lies halfway between the ghost and the first
    //      interior face, so the linear interpolant vanishes on it.
    void bc_vel() {
        R *u = fld(O_U), *v = fld(O_V), *w = fld(O_W);
        const R t = (flags & VOF_FLAG_NO_SLIP) ? R(-1) : R(1);
        // j pass (:959-987)
#pragma omp parallel for
        for (int k = 0; k < NZ; ++k)
            for (int i = 0; i < NX; ++i) {
                const size_t gb = id(i, jmin - 1, k), ib = id(i, jmin, k);
                const size_t gt = id(i, jmax + 1, k), it = id(i, jmax, k);
                u[gb] = t * u[ib]; v[ib] = R(0); w[gb] = t * w[ib];
                u[gt] = t * u[it]; v[gt] = R(0); w[gt] = t * w[it];
            }
        // i pass (:989-1019)
#pragma omp parallel for
        for (int k = 0; k < NZ; ++k)
            for (int j = 0; j < NY; ++j) {
                const size_t gl = id(imin - 1, j, k), il = id(imin, j, k);
                const size_t gr = id(imax + 1, j, k), ir = id(imax, j, k);
                u[il] = R(0); v[gl] = t * v[il]; w[gl] = t * w[il];
                u[gr] = R(0); v[gr] = t * v[ir]; w[gr] = t * w[ir];
            }
        // k pass (:1021-1051)
#pragma omp parallel for
        for (int j = 0; j < NY; ++j)
            for (int i = 0; i < NX; ++i) {
                const size_t gb = id(i, j, kmin - 1), ib = id(i, j, kmin);
                const size_t gf = id(i, j, kmax + 1), kf = id(i, j, kmax);
                u[gb] = t * u[ib]; v[gb] = t * v[ib]; w[ib] = R(0);
                u[gf] = t * u[kf]; v[gf] = t * v[kf]; w[gf] = R(0);
            }
    }

    // ---- apply_bc_F (:1122-1135): zero-gradient copies, j then i then k (:1053-1105).
    void bc_F() {
        R* F = fld(O_F);
#pragma omp parallel for
        for (int k = 0; k < NZ; ++k)
            for (int i = 0; i < NX; ++i) {
                F[id(i, jmin - 1, k)] = F[id(i, jmin, k)];
                F[id(i, jmax + 1, k)] = F[id(i, jmax, k)];
            }
#pragma omp parallel for
        for (int k = 0; k < NZ; ++k)
            for (int j = 0; j < NY; ++j) {
                F[id(imin - 1, j, k)] = F[id(imin, j, k)];
                F[id(imax + 1, j, k)] = F[id(imax, j, k)];
            }
#pragma omp parallel for
        for (int j = 0; j < NY; ++j)
            for (int i = 0; i < NX; ++i) {
                F[id(i, j, kmin - 1)] = F[id(i, j, kmin)];
                F[id(i, j, kmax + 1)] = F[id(i, j, kmax)];
            }
    }

    // ---- post_process_f_kernel (:502-507), all nTot entries.
    void clamp_F() {
        R* F = fld(O_F);
#pragma omp parallel for
        for (long long t = 0; t < (long long)ntot; ++t) F[t] = clamp01(F[t]);
    }

    // Donor-cell (low) and downwind (high) face fluxes shared by stages 1 and 2.
    // vel is the face velocity, fm/fp the cell values on the low/high side.
    static inline R flux_low(R vel, R dt_, R fm, R fp) { return (vel >= R(0)) ? vel * dt_ * fm : vel * dt_ * fp; }
    static inline R flux_high(R vel, R dt_, R fm, R fp) { return (vel <= R(0)) ? vel * dt_ * fm : vel * dt_ * fp; }

    // ================= FCT x sweep (:509-646, wrapper :1137-1148) =================
    void fct_x() {
        R *F = fld(O_F), *Ftd = fld(O_FTD), *u = fld(O_U), *ax = fld(O_AX), *ay = fld(O_AY),
          *cx = fld(O_CX), *cy = fld(O_CY), *rp = fld(O_RP), *rm = fld(O_RM);
        const R vol = dx * dy * dz;
        // stage 1 (:509-536)
#pragma omp parallel for collapse(2)
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    const size_t c = id(i, j, k), cp = id(i + 1, j, k), cm = id(i - 1, j, k);
                    const R dv = vol - dt * dy * dz * (u[cp] - u[c]);
                    const R fl = flux_low(u[c], dt, F[cm], F[c]);
                    const R fr = flux_low(u[cp], dt, F[c], F[cp]);
                    R out = (F[c] + (fl - fr) * dy * dz / vol) * vol / dv;
                    if (stats_on) a[O_S1][c] = out;
                    if (out > R(1) || out < R(0)) out = clamp01(out);
                    Ftd[c] = out;
                }
        // stage 2 (:538-586): faces written by the owning row only (race-free).
#pragma omp parallel for
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    const size_t c = id(i, j, k), cp = id(i + 1, j, k), cm = id(i - 1, j, k),
                                 cyp = id(i, j + 1, k);
                    const R fmx = std::fmax(Ftd[c], std::fmax(Ftd[cm], Ftd[cp]));
                    const R fmn = std::fmin(Ftd[c], std::fmin(Ftd[cm], Ftd[cp]));
                    const R flL = flux_low(u[c], dt, F[cm], F[c]);
                    const R frL = flux_low(u[cp], dt, F[c], F[cp]);
                    const R flH = flux_high(u[c], dt, F[cm], F[c]);
                    const R frH = flux_high(u[cp], dt, F[c], F[cp]);
                    ax[cp] = frH - frL;
                    ax[c] = flH - flL;
                    ay[cyp] = R(0);
                    ay[c] = R(0);
                    const R pp = std::fmax(R(0), ax[c]) - std::fmin(R(0), ax[cp]) +
                                 std::fmax(R(0), ay[c]) - std::fmin(R(0), ay[cyp]);
                    const R qp = (fmx - Ftd[c]) * dx;
                    rp[c] = (pp > R(0)) ? std::fmin(R(1), qp / pp) : R(0);
                    const R pm = std::fmax(R(0), ax[cp]) - std::fmin(R(0), ax[c]) +
                                 std::fmax(R(0), ay[cyp]) - std::fmin(R(0), ay[c]);
                    const R qm = (Ftd[c] - fmn) * dx;
                    rm[c] = (pm > R(0)) ? std::fmin(R(1), qm / pm) : R(0);
                }
        // stage 3 (:588-617)
#pragma omp parallel for collapse(2)
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    const size_t c = id(i, j, k), cxp = id(i + 1, j, k), cyp = id(i, j + 1, k);
                    cx[cxp] = (ax[cxp] >= R(0)) ? std::fmin(rp[cxp], rm[c]) : std::fmin(rp[c], rm[cxp]);
                    cy[cyp] = (ay[cyp] >= R(0)) ? std::fmin(rp[cyp], rm[c]) : std::fmin(rp[c], rm[cyp]);
                }
        // stage 4 (:619-646)
#pragma omp parallel for collapse(2)
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    const size_t c = id(i, j, k), cxp = id(i + 1, j, k), cyp = id(i, j + 1, k);
                    const R dv = vol - dt * dy * dz * (u[cxp] - u[c]);
                    const R out = Ftd[c] - ((ax[cxp] * cx[cxp] - ax[c] * cx[c] + ay[cyp] * cy[cyp] -
                                             ay[c] * cy[c]) / dy) * vol / dv;
                    if (stats_on) a[O_S4][c] = out;
                    F[c] = clamp01(out);
                }
    }

    // ================= FCT y sweep (:648-785, wrapper :1150-1161) =================
    void fct_y() {
        R *F = fld(O_F), *Ftd = fld(O_FTD), *v = fld(O_V), *ax = fld(O_AX), *ay = fld(O_AY),
          *cx = fld(O_CX), *cy = fld(O_CY), *rp = fld(O_RP), *rm = fld(O_RM);
        const R vol = dx * dy * dz;
        // stage 1 (:648-675)
#pragma omp parallel for collapse(2)
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    const size_t c = id(i, j, k), cp = id(i, j + 1, k), cm = id(i, j - 1, k);
                    const R dv = vol - dt * dx * dz * (v[cp] - v[c]);
                    const R ft = flux_low(v[cp], dt, F[c], F[cp]);
                    const R fb = flux_low(v[c], dt, F[cm], F[c]);
                    R out = (F[c] + (fb - ft) * dy / (dx * dy)) * vol / dv;
                    if (stats_on) a[O_S1][c] = out;
                    if (out > R(1) || out < R(0)) out = clamp01(out);
                    Ftd[c] = out;
                }
        // stage 2 (:677-725): ay faces j and j+1 lie in plane k, so split over k.
#pragma omp parallel for
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    const size_t c = id(i, j, k), cp = id(i, j + 1, k), cm = id(i, j - 1, k),
                                 cxp = id(i + 1, j, k);
                    const R fmx = std::fmax(Ftd[c], std::fmax(Ftd[cm], Ftd[cp]));
                    const R fmn = std::fmin(Ftd[c], std::fmin(Ftd[cm], Ftd[cp]));
                    const R ftL = flux_low(v[cp], dt, F[c], F[cp]);
                    const R fbL = flux_low(v[c], dt, F[cm], F[c]);
                    const R ftH = flux_high(v[cp], dt, F[c], F[cp]);
                    const R fbH = flux_high(v[c], dt, F[cm], F[c]);
                    ax[cxp] = R(0);
                    ax[c] = R(0);
                    ay[cp] = ftH - ftL;
                    ay[c] = fbH - fbL;
                    const R pp = std::fmax(R(0), ax[c]) - std::fmin(R(0), ax[cxp]) +
                                 std::fmax(R(0), ay[c]) - std::fmin(R(0), ay[cp]);
                    const R qp = (fmx - Ftd[c]) * dx;
                    rp[c] = (pp > R(0)) ? std::fmin(R(1), qp / pp) : R(0);
                    const R pm = std::fmax(R(0), ax[cxp]) - std::fmin(R(0), ax[c]) +
                                 std::fmax(R(0), ay[cp]) - std::fmin(R(0), ay[c]);
                    const R qm = (Ftd[c] - fmn) * dx;
                    rm[c] = (pm > R(0)) ? std::fmin(R(1), qm / pm) : R(0);
                }
        // stage 3 (:727-756)
#pragma omp parallel for collapse(2)
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    const size_t c = id(i, j, k), cxp = id(i + 1, j, k), cyp = id(i, j + 1, k);
                    cx[cxp] = (ax[cxp] >= R(0)) ? std::fmin(rp[cxp], rm[c]) : std::fmin(rp[c], rm[cxp]);
                    cy[cyp] = (ay[cyp] >= R(0)) ? std::fmin(rp[cyp], rm[c]) : std::fmin(rp[c], rm[cyp]);
                }
        // stage 4 (:758-785)
#pragma omp parallel for collapse(2)
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    const size_t c = id(i, j, k), cxp = id(i + 1, j, k), cyp = id(i, j + 1, k);
                    const R dv = vol - dt * dx * dz * (v[cyp] - v[c]);
                    const R out = Ftd[c] - ((ax[cxp] * cx[cxp] - ax[c] * cx[c] + ay[cyp] * cy[cyp] -
                                             ay[c] * cy[c]) / dy) * vol / dv;
                    if (stats_on) a[O_S4][c] = out;
                    F[c] = clamp01(out);
                }
    }

    // ================= FCT z sweep (:787-905, wrapper :1163-1174) =================
    void fct_z() {
        R *F = fld(O_F), *Ftd = fld(O_FTD), *w = fld(O_W), *az = fld(O_AZ), *cz = fld(O_CZ),
          *rp = fld(O_RP), *rm = fld(O_RM);
        const R vol = dx * dy * dz;
        // stage 1 (:787-814)
#pragma omp parallel for collapse(2)
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    const size_t c = id(i, j, k), cp = id(i, j, k + 1), cm = id(i, j, k - 1);
                    const R dv = vol - dt * dx * dy * (w[cp] - w[c]);
                    const R ff = flux_low(w[cp], dt, F[c], F[cp]);
                    const R fb = flux_low(w[c], dt, F[cm], F[c]);
                    R out = (F[c] + (fb - ff) * dy * dx / vol) * vol / dv;
                    if (stats_on) a[O_S1][c] = out;
                    if (out > R(1) || out < R(0)) out = clamp01(out);
                    Ftd[c] = out;
                }
        // stage 2 (:816-859): az faces k and k+1 lie in column (i,j), so split over j.
#pragma omp parallel for
        for (int j = jmin; j <= jmax; ++j)
            for (int k = kmin; k <= kmax; ++k)
                for (int i = imin; i <= imax; ++i) {
                    const size_t c = id(i, j, k), cp = id(i, j, k + 1), cm = id(i, j, k - 1);
                    const R fmx = std::fmax(Ftd[c], std::fmax(Ftd[cm], Ftd[cp]));
                    const R fmn = std::fmin(Ftd[c], std::fmin(Ftd[cm], Ftd[cp]));
                    const R ffL = flux_low(w[cp], dt, F[c], F[cp]);
                    const R fbL = flux_low(w[c], dt, F[cm], F[c]);
                    const R ffH = flux_high(w[cp], dt, F[c], F[cp]);
                    const R fbH = flux_high(w[c], dt, F[cm], F[c]);
                    az[cp] = ffH - ffL;
                    az[c] = fbH - fbL;
                    const R pp = std::fmax(R(0), az[c]) - std::fmin(R(0), az[cp]);
                    const R qp = (fmx - Ftd[c]) * dz;
                    rp[c] = (pp > R(0)) ? std::fmin(R(1), qp / pp) : R(0);
                    const R pm = std::fmax(R(0), az[cp]) - std::fmin(R(0), az[c]);
                    const R qm = (Ftd[c] - fmn) * dz;
                    rm[c] = (pm > R(0)) ? std::fmin(R(1), qm / pm) : R(0);
                }
        // stage 3 (:861-880)
#pragma omp parallel for collapse(2)
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    const size_t c = id(i, j, k), czp = id(i, j, k + 1);
                    cz[czp] = (az[czp] >= R(0)) ? std::fmin(rp[czp], rm[c]) : std::fmin(rp[c], rm[czp]);
                }
        // stage 4 (:882-905)
#pragma omp parallel for collapse(2)
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    const size_t c = id(i, j, k), czp = id(i, j, k + 1);
                    const R dv = vol - dt * dx * dy * (w[czp] - w[c]);
                    const R out = Ftd[c] - ((az[czp] * cz[czp] - az[c] * cz[c]) / dz) * vol / dv;
                    if (stats_on) a[O_S4][c] = out;
                    F[c] = clamp01(out);
                }
    }

    // ---- Volume accounting of one directional sweep d (0 x, 1 y, 2 z), after
    //      it ran with stats_on: with dv the sweep's dilated volume,
    //        clip1  = sum (Ftd - Ftd before the stage-1 clip) dv,
    //        clamp4 = sum (F - value before clamp01) dv  (clipping + rounding to
    //                 the 2^-23 grid of 1 + F in FP32),
    //        clip4  = the part of clamp4 a plain clip to [0, 1] accounts for,
    //        dil    = sum F (vol - dv).
    //      Stage 1 conserves sum Ftd dv and stage 4 sum F dv (telescoping
    //      fluxes), so sum F vol changes by clip1 + clamp4 + dil per sweep, up
    //      to the rounding of the sums themselves.
    void sweep_stats(int d) {
        const R *F = fld(O_F), *Ftd = fld(O_FTD), *s1 = fld(O_S1), *s4 = fld(O_S4);
        const R* V = fld(d == 0 ? O_U : d == 1 ? O_V : O_W);
        const R vol = dx * dy * dz;
        const R kdv = (d == 0) ? dt * dy * dz : (d == 1) ? dt * dx * dz : dt * dx * dy;
        const size_t off = (d == 0) ? 1 : (d == 1) ? size_t(NX) : size_t(NX) * size_t(NY);
        double c1 = 0, c4 = 0, k4 = 0, dl = 0;
#pragma omp parallel for collapse(2) reduction(+ : c1, c4, k4, dl)
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    const size_t c = id(i, j, k);
                    const double dv = double(vol - kdv * (V[c + off] - V[c]));
                    c1 += (double(Ftd[c]) - double(s1[c])) * dv;
                    c4 += (double(F[c]) - double(s4[c])) * dv;
                    k4 += (std::min(1.0, std::max(0.0, double(s4[c]))) - double(s4[c])) * dv;
                    dl += double(F[c]) * (double(vol) - dv);
                }
        st_clip1 += c1;
        st_clamp4 += c4;
        st_clip4 += k4;
        st_dil += dl;
    }
    // dt vol sum F div(u) over interior cells: the dilation a divergent
    // velocity field makes over one step (first order).
    double div_term() const {
        const R *F = a[O_F].data(), *u = a[O_U].data(), *v = a[O_V].data(), *w = a[O_W].data();
        double sdiv = 0;
#pragma omp parallel for collapse(2) reduction(+ : sdiv)
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    const size_t c = id(i, j, k);
                    const double dvg = (double(u[id(i + 1, j, k)]) - double(u[c])) / double(dx) +
                                       (double(v[id(i, j + 1, k)]) - double(v[c])) / double(dy) +
                                       (double(w[id(i, j, k + 1)]) - double(w[c])) / double(dz);
                    sdiv += double(F[c]) * dvg;
                }
        return double(dt) * double(dx) * double(dy) * double(dz) * sdiv;
    }

    // ---- CSF surface tension (extension, VOF_FLAG_CSF; sigma is stored but never
    //      used by the reference, :35, :1279).  F smoothed once by the separable
    //      (1/4, 1/2, 1/4) filter on interior cells.
    void smooth() {
        const R* F = fld(O_F);
        R* fs = fld(O_FSM);
#pragma omp parallel for collapse(2)
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    R ry[3];
                    for (int dz = -1; dz <= 1; ++dz) {
                        R rx[3];
                        for (int dy = -1; dy <= 1; ++dy)
                            rx[dy + 1] = (R(0.25) * F[id(i - 1, j + dy, k + dz)] + R(0.5) * F[id(i, j + dy, k + dz)]) +
                                         R(0.25) * F[id(i + 1, j + dy, k + dz)];
                        ry[dz + 1] = (R(0.25) * rx[0] + R(0.5) * rx[1]) + R(0.25) * rx[2];
                    }
                    fs[id(i, j, k)] = (R(0.25) * ry[0] + R(0.5) * ry[1]) + R(0.25) * ry[2];
                }
    }
    // Unit normal at the vertex whose low corner is cell (i,j,k), from the
    // smoothed field of the 2x2x2 cells around it (indices clamped to the interior).
    struct VN { R x, y, z; };
    VN vertex_normal(int i, int j, int k) const {
        const R* F = a[O_FSM].data();
        const int i0 = std::max(i, 1), i1 = std::min(i + 1, nx), j0 = std::max(j, 1), j1 = std::min(j + 1, ny);
        const int k0 = std::max(k, 1), k1 = std::min(k + 1, nz);
        const R f000 = F[id(i0, j0, k0)], f100 = F[id(i1, j0, k0)], f010 = F[id(i0, j1, k0)],
                f110 = F[id(i1, j1, k0)], f001 = F[id(i0, j0, k1)], f101 = F[id(i1, j0, k1)],
                f011 = F[id(i0, j1, k1)], f111 = F[id(i1, j1, k1)];
        VN n;
        n.x = (((f100 + f110) + (f101 + f111)) - ((f000 + f010) + (f001 + f011))) * (R(0.25) * dxi);
        n.y = (((f010 + f110) + (f011 + f111)) - ((f000 + f100) + (f001 + f101))) * (R(0.25) * dyi);
        n.z = (((f001 + f101) + (f011 + f111)) - ((f000 + f100) + (f010 + f110))) * (R(0.25) * dzi);
        const R mag = std::sqrt((n.x * n.x + n.y * n.y) + n.z * n.z);
        const R inv = R(1) / (mag + R(1e-8) * dxi);
        n.x = n.x * inv; n.y = n.y * inv; n.z = n.z * inv;
        return n;
    }
    // kappa = -div m over the cell's 8 vertices, interior cells.
    void curvature() {
        R* kap = fld(O_KAPPA);
#pragma omp parallel for collapse(2)
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    const VN v000 = vertex_normal(i - 1, j - 1, k - 1), v100 = vertex_normal(i, j - 1, k - 1);
                    const VN v010 = vertex_normal(i - 1, j, k - 1), v110 = vertex_normal(i, j, k - 1);
                    const VN v001 = vertex_normal(i - 1, j - 1, k), v101 = vertex_normal(i, j - 1, k);
                    const VN v011 = vertex_normal(i - 1, j, k), v111 = vertex_normal(i, j, k);
                    const R kx = (((v100.x + v110.x) + (v101.x + v111.x)) - ((v000.x + v010.x) + (v001.x + v011.x))) * (R(0.25) * dxi);
                    const R ky = (((v010.y + v110.y) + (v011.y + v111.y)) - ((v000.y + v100.y) + (v001.y + v101.y))) * (R(0.25) * dyi);
                    const R kz = (((v001.z + v101.z) + (v011.z + v111.z)) - ((v000.z + v100.z) + (v010.z + v110.z))) * (R(0.25) * dzi);
                    kap[id(i, j, k)] = -((kx + ky) + kz);
                }
    }
    // u* += dt sigma kappa_f dF/dn / rho_f on the faces the predictor writes.
    void csf() {
        const R *F = fld(O_F), *kap = fld(O_KAPPA), *rho = fld(O_RHO);
        R *us = fld(O_US), *vs = fld(O_VS), *ws = fld(O_WS);
        auto acc = [&](size_t c, size_t m, R h) {
            const R kf = R(0.5) * (kap[m] + kap[c]);
            const R gF = (F[c] - F[m]) * h;
            const R rf = R(0.5) * (rho[m] + rho[c]);
            return sigma * kf * gF / rf;
        };
#pragma omp parallel for collapse(2)
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    const size_t c = id(i, j, k);
                    if (i > imin) us[c] = us[c] + dt * acc(c, id(i - 1, j, k), dxi);
                    if (j > jmin) vs[c] = vs[c] + dt * acc(c, id(i, j - 1, k), dyi);
                    if (k > kmin) ws[c] = ws[c] + dt * acc(c, id(i, j, k - 1), dzi);
                }
    }

    // One time step, loop body :1362-1401 (istep is the already-incremented index).
    void step(long long istep) {
        nu_rho();                       // :1364
        advect_u(); advect_v(); advect_w();  // :1366-1368
        if (flags & VOF_FLAG_CSF) { smooth(); curvature(); csf(); }
        jacobi();                       // :1371-1375
        if (flags & VOF_FLAG_REMOVE_P_MEAN) remove_p_mean();
        update_uvw();                   // :1377-1379
        bc_vel();                       // :1382
        const int r = int(istep % 3);   // :1384-1396
        if (stats_on) st_div += div_term();
        static const int order[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};
        for (int q = 0; q < 3; ++q) {
            const int d = order[r][q];
            if (d == 0) fct_x(); else if (d == 1) fct_y(); else fct_z();
            if (stats_on) sweep_stats(d);
        }
        if (stats_on) {  // post_process_f's clamp01 on interior cells
            const R* F = fld(O_F);
            const R vol = dx * dy * dz;
            double sp = 0;
            for (int k = kmin; k <= kmax; ++k)
                for (int j = jmin; j <= jmax; ++j)
                    for (int i = imin; i <= imax; ++i) {
                        const R f = F[id(i, j, k)];
                        sp += (double(clamp01(f)) - double(f)) * double(vol);
                    }
            st_post += sp;
        }
        clamp_F();                      // :1398
        bc_F();                         // :1401
    }

    // FP64 diagnostics over interior cells.
    void diag(vof_diag* d) {
        const R* F = fld(O_F);
        const R* p = fld(pcur);
        double vs = 0, ps = 0, r2 = 0, b2 = 0, fmn = 1e300, fmx = -1e300;
        const double a_x = double(dxi * dxi), a_y = double(dyi * dyi), a_z = double(dzi * dzi);
        for (int k = kmin; k <= kmax; ++k)
            for (int j = jmin; j <= jmax; ++j)
                for (int i = imin; i <= imax; ++i) {
                    const size_t c = id(i, j, k);
                    vs += double(F[c]);
                    ps += double(p[c]);
                    fmn = std::min(fmn, double(F[c]));
                    fmx = std::max(fmx, double(F[c]));
                    const double ae = (i != imax) ? a_x : 0.0, aw = (i != imin) ? a_x : 0.0;
                    const double an = (j != jmax) ? a_y : 0.0, as = (j != jmin) ? a_y : 0.0;
                    const double af = (k != kmax) ? a_z : 0.0, ab = (k != kmin) ? a_z : 0.0;
                    const double ap = -(ae + aw + an + as + af + ab);
                    const double b = double(rhs_at(i, j, k));
                    const double Ap = ae * double(p[id(i + 1, j, k)]) + aw * double(p[id(i - 1, j, k)]) +
                                      an * double(p[id(i, j + 1, k)]) + as * double(p[id(i, j - 1, k)]) +
                                      af * double(p[id(i, j, k + 1)]) + ab * double(p[id(i, j, k - 1)]) +
                                      ap * double(p[c]);
                    r2 += (b - Ap) * (b - Ap);
                    b2 += b * b;
                }
        long long nonfin = 0;
        const int fl[5] = {O_U, O_V, O_W, pcur, O_F};
        for (int f : fl)
            for (size_t t = 0; t < ntot; ++t)
                if (!std::isfinite(double(a[f][t]))) ++nonfin;
        const R volR = dx * dy * dz;
        d->liquid_volume = double(volR) * vs;
        d->p_mean = ps / (double(nx) * ny * nz);
        d->residual_l2 = std::sqrt(r2);
        d->rhs_l2 = std::sqrt(b2);
        d->f_min = fmn;
        d->f_max = fmx;
        d->nonfinite = nonfin;
    }
};

template <class R>
Orc<R>* as(void* h) { return static_cast<Orc<R>*>(static_cast<OrcBase*>(h)); }

int field_index(int f) {
    // Public vof_field ids map onto oracle arrays; RHS is computed on demand.
    switch (f) {
        case VOF_FIELD_F: return O_F;
        case VOF_FIELD_U: return O_U;
        case VOF_FIELD_V: return O_V;
        case VOF_FIELD_W: return O_W;
        case VOF_FIELD_P: return O_P;  // resolved to pcur below
        case VOF_FIELD_USTAR: return O_US;
        case VOF_FIELD_VSTAR: return O_VS;
        case VOF_FIELD_WSTAR: return O_WS;
        case VOF_FIELD_RHO: return O_RHO;
        case VOF_FIELD_NU: return O_NU;
        case VOF_FIELD_KAPPA: return O_KAPPA;
        default: return -1;
    }
}

template <class R>
int get_field(Orc<R>* o, int f, double* out) {
    if (f == VOF_FIELD_RHS) {
        std::fill(out, out + o->ntot, 0.0);
        for (int k = o->kmin; k <= o->kmax; ++k)
            for (int j = o->jmin; j <= o->jmax; ++j)
                for (int i = o->imin; i <= o->imax; ++i) out[o->id(i, j, k)] = double(o->rhs_at(i, j, k));
        return 0;
    }
    int ix = field_index(f);
    if (f >= 100) ix = f - 100;  // raw oracle array index (scratch arrays)
    if (ix < 0 || ix >= O_NFIELDS) return 1;
    if (ix == O_P) ix = o->pcur;
    const R* src = o->fld(ix);
    for (size_t t = 0; t < o->ntot; ++t) out[t] = double(src[t]);
    return 0;
}

template <class R>
int set_field(Orc<R>* o, int f, const double* in) {
    int ix = field_index(f);
    if (f >= 100) ix = f - 100;
    if (ix < 0 || ix >= O_NFIELDS) return 1;
    if (ix == O_P) ix = o->pcur;
    R* dst = o->fld(ix);
    for (size_t t = 0; t < o->ntot; ++t) dst[t] = R(in[t]);
    return 0;
}

}  // namespace

extern "C" {

// Oracle C ABI (loaded with ctypes by tests and bench.py's cpu baseline only).
void* orc_create(const vof_params* p) {
    if (!p || p->nx < 1 || p->ny < 1 || p->nz < 1) return nullptr;
    try {
        if (p->precision == 8) return static_cast<OrcBase*>(new Orc<double>(*p));
        if (p->precision == 4) return static_cast<OrcBase*>(new Orc<float>(*p));
    } catch (const std::bad_alloc&) {
    }
    return nullptr;
}

void orc_destroy(void* h) { delete static_cast<OrcBase*>(h); }

#define ORC_DISPATCH(h, ...)                                    \
    do {                                                        \
        if (static_cast<OrcBase*>(h)->prec == 8) {              \
            auto* o = as<double>(h);                            \
            __VA_ARGS__;                                        \
        } else {                                                \
            auto* o = as<float>(h);                             \
            __VA_ARGS__;                                        \
        }                                                       \
    } while (0)

int orc_init_ic(void* h, int ic) { ORC_DISPATCH(h, o->init_ic(ic)); return 0; }
int orc_init_box(void* h, double bx, double by, double bz) { ORC_DISPATCH(h, o->init_box(bx, by, bz)); return 0; }

int orc_step(void* h, long long first_istep, int nsteps) {
    ORC_DISPATCH(h, for (int s = 0; s < nsteps; ++s) o->step(first_istep + s));
    return 0;
}

// Individual stages, for stage-level tests.  id: 0 nu_rho, 1 advect_u, 2 advect_v,
// 3 advect_w, 4 one Jacobi sweep (ping-pong), 5 update_uvw, 6 bc_vel, 7 fct_x,
// 8 fct_y, 9 fct_z, 10 clamp_F, 11 bc_F, 12 remove_p_mean, 13 CSF smoothing and
// curvature, 14 CSF face acceleration.
int orc_stage(void* h, int stage) {
    int rc = 0;
    ORC_DISPATCH(h, {
        switch (stage) {
            case 0: o->nu_rho(); break;
            case 1: o->advect_u(); break;
            case 2: o->advect_v(); break;
            case 3: o->advect_w(); break;
            case 4: {
                const int src = o->pcur, dst = (src == O_P) ? O_PT : O_P;
                o->jacobi_sweep(o->fld(src), o->fld(dst));
                o->pcur = dst;
                break;
            }
            case 5: o->update_uvw(); break;
            case 6: o->bc_vel(); break;
            case 7: o->fct_x(); break;
            case 8: o->fct_y(); break;
            case 9: o->fct_z(); break;
            case 10: o->clamp_F(); break;
            case 11: o->bc_F(); break;
            case 12: o->remove_p_mean(); break;
            case 13: o->smooth(); o->curvature(); break;
            case 14: o->csf(); break;
            default: rc = 1;
        }
    });
    return rc;
}

// Volume accounting (see sweep_stats): on != 0 starts it (sums reset);
// orc_stats writes {clip1, clamp4, clip4, dil, post, div} summed over the
// steps since.
int orc_stats_enable(void* h, int on) {
    ORC_DISPATCH(h, o->stats_on = (on != 0); o->st_clip1 = o->st_clamp4 = o->st_clip4 = o->st_dil = o->st_post =
                                              o->st_div = 0.0);
    return 0;
}
int orc_stats(void* h, double* out) {
    ORC_DISPATCH(h, out[0] = o->st_clip1; out[1] = o->st_clamp4; out[2] = o->st_clip4; out[3] = o->st_dil;
                 out[4] = o->st_post; out[5] = o->st_div);
    return 0;
}
int orc_get(void* h, int f, double* out) { int rc = 0; ORC_DISPATCH(h, rc = get_field(o, f, out)); return rc; }
int orc_set(void* h, int f, const double* in) { int rc = 0; ORC_DISPATCH(h, rc = set_field(o, f, in)); return rc; }
int orc_diag(void* h, vof_diag* d) { ORC_DISPATCH(h, o->diag(d)); return 0; }
long long orc_total(void* h) { long long n = 0; ORC_DISPATCH(h, n = (long long)o->ntot); return n; }

// Raw single-precision/double copy of a field in its native type (bit patterns).
int orc_get_raw(void* h, int f, void* out) {
    int rc = 0;
    ORC_DISPATCH(h, {
        int ix = field_index(f);
        if (f >= 100) ix = f - 100;
        if (ix < 0 || ix >= O_NFIELDS) { rc = 1; }
        else {
            if (ix == O_P) ix = o->pcur;
            std::memcpy(out, o->fld(ix), o->ntot * sizeof(o->a[0][0]));
        }
    });
    return rc;
}

// write_legacy_vtk (:907-941) restated with std::ofstream on a float buffer:
// header before setprecision(8), values x-fastest one per line.
int orc_write_vtk_float(const char* path, const float* field, int nxTot, int nyTot, int nzTot) {
    std::ofstream out(path, std::ios::out);
    if (!out) return 1;
    const float sx = 1.0f / static_cast<float>(nxTot - 1);
    const float sy = 1.0f / static_cast<float>(nyTot - 1);
    const float sz = 1.0f / static_cast<float>(nzTot - 1);
    out << "# vtk DataFile Version 3.0\n" << "3D VOF field\n" << "ASCII\n" << "DATASET STRUCTURED_POINTS\n";
    out << "DIMENSIONS " << nxTot << " " << nyTot << " " << nzTot << "\n";
    out << "ORIGIN 0 0 0\n";
    out << "SPACING " << sx << " " << sy << " " << sz << "\n";
    out << "POINT_DATA " << static_cast<long long>(nxTot) * nyTot * nzTot << "\n";
    out << "SCALARS VOF float 1\n" << "LOOKUP_TABLE default\n";
    out << std::setprecision(8);
    const size_t n = static_cast<size_t>(nxTot) * nyTot * nzTot;
    for (size_t t = 0; t < n; ++t) out << field[t] << "\n";
    return out.good() ? 0 : 1;
}

// OpenMP threads the oracle's loops use (the cores of the CPU baseline).
int orc_threads(void) {
    int n = 1;
#pragma omp parallel
    {
#pragma omp single
        n = omp_get_num_threads();
    }
    return n;
}

// Standalone helpers for known-answer tests.
float orc_clamp01f(float v) { return clamp01<float>(v); }
double orc_clamp01d(double v) { return clamp01<double>(v); }

}  // extern "C"
