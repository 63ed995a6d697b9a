// vof_adjoint.hip — hand-written reverse-mode kernels of the VOF step.
//
// Each forward stage has a kernel that maps the adjoint of its outputs to the
// adjoint of its inputs (a vector-Jacobian product) in FP64, evaluated at the
// forward values of the step being reversed (FP32 or FP64 fields, recomputed
// by the forward kernels themselves).  The switches of the scheme (upwind
// side, limiter min/max, clips) are differentiated along the branch the
// forward took; where a min/max has two equal operands the derivative is
// split evenly between them, which is what a central difference sees at such
// a kink (clamp01 at exactly 0 or 1: 1/2).  Scatter-type transposes use FP64
// atomic adds; gathers are written per output cell.
#include "vof_adjoint.h"

namespace vof {

namespace {

__device__ __forceinline__ double tie_w(bool win, bool tie) { return win ? 1.0 : (tie ? 0.5 : 0.0); }

// d clamp01(v)/dv: 1 inside (0,1), 1/2 at exactly 0 or 1, 0 outside.
template <class R>
__device__ __forceinline__ double dclamp(R v) {
    if (v > R(0) && v < R(1)) return 1.0;
    if (v == R(0) || v == R(1)) return 0.5;
    return 0.0;
}

__device__ __forceinline__ void aadd(double* p, double x) {
    if (x != 0.0) atomicAdd(p, x);
}

struct Cell {
    int i, j, k, c;
};

// Thread t -> stored cell (with_ghosts) or interior cell (x fastest).
template <class R>
__device__ __forceinline__ bool cell_of(const Inst<R>& I, long t, bool with_ghosts, Cell& q) {
    const int ex = with_ghosts ? 0 : 1;
    const long nxr = I.nx + 2 - 2 * ex, nyr = I.ny + 2 - 2 * ex, nzr = I.nz + 2 - 2 * ex;
    if (t >= nxr * nyr * nzr) return false;
    q.i = int(t % nxr) + ex;
    q.j = int((t / nxr) % nyr) + ex;
    q.k = int(t / (nxr * nyr)) + ex;
    q.c = q.k * I.sy + q.j * I.sx + q.i;
    return true;
}

template <class R>
__device__ __forceinline__ bool interior(const Inst<R>& I, int i, int j, int k) {
    return i >= 1 && i <= I.nx && j >= 1 && j <= I.ny && k >= 1 && k <= I.nz;
}

// Sum NQ per-lane values over the 256 lanes of a work-group into part[blk*NQ + q].
// Every lane of the work-group must call it.
template <int NQ>
__device__ __forceinline__ void block_parts(const double (&v)[NQ], double* part) {
    __shared__ double sh[NQ][256];
    for (int q = 0; q < NQ; ++q) sh[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int q = 0; q < NQ; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < NQ; ++q) part[size_t(blockIdx.x) * NQ + q] = sh[q][0];
}

}  // namespace

// Adds the NQ column sums of part (nb rows) to acc[0..NQ), one work-group, in
// a fixed order (deterministic).
template <int NQ>
__global__ __launch_bounds__(256) void k_sum_parts(const double* __restrict__ part, int nb, double* __restrict__ acc) {
    double v[NQ];
    for (int q = 0; q < NQ; ++q) v[q] = 0.0;
    for (int b = (int)threadIdx.x; b < nb; b += 256)
        for (int q = 0; q < NQ; ++q) v[q] += part[size_t(b) * NQ + q];
    __shared__ double sh[NQ][256];
    for (int q = 0; q < NQ; ++q) sh[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int q = 0; q < NQ; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < NQ; ++q) acc[q] += sh[q][0];
}

// ---------------------------------------------------------------------------
// Objective J = sum(F z) / sum(F) over interior cells (z = cell-centre height
// (k - 1/2) dz): per-work-group partial sums (FP64), added on the host in order.
template <class R>
__global__ __launch_bounds__(256) void k_centroid_parts(const Inst<R>* __restrict__ tab, const R* __restrict__ F,
                                                        double* __restrict__ part) {
    const Inst<R>& I = tab[0];
    Cell q;
    double v[2] = {0.0, 0.0};
    if (cell_of(I, long(blockIdx.x) * 256 + threadIdx.x, false, q)) {
        const double f = double(F[q.c]);
        v[0] = f;
        v[1] = f * ((double(q.k) - 0.5) * double(I.dz));
    }
    block_parts<2>(v, part);
}

// aF = gJ * dJ/dF (+ seed): gJ * (z - J) / sum(F) on interior cells, 0 on ghosts.
template <class R>
__global__ __launch_bounds__(256) void k_centroid_adj(const Inst<R>* __restrict__ tab, double J, double sumF,
                                                      double gJ, const double* __restrict__ seed,
                                                      double* __restrict__ aF) {
    const Inst<R>& I = tab[0];
    Cell q;
    if (!cell_of(I, long(blockIdx.x) * 256 + threadIdx.x, true, q)) return;
    double g = interior(I, q.i, q.j, q.k) ? gJ * (((double(q.k) - 0.5) * double(I.dz)) - J) / sumF : 0.0;
    if (seed) g += seed[q.c];
    aF[q.c] = g;
}

// Transpose of the F ghost fill (every ghost copies the nearest interior
// cell): ghost adjoints are added to their source cell and cleared.
template <class R>
__global__ __launch_bounds__(256) void k_ghostF_adj(const Inst<R>* __restrict__ tab, double* __restrict__ aF) {
    const Inst<R>& I = tab[0];
    Cell q;
    if (!cell_of(I, long(blockIdx.x) * 256 + threadIdx.x, true, q)) return;
    const int ci = min(max(q.i, 1), I.nx), cj = min(max(q.j, 1), I.ny), ck = min(max(q.k, 1), I.nz);
    if (ci == q.i && cj == q.j && ck == q.k) return;
    const double g = aF[q.c];
    aF[q.c] = 0.0;
    aadd(&aF[ck * I.sy + cj * I.sx + ci], g);
}

// ---------------------------------------------------------------------------
// FCT sweep reverse, one lane per line along the sweep direction.  The lane
// re-runs the forward sweep of its line with the forward's association
// (intermediates kept in the line's slots of the tape arrays), then walks the
// four stages backwards, accumulating adjoints of its own line only (no races).
// F: sweep input; V: face velocity along the line; aF: on entry the adjoint of
// the sweep output, on exit that of the sweep input (in place: ghost entries
// pass through, their only use is a flux through a wall face, whose velocity
// is zero); aV += the face-velocity adjoints.  fin: this is the last sweep of
// the step, whose output went through clamp01 a second time.
template <class R, int DIR>
__global__ __launch_bounds__(256) void k_fct_adj(const Inst<R>* __restrict__ tab, const R* __restrict__ F,
                                                 const R* __restrict__ V, double* __restrict__ aF,
                                                 double* __restrict__ aV, FctTape<R> tp, int fin) {
    const Inst<R>& I = tab[0];
    const int n = (DIR == 0) ? I.nx : (DIR == 1) ? I.ny : I.nz;
    const int na = (DIR == 0) ? I.ny : I.nx;  // transverse axis along the lanes
    const int nb = (DIR == 2) ? I.ny : I.nz;
    const long t = long(blockIdx.x) * 256 + threadIdx.x;
    if (t >= long(na) * nb) return;
    const int ta = 1 + int(t % na), tb = 1 + int(t / na);
    int base, S;
    if (DIR == 0) { base = tb * I.sy + ta * I.sx; S = 1; }     // x line at (y, z) = (ta, tb)
    else if (DIR == 1) { base = tb * I.sy + ta; S = I.sx; }    // y line at (x, z) = (ta, tb)
    else { base = tb * I.sx + ta; S = I.sy; }                  // z line at (x, y) = (ta, tb)
    const R dt = I.dt, vol = I.vol;
    const R kdv = (DIR == 0) ? I.dtdydz : (DIR == 1) ? I.dtdxdz : I.dtdxdy;
    const R qf = (DIR == 2) ? I.dz : I.dx;
    const R s4 = (DIR == 2) ? I.dz : I.dy;
    auto s1 = [&](R d) -> R {
        if (DIR == 0) return d * I.dy * I.dz / I.vol;
        if (DIR == 1) return d * I.dy / I.dxdy;
        return d * I.dy * I.dx / I.vol;
    };
    const double ds1 = (DIR == 0) ? double(I.dy) * double(I.dz) / double(I.vol)
                     : (DIR == 1) ? double(I.dy) / double(I.dxdy)
                                  : double(I.dy) * double(I.dx) / double(I.vol);
    auto at = [&](int m) { return base + m * S; };
    // ---- forward re-run (face f between cells f-1 and f, stored at slot f)
    for (int f = 1; f <= n + 1; ++f) {
        const R vf = V[at(f)], vdt = vf * dt, flo = F[at(f - 1)], fhi = F[at(f)];
        const R L = (vf >= R(0)) ? vdt * flo : vdt * fhi;
        const R H = (vf <= R(0)) ? vdt * flo : vdt * fhi;
        tp.L[at(f)] = L;
        tp.a[at(f)] = H - L;
    }
    for (int c = 1; c <= n; ++c) {
        const R dv = vol - kdv * (V[at(c + 1)] - V[at(c)]);
        R o = (F[at(c)] + s1(tp.L[at(c)] - tp.L[at(c + 1)])) * vol / dv;
        tp.y[at(c)] = o;
        if (o > R(1) || o < R(0)) o = clamp01(o);
        tp.ftd[at(c)] = o;
        tp.dv[at(c)] = dv;
    }
    auto ftd = [&](int c) { return (c >= 1 && c <= n) ? tp.ftd[at(c)] : R(0); };
    for (int c = 1; c <= n; ++c) {
        const R fc = ftd(c), fl = ftd(c - 1), fr = ftd(c + 1);
        const R fmx = vmax(fc, vmax(fl, fr)), fmn = vmin(fc, vmin(fl, fr));
        const R alo = tp.a[at(c)], ahi = tp.a[at(c + 1)];
        const R pp = vmax(R(0), alo) - vmin(R(0), ahi);
        const R pm = vmax(R(0), ahi) - vmin(R(0), alo);
        tp.rp[at(c)] = (pp > R(0)) ? vmin(R(1), (fmx - fc) * qf / pp) : R(0);
        tp.rm[at(c)] = (pm > R(0)) ? vmin(R(1), (fc - fmn) * qf / pm) : R(0);
    }
    auto rp = [&](int c) { return (c >= 1 && c <= n) ? tp.rp[at(c)] : R(0); };
    auto rm = [&](int c) { return (c >= 1 && c <= n) ? tp.rm[at(c)] : R(0); };
    tp.C[at(1)] = R(0);
    for (int f = 2; f <= n + 1; ++f)
        tp.C[at(f)] = (tp.a[at(f)] >= R(0)) ? vmin(rp(f), rm(f - 1)) : vmin(rp(f - 1), rm(f));
    for (int m = 0; m <= n + 1; ++m) {
        tp.gftd[at(m)] = 0.0; tp.ga[at(m)] = 0.0; tp.gL[at(m)] = 0.0;
        tp.grp[at(m)] = 0.0; tp.grm[at(m)] = 0.0; tp.gdv[at(m)] = 0.0;
    }
    // ---- stage 4: Fn = clamp01(yv), yv = ftd - ((aC(c+1) - aC(c)) / s4) * vol / dv, aC = a*C
    for (int c = 1; c <= n; ++c) {
        const R aC1 = tp.a[at(c + 1)] * tp.C[at(c + 1)], aC0 = tp.a[at(c)] * tp.C[at(c)];
        const R dv = tp.dv[at(c)];
        const R num = (aC1 - aC0) / s4;
        const R yv = tp.ftd[at(c)] - num * vol / dv;
        double g = aF[at(c)] * dclamp(yv);
        if (fin) g *= dclamp(clamp01(yv));
        aF[at(c)] = 0.0;  // the sweep overwrote the interior value
        if (g == 0.0) continue;
        const double k = double(vol) / (double(s4) * double(dv));
        tp.gftd[at(c)] += g;
        const double gaC1 = -g * k, gaC0 = g * k;
        tp.ga[at(c + 1)] += gaC1 * double(tp.C[at(c + 1)]);
        tp.ga[at(c)] += gaC0 * double(tp.C[at(c)]);
        tp.gdv[at(c)] += g * double(num) * double(vol) / (double(dv) * double(dv));
        for (int e = 0; e < 2; ++e) {  // limiter of faces c+1 and c
            const int f = c + 1 - e;
            const double gC = (e == 0 ? gaC1 : gaC0) * double(tp.a[at(f)]);
            if (f < 2 || gC == 0.0) continue;
            R x, y2;
            int cx, cy;  // C = min(x, y2) with x = rp(cx), y2 = rm(cy)
            if (tp.a[at(f)] >= R(0)) { x = rp(f); cx = f; y2 = rm(f - 1); cy = f - 1; }
            else { x = rp(f - 1); cx = f - 1; y2 = rm(f); cy = f; }
            if (cx >= 1 && cx <= n) tp.grp[at(cx)] += gC * tie_w(x < y2, x == y2);
            if (cy >= 1 && cy <= n) tp.grm[at(cy)] += gC * tie_w(y2 < x, x == y2);
        }
    }
    // ---- stage 2: rp = min(1, Q+/P+), rm = min(1, Q-/P-)
    for (int c = 1; c <= n; ++c) {
        const R fc = ftd(c), fl = ftd(c - 1), fr = ftd(c + 1);
        const R alo = tp.a[at(c)], ahi = tp.a[at(c + 1)];
        for (int s = 0; s < 2; ++s) {
            const double gr = (s == 0) ? tp.grp[at(c)] : tp.grm[at(c)];
            if (gr == 0.0) continue;
            const R x = (s == 0) ? alo : ahi, y = (s == 0) ? ahi : alo;  // P = max(0,x) - min(0,y)
            const R P = vmax(R(0), x) - vmin(R(0), y);
            if (!(P > R(0))) continue;
            const R m1 = (s == 0) ? vmax(fl, fr) : vmin(fl, fr);
            const R ext = (s == 0) ? vmax(fc, m1) : vmin(fc, m1);
            const R Q = ((s == 0) ? (ext - fc) : (fc - ext)) * qf;
            const R ratio = Q / P;
            const double wr = tie_w(ratio < R(1), ratio == R(1));
            if (wr == 0.0) continue;
            const double gratio = gr * wr;
            const double gQ = gratio / double(P);
            const double gP = -gratio * double(Q) / (double(P) * double(P));
            const double gext = ((s == 0) ? gQ : -gQ) * double(qf);
            tp.gftd[at(c)] += ((s == 0) ? -gQ : gQ) * double(qf);
            const double wc = tie_w((s == 0) ? (fc > m1) : (fc < m1), fc == m1);
            tp.gftd[at(c)] += gext * wc;
            const double wl = tie_w((s == 0) ? (fl > fr) : (fl < fr), fl == fr);
            if (c - 1 >= 1) tp.gftd[at(c - 1)] += gext * (1.0 - wc) * wl;
            if (c + 1 <= n) tp.gftd[at(c + 1)] += gext * (1.0 - wc) * (1.0 - wl);
            const double dx = tie_w(x > R(0), x == R(0)), dy = -tie_w(y < R(0), y == R(0));
            tp.ga[at(s == 0 ? c : c + 1)] += gP * dx;
            tp.ga[at(s == 0 ? c + 1 : c)] += gP * dy;
        }
    }
    // ---- a = H - L; stage 1 (with its clip) on cells
    for (int f = 1; f <= n + 1; ++f) tp.gL[at(f)] -= tp.ga[at(f)];
    for (int c = 1; c <= n; ++c) {
        const double gf = tp.gftd[at(c)];
        if (gf == 0.0) continue;
        const R o = tp.y[at(c)];
        const double go = gf * ((o > R(1) || o < R(0)) ? dclamp(o) : 1.0);
        if (go == 0.0) continue;
        const double dv = double(tp.dv[at(c)]);
        const double k = double(vol) / dv;
        aF[at(c)] += go * k;
        tp.gL[at(c)] += go * k * ds1;
        tp.gL[at(c + 1)] -= go * k * ds1;
        tp.gdv[at(c)] -= go * double(o) / dv;
    }
    // ---- fluxes L (donor) and H (downwind): L = vdt * F(up), H = vdt * F(down)
    for (int f = 1; f <= n + 1; ++f) {
        const double gL = tp.gL[at(f)], gH = tp.ga[at(f)];
        const R vf = V[at(f)], flo = F[at(f - 1)], fhi = F[at(f)];
        const double vdt = double(vf * dt);
        const bool lo_L = (vf >= R(0)), lo_H = (vf <= R(0));
        double gv = 0.0;
        if (gL != 0.0) {
            aF[at(lo_L ? f - 1 : f)] += gL * vdt;
            gv += gL * double(dt) * double(lo_L ? flo : fhi);
        }
        if (gH != 0.0) {
            aF[at(lo_H ? f - 1 : f)] += gH * vdt;
            gv += gH * double(dt) * double(lo_H ? flo : fhi);
        }
        aV[at(f)] += gv;
    }
    // ---- dv = vol - kdv * (V(c+1) - V(c))
    for (int c = 1; c <= n; ++c) {
        const double g = tp.gdv[at(c)];
        if (g == 0.0) continue;
        aV[at(c + 1)] -= g * double(kdv);
        aV[at(c)] += g * double(kdv);
    }
}

// ---------------------------------------------------------------------------
// Transpose of the velocity correction and its boundary closed form: stored
// u(i,j,k) = u_in(i, cj, ck) for 2 <= i <= nx, 0 on the walls i = 1, nx+1, and
// u_old(0, cj, ck) on plane 0 (likewise v, w).  u_in = u* - dt/r (p0 - p1) / h
// with r = (rho0 + rho1)/2 feeds u*, p and rho; plane 0 feeds u_old.
template <class R>
__global__ __launch_bounds__(256) void k_correct_adj(const Inst<R>* __restrict__ tab, const R* __restrict__ F,
                                                     const R* __restrict__ p, const double* __restrict__ au,
                                                     const double* __restrict__ av, const double* __restrict__ aw,
                                                     double* __restrict__ aus, double* __restrict__ avs,
                                                     double* __restrict__ aws, double* __restrict__ ap,
                                                     double* __restrict__ arho, double* __restrict__ au0,
                                                     double* __restrict__ av0, double* __restrict__ aw0) {
    const Inst<R>& I = tab[0];
    Cell q;
    if (!cell_of(I, long(blockIdx.x) * 256 + threadIdx.x, true, q)) return;
    const int nx = I.nx, ny = I.ny, nz = I.nz, sx = I.sx, sy = I.sy;
    const int ci = min(max(q.i, 1), nx), cj = min(max(q.j, 1), ny), ck = min(max(q.k, 1), nz);
    auto corr = [&](double g, int c, int off, R hi, double* ast) {
        if (g == 0.0) return;
        aadd(&ast[c], g);
        const double r = 0.5 * (double(rho_of(I, F[c])) + double(rho_of(I, F[c - off])));
        const double dp = double(p[c]) - double(p[c - off]);
        const double kk = double(I.dt) / r * double(hi);
        aadd(&ap[c], -g * kk);
        aadd(&ap[c - off], g * kk);
        const double gr = g * double(I.dt) * dp * double(hi) / (r * r);
        aadd(&arho[c], 0.5 * gr);
        aadd(&arho[c - off], 0.5 * gr);
    };
    const double gu = au[q.c], gv = av[q.c], gw = aw[q.c];
    if (q.i >= 2 && q.i <= nx) corr(gu, ck * sy + cj * sx + q.i, 1, I.dxi, aus);
    else if (q.i == 0) aadd(&au0[ck * sy + cj * sx], gu);
    if (q.j >= 2 && q.j <= ny) corr(gv, ck * sy + q.j * sx + ci, sx, I.dyi, avs);
    else if (q.j == 0) aadd(&av0[ck * sy + ci], gv);
    if (q.k >= 2 && q.k <= nz) corr(gw, q.k * sy + cj * sx + ci, sy, I.dzi, aws);
    else if (q.k == 0) aadd(&aw0[cj * sx + ci], gw);
}

// Transpose of one Jacobi sweep p' = (rhs - N p) / ap on interior cells:
// gout(c) = -sum over the interior neighbours m that read c of coef(m->c) gin(m) / ap(m),
// arhs += gin / ap.  Ghost entries of gout come out zero (zero coefficients).
template <class R>
__global__ __launch_bounds__(256) void k_jacobi_adj(const Inst<R>* __restrict__ tab, const double* __restrict__ gin,
                                                    double* __restrict__ gout, double* __restrict__ arhs) {
    const Inst<R>& I = tab[0];
    Cell q;
    if (!cell_of(I, long(blockIdx.x) * 256 + threadIdx.x, true, q)) return;
    const int nx = I.nx, ny = I.ny, nz = I.nz, sx = I.sx, sy = I.sy;
    auto apc = [&](int i, int j, int k) {  // diagonal, summed in the forward's order
        const R ae = (i != nx) ? I.axx : R(0), aw = (i != 1) ? I.axx : R(0);
        const R an = (j != ny) ? I.ayy : R(0), as = (j != 1) ? I.ayy : R(0);
        const R af = (k != nz) ? I.azz : R(0), ab = (k != 1) ? I.azz : R(0);
        return double(-(ae + aw + an + as + af + ab));
    };
    const int i = q.i, j = q.j, k = q.k, c = q.c;
    double s = 0.0;
    if (interior(I, i + 1, j, k) && i + 1 != 1) s -= double(I.axx) * gin[c + 1] / apc(i + 1, j, k);
    if (interior(I, i - 1, j, k) && i - 1 != nx) s -= double(I.axx) * gin[c - 1] / apc(i - 1, j, k);
    if (interior(I, i, j + 1, k) && j + 1 != 1) s -= double(I.ayy) * gin[c + sx] / apc(i, j + 1, k);
    if (interior(I, i, j - 1, k) && j - 1 != ny) s -= double(I.ayy) * gin[c - sx] / apc(i, j - 1, k);
    if (interior(I, i, j, k + 1) && k + 1 != 1) s -= double(I.azz) * gin[c + sy] / apc(i, j, k + 1);
    if (interior(I, i, j, k - 1) && k - 1 != nz) s -= double(I.azz) * gin[c - sy] / apc(i, j, k - 1);
    gout[c] = s;
    if (interior(I, i, j, k)) arhs[c] += gin[c] / apc(i, j, k);
}

// Transpose of rhs = rho/dt * div(u*): each face gathers from its two cells.
template <class R>
__global__ __launch_bounds__(256) void k_rhs_adj(const Inst<R>* __restrict__ tab, const R* __restrict__ F,
                                                 const R* __restrict__ us, const R* __restrict__ vs,
                                                 const R* __restrict__ ws, const double* __restrict__ arhs,
                                                 double* __restrict__ aus, double* __restrict__ avs,
                                                 double* __restrict__ aws, double* __restrict__ arho) {
    const Inst<R>& I = tab[0];
    Cell q;
    if (!cell_of(I, long(blockIdx.x) * 256 + threadIdx.x, true, q)) return;
    const int sx = I.sx, sy = I.sy, c = q.c;
    const double rdt = 1.0 / double(I.dt);
    auto rf = [&](int cc) { return double(rho_of(I, F[cc])) * rdt; };
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (interior(I, q.i, q.j, q.k)) {
        const double g = arhs[c] * rf(c);
        gx -= g * double(I.dxi);
        gy -= g * double(I.dyi);
        gz -= g * double(I.dzi);
        const double div = (double(us[c + 1]) - double(us[c])) * double(I.dxi) +
                           (double(vs[c + sx]) - double(vs[c])) * double(I.dyi) +
                           (double(ws[c + sy]) - double(ws[c])) * double(I.dzi);
        arho[c] += arhs[c] * div * rdt;
    }
    if (interior(I, q.i - 1, q.j, q.k)) gx += arhs[c - 1] * rf(c - 1) * double(I.dxi);
    if (interior(I, q.i, q.j - 1, q.k)) gy += arhs[c - sx] * rf(c - sx) * double(I.dyi);
    if (interior(I, q.i, q.j, q.k - 1)) gz += arhs[c - sy] * rf(c - sy) * double(I.dzi);
    aus[c] += gx;
    avs[c] += gy;
    aws[c] += gz;
}

// Partial derivatives of one predictor face value
//   out = c + dt * (visc + adv + g),  visc = nu sum_d (m_d - 2c + p_d) / h_d^2,
//   adv = -sum_d a_d * (a_d > 0 ? (c - m_d) : (p_d - c)) / h_d
// times g_out, accumulated into the operand adjoints.  own: the axis whose
// advecting velocity is the face value c itself.
struct MomAdj {
    double c = 0, xm = 0, xp = 0, ym = 0, yp = 0, zm = 0, zp = 0, ax = 0, ay = 0, az = 0;
};

template <class R>
__device__ __forceinline__ void mom_adj(const Inst<R>& I, double g, R c, R xm, R xp, R ym, R yp, R zm, R zp, R ax,
                                        R ay, R az, R n0, int own, MomAdj& d, double& dnu) {
    const double hx = double(I.dxi), hy = double(I.dyi), hz = double(I.dzi);
    const double gd = g * double(I.dt), n = double(n0);
    dnu += gd * ((double(xm) - 2.0 * double(c) + double(xp)) * hx * hx +
                 (double(ym) - 2.0 * double(c) + double(yp)) * hy * hy +
                 (double(zm) - 2.0 * double(c) + double(zp)) * hz * hz);
    d.c += g - 2.0 * gd * n * (hx * hx + hy * hy + hz * hz);
    d.xm += gd * n * hx * hx; d.xp += gd * n * hx * hx;
    d.ym += gd * n * hy * hy; d.yp += gd * n * hy * hy;
    d.zm += gd * n * hz * hz; d.zp += gd * n * hz * hz;
    const double axd = double(ax), ayd = double(ay), azd = double(az);
    double ddx, ddy, ddz;
    if (ax > R(0)) { ddx = (double(c) - double(xm)) * hx; d.c -= gd * axd * hx; d.xm += gd * axd * hx; }
    else { ddx = (double(xp) - double(c)) * hx; d.c += gd * axd * hx; d.xp -= gd * axd * hx; }
    if (ay > R(0)) { ddy = (double(c) - double(ym)) * hy; d.c -= gd * ayd * hy; d.ym += gd * ayd * hy; }
    else { ddy = (double(yp) - double(c)) * hy; d.c += gd * ayd * hy; d.yp -= gd * ayd * hy; }
    if (az > R(0)) { ddz = (double(c) - double(zm)) * hz; d.c -= gd * azd * hz; d.zm += gd * azd * hz; }
    else { ddz = (double(zp) - double(c)) * hz; d.c += gd * azd * hz; d.zp -= gd * azd * hz; }
    if (own == 0) d.c -= gd * ddx; else d.ax -= gd * ddx;
    if (own == 1) d.c -= gd * ddy; else d.ay -= gd * ddy;
    if (own == 2) d.c -= gd * ddz; else d.az -= gd * ddz;
}

// Transpose of the momentum predictor: the lane of interior cell c scatters the
// adjoints of u* (face i >= 2), v* (face j >= 2) and w* (face k >= 2) into the
// velocity adjoints (atomics) and nu(c); it also sums dJ/dg = dt * sum a(u*).
template <class R>
__global__ __launch_bounds__(256) void k_predict_adj(const Inst<R>* __restrict__ tab, const R* __restrict__ F,
                                                     const R* __restrict__ u, const R* __restrict__ v,
                                                     const R* __restrict__ w, const double* __restrict__ aus,
                                                     const double* __restrict__ avs, const double* __restrict__ aws,
                                                     double* __restrict__ au, double* __restrict__ av,
                                                     double* __restrict__ aw, double* __restrict__ anu,
                                                     double* __restrict__ part) {
    const Inst<R>& I = tab[0];
    Cell q;
    double gsum[3] = {0.0, 0.0, 0.0};
    if (cell_of(I, long(blockIdx.x) * 256 + threadIdx.x, false, q)) {
        const int sx = I.sx, sy = I.sy, c = q.c;
        const R n0 = nu_of(I, F[c]);
        double gn = 0.0;
        auto scatter7 = [&](double* a, const MomAdj& d) {
            aadd(&a[c], d.c); aadd(&a[c - 1], d.xm); aadd(&a[c + 1], d.xp); aadd(&a[c - sx], d.ym);
            aadd(&a[c + sx], d.yp); aadd(&a[c - sy], d.zm); aadd(&a[c + sy], d.zp);
        };
        auto scatter4 = [&](double* a, double g, int o0, int o1, int o2, int o3) {
            const double h = 0.25 * g;
            aadd(&a[c + o0], h); aadd(&a[c + o1], h); aadd(&a[c + o2], h); aadd(&a[c + o3], h);
        };
        if (q.i >= 2 && aus[c] != 0.0) {  // u*: advecting (u, vh, wh)
            const R vh = R(0.25) * (v[c - 1] + v[c - 1 + sx] + v[c] + v[c + sx]);
            const R wh = R(0.25) * (w[c - 1] + w[c - 1 + sy] + w[c] + w[c + sy]);
            MomAdj d;
            mom_adj(I, aus[c], u[c], u[c - 1], u[c + 1], u[c - sx], u[c + sx], u[c - sy], u[c + sy], u[c], vh, wh, n0,
                    0, d, gn);
            scatter7(au, d);
            scatter4(av, d.ay, -1, -1 + sx, 0, sx);
            scatter4(aw, d.az, -1, -1 + sy, 0, sy);
            gsum[0] = aus[c] * double(I.dt);
        }
        if (q.j >= 2 && avs[c] != 0.0) {  // v*: advecting (uh, v, wh)
            const R uh = R(0.25) * (u[c - sx] + u[c] + u[c + 1 - sx] + u[c + 1]);
            const R wh = R(0.25) * (w[c - sx + sy] + w[c - sx] + w[c] + w[c + sy]);
            MomAdj d;
            mom_adj(I, avs[c], v[c], v[c - 1], v[c + 1], v[c - sx], v[c + sx], v[c - sy], v[c + sy], uh, v[c], wh, n0,
                    1, d, gn);
            scatter7(av, d);
            scatter4(au, d.ax, -sx, 0, 1 - sx, 1);
            scatter4(aw, d.az, -sx + sy, -sx, 0, sy);
            gsum[1] = avs[c] * double(I.dt);
        }
        if (q.k >= 2 && aws[c] != 0.0) {  // w*: advecting (uh, vh, w)
            const R uh = R(0.25) * (u[c + 1 - sy] + u[c - sy] + u[c + 1] + u[c]);
            const R vh = R(0.25) * (v[c + sx - sy] + v[c - sy] + v[c] + v[c + sx]);
            MomAdj d;
            mom_adj(I, aws[c], w[c], w[c - 1], w[c + 1], w[c - sx], w[c + sx], w[c - sy], w[c + sy], uh, vh, w[c], n0,
                    2, d, gn);
            scatter7(aw, d);
            scatter4(au, d.ax, 1 - sy, -sy, 1, 0);
            scatter4(av, d.ay, sx - sy, -sy, 0, sx);
            gsum[2] = aws[c] * double(I.dt);
        }
        anu[c] += gn;
    }
    block_parts<3>(gsum, part);
}

// rho = rho_g (1 - f) + rho_l f, nu = nu_l f + nu_g (1 - f), f = clamp01(F):
// aF += (arho (rho_l - rho_g) + anu (nu_l - nu_g)) dclamp(F); work-group sums of
// arho f, arho (1-f), anu f, anu (1-f) for the parameter gradients.
template <class R>
__global__ __launch_bounds__(256) void k_rhonu_adj(const Inst<R>* __restrict__ tab, const R* __restrict__ F,
                                                   const double* __restrict__ arho, const double* __restrict__ anu,
                                                   double* __restrict__ aF, double* __restrict__ part) {
    const Inst<R>& I = tab[0];
    Cell q;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    if (cell_of(I, long(blockIdx.x) * 256 + threadIdx.x, true, q)) {
        const R f = clamp01(F[q.c]);
        const double gr = arho[q.c], gn = anu[q.c];
        if (gr != 0.0 || gn != 0.0) {
            const double gf = gr * (double(I.rho_l) - double(I.rho_g)) + gn * (double(I.nu_l) - double(I.nu_g));
            aF[q.c] += gf * dclamp(F[q.c]);
            s[0] = gr * double(f);
            s[1] = gr * double(R(1) - f);
            s[2] = gn * double(f);
            s[3] = gn * double(R(1) - f);
        }
    }
    block_parts<4>(s, part);
}

__global__ __launch_bounds__(256) void k_axpy(const double* __restrict__ src, double* __restrict__ dst, long n) {
    const long t = long(blockIdx.x) * 256 + threadIdx.x;
    if (t < n) dst[t] += src[t];
}

// ---------------------------------------------------------------------------
// Host launchers.
template <class R>
void adj_centroid_parts(const Inst<R>* tab, const Inst<R>& h, const R* F, double* part, hipStream_t st) {
    hipLaunchKernelGGL(k_centroid_parts<R>, dim3(adj_blocks(long(h.nx) * h.ny * h.nz)), dim3(256), 0, st, tab, F,
                       part);
}
template <class R>
void adj_centroid(const Inst<R>* tab, const Inst<R>& h, double J, double sumF, double gJ, const double* seed,
                  double* aF, hipStream_t st) {
    hipLaunchKernelGGL(k_centroid_adj<R>, dim3(adj_blocks(h.ntot)), dim3(256), 0, st, tab, J, sumF, gJ, seed, aF);
}
template <class R>
void adj_ghostF(const Inst<R>* tab, const Inst<R>& h, double* aF, hipStream_t st) {
    hipLaunchKernelGGL(k_ghostF_adj<R>, dim3(adj_blocks(h.ntot)), dim3(256), 0, st, tab, aF);
}
template <class R>
void adj_fct(const Inst<R>* tab, const Inst<R>& h, int dir, const R* F, const R* V, double* aF, double* aV,
             const FctTape<R>& tp, int fin, hipStream_t st) {
    const long lines = (dir == 0) ? long(h.ny) * h.nz : (dir == 1) ? long(h.nx) * h.nz : long(h.nx) * h.ny;
    const dim3 grid(adj_blocks(lines));
    if (dir == 0) hipLaunchKernelGGL((k_fct_adj<R, 0>), grid, dim3(256), 0, st, tab, F, V, aF, aV, tp, fin);
    else if (dir == 1) hipLaunchKernelGGL((k_fct_adj<R, 1>), grid, dim3(256), 0, st, tab, F, V, aF, aV, tp, fin);
    else hipLaunchKernelGGL((k_fct_adj<R, 2>), grid, dim3(256), 0, st, tab, F, V, aF, aV, tp, fin);
}
template <class R>
void adj_correct(const Inst<R>* tab, const Inst<R>& h, const R* F, const R* p, const double* au, const double* av,
                 const double* aw, double* aus, double* avs, double* aws, double* ap, double* arho, double* au0,
                 double* av0, double* aw0, hipStream_t st) {
    hipLaunchKernelGGL(k_correct_adj<R>, dim3(adj_blocks(h.ntot)), dim3(256), 0, st, tab, F, p, au, av, aw, aus, avs,
                       aws, ap, arho, au0, av0, aw0);
}
template <class R>
void adj_jacobi(const Inst<R>* tab, const Inst<R>& h, const double* gin, double* gout, double* arhs, hipStream_t st) {
    hipLaunchKernelGGL(k_jacobi_adj<R>, dim3(adj_blocks(h.ntot)), dim3(256), 0, st, tab, gin, gout, arhs);
}
template <class R>
void adj_rhs(const Inst<R>* tab, const Inst<R>& h, const R* F, const R* us, const R* vs, const R* ws,
             const double* arhs, double* aus, double* avs, double* aws, double* arho, hipStream_t st) {
    hipLaunchKernelGGL(k_rhs_adj<R>, dim3(adj_blocks(h.ntot)), dim3(256), 0, st, tab, F, us, vs, ws, arhs, aus, avs,
                       aws, arho);
}
template <class R>
void adj_predict(const Inst<R>* tab, const Inst<R>& h, const R* F, const R* u, const R* v, const R* w,
                 const double* aus, const double* avs, const double* aws, double* au, double* av, double* aw,
                 double* anu, double* part, double* acc, hipStream_t st) {
    const int nb = adj_blocks(long(h.nx) * h.ny * h.nz);
    hipLaunchKernelGGL(k_predict_adj<R>, dim3(nb), dim3(256), 0, st, tab, F, u, v, w, aus, avs, aws, au, av, aw, anu,
                       part);
    hipLaunchKernelGGL(k_sum_parts<3>, dim3(1), dim3(256), 0, st, part, nb, acc + ADJ_GX);
}
template <class R>
void adj_rhonu(const Inst<R>* tab, const Inst<R>& h, const R* F, const double* arho, const double* anu, double* aF,
               double* part, double* acc, hipStream_t st) {
    const int nb = adj_blocks(h.ntot);
    hipLaunchKernelGGL(k_rhonu_adj<R>, dim3(nb), dim3(256), 0, st, tab, F, arho, anu, aF, part);
    hipLaunchKernelGGL(k_sum_parts<4>, dim3(1), dim3(256), 0, st, part, nb, acc + ADJ_RHO_L);
}
void adj_axpy(const double* src, double* dst, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_axpy, dim3(adj_blocks(n)), dim3(256), 0, st, src, dst, n);
}

#define VOF_ADJ_INSTANTIATE(R)                                                                                     \
    template void adj_centroid_parts<R>(const Inst<R>*, const Inst<R>&, const R*, double*, hipStream_t);          \
    template void adj_centroid<R>(const Inst<R>*, const Inst<R>&, double, double, double, const double*, double*, \
                                  hipStream_t);                                                                    \
    template void adj_ghostF<R>(const Inst<R>*, const Inst<R>&, double*, hipStream_t);                            \
    template void adj_fct<R>(const Inst<R>*, const Inst<R>&, int, const R*, const R*, double*, double*,           \
                             const FctTape<R>&, int, hipStream_t);                                                 \
    template void adj_correct<R>(const Inst<R>*, const Inst<R>&, const R*, const R*, const double*, const double*, \
                                 const double*, double*, double*, double*, double*, double*, double*, double*,     \
                                 double*, hipStream_t);                                                            \
    template void adj_jacobi<R>(const Inst<R>*, const Inst<R>&, const double*, double*, double*, hipStream_t);    \
    template void adj_rhs<R>(const Inst<R>*, const Inst<R>&, const R*, const R*, const R*, const R*,              \
                             const double*, double*, double*, double*, double*, hipStream_t);                      \
    template void adj_predict<R>(const Inst<R>*, const Inst<R>&, const R*, const R*, const R*, const R*,          \
                                 const double*, const double*, const double*, double*, double*, double*, double*,  \
                                 double*, double*, hipStream_t);                                                   \
    template void adj_rhonu<R>(const Inst<R>*, const Inst<R>&, const R*, const double*, const double*, double*,    \
                               double*, double*, hipStream_t);

VOF_ADJ_INSTANTIATE(float)
VOF_ADJ_INSTANTIATE(double)

}  // namespace vof
