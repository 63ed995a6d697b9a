// vof_kernels.hip — hand-written gfx950 kernels of the VOF timestep.
//
// Every kernel works on an ensemble: a table of `Inst` records plus a prefix
// table of work-group counts (find_inst), so the plain solver (one instance)
// and the batched solver (many instances of uneven size) run the same code
// and produce the same bits.
//
// Floating-point contract: built with -ffp-contract=off and IEEE division, and
// every expression keeps the association of the reference kernel it replaces,
// so each field is bit-identical to the reference's arithmetic (the CPU oracle
// checks this).  What changes is data movement: stages are fused, derived
// fields (rho, nu) are recomputed from F in registers, the Jacobi right-hand
// side is formed once per step, and each FCT direction is one kernel.
#include "vof_device.h"
#include "vof_kernels.h"

namespace vof {

// ---------------------------------------------------------------------------
// Shared launch prologue: logical work-group, instance and local work-group.
#define VOF_PROLOGUE(R)                                                   \
    const int b_ = xcd_remap(blockIdx.x, gridDim.x);                      \
    const int inst_ = (ninst == 1) ? 0 : find_inst(boff, ninst, b_);      \
    const int lb = b_ - boff[inst_];                                      \
    const Inst<R>& I = tab[inst_];

// Column mapping: 256 lanes over the (nx+2)(ny+2) plane, x fastest, each lane
// walks KZ consecutive planes.  Returns false for lanes past the plane.
struct Col {
    int i, j, k0, k1;
};

template <class R>
__device__ __forceinline__ bool col_map(const Inst<R>& I, int lb, int kz, int klo, int khi, Col& c) {
    const int plane = I.sx * (I.ny + 2);
    const int npb = (plane + 255) >> 8;
    const int pb = lb % npb, zb = lb / npb;
    const int t = pb * 256 + (int)threadIdx.x;
    if (t >= plane) return false;
    c.j = t / I.sx;
    c.i = t - c.j * I.sx;
    c.k0 = klo + zb * kz;
    c.k1 = min(khi, c.k0 + kz - 1);
    return true;
}

// ---------------------------------------------------------------------------
// Momentum predictor (advect_{u,v,w}_kernel, voftrial1.cu:230-399), fused: the
// thread of cell (i,j,k) forms u* on face i, v* on face j and w* on face k, each
// only on interior faces (i>=2, j>=2, k>=2 — faces on the walls are never
// written, as in the reference).  nu comes from F in registers (:95-104).
template <class R>
__device__ __forceinline__ R predict_u(const Inst<R>& I, const R* __restrict__ u, const R* __restrict__ v,
                                       const R* __restrict__ w, int c, R n0) {
    const int sx = I.sx, sy = I.sy;
    const R vh = R(0.25) * (v[c - 1] + v[c - 1 + sx] + v[c] + v[c + sx]);
    const R wh = R(0.25) * (w[c - 1] + w[c - 1 + sy] + w[c] + w[c + sy]);
    const R uc = u[c];
    const R uxm = u[c - 1], uxp = u[c + 1], uym = u[c - sx], uyp = u[c + sx], uzm = u[c - sy], uzp = u[c + sy];
    const R ddx = (uc > R(0)) ? (uc - uxm) * I.dxi : (uxp - uc) * I.dxi;
    const R ddy = (vh > R(0)) ? (uc - uym) * I.dyi : (uyp - uc) * I.dyi;
    const R ddz = (wh > R(0)) ? (uc - uzm) * I.dzi : (uzp - uc) * I.dzi;
    const R visc = n0 * (uxm - R(2) * uc + uxp) * I.dxi * I.dxi + n0 * (uym - R(2) * uc + uyp) * I.dyi * I.dyi +
                   n0 * (uzm - R(2) * uc + uzp) * I.dzi * I.dzi;
    const R adv = -uc * ddx - vh * ddy - wh * ddz;
    return uc + I.dt * (visc + adv + I.gx);
}

template <class R>
__device__ __forceinline__ R predict_v(const Inst<R>& I, const R* __restrict__ u, const R* __restrict__ v,
                                       const R* __restrict__ w, int c, R n0) {
    const int sx = I.sx, sy = I.sy;
    const R uh = R(0.25) * (u[c - sx] + u[c] + u[c + 1 - sx] + u[c + 1]);
    const R wh = R(0.25) * (w[c - sx + sy] + w[c - sx] + w[c] + w[c + sy]);
    const R vc = v[c];
    const R vxm = v[c - 1], vxp = v[c + 1], vym = v[c - sx], vyp = v[c + sx], vzm = v[c - sy], vzp = v[c + sy];
    const R ddx = (uh > R(0)) ? (vc - vxm) * I.dxi : (vxp - vc) * I.dxi;
    const R ddy = (vc > R(0)) ? (vc - vym) * I.dyi : (vyp - vc) * I.dyi;
    const R ddz = (wh > R(0)) ? (vc - vzm) * I.dzi : (vzp - vc) * I.dzi;
    const R visc = n0 * (vxm - R(2) * vc + vxp) * I.dxi * I.dxi + n0 * (vym - R(2) * vc + vyp) * I.dyi * I.dyi +
                   n0 * (vzm - R(2) * vc + vzp) * I.dzi * I.dzi;
    const R adv = -uh * ddx - vc * ddy - wh * ddz;
    return vc + I.dt * (visc + adv + I.gy);
}

template <class R>
__device__ __forceinline__ R predict_w(const Inst<R>& I, const R* __restrict__ u, const R* __restrict__ v,
                                       const R* __restrict__ w, int c, R n0) {
    const int sx = I.sx, sy = I.sy;
    const R uh = R(0.25) * (u[c + 1 - sy] + u[c - sy] + u[c + 1] + u[c]);
    const R vh = R(0.25) * (v[c + sx - sy] + v[c - sy] + v[c] + v[c + sx]);
    const R wc = w[c];
    const R wxm = w[c - 1], wxp = w[c + 1], wym = w[c - sx], wyp = w[c + sx], wzm = w[c - sy], wzp = w[c + sy];
    const R ddx = (uh > R(0)) ? (wc - wxm) * I.dxi : (wxp - wc) * I.dxi;
    const R ddy = (vh > R(0)) ? (wc - wym) * I.dyi : (wyp - wc) * I.dyi;
    const R ddz = (wc > R(0)) ? (wc - wzm) * I.dzi : (wzp - wc) * I.dzi;
    const R visc = n0 * (wxm - R(2) * wc + wxp) * I.dxi * I.dxi + n0 * (wym - R(2) * wc + wyp) * I.dyi * I.dyi +
                   n0 * (wzm - R(2) * wc + wzp) * I.dzi * I.dzi;
    const R adv = -uh * ddx - vh * ddy - wc * ddz;
    return wc + I.dt * (visc + adv + I.gz);
}

template <class R>
__global__ __launch_bounds__(256) void k_predict(const Inst<R>* __restrict__ tab, const int* __restrict__ boff,
                                                 int ninst, int fsel, int kz) {
    VOF_PROLOGUE(R)
    Col cm;
    if (!col_map(I, lb, kz, 1, I.nz, cm)) return;
    const int i = cm.i, j = cm.j;
    if (i < 1 || i > I.nx || j < 1 || j > I.ny) return;
    const R* __restrict__ F = I.F[fsel];
    const R* __restrict__ u = I.u;
    const R* __restrict__ v = I.v;
    const R* __restrict__ w = I.w;
    for (int k = cm.k0; k <= cm.k1; ++k) {
        const int c = k * I.sy + j * I.sx + i;
        const R n0 = nu_of(I, F[c]);
        if (i >= 2) I.us[c] = predict_u(I, u, v, w, c, n0);
        if (j >= 2) I.vs[c] = predict_v(I, u, v, w, c, n0);
        if (k >= 2) I.ws[c] = predict_w(I, u, v, w, c, n0);
    }
}

// ---------------------------------------------------------------------------
// Poisson right-hand side rho/dt * div(u*) (the rhs term of
// solve_p_jacobi_kernel, :418-422), formed once per step instead of in every
// sweep; it depends only on u* and rho, so the sweeps see the same bits.
template <class R>
__global__ __launch_bounds__(256) void k_rhs(const Inst<R>* __restrict__ tab, const int* __restrict__ boff,
                                             int ninst, int fsel, int kz) {
    VOF_PROLOGUE(R)
    Col cm;
    if (!col_map(I, lb, kz, 1, I.nz, cm)) return;
    const int i = cm.i, j = cm.j;
    if (i < 1 || i > I.nx || j < 1 || j > I.ny) return;
    const R* __restrict__ F = I.F[fsel];
    const R* __restrict__ us = I.us;
    const R* __restrict__ vs = I.vs;
    const R* __restrict__ ws = I.ws;
    for (int k = cm.k0; k <= cm.k1; ++k) {
        const int c = k * I.sy + j * I.sx + i;
        const R rho = rho_of(I, F[c]);
        I.rhs[c] = rho / I.dt *
                   ((us[c + 1] - us[c]) * I.dxi + (vs[c + I.sx] - vs[c]) * I.dyi + (ws[c + I.sy] - ws[c]) * I.dzi);
    }
}

// ---------------------------------------------------------------------------
// One Jacobi sweep (solve_p_jacobi_kernel, :401-439) with the precomputed rhs.
// Boundary neighbours carry a zero coefficient times the (zero) ghost p, kept
// in the expression so the rounding sequence is the reference's.
template <class R>
__device__ __forceinline__ R jacobi_cell(const Inst<R>& I, const R* __restrict__ p, R rhs, int c, int i, int j,
                                         int k) {
    const R ae = (i != I.nx) ? I.axx : R(0);
    const R aw = (i != 1) ? I.axx : R(0);
    const R an = (j != I.ny) ? I.ayy : R(0);
    const R as = (j != 1) ? I.ayy : R(0);
    const R af = (k != I.nz) ? I.azz : R(0);
    const R ab = (k != 1) ? I.azz : R(0);
    const R ap = -(ae + aw + an + as + af + ab);
    return (rhs - ae * p[c + 1] - aw * p[c - 1] - an * p[c + I.sx] - as * p[c - I.sx] - af * p[c + I.sy] -
            ab * p[c - I.sy]) / ap;
}

template <class R>
__global__ __launch_bounds__(256) void k_jacobi(const Inst<R>* __restrict__ tab, const int* __restrict__ boff,
                                                int ninst, int psrc, int kz) {
    VOF_PROLOGUE(R)
    Col cm;
    if (!col_map(I, lb, kz, 1, I.nz, cm)) return;
    const int i = cm.i, j = cm.j;
    if (i < 1 || i > I.nx || j < 1 || j > I.ny) return;
    const R* __restrict__ p = I.p[psrc];
    R* __restrict__ pt = I.p[psrc ^ 1];
    const R* __restrict__ rhs = I.rhs;
    for (int k = cm.k0; k <= cm.k1; ++k) {
        const int c = k * I.sy + j * I.sx + i;
        pt[c] = jacobi_cell(I, p, rhs[c], c, i, j, k);
    }
}

// ---------------------------------------------------------------------------
// Velocity correction (update_{u,v,w}_kernel, :441-493) fused with the
// velocity boundary conditions (apply_bc_velocity, :1107-1120).  The three BC
// passes (j, then i, then k, over full planes) reduce to closed forms:
//   u(i,j,k) = 0                        for i in {1, nx+1} (normal wall faces)
//            = u(0, cj, ck)             for i = 0 (that plane is never computed)
//            = u_corr(i, cj, ck)        for 2 <= i <= nx,
// with cj = clamp(j,1,ny), ck = clamp(k,1,nz) — the tangential ghosts copy
// the nearest interior face (zero gradient).  v and w are the same with the
// roles of the axes exchanged.  The work-item of every stored cell writes its
// own final value, so no pass ordering is needed.
template <class R>
__device__ __forceinline__ R corr(const Inst<R>& I, const R* __restrict__ F, const R* __restrict__ st,
                                  const R* __restrict__ p, int c, int off, R hi) {
    const R r = R(0.5) * (rho_of(I, F[c]) + rho_of(I, F[c - off]));
    return st[c] - I.dt / r * (p[c] - p[c - off]) * hi;
}

template <class R>
__global__ __launch_bounds__(256) void k_correct(const Inst<R>* __restrict__ tab, const int* __restrict__ boff,
                                                 int ninst, int fsel, int psel, int kz) {
    VOF_PROLOGUE(R)
    Col cm;
    if (!col_map(I, lb, kz, 0, I.nz + 1, cm)) return;
    const int i = cm.i, j = cm.j;
    const int nx = I.nx, ny = I.ny, nz = I.nz, sx = I.sx, sy = I.sy;
    const R* __restrict__ F = I.F[fsel];
    const R* __restrict__ p = I.p[psel];
    const int ci = min(max(i, 1), nx), cj = min(max(j, 1), ny);
    const bool jin = (j >= 1 && j <= ny), iin = (i >= 1 && i <= nx);
    for (int k = cm.k0; k <= cm.k1; ++k) {
        const int ck = min(max(k, 1), nz);
        const bool kin = (k >= 1 && k <= nz);
        const int c = k * sy + j * sx + i;
        // u
        if (i == 1 || i == nx + 1) {
            I.u[c] = R(0);
        } else if (i == 0) {
            if (!(jin && kin)) I.u[c] = I.u[ck * sy + cj * sx];
        } else {
            I.u[c] = corr(I, F, I.us, p, ck * sy + cj * sx + i, 1, I.dxi);
        }
        // v
        if (j == 1 || j == ny + 1) {
            I.v[c] = R(0);
        } else if (j == 0) {
            if (!(iin && kin)) I.v[c] = I.v[ck * sy + ci];
        } else {
            I.v[c] = corr(I, F, I.vs, p, ck * sy + j * sx + ci, sx, I.dyi);
        }
        // w
        if (k == 1 || k == nz + 1) {
            I.w[c] = R(0);
        } else if (k == 0) {
            if (!(iin && jin)) I.w[c] = I.w[cj * sx + ci];
        } else {
            I.w[c] = corr(I, F, I.ws, p, k * sy + cj * sx + ci, sy, I.dzi);
        }
    }
}

// ---------------------------------------------------------------------------
// FCT, one fused kernel per direction (fct_{x,y,z}_sweep_stage1..4,
// :509-905, plus the wrappers :1137-1174).  Along the sweep line, with face f
// the low face of cell f and vel(f) its velocity:
//   L(f), H(f)  donor-cell and downwind fluxes  vel*dt*F(upwind|downwind)
//   a(f)        = H(f) - L(f)
//   Ftd(c)      = ((F(c) + <(L(c) - L(c+1)) scaled>) * vol) / dv(c),  clipped if outside [0,1]
//   rp/rm(c)    Zalesak ratios from Ftd(c-1..c+1), a(c), a(c+1)
//   C(f)        = a(f) >= 0 ? min(rp(f), rm(f-1)) : min(rp(f-1), rm(f))
//   Fnew(c)     = clamp01(Ftd(c) - ((<a*C>(c+1) - <a*C>(c)) / s4) * vol / dv(c))
// The reference's never-written entries read zero (Ftd, rp, rm outside the
// interior; C on the first face), and the always-zero cross-direction scratch
// terms (ay/cy in the x sweep, ax/cx in the y sweep) are dropped: they only
// add signed zeros, which the final clamp01 maps to +0.
//
// Per-direction scale factors, as the reference associates them:
//   stage 1   x: (((d*dy)*dz)/vol)     y: ((d*dy)/(dx*dy))    z: (((d*dy)*dx)/vol)
//   Q factor  x: dx   y: dx   z: dz   (:579, :718, :853)
//   stage 4   x: /dy  y: /dy  z: /dz  (:643, :782, :902)

template <class R, int DIR>
__device__ __forceinline__ R fct_stage1_scale(const Inst<R>& I, R d) {
    if (DIR == 0) return d * I.dy * I.dz / I.vol;
    if (DIR == 1) return d * I.dy / I.dxdy;
    return d * I.dy * I.dx / I.vol;
}

// Gas cells stay exactly zero: if F vanishes on the whole 7-cell stencil of a
// cell, every flux is a signed zero, Ftd = +0, rp = rm = 0 and Fnew = clamp01(+-0)
// = +0, so the divisions are skipped (wave-uniformly) with identical results.

// x sweep: lanes run along x over the flattened (x,y) plane of one k slab.
// Neighbour values move across lanes (ds_bpermute shuffles); a wave of 64 lanes
// covers 64 consecutive cells and stores the middle 58, which have their whole
// +-3 stencil inside the wave.  Row ends need no care: ghost lanes carry the
// reference's zero conventions, which decouple consecutive rows.
template <class R>
__global__ __launch_bounds__(256) void k_fct_x(const Inst<R>* __restrict__ tab, const int* __restrict__ boff,
                                               int ninst, int src) {
    VOF_PROLOGUE(R)
    const int P = I.sx * (I.ny + 2);
    const int wpp = (P + 57) / 58;
    const int bpp = (wpp + 3) >> 2;
    const int k = 1 + lb / bpp;
    const int wv = (lb % bpp) * 4 + ((int)threadIdx.x >> 6);
    const int lane = (int)threadIdx.x & 63;
    const int t = wv * 58 - 3 + lane;
    const bool inpl = (t >= 0 && t < P);
    const int tc = t < 0 ? 0 : (t >= P ? P - 1 : t);
    const int j = tc / I.sx, i = tc - j * I.sx;
    const bool cell = inpl && i >= 1 && i <= I.nx && j >= 1 && j <= I.ny;
    const bool store = cell && lane >= 3 && lane <= 60;
    const int c = k * I.sy + tc;
    const R f0 = I.F[src][c];
    const R u0 = I.u[c];
    R* __restrict__ Fo = I.F[src ^ 1];
    if (__all(f0 == R(0))) {
        if (store) Fo[c] = R(0);
        return;
    }
    const R fm = __shfl_up(f0, 1);
    const R fp = __shfl_down(f0, 1);
    const R up = __shfl_down(u0, 1);
    const R vdt = u0 * I.dt;
    const R L0 = (u0 >= R(0)) ? vdt * fm : vdt * f0;
    const R H0 = (u0 <= R(0)) ? vdt * fm : vdt * f0;
    const R a0 = H0 - L0;
    const R L1 = __shfl_down(L0, 1);
    const R a1 = __shfl_down(a0, 1);
    const R dv = I.vol - I.dtdydz * (up - u0);
    R ftd = R(0);
    if (cell) {
        R o = (f0 + fct_stage1_scale<R, 0>(I, L0 - L1)) * I.vol / dv;
        if (o > R(1) || o < R(0)) o = clamp01(o);
        ftd = o;
    }
    const R ftm = __shfl_up(ftd, 1);
    const R ftp = __shfl_down(ftd, 1);
    R rp = R(0), rm = R(0);
    if (cell) {
        const R fmx = vmax(ftd, vmax(ftm, ftp));
        const R fmn = vmin(ftd, vmin(ftm, ftp));
        const R pp = vmax(R(0), a0) - vmin(R(0), a1);
        if (pp > R(0)) rp = vmin(R(1), (fmx - ftd) * I.dx / pp);
        const R pm = vmax(R(0), a1) - vmin(R(0), a0);
        if (pm > R(0)) rm = vmin(R(1), (ftd - fmn) * I.dx / pm);
    }
    const R rp1 = __shfl_down(rp, 1);
    const R rm1 = __shfl_down(rm, 1);
    const R c1 = cell ? ((a1 >= R(0)) ? vmin(rp1, rm) : vmin(rp, rm1)) : R(0);
    const R ac1 = a1 * c1;
    const R ac0 = __shfl_up(ac1, 1);
    if (store) Fo[c] = clamp01(ftd - ((ac1 - ac0) / I.dy) * I.vol / dv);
}

// y and z sweeps: lanes run along x (coalesced), each lane marches along the
// sweep line through a register window, producing one Fnew per step two cells
// behind the Ftd front.  Lines are cut into chunks of ch cells with a 5-cell
// warm-up, so that there are enough waves to cover memory latency.
template <class R, int DIR>
__global__ __launch_bounds__(256) void k_fct_march(const Inst<R>* __restrict__ tab, const int* __restrict__ boff,
                                                   int ninst, int src, int ch) {
    VOF_PROLOGUE(R)
    const int nd = (DIR == 1) ? I.ny : I.nz;
    const int no = (DIR == 1) ? I.nz : I.ny;
    const int ntr = no * I.sx;
    const int nbt = (ntr + 255) >> 8;
    const int tb = lb % nbt, chn = lb / nbt;
    const int t = tb * 256 + (int)threadIdx.x;
    if (t >= ntr) return;
    const int o = 1 + t / I.sx;
    const int i = t - (o - 1) * I.sx;
    if (i < 1 || i > I.nx) return;
    const int S = (DIR == 1) ? I.sx : I.sy;
    const int base = (DIR == 1) ? (o * I.sy + i) : (o * I.sx + i);
    const R* __restrict__ F = I.F[src] + base;
    R* __restrict__ Fo = I.F[src ^ 1] + base;
    const R* __restrict__ V = ((DIR == 1) ? I.v : I.w) + base;
    const R kdv = (DIR == 1) ? I.dtdxdz : I.dtdxdy;
    const R qf = (DIR == 1) ? I.dx : I.dz;
    const R s4 = (DIR == 1) ? I.dy : I.dz;
    const R dt = I.dt, vol = I.vol;
    const int m0 = 1 + chn * ch;
    const int m1 = min(nd, m0 + ch - 1);
    auto cl = [nd](int m) { return m < 0 ? 0 : (m > nd + 1 ? nd + 1 : m); };

    int c = m0 - 2;
    const R Fcm = F[cl(c - 1) * S];
    R Fc = F[cl(c) * S];
    R Vc = V[cl(c) * S];
    R Lc, ac;
    {
        const R vdt = Vc * dt;
        Lc = (Vc >= R(0)) ? vdt * Fcm : vdt * Fc;
        const R Hc = (Vc <= R(0)) ? vdt * Fcm : vdt * Fc;
        ac = Hc - Lc;
    }
    int zr = (Fc == R(0)) ? ((Fcm == R(0)) ? 2 : 1) : 0;
    R am1 = R(0), am2 = R(0), ftm1 = R(0), ftm2 = R(0), dvm1 = R(1), dvm2 = R(1);
    R rpm1 = R(0), rmm1 = R(0), rpm2 = R(0), rmm2 = R(0), cm2 = R(0);
    R Fn = F[cl(c + 1) * S];
    R Vn = V[cl(c + 1) * S];
    for (; c <= m1 + 2; ++c) {
        const R Fcp = Fn, Vcp = Vn;
        Fn = F[cl(c + 2) * S];
        Vn = V[cl(c + 2) * S];
        zr = (Fcp == R(0)) ? zr + 1 : 0;
        const R dvc = vol - kdv * (Vcp - Vc);
        R Lcp, acp, ftc, cm1;
        if (zr >= 7) {
            // F(c-5..c+1) all zero: every value formed this step is (signed) zero.
            Lcp = R(0); acp = R(0); ftc = R(0);
            rpm1 = R(0); rmm1 = R(0); cm1 = R(0);
            if (c - 2 >= m0) Fo[(c - 2) * S] = R(0);
        } else {
            const R vdt = Vcp * dt;
            Lcp = (Vcp >= R(0)) ? vdt * Fc : vdt * Fcp;
            const R Hcp = (Vcp <= R(0)) ? vdt * Fc : vdt * Fcp;
            acp = Hcp - Lcp;
            ftc = R(0);
            if (c >= 1 && c <= nd) {
                R oo = (Fc + fct_stage1_scale<R, DIR>(I, Lc - Lcp)) * vol / dvc;
                if (oo > R(1) || oo < R(0)) oo = clamp01(oo);
                ftc = oo;
            }
            // Zalesak ratios of cell c-1
            rpm1 = R(0); rmm1 = R(0);
            if (c - 1 >= 1 && c - 1 <= nd) {
                const R fmx = vmax(ftm1, vmax(ftm2, ftc));
                const R fmn = vmin(ftm1, vmin(ftm2, ftc));
                const R pp = vmax(R(0), am1) - vmin(R(0), ac);
                if (pp > R(0)) rpm1 = vmin(R(1), (fmx - ftm1) * qf / pp);
                const R pm = vmax(R(0), ac) - vmin(R(0), am1);
                if (pm > R(0)) rmm1 = vmin(R(1), (ftm1 - fmn) * qf / pm);
            }
            // limiter on face c-1
            cm1 = R(0);
            if (c - 1 >= 2 && c - 1 <= nd + 1) cm1 = (am1 >= R(0)) ? vmin(rpm1, rmm2) : vmin(rpm2, rmm1);
            // corrected value of cell c-2
            if (c - 2 >= m0) Fo[(c - 2) * S] = clamp01(ftm2 - ((am1 * cm1 - am2 * cm2) / s4) * vol / dvm2);
        }
        // shift the window
        Fc = Fcp; Vc = Vcp; Lc = Lcp;
        am2 = am1; am1 = ac; ac = acp;
        ftm2 = ftm1; ftm1 = ftc;
        dvm2 = dvm1; dvm1 = dvc;
        rpm2 = rpm1; rmm2 = rmm1;
        cm2 = cm1;
    }
}

// ---------------------------------------------------------------------------
// F ghost layer (apply_bc_F, :1122-1135 -> set_bc_F_{j,i,k}, :1053-1105).  The
// three passes leave every ghost equal to the nearest interior cell,
// F(clamp(i), clamp(j), clamp(k)); this kernel writes that directly, each
// ghost once: k faces (full planes), then j faces and i faces of the interior
// slabs.  post_process_f (:502-507) is the identity here: every interior value
// already came out of clamp01 in the last sweep, and clamp01 is idempotent on
// its own outputs; ghosts are overwritten.
template <class R>
__global__ __launch_bounds__(256) void k_ghost_F(const Inst<R>* __restrict__ tab, const int* __restrict__ boff,
                                                 int ninst, int fsel) {
    VOF_PROLOGUE(R)
    const int nx = I.nx, ny = I.ny, nz = I.nz, sx = I.sx, sy = I.sy;
    const int nk = 2 * sx * (ny + 2), nj = 2 * sx * nz, ni = 2 * ny * nz;
    int t = lb * 256 + (int)threadIdx.x;
    int i, j, k;
    if (t < nk) {
        const int side = t / (sx * (ny + 2));
        const int r = t - side * sx * (ny + 2);
        j = r / sx; i = r - j * sx; k = side ? nz + 1 : 0;
    } else if ((t -= nk) < nj) {
        const int side = t / (sx * nz);
        const int r = t - side * sx * nz;
        k = 1 + r / sx; i = r - (k - 1) * sx; j = side ? ny + 1 : 0;
    } else if ((t -= nj) < ni) {
        const int side = t / (ny * nz);
        const int r = t - side * ny * nz;
        k = 1 + r / ny; j = 1 + r - (k - 1) * ny; i = side ? nx + 1 : 0;
    } else {
        return;
    }
    R* __restrict__ F = I.F[fsel];
    const int ci = min(max(i, 1), nx), cj = min(max(j, 1), ny), ck = min(max(k, 1), nz);
    F[k * sy + j * sx + i] = F[ck * sy + cj * sx + ci];
}

// ---------------------------------------------------------------------------
// Mean-pressure removal (extension; no reference counterpart): subtract the
// FP64 interior mean, stored by the reduction, from interior p.
template <class R>
__global__ __launch_bounds__(256) void k_sub_mean(const Inst<R>* __restrict__ tab, const int* __restrict__ boff,
                                                  int ninst, int psel, const double* __restrict__ means, int kz) {
    VOF_PROLOGUE(R)
    Col cm;
    if (!col_map(I, lb, kz, 1, I.nz, cm)) return;
    const int i = cm.i, j = cm.j;
    if (i < 1 || i > I.nx || j < 1 || j > I.ny) return;
    const R m = R(means[inst_]);
    R* __restrict__ p = I.p[psel];
    for (int k = cm.k0; k <= cm.k1; ++k) {
        const int c = k * I.sy + j * I.sx + i;
        p[c] = p[c] - m;
    }
}

// ---------------------------------------------------------------------------
// Deterministic FP64 reductions.  Pass 1: each work-group reduces its cells
// (wave shuffles, then LDS across the 4 waves) and writes one partial per
// quantity; pass 2: one work-group per instance sums the partials in a fixed
// order.  Quantities: 0 sum F, 1 sum p, 2 sum (rhs - A p)^2, 3 sum rhs^2,
// 4 min F, 5 max F, 6 non-finite count (u,v,w,p,F over all stored cells).
constexpr int NRED = 7;

__device__ __forceinline__ double wave_sum(double x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}
__device__ __forceinline__ double wave_min(double x) {
    for (int o = 32; o > 0; o >>= 1) x = fmin(x, __shfl_xor(x, o));
    return x;
}
__device__ __forceinline__ double wave_max(double x) {
    for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o));
    return x;
}

template <class R>
__global__ __launch_bounds__(256) void k_reduce1(const Inst<R>* __restrict__ tab, const int* __restrict__ boff,
                                                 int ninst, int fsel, int psel, int kz, int mode,
                                                 double* __restrict__ part) {
    const int b_ = blockIdx.x;
    const int inst_ = (ninst == 1) ? 0 : find_inst(boff, ninst, b_);
    const int lb = b_ - boff[inst_];
    const Inst<R>& I = tab[inst_];
    double s[NRED] = {0.0, 0.0, 0.0, 0.0, 1e300, -1e300, 0.0};
    Col cm;
    if (col_map(I, lb, kz, 0, I.nz + 1, cm)) {
        const int i = cm.i, j = cm.j;
        const R* __restrict__ F = I.F[fsel];
        const R* __restrict__ p = I.p[psel];
        const bool ij_in = (i >= 1 && i <= I.nx && j >= 1 && j <= I.ny);
        const double ax = double(I.axx), ay = double(I.ayy), az = double(I.azz);
        for (int k = cm.k0; k <= cm.k1; ++k) {
            const int c = k * I.sy + j * I.sx + i;
            if (mode == 0) {
                s[6] += (!isfinite(double(F[c]))) + (!isfinite(double(p[c]))) + (!isfinite(double(I.u[c]))) +
                        (!isfinite(double(I.v[c]))) + (!isfinite(double(I.w[c])));
            }
            if (!ij_in || k < 1 || k > I.nz) continue;
            const double f = double(F[c]);
            const double pc = double(p[c]);
            s[0] += f;
            s[1] += pc;
            if (mode == 0) {
                s[4] = fmin(s[4], f);
                s[5] = fmax(s[5], f);
                const double ae = (i != I.nx) ? ax : 0.0, aw = (i != 1) ? ax : 0.0;
                const double an = (j != I.ny) ? ay : 0.0, as = (j != 1) ? ay : 0.0;
                const double af = (k != I.nz) ? az : 0.0, ab = (k != 1) ? az : 0.0;
                const double ap = -(ae + aw + an + as + af + ab);
                const double b = double(I.rhs[c]);
                const double Ap = ae * double(p[c + 1]) + aw * double(p[c - 1]) + an * double(p[c + I.sx]) +
                                  as * double(p[c - I.sx]) + af * double(p[c + I.sy]) + ab * double(p[c - I.sy]) +
                                  ap * pc;
                s[2] += (b - Ap) * (b - Ap);
                s[3] += b * b;
            }
        }
    }
    __shared__ double sh[4][NRED];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NRED; ++q) {
        const double r = (q == 4) ? wave_min(s[q]) : (q == 5) ? wave_max(s[q]) : wave_sum(s[q]);
        if (lane == 0) sh[wid][q] = r;
    }
    __syncthreads();
    if (threadIdx.x < NRED) {
        const int q = threadIdx.x;
        double r = sh[0][q];
        for (int w2 = 1; w2 < 4; ++w2) r = (q == 4) ? fmin(r, sh[w2][q]) : (q == 5) ? fmax(r, sh[w2][q]) : r + sh[w2][q];
        part[(size_t)b_ * NRED + q] = r;
    }
}

// Pass 2: one work-group per instance.  out[inst*NRED + q].  If means is not
// null, also stores the p mean (sum p / n^3) for k_sub_mean.
__global__ __launch_bounds__(256) void k_reduce2(const int* __restrict__ boff, const double* __restrict__ part,
                                                 const long long* __restrict__ ncell, double* __restrict__ out,
                                                 double* __restrict__ means) {
    const int inst = blockIdx.x;
    const int b0 = boff[inst], b1 = boff[inst + 1];
    __shared__ double sh[4][NRED];
    double s[NRED] = {0.0, 0.0, 0.0, 0.0, 1e300, -1e300, 0.0};
    for (int b = b0 + (int)threadIdx.x; b < b1; b += 256) {
#pragma unroll
        for (int q = 0; q < NRED; ++q) {
            const double x = part[(size_t)b * NRED + q];
            s[q] = (q == 4) ? fmin(s[q], x) : (q == 5) ? fmax(s[q], x) : s[q] + x;
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NRED; ++q) {
        const double r = (q == 4) ? wave_min(s[q]) : (q == 5) ? wave_max(s[q]) : wave_sum(s[q]);
        if (lane == 0) sh[wid][q] = r;
    }
    __syncthreads();
    if (threadIdx.x < NRED) {
        const int q = threadIdx.x;
        double r = sh[0][q];
        for (int w2 = 1; w2 < 4; ++w2) r = (q == 4) ? fmin(r, sh[w2][q]) : (q == 5) ? fmax(r, sh[w2][q]) : r + sh[w2][q];
        out[(size_t)inst * NRED + q] = r;
        if (q == 1 && means) means[inst] = r / double(ncell[inst]);
    }
}

// ---------------------------------------------------------------------------
// Host launchers.
template <class R>
void launch_predict(const Ens<R>& e, const Geo& g, int fsel, hipStream_t st) {
    hipLaunchKernelGGL(k_predict<R>, dim3(g.nblk), dim3(256), 0, st, e.tab, g.boff, e.ninst, fsel, g.kz);
}
template <class R>
void launch_rhs(const Ens<R>& e, const Geo& g, int fsel, hipStream_t st) {
    hipLaunchKernelGGL(k_rhs<R>, dim3(g.nblk), dim3(256), 0, st, e.tab, g.boff, e.ninst, fsel, g.kz);
}
template <class R>
void launch_jacobi(const Ens<R>& e, const Geo& g, int psrc, hipStream_t st) {
    hipLaunchKernelGGL(k_jacobi<R>, dim3(g.nblk), dim3(256), 0, st, e.tab, g.boff, e.ninst, psrc, g.kz);
}
template <class R>
void launch_correct(const Ens<R>& e, const Geo& g, int fsel, int psel, hipStream_t st) {
    hipLaunchKernelGGL(k_correct<R>, dim3(g.nblk), dim3(256), 0, st, e.tab, g.boff, e.ninst, fsel, psel, g.kz);
}
template <class R>
void launch_fct(const Ens<R>& e, const Geo& g, int dir, int src, hipStream_t st) {
    if (dir == 0)
        hipLaunchKernelGGL(k_fct_x<R>, dim3(g.nblk), dim3(256), 0, st, e.tab, g.boff, e.ninst, src);
    else if (dir == 1)
        hipLaunchKernelGGL((k_fct_march<R, 1>), dim3(g.nblk), dim3(256), 0, st, e.tab, g.boff, e.ninst, src, g.kz);
    else
        hipLaunchKernelGGL((k_fct_march<R, 2>), dim3(g.nblk), dim3(256), 0, st, e.tab, g.boff, e.ninst, src, g.kz);
}
template <class R>
void launch_ghost_F(const Ens<R>& e, const Geo& g, int fsel, hipStream_t st) {
    hipLaunchKernelGGL(k_ghost_F<R>, dim3(g.nblk), dim3(256), 0, st, e.tab, g.boff, e.ninst, fsel);
}
template <class R>
void launch_sub_mean(const Ens<R>& e, const Geo& g, int psel, const double* means, hipStream_t st) {
    hipLaunchKernelGGL(k_sub_mean<R>, dim3(g.nblk), dim3(256), 0, st, e.tab, g.boff, e.ninst, psel, means, g.kz);
}
template <class R>
void launch_reduce(const Ens<R>& e, const Geo& g, int fsel, int psel, int mode, double* part,
                   const long long* ncell, double* out, double* means, hipStream_t st) {
    hipLaunchKernelGGL(k_reduce1<R>, dim3(g.nblk), dim3(256), 0, st, e.tab, g.boff, e.ninst, fsel, psel, g.kz, mode,
                       part);
    hipLaunchKernelGGL(k_reduce2, dim3(e.ninst), dim3(256), 0, st, g.boff, (const double*)part, ncell, out, means);
}

#define VOF_INSTANTIATE(R)                                                                               \
    template void launch_predict<R>(const Ens<R>&, const Geo&, int, hipStream_t);                        \
    template void launch_rhs<R>(const Ens<R>&, const Geo&, int, hipStream_t);                            \
    template void launch_jacobi<R>(const Ens<R>&, const Geo&, int, hipStream_t);                         \
    template void launch_correct<R>(const Ens<R>&, const Geo&, int, int, hipStream_t);                   \
    template void launch_fct<R>(const Ens<R>&, const Geo&, int, int, hipStream_t);                       \
    template void launch_ghost_F<R>(const Ens<R>&, const Geo&, int, hipStream_t);                        \
    template void launch_sub_mean<R>(const Ens<R>&, const Geo&, int, const double*, hipStream_t);        \
    template void launch_reduce<R>(const Ens<R>&, const Geo&, int, int, int, double*, const long long*, \
                                   double*, double*, hipStream_t);

VOF_INSTANTIATE(float)
VOF_INSTANTIATE(double)

}  // namespace vof
