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
// checks this).  What changes is data movement:
//  * column kernels (predictor, rhs, Jacobi, correction) march along z with the
//    z-neighbours carried in registers; x-neighbours come from the adjacent
//    lane (wave shuffles) in overlapped waves of 64 lanes that store the middle
//    64-2H; y-neighbours are one extra coalesced row load;
//  * most kernels have a register-blocked form for chunks of 4/8/16 planes or
//    cells: every load of a chunk is issued before its first point is formed;
//  * derived fields (rho, nu) are recomputed from F in registers;
//  * the Jacobi right-hand side is formed once per step, and two to five
//    sweeps can share a launch (LDS ring between the rows of a 16-row tile);
//  * each FCT direction is one fused kernel, the last one of a step also
//    applies post_process_f and writes the F ghost layer;
//  * work-groups go to XCDs in contiguous ranges (neighbouring tiles share an
//    L2), except in the FCT sweeps, whose gas skip makes the work depend on
//    the data and which use the dispatcher's round robin.
#include <algorithm>
#include <mutex>
#include <type_traits>

#include "../../include/vof_capi.h"
#include "vof_device.h"
#include "vof_kernels.h"

namespace vof {

// ---------------------------------------------------------------------------
// Shared launch prologue: logical work-group, instance and local work-group.
// The instance record is copied into (scalar) registers: buffer stores are
// opaque to alias analysis, and a reference would be re-read after each one.
#define VOF_PROLOGUE(R) VOF_PROLOGUE_B(R, xcd_remap(blockIdx.x, gridDim.x))
// Work-group b on XCD b % 8 (the dispatcher's own round robin), for kernels
// whose work depends on the data: the FCT sweeps skip gas wave-uniformly, and
// a contiguous range per XCD would give the XCDs that own the liquid region
// all the limiter arithmetic.
#define VOF_PROLOGUE_RR(R) VOF_PROLOGUE_B(R, (int)blockIdx.x)
#define VOF_PROLOGUE_B(R, BLOCK)                                          \
    const int b_ = (BLOCK);                                               \
    int lb = b_, inst_ = 0;                                               \
    Inst<R> I = ea.one;                                                   \
    if (ea.ninst != 1) {                                                  \
        inst_ = find_inst(ea.boff, ea.ninst, b_);                         \
        lb = b_ - ea.boff[inst_];                                         \
        I = ea.tab[inst_];                                                \
    }

template <class R> __device__ __forceinline__ R sh_up(R x) { return lane_up(x); }  // value of lane-1
template <class R> __device__ __forceinline__ R sh_dn(R x) { return lane_dn(x); }  // value of lane+1

// Overlapped-wave column mapping.  The (nx+2)(ny+2) plane is flattened x
// fastest; wave w covers positions [w*S - H, w*S - H + 64) with S = 64 - 2H and
// stores only its middle S lanes, whose +-H neighbours along x are in the wave.
// Each wave walks the planes k0..k1 of its z chunk.  All lanes of a wave stay
// active (clamped positions) so that every shuffle reads a live lane.
struct WMap {
    int lane, t, tc, i, j, k0, k1;
    bool own;  // lane's cell is stored by this wave
};

// PAD8: the work-groups of one plane chunk are padded to a multiple of 8, so
// that under the dispatcher's round robin the chunks of one column run on one
// XCD (for kernels whose z windows overlap between chunks).
template <class R, int H, bool PAD8 = false>
__device__ __forceinline__ bool wmap(const Inst<R>& I, int lb, int kz, int klo, int khi, WMap& m) {
    constexpr int S = 64 - 2 * H;
    const int P = I.sx * (I.ny + 2);
    const int wpp = (P + S - 1) / S;
    const int bpp = PAD8 ? ((((wpp + 3) >> 2) + 7) & ~7) : ((wpp + 3) >> 2);
    const int zb = lb / bpp;
    const int wv = (lb - zb * bpp) * 4 + ((int)threadIdx.x >> 6);
    if (wv >= wpp) return false;  // wave-uniform
    m.lane = (int)threadIdx.x & 63;
    m.t = wv * S - H + m.lane;
    m.tc = min(max(m.t, 0), P - 1);
    m.j = m.tc / I.sx;
    m.i = m.tc - m.j * I.sx;
    m.own = (m.lane >= H) && (m.lane < 64 - H) && (m.t >= 0) && (m.t < P);
    m.k0 = klo + zb * kz;
    m.k1 = min(khi, m.k0 + kz - 1);
    return true;
}

// ---------------------------------------------------------------------------
// Momentum predictor (advect_{u,v,w}_kernel, voftrial1.cu:230-399) for one
// face: first-order upwind advection by (ax, ay, az), explicit viscosity with
// the cell's nu, plus gravity.  For u* the x advecting velocity is u itself,
// for v* the y one is v, for w* the z one is w; the other two are the 4-point
// averages of the reference.  Same association as :266-284.
template <class R>
__device__ __forceinline__ R mom(const Inst<R>& I, R c, R xm, R xp, R ym, R yp, R zm, R zp, R ax, R ay, R az,
                                 R n0, R g) {
    const R ddx = (ax > R(0)) ? (c - xm) * I.dxi : (xp - c) * I.dxi;
    const R ddy = (ay > R(0)) ? (c - ym) * I.dyi : (yp - c) * I.dyi;
    const R ddz = (az > R(0)) ? (c - zm) * I.dzi : (zp - c) * I.dzi;
    const R visc = n0 * (xm - R(2) * c + xp) * I.dxi * I.dxi + n0 * (ym - R(2) * c + yp) * I.dyi * I.dyi +
                   n0 * (zm - R(2) * c + zp) * I.dzi * I.dzi;
    const R adv = -ax * ddx - ay * ddy - az * ddz;
    return c + I.dt * (visc + adv + g);
}

// Fused predictor: the lane of cell (i,j,k) forms u* on face i, v* on face j
// and w* on face k (interior faces only: i>=2, j>=2, k>=2; wall faces are never
// written, as in the reference).  Per plane: 10 coalesced loads (u,v,w at k+1
// and at rows j-1, j+1 of plane k+1, F at k), 8 lane shuffles.
// KZ > 0: the z chunk is KZ planes and the plane loop is unrolled, so the
// register window moves by renaming instead of copies; KZ = 0: any chunk.
template <class R, int KZ>
__global__ __launch_bounds__(256) void k_predict(const EnsArg<R> ea,
                                                 int fsel, int kz) {
    VOF_PROLOGUE(R)
    WMap m;
    if (!wmap<R, 1>(I, lb, kz, 1, I.nz, m)) return;
    const int i = m.i, j = m.j, sx = I.sx, sy = I.sy;
    const int P = sx * (I.ny + 2);
    const bool out = m.own && i >= 1 && i <= I.nx && j >= 1 && j <= I.ny;
    const int tc = m.tc;
    const int tjm = max(tc - sx, 0), tjp = min(tc + sx, P - 1);
    const Buf<R> F(sel2(I.F, fsel), I.ntot), u(I.u, I.ntot), v(I.v, I.ntot), w(I.w, I.ntot);
    const Buf<R> us(I.us, I.ntot), vs(I.vs, I.ntot), ws(I.ws, I.ntot);
    int k = m.k0;
    int o = k * sy;
    R u_km = u.ld(tc, o - sy), u_k = u.ld(tc, o);
    R v_km = v.ld(tc, o - sy), v_k = v.ld(tc, o);
    R w_km = w.ld(tc, o - sy), w_k = w.ld(tc, o);
    R ujm = u.ld(tjm, o), ujp = u.ld(tjp, o);
    R vjm = v.ld(tjm, o), vjp = v.ld(tjp, o), vjp_km = v.ld(tjp, o - sy);
    R wjm = w.ld(tjm, o), wjp = w.ld(tjp, o);
    R udn_km = sh_dn(u_km);  // u(i+1, j, k-1)
    R wup_k = sh_up(w_k);    // w(i-1, j, k)
    // plane k+1 operands (and F of plane k), fetched one plane ahead
    const int olim = I.nz * sy;
    R nu = u.ld(tc, o + sy), nv = v.ld(tc, o + sy), nw = w.ld(tc, o + sy);
    R nujm = u.ld(tjm, o + sy), nujp = u.ld(tjp, o + sy), nvjm = v.ld(tjm, o + sy), nvjp = v.ld(tjp, o + sy);
    R nwjm = w.ld(tjm, o + sy), nwjp = w.ld(tjp, o + sy), nf = F.ld(tc, o);
    auto plane = [&]() __attribute__((always_inline)) {
        const R u_kp = nu, v_kp = nv, w_kp = nw;
        const R ujm_n = nujm, ujp_n = nujp, vjm_n = nvjm, vjp_n = nvjp, wjm_n = nwjm, wjp_n = nwjp;
        const R f = nf;
        {
            const int o2 = min(o + sy, olim), o3 = o2 + sy;
            nu = u.ld(tc, o3); nv = v.ld(tc, o3); nw = w.ld(tc, o3);
            nujm = u.ld(tjm, o3); nujp = u.ld(tjp, o3);
            nvjm = v.ld(tjm, o3); nvjp = v.ld(tjp, o3);
            nwjm = w.ld(tjm, o3); nwjp = w.ld(tjp, o3);
            nf = F.ld(tc, o2);
        }
        const R uup = sh_up(u_k), udn = sh_dn(u_k);
        const R vup = sh_up(v_k), vdn = sh_dn(v_k);
        const R wdn = sh_dn(w_k);
        const R vjp_up = sh_up(vjp);   // v(i-1, j+1, k)
        const R ujm_dn = sh_dn(ujm);   // u(i+1, j-1, k)
        const R wup_kp = sh_up(w_kp);  // w(i-1, j, k+1)
        if (out) {
            const R n0 = nu_of(I, f);
            if (i >= 2) {  // u* (:253-284)
                const R vh = R(0.25) * (vup + vjp_up + v_k + vjp);
                const R wh = R(0.25) * (wup_k + wup_kp + w_k + w_kp);
                us.st(mom(I, u_k, uup, udn, ujm, ujp, u_km, u_kp, u_k, vh, wh, n0, I.gx), tc, o);
            }
            if (j >= 2) {  // v* (:310-341)
                const R uh = R(0.25) * (ujm + u_k + ujm_dn + udn);
                const R wh = R(0.25) * (wjm_n + wjm + w_k + w_kp);
                vs.st(mom(I, v_k, vup, vdn, vjm, vjp, v_km, v_kp, uh, v_k, wh, n0, I.gy), tc, o);
            }
            if (k >= 2) {  // w* (:367-398)
                const R uh = R(0.25) * (udn_km + u_km + udn + u_k);
                const R vh = R(0.25) * (vjp_km + v_km + v_k + vjp);
                ws.st(mom(I, w_k, wup_k, wdn, wjm, wjp, w_km, w_kp, uh, vh, w_k, n0, I.gz), tc, o);
            }
        }
        u_km = u_k; u_k = u_kp;
        v_km = v_k; v_k = v_kp;
        w_km = w_k; w_k = w_kp;
        ujm = ujm_n; ujp = ujp_n;
        vjp_km = vjp; vjm = vjm_n; vjp = vjp_n;
        wjm = wjm_n; wjp = wjp_n;
        udn_km = udn; wup_k = wup_kp;
    };
    if constexpr (KZ > 0) {
        if (m.k1 - m.k0 + 1 == KZ) {  // every chunk but a short last one
#pragma unroll
            for (int e = 0; e < KZ; ++e, ++k, o += sy) plane();
            return;
        }
    }
    for (; k <= m.k1; ++k, o += sy) plane();
}

// The same predictor with each lane holding a band of RB rows (j = 1 + J*RB +
// r, r = 0..RB-1) of one column position: the y neighbours of the band's rows
// are registers, and only the two rows bordering the band are extra loads
// (3 + 6/RB loads of u, v, w per cell and plane instead of 9).  Waves overlap
// along the flattened (i, J) plane of band rows as k_predict's do along (i, j),
// storing the middle 62 lanes; a row's offset from the lane's band base is the
// wave-uniform r*sx, added to the plane offset.  Each face is formed by the
// same mom() with the same operands as k_predict, so the results are equal bit
// for bit.
template <class R, int RB, int KZ>
__global__ __launch_bounds__(256) void k_predict_band(const EnsArg<R> ea, int fsel) {
    VOF_PROLOGUE(R)
    const int nx = I.nx, ny = I.ny, sx = I.sx, sy = I.sy;
    const int nb = (ny + RB - 1) / RB;
    const int P = sx * nb;  // positions of the (i, J) plane
    const int wpp = (P + 61) / 62, bpp = (wpp + 3) >> 2;
    const int zb = lb / bpp;
    const int wv = (lb - zb * bpp) * 4 + ((int)threadIdx.x >> 6);
    if (wv >= wpp) return;  // wave-uniform
    const int lane = (int)threadIdx.x & 63;
    const int t = wv * 62 - 1 + lane;
    const int tc = min(max(t, 0), P - 1);
    const int J = tc / sx, i = tc - J * sx;
    const int j0 = 1 + J * RB;
    const bool own = lane >= 1 && lane < 63 && t >= 0 && t < P && i >= 1 && i <= nx;
    const int vo = j0 * sx + i;  // lane offset of row r = 0 within a plane
    const int k0 = 1 + zb * KZ, k1 = min(I.nz, k0 + KZ - 1);
    const Buf<R> F(sel2(I.F, fsel), I.ntot), u(I.u, I.ntot), v(I.v, I.ntot), w(I.w, I.ntot);
    const Buf<R> us(I.us, I.ntot), vs(I.vs, I.ntot), ws(I.ws, I.ntot);
    // rows q = 0 .. RB+1 are j0-1 .. j0+RB (row offset (q-1)*sx)
    int o = k0 * sy;
    R uk[RB + 2], vk[RB + 2], wk[RB + 2];  // plane k
    R um[RB], vm[RB + 1], wm[RB];          // plane k-1 (v: one row more, v(j+1, k-1))
    R nu[RB + 2], nv[RB + 2], nw[RB + 2];  // plane k+1, fetched one plane ahead
    R nf[RB];                              // F of plane k, fetched one plane ahead
#pragma unroll
    for (int q = 0; q < RB + 2; ++q) {
        const int so = o + (q - 1) * sx;
        uk[q] = u.ld(vo, so); vk[q] = v.ld(vo, so); wk[q] = w.ld(vo, so);
        nu[q] = u.ld(vo, so + sy); nv[q] = v.ld(vo, so + sy); nw[q] = w.ld(vo, so + sy);
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
        const int so = o - sy + r * sx;
        um[r] = u.ld(vo, so); vm[r] = v.ld(vo, so); wm[r] = w.ld(vo, so);
        nf[r] = F.ld(vo, o + r * sx);
    }
    vm[RB] = v.ld(vo, o - sy + RB * sx);
    const int olim = I.nz * sy;
    int k = k0;
    auto plane = [&]() __attribute__((always_inline)) {
        R up[RB + 2], vp[RB + 2], wp[RB + 2], f[RB];
#pragma unroll
        for (int q = 0; q < RB + 2; ++q) { up[q] = nu[q]; vp[q] = nv[q]; wp[q] = nw[q]; }
#pragma unroll
        for (int r = 0; r < RB; ++r) f[r] = nf[r];
        {
            const int o2 = min(o + sy, olim), o3 = o2 + sy;
#pragma unroll
            for (int q = 0; q < RB + 2; ++q) {
                const int so = o3 + (q - 1) * sx;
                nu[q] = u.ld(vo, so); nv[q] = v.ld(vo, so); nw[q] = w.ld(vo, so);
            }
#pragma unroll
            for (int r = 0; r < RB; ++r) nf[r] = F.ld(vo, o2 + r * sx);
        }
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const int q = r + 1;  // row j = j0 + r
            const R u_k = uk[q], v_k = vk[q], w_k = wk[q];
            const R ujm = uk[q - 1], ujp = uk[q + 1], vjm = vk[q - 1], vjp = vk[q + 1];
            const R wjm = wk[q - 1], wjp = wk[q + 1];
            const R u_km = um[r], v_km = vm[r], w_km = wm[r], vjp_km = vm[r + 1];
            const R u_kp = up[q], v_kp = vp[q], w_kp = wp[q], wjm_n = wp[q - 1];
            const R uup = sh_up(u_k), udn = sh_dn(u_k);
            const R vup = sh_up(v_k), vdn = sh_dn(v_k);
            const R wup_k = sh_up(w_k), wdn = sh_dn(w_k);
            const R vjp_up = sh_up(vjp);   // v(i-1, j+1, k)
            const R ujm_dn = sh_dn(ujm);   // u(i+1, j-1, k)
            const R wup_kp = sh_up(w_kp);  // w(i-1, j, k+1)
            const R udn_km = sh_dn(u_km);  // u(i+1, j, k-1)
            const int j = j0 + r;
            if (own && j <= ny) {
                const R n0 = nu_of(I, f[r]);
                if (i >= 2) {  // u* (:253-284)
                    const R vh = R(0.25) * (vup + vjp_up + v_k + vjp);
                    const R wh = R(0.25) * (wup_k + wup_kp + w_k + w_kp);
                    us.st(mom(I, u_k, uup, udn, ujm, ujp, u_km, u_kp, u_k, vh, wh, n0, I.gx), vo, o + r * sx);
                }
                if (j >= 2) {  // v* (:310-341)
                    const R uh = R(0.25) * (ujm + u_k + ujm_dn + udn);
                    const R wh = R(0.25) * (wjm_n + wjm + w_k + w_kp);
                    vs.st(mom(I, v_k, vup, vdn, vjm, vjp, v_km, v_kp, uh, v_k, wh, n0, I.gy), vo, o + r * sx);
                }
                if (k >= 2) {  // w* (:367-398)
                    const R uh = R(0.25) * (udn_km + u_km + udn + u_k);
                    const R vh = R(0.25) * (vjp_km + v_km + v_k + vjp);
                    ws.st(mom(I, w_k, wup_k, wdn, wjm, wjp, w_km, w_kp, uh, vh, w_k, n0, I.gz), vo, o + r * sx);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < RB; ++r) { um[r] = uk[r + 1]; wm[r] = wk[r + 1]; }
#pragma unroll
        for (int r = 0; r <= RB; ++r) vm[r] = vk[r + 1];
#pragma unroll
        for (int q = 0; q < RB + 2; ++q) { uk[q] = up[q]; vk[q] = vp[q]; wk[q] = wp[q]; }
    };
    if (k1 - k0 + 1 == KZ) {  // every chunk but a short last one
#pragma unroll
        for (int e = 0; e < KZ; ++e, ++k, o += sy) plane();
        return;
    }
    for (; k <= k1; ++k, o += sy) plane();
}

// ---------------------------------------------------------------------------
// Predictor with the density-weighted viscous term (VOF_FLAG_RHO_PREDICTOR;
// extension: the reference's predictor takes rho and never reads it, :233,
// :290, :347, and uses the nu of one cell, :277-280).  The viscous
// acceleration of a face between cells a (low) and b (high) along its axis d
// is div(mu grad phi) / rho_f in flux form, mu = rho nu per cell:
//   visc = ((T_x + T_y) + T_z) / (0.5 (rho_a + rho_b)),
//   T_t  = (m+ (phi(+e_t) - phi) - m- (phi - phi(-e_t))) * h_t^-1 * h_t^-1,
// with m+ = mu_b, m- = mu_a along d, and across it the edge averages
// m+ = 0.25 ((mu_a + mu_b) + (mu_a(+e_t) + mu_b(+e_t))),
// m- = 0.25 ((mu_a(-e_t) + mu_b(-e_t)) + (mu_a + mu_b)).  Advection and
// gravity are the reference's (:253-284 and the v, w forms), in the same
// association as the CPU oracle.  One lane per interior cell forms the faces
// it owns (u* on face i >= 2, v* on j >= 2, w* on k >= 2); operands are read
// straight from memory (this form is not on the measured path).
template <class R>
__device__ __forceinline__ R mu_of(const Inst<R>& I, R F) {
    return rho_of(I, F) * nu_of(I, F);
}

template <class R, int D>
__device__ __forceinline__ R visc_rho(const Inst<R>& I, const R* __restrict__ F, const R* __restrict__ phi, int c) {
    const int off[3] = {1, I.sx, I.sy};
    const R hi[3] = {I.dxi, I.dyi, I.dzi};
    const int ca = c - off[D], cb = c;
    const R Fa = F[ca], Fb = F[cb];
    const R ma = mu_of(I, Fa), mb = mu_of(I, Fb);
    R T[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const int o = off[t];
        R mp, mm;
        if (t == D) {
            mp = mb;
            mm = ma;
        } else {
            mp = R(0.25) * ((ma + mb) + (mu_of(I, F[ca + o]) + mu_of(I, F[cb + o])));
            mm = R(0.25) * ((mu_of(I, F[ca - o]) + mu_of(I, F[cb - o])) + (ma + mb));
        }
        T[t] = (mp * (phi[c + o] - phi[c]) - mm * (phi[c] - phi[c - o])) * hi[t] * hi[t];
    }
    const R rf = R(0.5) * (rho_of(I, Fa) + rho_of(I, Fb));
    return ((T[0] + T[1]) + T[2]) / rf;
}

// Advected, viscous and gravity update of face c of component D (u, v, w).
template <class R, int D>
__device__ __forceinline__ R predict_face_rho(const Inst<R>& I, const R* __restrict__ F, const R* __restrict__ u,
                                              const R* __restrict__ v, const R* __restrict__ w, int c) {
    const int sx = I.sx, sy = I.sy;
    const R* phi = (D == 0) ? u : (D == 1) ? v : w;
    const R pc = phi[c];
    R ax, ay, az, g;
    if (D == 0) {  // :253-264
        ax = pc;
        ay = R(0.25) * (v[c - 1] + v[c - 1 + sx] + v[c] + v[c + sx]);
        az = R(0.25) * (w[c - 1] + w[c - 1 + sy] + w[c] + w[c + sy]);
        g = I.gx;
    } else if (D == 1) {  // :310-321
        ax = R(0.25) * (u[c - sx] + u[c] + u[c + 1 - sx] + u[c + 1]);
        ay = pc;
        az = R(0.25) * (w[c - sx + sy] + w[c - sx] + w[c] + w[c + sy]);
        g = I.gy;
    } else {  // :367-378
        ax = R(0.25) * (u[c + 1 - sy] + u[c - sy] + u[c + 1] + u[c]);
        ay = R(0.25) * (v[c + sx - sy] + v[c - sy] + v[c] + v[c + sx]);
        az = pc;
        g = I.gz;
    }
    const R ddx = (ax > R(0)) ? (pc - phi[c - 1]) * I.dxi : (phi[c + 1] - pc) * I.dxi;
    const R ddy = (ay > R(0)) ? (pc - phi[c - sx]) * I.dyi : (phi[c + sx] - pc) * I.dyi;
    const R ddz = (az > R(0)) ? (pc - phi[c - sy]) * I.dzi : (phi[c + sy] - pc) * I.dzi;
    const R visc = visc_rho<R, D>(I, F, phi, c);
    const R adv = -ax * ddx - ay * ddy - az * ddz;
    return pc + I.dt * (visc + adv + g);
}

template <class R>
__global__ __launch_bounds__(256) void k_predict_rho(const EnsArg<R> ea, int fsel) {
    VOF_PROLOGUE(R)
    const int q = lb * 256 + (int)threadIdx.x;
    const int nxy = I.nx * I.ny;
    if (q >= nxy * I.nz) return;
    const int k = 1 + q / nxy, r = q - (k - 1) * nxy;
    const int j = 1 + r / I.nx, i = 1 + r - (j - 1) * I.nx;
    const int c = k * I.sy + j * I.sx + i;
    const R* __restrict__ F = sel2(I.F, fsel);
    if (i >= 2) I.us[c] = predict_face_rho<R, 0>(I, F, I.u, I.v, I.w, c);
    if (j >= 2) I.vs[c] = predict_face_rho<R, 1>(I, F, I.u, I.v, I.w, c);
    if (k >= 2) I.ws[c] = predict_face_rho<R, 2>(I, F, I.u, I.v, I.w, c);
}

// ---------------------------------------------------------------------------
// Poisson right-hand side rho/dt * div(u*) (the rhs term of
// solve_p_jacobi_kernel, :418-422), formed once per step instead of in every
// sweep; it depends only on u* and rho, so the sweeps see the same bits.
template <class R>
__global__ __launch_bounds__(256) void k_rhs(const EnsArg<R> ea,
                                             int fsel, int kz) {
    VOF_PROLOGUE(R)
    WMap m;
    if (!wmap<R, 1>(I, lb, kz, 1, I.nz, m)) return;
    const int i = m.i, j = m.j, sx = I.sx, sy = I.sy;
    const int P = sx * (I.ny + 2);
    const bool out = m.own && i >= 1 && i <= I.nx && j >= 1 && j <= I.ny;
    const int tc = m.tc, tjp = min(tc + sx, P - 1);
    const Buf<R> F(sel2(I.F, fsel), I.ntot), us(I.us, I.ntot), vs(I.vs, I.ntot), ws(I.ws, I.ntot), rhs(I.rhs, I.ntot);
    int o = m.k0 * sy;
    const int olim = I.nz * sy;
    R ws_k = ws.ld(tc, o);
    R n_us = us.ld(tc, o), n_vs = vs.ld(tc, o), n_vsp = vs.ld(tjp, o), n_ws = ws.ld(tc, o + sy), n_f = F.ld(tc, o);
    for (int k = m.k0; k <= m.k1; ++k, o += sy) {
        const R usk = n_us, vsk = n_vs, vsp = n_vsp, ws_kp = n_ws, f = n_f;
        const int o2 = min(o + sy, olim);
        n_us = us.ld(tc, o2);
        n_vs = vs.ld(tc, o2);
        n_vsp = vs.ld(tjp, o2);
        n_ws = ws.ld(tc, o2 + sy);
        n_f = F.ld(tc, o2);
        const R usp = sh_dn(usk);
        if (out)
            rhs.st(rho_of(I, f) / I.dt * ((usp - usk) * I.dxi + (vsp - vsk) * I.dyi + (ws_kp - ws_k) * I.dzi), tc, o);
        ws_k = ws_kp;
    }
}

// The same for z-chunks of KZ planes with the chunk's operands (u*, v* on the
// lane's and the next row, F on KZ planes, w* on KZ+1) loaded first.
template <class R, int KZ>
__global__ __launch_bounds__(256) void k_rhs_blk(const EnsArg<R> ea, int fsel) {
    VOF_PROLOGUE(R)
    WMap m;
    if (!wmap<R, 1>(I, lb, KZ, 1, I.nz, m)) return;
    const int i = m.i, j = m.j, sx = I.sx, sy = I.sy, nz = I.nz, k0 = m.k0, k1 = m.k1;
    const int P = sx * (I.ny + 2);
    const bool out = m.own && i >= 1 && i <= I.nx && j >= 1 && j <= I.ny;
    const int tc = m.tc, tjp = min(tc + sx, P - 1);
    const Buf<R> F(sel2(I.F, fsel), I.ntot), us(I.us, I.ntot), vs(I.vs, I.ntot), ws(I.ws, I.ntot), rhs(I.rhs, I.ntot);
    R usw[KZ], vsw[KZ], vpw[KZ], fw[KZ], wsw[KZ + 1];  // planes k0+e
#pragma unroll
    for (int e = 0; e < KZ; ++e) {
        const int o = min(k0 + e, nz) * sy;
        usw[e] = us.ld(tc, o);
        vsw[e] = vs.ld(tc, o);
        vpw[e] = vs.ld(tjp, o);
        fw[e] = F.ld(tc, o);
    }
#pragma unroll
    for (int e = 0; e < KZ + 1; ++e) wsw[e] = ws.ld(tc, min(k0 + e, nz + 1) * sy);
#pragma unroll
    for (int e = 0; e < KZ; ++e) {
        const int k = k0 + e;
        if (k > k1) break;
        const R usp = sh_dn(usw[e]);
        if (out)
            rhs.st(rho_of(I, fw[e]) / I.dt *
                       ((usp - usw[e]) * I.dxi + (vpw[e] - vsw[e]) * I.dyi + (wsw[e + 1] - wsw[e]) * I.dzi),
                   tc, k * sy);
    }
}

// ---------------------------------------------------------------------------
// Jacobi point (solve_p_jacobi_kernel, :424-438): the numerator in the
// reference's order, divided by ap = -(ae+aw+an+as+af+ab).  The divisor of a
// lane takes only four values, one per kind of plane (interior, k = 1,
// k = nz, nz = 1), so its reciprocal is formed once and every point divides
// by Markstein's correction (div_rn: the IEEE quotient's bits, no per-point
// reciprocal).  Boundary coefficients are zero times the (zero) ghost p,
// kept in the expression so the rounding sequence is the reference's.
template <class R>
__device__ __forceinline__ R jac_num(R b, R ae, R aw, R an, R as, R af, R ab, R pe, R pw, R pn, R ps, R pf,
                                     R pb) {
    return b - ae * pe - aw * pw - an * pn - as * ps - af * pf - ab * pb;
}

template <class R>
struct JacAp {
    R in, lo, hi, one;  // ap per kind of plane, summed in the order of :430
    R rin;              // RN(1/in)
    bool kin;           // rin usable (div_b_ok)
    __device__ __forceinline__ JacAp(R ae, R aw, R an, R as, R azz, int nz) {
        in = -(ae + aw + an + as + azz + azz);
        lo = -(ae + aw + an + as + ((nz != 1) ? azz : R(0)) + R(0));
        hi = -(ae + aw + an + as + R(0) + ((nz != 1) ? azz : R(0)));
        one = -(ae + aw + an + as + R(0) + R(0));
        rin = R(1) / in;
        kin = div_b_ok(in);
    }
    // num / ap on an interior plane (1 < s < nz): reused reciprocal
    __device__ __forceinline__ R div_in(R num) const { return div_rn(num, in, rin, kin); }
    // the same with the wave-uniform fallback test (div_rn_wave) for the
    // lanes of the wave mask need (formed by need_mask)
    struct Need {
        uint64_t must, need;
    };
    __device__ __forceinline__ Need need_mask(bool need) const {
        return Need{__builtin_amdgcn_ballot_w64(need && !kin), __builtin_amdgcn_ballot_w64(need)};
    }
    __device__ __forceinline__ R div_in_wave(R num, const Need& m) const {
        return div_rn_wave(num, in, rin, m.must, m.need);
    }
    // num / ap of plane s (s wave-uniform); the first and last planes (two
    // of every column's nz) divide the IEEE way.  Bit-identical either way.
    __device__ __forceinline__ R div(R num, int s, int nz) const {
        if (s != 1 && s != nz) return div_in(num);
        const R a_one = one, a_lo = lo, a_hi = hi;  // selected as values, not member addresses
        return num / (nz == 1 ? a_one : (s == 1 ? a_lo : a_hi));
    }
};

// ---------------------------------------------------------------------------
// One Jacobi sweep (solve_p_jacobi_kernel, :401-439) with the precomputed rhs.
// Boundary neighbours carry a zero coefficient times the (zero) ghost p, kept
// in the expression so the rounding sequence is the reference's.
template <class R>
__global__ __launch_bounds__(256) void k_jacobi(const EnsArg<R> ea,
                                                int psrc, int kz) {
    VOF_PROLOGUE(R)
    WMap m;
    if (!wmap<R, 1>(I, lb, kz, 1, I.nz, m)) return;
    const int i = m.i, j = m.j, sx = I.sx, sy = I.sy;
    const int P = sx * (I.ny + 2);
    const bool out = m.own && i >= 1 && i <= I.nx && j >= 1 && j <= I.ny;
    const int tc = m.tc, tjm = max(tc - sx, 0), tjp = min(tc + sx, P - 1);
    const Buf<R> p(sel2(I.p, psrc), I.ntot), pt(sel2(I.p, psrc ^ 1), I.ntot), rhs(I.rhs, I.ntot);
    const R ae = (i != I.nx) ? I.axx : R(0);
    const R aw = (i != 1) ? I.axx : R(0);
    const R an = (j != I.ny) ? I.ayy : R(0);
    const R as = (j != 1) ? I.ayy : R(0);
    const R azz = I.azz;
    const JacAp<R> apk(ae, aw, an, as, azz, I.nz);
    int o = m.k0 * sy;
    const int olim = I.nz * sy;  // last plane offset whose +1 neighbour is stored
    R p_km = p.ld(tc, o - sy), p_k = p.ld(tc, o);
    // operands of the current plane, fetched one plane ahead
    R n_kp = p.ld(tc, o + sy), n_jm = p.ld(tjm, o), n_jp = p.ld(tjp, o), n_b = rhs.ld(tc, o);
    for (int k = m.k0; k <= m.k1; ++k, o += sy) {
        const R p_kp = n_kp, pjm = n_jm, pjp = n_jp, b = n_b;
        const int o2 = min(o + sy, olim);
        n_kp = p.ld(tc, o2 + sy);
        n_jm = p.ld(tjm, o2);
        n_jp = p.ld(tjp, o2);
        n_b = rhs.ld(tc, o2);
        const R pim = sh_up(p_k), pip = sh_dn(p_k);
        if (out) {
            const R af = (k != I.nz) ? azz : R(0);
            const R ab = (k != 1) ? azz : R(0);
            pt.st(apk.div(jac_num(b, ae, aw, an, as, af, ab, pip, pim, pjp, pjm, p_kp, p_km), k, I.nz), tc, o);
        }
        p_km = p_k;
        p_k = p_kp;
    }
}

// The same sweep for z-chunks of KZ planes, with every operand of the chunk
// (p on KZ+2 planes of the column, p on the two neighbouring rows and rhs on
// KZ planes) loaded before the first point is formed: a wave has all its
// loads in flight at once instead of one plane's worth per memory latency.
template <class R, int KZ>
__global__ __launch_bounds__(256) void k_jacobi_blk(const EnsArg<R> ea, int psrc) {
    VOF_PROLOGUE(R)
    WMap m;
    if (!wmap<R, 1>(I, lb, KZ, 1, I.nz, m)) return;
    const int i = m.i, j = m.j, sx = I.sx, sy = I.sy;
    const int P = sx * (I.ny + 2);
    const bool out = m.own && i >= 1 && i <= I.nx && j >= 1 && j <= I.ny;
    const int tc = m.tc, tjm = max(tc - sx, 0), tjp = min(tc + sx, P - 1);
    const Buf<R> p(sel2(I.p, psrc), I.ntot), pt(sel2(I.p, psrc ^ 1), I.ntot), rhs(I.rhs, I.ntot);
    const R ae = (i != I.nx) ? I.axx : R(0);
    const R aw = (i != 1) ? I.axx : R(0);
    const R an = (j != I.ny) ? I.ayy : R(0);
    const R as = (j != 1) ? I.ayy : R(0);
    const R azz = I.azz;
    const int k0 = m.k0, k1 = m.k1, nz = I.nz;
    const bool interior = (k0 >= 2 && k0 + KZ - 1 <= nz - 1);  // full chunk of planes 2 .. nz-1
    const JacAp<R> apk(ae, aw, an, as, azz, nz);
    R pc[KZ + 2], pjm[KZ], pjp[KZ], b[KZ];  // pc[e]: plane k0-1+e; the others: plane k0+e
#pragma unroll
    for (int e = 0; e < KZ + 2; ++e) pc[e] = p.ld(tc, min(k0 - 1 + e, nz + 1) * sy);
#pragma unroll
    for (int e = 0; e < KZ; ++e) {
        const int o = min(k0 + e, nz) * sy;
        pjm[e] = p.ld(tjm, o);
        pjp[e] = p.ld(tjp, o);
        b[e] = rhs.ld(tc, o);
    }
    auto march = [&](auto in_tag) __attribute__((always_inline)) {
        constexpr bool IN = decltype(in_tag)::value;
#pragma unroll
        for (int e = 0; e < KZ; ++e) {
            const int k = k0 + e;
            if (!IN && k > k1) break;
            const R p_k = pc[e + 1];
            const R pim = sh_up(p_k), pip = sh_dn(p_k);
            if (out) {
                const R af = (IN || k != nz) ? azz : R(0);
                const R ab = (IN || k != 1) ? azz : R(0);
                const R num = jac_num(b[e], ae, aw, an, as, af, ab, pip, pim, pjp[e], pjm[e], pc[e + 2], pc[e]);
                pt.st(IN ? apk.div_in(num) : apk.div(num, k, nz), tc, k * sy);
            }
        }
    };
    if (interior) march(std::true_type{});
    else march(std::false_type{});
}

// ---------------------------------------------------------------------------
// Two Jacobi sweeps per launch (temporal blocking; the reference runs one
// sweep per launch, :1371-1375).  A work-group of 16 waves covers a tile of
// 16 rows x 64 lanes: wave w holds row j = 14*band + w, lane l holds
// i = 60*xtile - 1 + l.  Sweep 1 is formed on every lane of every row (zero
// outside the interior, as the never-written ghost p of the reference), and
// shared across rows through a 3-plane LDS ring; sweep 2 is stored on lanes
// 2..61 of rows 1..14, which have all their sweep-1 neighbours.  The tile
// marches along z with sweep 2 one plane behind sweep 1, the z neighbours of
// both sweeps held in registers.  Each sweep keeps the reference's operation
// order, so two launches of this kernel equal four single sweeps bit for bit.
constexpr int kJ2Rows = 16, kJ2OutRows = 14, kJ2OutLanes = 60;

// Streamed form for any chunk length: the operands of plane s + D (the p
// column one plane further, p on the two neighbouring rows and rhs) are
// requested while plane s is formed, so D planes of loads are in flight
// behind the march; the prefetch ring is indexed by the plane modulo D with
// the march unrolled D-fold, so every slot is a fixed register.
template <class R, int D>
__global__ __launch_bounds__(1024) void k_jacobi2(const EnsArg<R> ea,
                                                  int psrc, int kz) {
    VOF_PROLOGUE(R)
    __shared__ R ring[3][kJ2Rows][64];
    const int nx = I.nx, ny = I.ny, nz = I.nz, sx = I.sx, sy = I.sy;
    // tiles over the flattened plane of (i, row band) positions, as in
    // k_jacobi2_dma: lane l of tile xt holds position 60 xt - 2 + l
    const int nbd = (ny + kJ2OutRows - 1) / kJ2OutRows;
    const int npos = sx * nbd;
    const int per_z = (npos + kJ2OutLanes - 1) / kJ2OutLanes;
    const int zc = lb / per_z, xt = lb - zc * per_z;
    const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
    const int t = min(max(kJ2OutLanes * xt - 2 + lane, 0), npos - 1);
    const int bd = t / sx;
    const int i = t - bd * sx;
    const int j = kJ2OutRows * bd + w;
    const int k0 = 1 + zc * kz, k1 = min(nz, k0 + kz - 1);
    const bool in1 = (i >= 1 && i <= nx && j >= 1 && j <= ny);
    const bool st2 = in1 && lane >= 2 && lane <= 61 && w >= 1 && w <= kJ2OutRows;
    const int ic = min(max(i, 0), nx + 1), jc = min(max(j, 0), ny + 1);
    const int tc = jc * sx + ic;
    const int tjm = max(jc - 1, 0) * sx + ic, tjp = min(jc + 1, ny + 1) * sx + ic;
    const Buf<R> p(sel2(I.p, psrc), I.ntot), pt(sel2(I.p, psrc ^ 1), I.ntot), rhs(I.rhs, I.ntot);
    const R ae = (i != nx) ? I.axx : R(0), aw = (i != 1) ? I.axx : R(0);
    const R an = (j != ny) ? I.ayy : R(0), as = (j != 1) ? I.ayy : R(0);
    const R azz = I.azz;
    const JacAp<R> apk(ae, aw, an, as, azz, nz);
    const int olast = (nz + 1) * sy;
    auto po = [&](int s) { return min(max(s, 0) * sy, olast); };  // clamped plane offset
    // march state: sweep-0 column at s-1, s; sweep-1 column at s-2, s-1; rhs at s-1
    const int s_beg = k0 - 1, s_end = k1 + 1;
    R q0m = p.ld(tc, po(s_beg - 1)), q0 = p.ld(tc, po(s_beg));
    R q1mm = R(0), q1m = R(0), bm = R(0);
    R fq[D], fjm[D], fjp[D], fb[D];  // slot d: plane s_beg + d (column: one plane further)
#pragma unroll
    for (int d = 0; d < D; ++d) {
        fq[d] = p.ld(tc, po(s_beg + d + 1));
        fjm[d] = p.ld(tjm, po(s_beg + d));
        fjp[d] = p.ld(tjp, po(s_beg + d));
        fb[d] = rhs.ld(tc, po(s_beg + d));
    }
    for (int s0 = s_beg; s0 <= s_end; s0 += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int s = s0 + d;
            if (s > s_end) break;  // uniform over the work-group
            const R q0p = fq[d], q0jm = fjm[d], q0jp = fjp[d], b = fb[d];
            fq[d] = p.ld(tc, po(s + D + 1));
            fjm[d] = p.ld(tjm, po(s + D));
            fjp[d] = p.ld(tjp, po(s + D));
            fb[d] = rhs.ld(tc, po(s + D));
            // sweep 1 at plane s
            const R q0w = sh_up(q0), q0e = sh_dn(q0);
            R q1 = R(0);
            if (in1 && s >= 1 && s <= nz) {
                const R af = (s != nz) ? azz : R(0), ab = (s != 1) ? azz : R(0);
                q1 = apk.div(jac_num(b, ae, aw, an, as, af, ab, q0e, q0w, q0jp, q0jm, q0p, q0m), s, nz);
            }
            const int slot = s % 3, pslot = (s + 2) % 3;
            ring[slot][w][lane] = q1;
            __syncthreads();
            // sweep 2 at plane s-1
            const int s2 = s - 1;
            const R q1w = sh_up(q1m), q1e = sh_dn(q1m);
            if (st2 && s2 >= k0) {
                const R q1s = ring[pslot][w - 1][lane], q1n = ring[pslot][w + 1][lane];
                const R af = (s2 != nz) ? azz : R(0), ab = (s2 != 1) ? azz : R(0);
                pt.st(apk.div(jac_num(bm, ae, aw, an, as, af, ab, q1e, q1w, q1n, q1s, q1, q1mm), s2, nz), tc,
                      s2 * sy);
            }
            q0m = q0; q0 = q0p;
            q1mm = q1m; q1m = q1;
            bm = b;
        }
    }
}

// ---------------------------------------------------------------------------
// T Jacobi sweeps per launch, T = 3..5 (temporal blocking deeper than the
// two-sweep kernels).  The tile is the two-sweep one with the halo of T
// sweeps: 16 waves x 64 lanes, wave w on row j = R*band + w - T + 2, lane l on
// i = L*xtile + l - T + 1, storing lanes T .. 63-T (L = 64 - 2T) of rows
// T-1 .. 16-T (R = 18 - 2T).  While the tile marches along z (input plane s),
// sweep t forms plane s - t + 1: its x neighbours by lane shifts, its own z
// neighbours from the registers of sweep t-1 (the planes that sweep formed in
// this step and the two before), its y neighbours from an LDS ring of sweep
// t-1 (two plane slots per sweep: a slot written in one step is read in the
// next, with one barrier per step between).  Sweep t is formed on rows
// t-1 .. 16-t, those whose y neighbours sweep t-1 holds; points outside the
// interior (ghost rows, lanes and planes) are zero, as the never-written
// ghost p of the reference.  Each sweep keeps the reference's operation
// order, so one launch equals T single sweeps bit for bit.
template <int T>
struct JacTile {
    static constexpr int lanes = 64 - 2 * T, rows = 18 - 2 * T;
};

template <class R, int T>
__global__ __launch_bounds__(1024) void k_jacobiT(const EnsArg<R> ea, int psrc, int kz) {
    static_assert(T >= 3 && T <= kJacobiMaxT, "sweeps per launch");
    VOF_PROLOGUE(R)
    constexpr int OL = JacTile<T>::lanes, OR = JacTile<T>::rows;
    __shared__ R ring[T - 1][2][kJ2Rows][64];
    const int nx = I.nx, ny = I.ny, nz = I.nz, sx = I.sx, sy = I.sy;
    const int nxt = (nx + OL - 1) / OL;
    const int nbd = (ny + OR - 1) / OR;
    const int per_z = nxt * nbd;
    const int zc = lb / per_z, rem = lb - zc * per_z;
    const int bd = rem / nxt, xt = rem - bd * nxt;
    const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
    const int i = OL * xt + lane - T + 1;
    const int j = OR * bd + w - T + 2;
    const int k0 = 1 + zc * kz, k1 = min(nz, k0 + kz - 1);
    const bool in1 = (i >= 1 && i <= nx && j >= 1 && j <= ny);
    const bool stT = in1 && lane >= T && lane <= 63 - T && w >= T - 1 && w <= 16 - T;
    const int ic = min(max(i, 0), nx + 1), jc = min(max(j, 0), ny + 1);
    const int tc = jc * sx + ic;
    const int tjm = max(jc - 1, 0) * sx + ic, tjp = min(jc + 1, ny + 1) * sx + ic;
    const Buf<R> p(sel2(I.p, psrc), I.ntot), pt(sel2(I.p, psrc ^ 1), I.ntot), rhs(I.rhs, I.ntot);
    const R ae = (i != nx) ? I.axx : R(0), aw = (i != 1) ? I.axx : R(0);
    const R an = (j != ny) ? I.ayy : R(0), as = (j != 1) ? I.ayy : R(0);
    const R azz = I.azz;
    const JacAp<R> apk(ae, aw, an, as, azz, nz);
    const int olast = (nz + 1) * sy;
    auto po = [&](int s) { return min(max(s, 0) * sy, olast); };  // clamped plane offset
    const int s_beg = k0 - T + 1, s_end = k1 + T - 1;
    // input p: the column at s-1, s; plane s+1 of the column, rows j-1, j+1
    // and rhs of plane s fetched one step ahead
    R q0m = p.ld(tc, po(s_beg - 1)), q0 = p.ld(tc, po(s_beg));
    R fq = p.ld(tc, po(s_beg + 1)), fjm = p.ld(tjm, po(s_beg)), fjp = p.ld(tjp, po(s_beg));
    R fb = rhs.ld(tc, po(s_beg));
    // sweep t (1 <= t < T): lc[t-1] its value at the plane it formed in the
    // last step, lm[t-1] at the one before; bh[t-1] rhs of the plane sweep t
    // forms in this step
    R lc[T - 1], lm[T - 1], bh[T];
#pragma unroll
    for (int t = 0; t < T - 1; ++t) lc[t] = lm[t] = R(0);
#pragma unroll
    for (int t = 0; t < T; ++t) bh[t] = R(0);
    for (int s = s_beg; s <= s_end; ++s) {
        const R q0p = fq, q0jm = fjm, q0jp = fjp;
#pragma unroll
        for (int t = T - 1; t > 0; --t) bh[t] = bh[t - 1];
        bh[0] = fb;
        fq = p.ld(tc, po(s + 2));
        fjm = p.ld(tjm, po(s + 1));
        fjp = p.ld(tjp, po(s + 1));
        fb = rhs.ld(tc, po(s + 1));
        R cur[T];
        {  // sweep 1 at plane s
            const R qw = sh_up(q0), qe = sh_dn(q0);
            R q = R(0);
            if (in1 && s >= 1 && s <= nz) {
                const R af = (s != nz) ? azz : R(0), ab = (s != 1) ? azz : R(0);
                q = apk.div(jac_num(bh[0], ae, aw, an, as, af, ab, qe, qw, q0jp, q0jm, q0p, q0m), s, nz);
            }
            cur[0] = q;
            ring[0][s & 1][w][lane] = q;
        }
#pragma unroll
        for (int t = 2; t <= T; ++t) {  // sweep t at plane z = s - t + 1
            const int z = s - t + 1;
            const R c = lc[t - 2];  // sweep t-1 at plane z
            const R cw = sh_up(c), ce = sh_dn(c);
            R q = R(0);
            if (w >= t - 1 && w <= 16 - t && in1 && z >= 1 && z <= nz) {
                const R cs = ring[t - 2][z & 1][w - 1][lane], cn = ring[t - 2][z & 1][w + 1][lane];
                const R af = (z != nz) ? azz : R(0), ab = (z != 1) ? azz : R(0);
                q = apk.div(jac_num(bh[t - 1], ae, aw, an, as, af, ab, ce, cw, cn, cs, cur[t - 2], lm[t - 2]), z, nz);
            }
            cur[t - 1] = q;
            if (t < T) ring[t - 1][z & 1][w][lane] = q;
            else if (stT && z >= k0 && z <= k1) pt.st(q, tc, z * sy);
        }
        __syncthreads();
        q0m = q0;
        q0 = q0p;
#pragma unroll
        for (int t = 0; t < T - 1; ++t) {
            lm[t] = lc[t];
            lc[t] = cur[t];
        }
    }
}

// The same two sweeps for z-chunks of KZ planes, with every operand of the
// chunk loaded before the march (p on the KZ+4 planes of the lane's column,
// p on the two neighbouring rows and rhs on KZ+2 planes): the march then waits
// on memory once per chunk instead of once per plane between its barriers.
// The chunk bounds are uniform over the work-group, so every wave reaches
// every barrier.
template <class R, int KZ>
__global__ __launch_bounds__(1024) void k_jacobi2_blk(const EnsArg<R> ea, int psrc) {
    VOF_PROLOGUE(R)
    // Sweep-1 planes shared across rows: a ring of five slots (plane mod 5)
    // in an interior chunk, of three (plane mod 3) in the others.  Where it
    // fits beside the ring within half a CU's LDS (two work-groups per CU),
    // the p rows of an interior chunk's planes, with the row beyond each tile
    // edge, are shared too (prow).
    constexpr bool kRowsLds = sizeof(R) * 64 * ((KZ + 2) * (kJ2Rows + 2) + 5 * kJ2Rows) <= 80 * 1024;
    __shared__ R ring[5][kJ2Rows][64];
    __shared__ R prow[kRowsLds ? KZ + 2 : 1][kJ2Rows + 2][64];
    const int nx = I.nx, ny = I.ny, nz = I.nz, sx = I.sx, sy = I.sy;
    // tiles over the flattened plane of (i, row band) positions, as in
    // k_jacobi2_dma: lane l of tile xt holds position 60 xt - 2 + l
    const int nbd = (ny + kJ2OutRows - 1) / kJ2OutRows;
    const int npos = sx * nbd;
    const int per_z = (npos + kJ2OutLanes - 1) / kJ2OutLanes;
    const int zc = lb / per_z, xt = lb - zc * per_z;
    const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
    const int t = min(max(kJ2OutLanes * xt - 2 + lane, 0), npos - 1);
    const int bd = t / sx;
    const int i = t - bd * sx;
    const int j = kJ2OutRows * bd + w;
    const int k0 = 1 + zc * KZ, k1 = min(nz, k0 + KZ - 1);
    const bool in1 = (i >= 1 && i <= nx && j >= 1 && j <= ny);
    const bool st2 = in1 && lane >= 2 && lane <= 61 && w >= 1 && w <= kJ2OutRows;
    const int ic = min(max(i, 0), nx + 1), jc = min(max(j, 0), ny + 1);
    const int tc = jc * sx + ic;
    const int tjm = max(jc - 1, 0) * sx + ic, tjp = min(jc + 1, ny + 1) * sx + ic;
    const Buf<R> p(sel2(I.p, psrc), I.ntot), pt(sel2(I.p, psrc ^ 1), I.ntot), rhs(I.rhs, I.ntot);
    const R ae = (i != nx) ? I.axx : R(0), aw = (i != 1) ? I.axx : R(0);
    const R an = (j != ny) ? I.ayy : R(0), as = (j != 1) ? I.ayy : R(0);
    const R azz = I.azz;
    // interior chunk: every plane the march forms, k0-1 .. k1+1, lies in 2 .. nz-1
    const bool interior = (k0 >= 3 && k0 + KZ - 1 <= nz - 2);
    const JacAp<R> apk(ae, aw, an, as, azz, nz);
    const int olast = (nz + 1) * sy;
    auto po = [&](int s) { return min(max(s, 0) * sy, olast); };  // clamped plane offset
    R pc[KZ + 4];                          // column, planes k0-2+e
    R pjm[KZ + 2], pjp[KZ + 2], bb[KZ + 2];  // planes k0-1+e
    auto load_rows = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int e = 0; e < KZ + 2; ++e) {
            const int o = po(k0 - 1 + e);
            pjm[e] = p.ld(tjm, o);
            pjp[e] = p.ld(tjp, o);
        }
    };
#pragma unroll
    for (int e = 0; e < KZ + 4; ++e) pc[e] = p.ld(tc, po(k0 - 2 + e));
#pragma unroll
    for (int e = 0; e < KZ + 2; ++e) bb[e] = rhs.ld(tc, po(k0 - 1 + e));
    if (interior) {
        // Every plane the chunk forms is interior: coefficients and divisor
        // fixed, sweep 1 formed on every lane and masked by a select, and one
        // barrier per two planes.  Sweep 1 of planes 2t, 2t+1 (relative index
        // e = plane - (k0-1)) is written to ring slot e mod 5, then after the
        // barrier sweep 2 of planes 2t-1 and 2t reads the rows around it; its
        // z neighbours are in registers (q1v).  Five slots keep the pair a
        // wave may write before the next barrier (2t+2, 2t+3) apart from the
        // pair other waves may still read (2t-1, 2t).
        if constexpr (kRowsLds) {
            // Each wave stores its own column of p; the first and last waves
            // also load the row beyond the tile.  A lane then issues 22 loads
            // (32 on those two waves) instead of 42.
            const bool edge = (w == 0 || w == kJ2Rows - 1);  // uniform per wave
#pragma unroll
            for (int e = 0; e < KZ + 2; ++e) prow[e][w + 1][lane] = pc[e + 1];
            if (edge) {
                const int th = (w == 0) ? tjm : tjp, hrow = (w == 0) ? 0 : kJ2Rows + 1;
#pragma unroll
                for (int e = 0; e < KZ + 2; ++e) prow[e][hrow][lane] = p.ld(th, po(k0 - 1 + e));
            }
            __syncthreads();
        } else {
            load_rows();
        }
        R q1v[KZ + 2];
        const auto nd1 = apk.need_mask(in1), nd2 = apk.need_mask(st2);
#pragma unroll
        for (int t = 0; t < (KZ + 2) / 2; ++t) {
#pragma unroll
            for (int e = 2 * t; e < 2 * t + 2; ++e) {
                const R q0 = pc[e + 1];
                const R q0w = sh_up(q0), q0e = sh_dn(q0);
                const R qs = kRowsLds ? prow[e][w][lane] : pjm[e];
                const R qn = kRowsLds ? prow[e][w + 2][lane] : pjp[e];
                const R num = jac_num(bb[e], ae, aw, an, as, azz, azz, q0e, q0w, qn, qs, pc[e + 2], pc[e]);
                const R q = apk.div_in_wave(num, nd1);
                q1v[e] = in1 ? q : R(0);
                ring[e % 5][w][lane] = q1v[e];
            }
            __syncthreads();
#pragma unroll
            for (int e = 2 * t - 1; e <= 2 * t; ++e) {
                if (e < 1 || e > KZ) continue;  // sweep-2 planes k0 .. k1
                const R q1c = q1v[e];
                const R q1w = sh_up(q1c), q1e = sh_dn(q1c);
                const R q1s = ring[e % 5][w == 0 ? 0 : w - 1][lane];
                const R q1n = ring[e % 5][w == kJ2Rows - 1 ? w : w + 1][lane];
                const R num = jac_num(bb[e], ae, aw, an, as, azz, azz, q1e, q1w, q1n, q1s, q1v[e + 1], q1v[e - 1]);
                const R q = apk.div_in_wave(num, nd2);
                if (st2) pt.st(q, tc, (k0 - 1 + e) * sy);
            }
        }
        return;
    }
    // Chunks touching the first or last planes: sweep 1 at plane s, sweep 2
    // at plane s-1, for s = k0-1 .. k1+1, coefficients and divisor selected
    // per plane.
    load_rows();
    auto march = [&](auto in_tag) __attribute__((always_inline)) {
        constexpr bool IN = decltype(in_tag)::value;
        R q1mm = R(0), q1m = R(0);
#pragma unroll
        for (int e = 0; e < KZ + 2; ++e) {
            const int s = k0 - 1 + e;
            if (!IN && s > k1 + 1) break;  // uniform over the work-group
            const R q0 = pc[e + 1];
            const R q0w = sh_up(q0), q0e = sh_dn(q0);
            R q1 = R(0);
            if (in1 && (IN || (s >= 1 && s <= nz))) {
                const R af = (IN || s != nz) ? azz : R(0), ab = (IN || s != 1) ? azz : R(0);
                const R num = jac_num(bb[e], ae, aw, an, as, af, ab, q0e, q0w, pjp[e], pjm[e], pc[e + 2], pc[e]);
                q1 = IN ? apk.div_in(num) : apk.div(num, s, nz);
            }
            const int slot = s % 3, pslot = (s + 2) % 3;
            ring[slot][w][lane] = q1;
            __syncthreads();
            const int s2 = s - 1;
            const R q1w = sh_up(q1m), q1e = sh_dn(q1m);
            if (st2 && (IN ? e >= 2 : (e >= 1 && s2 >= k0))) {  // plane k0-1 is the previous chunk's
                const R q1s = ring[pslot][w - 1][lane], q1n = ring[pslot][w + 1][lane];
                const R af = (IN || s2 != nz) ? azz : R(0), ab = (IN || s2 != 1) ? azz : R(0);
                const R num = jac_num(bb[e - 1], ae, aw, an, as, af, ab, q1e, q1w, q1n, q1s, q1, q1mm);
                pt.st(IN ? apk.div_in(num) : apk.div(num, s2, nz), tc, s2 * sy);
            }
            q1mm = q1m;
            q1m = q1;
        }
    };
    march(std::false_type{});
}

// Streamed two sweeps (FP32) with the p planes and the rhs brought into LDS by
// the buffer unit (buffer_load ... lds): a request for a plane ahead holds no
// register, so kJ2DmaAhead groups of planes stay in flight behind the march
// at no cost in occupancy, and each wave requests only its own row (the
// first and last waves also the row beyond the tile) instead of its row and
// both neighbours.  The tile marches along z through a chunk of KZ planes;
// sweep 1 is formed at plane s and sweep 2 at plane s-1 between the same two
// barriers.  Every chunk has KZ planes (the last one starts at nz-KZ+1 and
// recomputes planes of the one before it with the same inputs, so both write
// the same values), which makes the whole march, its ring slots, requests
// and waits, known at compile time; only chunks that reach plane 1 or nz
// test the plane per step.
//
// Ordering.  A request completes in LDS when the wave's vector-memory count
// says so (loads retire in issue order; the count is taken as the number of
// loads issued after the awaited ones, so stores still in flight only make
// the wait longer), and the barrier then publishes every wave's rows.  The
// barrier is the bare s_barrier with LDS writes waited for before it: a
// work-group fence would also wait for the requests still in flight.  Step e
// requests p plane e+A+3 (A = kJ2DmaAhead) into the slot of plane
// e+A+3-kJ2DmaSlots <= e.  Plane r is read at step r-2 on the wave's own row,
// which only that wave writes, and at step r-1 on the rows of the
// neighbouring waves, which have all passed the barrier of step r-1 before
// any wave reaches step r.  Rhs rows are read only by the wave that requests
// them, at step r, and replaced at step r+kJ2DmaSlots-A-1 > r.
constexpr int kJ2DmaAhead = 3;  // groups requested beyond the one awaited
constexpr int kJ2DmaSlots = 6;  // p and rhs ring slots
static_assert(kJ2DmaSlots >= kJ2DmaAhead + 3, "p slot reused before its last read");

typedef __attribute__((address_space(3))) void lds_void;

// Lane l's element (per-lane offset vo, plane offset so) into row[l] in LDS.
// Written as the instruction itself: the compiler's own wait for a request
// into LDS is for every request in flight, which would drain the ring.  The
// asm has no register result; it waits 5 states for an SGPR operand written
// by a vector instruction and 1 after writing M0.
__device__ __forceinline__ void row_to_lds(const Buf<float>& b, float* row, int vo, int so) {
    const unsigned lds = (unsigned)(uintptr_t)(lds_void*)row;
    asm volatile(
        "s_nop 4\n\t"
        "s_mov_b32 m0, %0\n\t"
        "s_nop 0\n\t"
        "buffer_load_dword %1, %2, %3 offen lds"
        :
        : "s"(lds), "v"(byte_off(vo, 4u)), "s"(b.r), "s"(byte_off(so, 4u))
        : "memory", "m0");
}

// s_waitcnt vmcnt(N), other counters not waited on (gfx9 encoding).
template <int N>
__device__ __forceinline__ void wait_vm() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | (0x7 << 4) | (0xF << 8) | ((N >> 4) << 14));
}

// This wave's requests older than its last m groups (three each) complete.
__device__ __forceinline__ void wait_groups(int m) {
    switch (m) {  // a constant in the unrolled march
        case 0: wait_vm<0>(); break;
        case 1: wait_vm<3>(); break;
        case 2: wait_vm<6>(); break;
        default: wait_vm<9>(); break;
    }
}
static_assert(kJ2DmaAhead <= 3, "wait_groups covers three groups ahead");

// LDS writes of this wave done, then the work-group barrier; no code moves
// across it.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_s_waitcnt(0xF | (0x7 << 4) | (0x0 << 8) | (0x3 << 14));  // lgkmcnt(0)
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

template <int KZ>
__global__ __launch_bounds__(1024) void k_jacobi2_dma(const EnsArg<float> ea, int psrc) {
    using R = float;
    VOF_PROLOGUE(R)
    constexpr int NS = kJ2DmaSlots, A = kJ2DmaAhead;
    static_assert(KZ >= A, "prologue requests fit in the chunk");
    __shared__ R P[NS][kJ2Rows + 2][64];  // p, rows j0-1 .. j0+kJ2Rows of the tile
    __shared__ R B[NS][kJ2Rows][64];      // rhs
    __shared__ R Q[3][kJ2Rows][64];       // sweep 1
    const int nx = I.nx, ny = I.ny, nz = I.nz, sx = I.sx, sy = I.sy;
    // Lane l of tile xt holds position t = 60 xt - 2 + l of the flattened
    // plane of (i, band) positions, t = band * sx + i, and wave w row
    // j = 14 band + w of that band, so a tile may run from the end of one
    // band's row into the next band's (lanes 2..61 store; the ghost columns
    // i = 0, nx+1 between them are zero in sweep 1, as in the reference, and
    // the x neighbours of every interior point lie in the same band row).
    const int nbd = (ny + kJ2OutRows - 1) / kJ2OutRows;
    const int npos = sx * nbd;
    const int per_z = (npos + kJ2OutLanes - 1) / kJ2OutLanes;
    const int zc = lb / per_z, xt = lb - zc * per_z;
    const int lane = (int)threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int t = min(max(kJ2OutLanes * xt - 2 + lane, 0), npos - 1);
    const int bd = t / sx;
    const int i = t - bd * sx;
    const int j = kJ2OutRows * bd + w;
    const int k0 = min(1 + zc * KZ, nz - KZ + 1);  // planes k0 .. k0+KZ-1 (the launch has nz >= KZ)
    const bool in1 = (i >= 1 && i <= nx && j >= 1 && j <= ny);
    const bool st2 = in1 && lane >= 2 && lane <= 61 && w >= 1 && w <= kJ2OutRows;
    const int ic = min(max(i, 0), nx + 1), jc = min(max(j, 0), ny + 1);
    const int tc = jc * sx + ic;
    // third request of a group: the row beyond the tile on the first and last
    // waves, the wave's own row again (same values) on the others, so that
    // every wave counts three requests per group
    const int th = (w == 0) ? max(jc - 1, 0) * sx + ic : (w == kJ2Rows - 1) ? min(jc + 1, ny + 1) * sx + ic : tc;
    const int hrow = (w == 0) ? 0 : (w == kJ2Rows - 1) ? kJ2Rows + 1 : w + 1;
    const Buf<R> p(sel2(I.p, psrc), I.ntot), pt(sel2(I.p, psrc ^ 1), I.ntot), rhs(I.rhs, I.ntot);
    const R ae = (i != nx) ? I.axx : R(0), aw = (i != 1) ? I.axx : R(0);
    const R an = (j != ny) ? I.ayy : R(0), as = (j != 1) ? I.ayy : R(0);
    const R azz = I.azz;
    const JacAp<R> apk(ae, aw, an, as, azz, nz);
    const auto nd1 = apk.need_mask(in1), nd2 = apk.need_mask(st2);
    // interior chunk: every plane the march forms, k0-1 .. k0+KZ, lies in 2 .. nz-1
    const bool interior = (k0 >= 3 && k0 + KZ - 1 <= nz - 2);
    const int olast = (nz + 1) * sy;
    auto po = [&](int s) { return min(max(s, 0) * sy, olast); };  // clamped plane offset
    // p plane r of the chunk is plane k0-2+r; rhs plane r is plane k0-1+r.
    // Group x: p plane x+2 and rhs plane x; step e needs groups 0..e and p
    // planes 0, 1.
    auto req_p = [&](int slot, int o) __attribute__((always_inline)) {
        row_to_lds(p, &P[slot][w + 1][0], tc, o);
        row_to_lds(p, &P[slot][hrow][0], th, o);
    };
    auto req_b = [&](int slot, int o) __attribute__((always_inline)) { row_to_lds(rhs, &B[slot][w][0], tc, o); };
    req_p(0, po(k0 - 2));
    req_p(1, po(k0 - 1));
#pragma unroll
    for (int x = 0; x <= A; ++x) {
        req_p((x + 2) % NS, po(k0 + x));
        req_b(x % NS, po(k0 - 1 + x));
    }
    wait_vm<3 * A>();  // groups 1 .. A may stay in flight
    lds_barrier();
    const R q0m0 = P[0][w + 1][lane], q00 = P[1][w + 1][lane];  // p column at k0-2, k0-1
    // Chunk kinds: 0 every plane the march forms is interior; 1 the first
    // chunk (k0 = 1), 2 the last one (ending at nz): their two boundary planes
    // (the ghost plane 0 and plane 1, or plane nz and the ghost plane nz+1)
    // sit at fixed steps, so kinds 0-2 share one march whose code differs
    // only at those steps, by a wave-uniform branch on the kind (one copy of
    // the march: the instruction cache a CU pair shares holds the code of all
    // the work-groups it runs).  3: any other chunk (short grids), with the
    // plane tested per step.
    constexpr int kInterior = 0, kFirst = 1, kLast = 2, kAny = 3;
    const int kind = interior ? kInterior
                     : (k0 == 1 && nz >= KZ + 2)          ? kFirst   // planes 2 .. KZ+1 <= nz-1 interior
                     : (k0 + KZ - 1 == nz && k0 >= 3)     ? kLast    // planes k0-1 >= 2 .. nz-1 interior
                                                          : kAny;
    auto march = [&](auto any_tag) __attribute__((always_inline)) {
        constexpr bool ANY = decltype(any_tag)::value;
        R q0m = q0m0, q0 = q00;                // p column at s-1, s
        R q1mm = R(0), q1m = R(0), bm = R(0);  // sweep 1 at s-2, s-1; rhs at s-1
#pragma unroll
        for (int e = 0; e <= KZ + 1; ++e) {
            const int s = k0 - 1 + e;
            // sweep 1 at plane s
            const R q0p = P[(e + 2) % NS][w + 1][lane];
            const R pjm = P[(e + 1) % NS][w][lane], pjp = P[(e + 1) % NS][w + 2][lane];
            const R b = B[e % NS][w][lane];
            const R q0w = sh_up(q0), q0e = sh_dn(q0);
            R q1;
            if (ANY) {
                q1 = R(0);
                if (in1 && s >= 1 && s <= nz) {
                    const R af = (s != nz) ? azz : R(0), ab = (s != 1) ? azz : R(0);
                    q1 = apk.div(jac_num(b, ae, aw, an, as, af, ab, q0e, q0w, pjp, pjm, q0p, q0m), s, nz);
                }
            } else if ((e == 0 && kind == kFirst) || (e == KZ + 1 && kind == kLast)) {
                q1 = R(0);  // ghost plane 0 / nz+1
            } else if (e == 1 && kind == kFirst) {  // plane 1: no bottom neighbour (:428-429)
                const R q = jac_num(b, ae, aw, an, as, azz, R(0), q0e, q0w, pjp, pjm, q0p, q0m) / apk.lo;
                q1 = in1 ? q : R(0);
            } else if (e == KZ && kind == kLast) {  // plane nz: no top neighbour
                const R q = jac_num(b, ae, aw, an, as, R(0), azz, q0e, q0w, pjp, pjm, q0p, q0m) / apk.hi;
                q1 = in1 ? q : R(0);
            } else {
                const R q = apk.div_in_wave(jac_num(b, ae, aw, an, as, azz, azz, q0e, q0w, pjp, pjm, q0p, q0m), nd1);
                q1 = in1 ? q : R(0);
            }
            Q[e % 3][w][lane] = q1;
            // request group e+A+1 (p plane k0+e+A+1, rhs plane k0+e+A), then
            // wait for group e+1, the next step's.  Plane offsets are formed
            // unclamped where they lie in 0 .. nz+1 (kinds 0 and 1; the last
            // chunk's final p plane would be nz+2).
            if (e + A + 1 <= KZ + 1) {
                const int pp = k0 + e + A + 1, pb = pp - 1;
                const bool clamp = ANY || (e + A + 1 == KZ + 1 && kind == kLast);
                req_p((e + A + 3) % NS, clamp ? po(pp) : pp * sy);
                req_b((e + A + 1) % NS, clamp ? po(pb) : pb * sy);
            }
            if (e <= KZ) wait_groups(min(A, KZ - e));
            lds_barrier();
            // sweep 2 at plane s-1 (planes k0 .. k0+KZ-1: steps 2 .. KZ+1)
            if (e >= 2) {
                const int s2 = s - 1;
                const R q1w = sh_up(q1m), q1e = sh_dn(q1m);
                const R q1s = Q[(e + 2) % 3][w == 0 ? 0 : w - 1][lane];
                const R q1n = Q[(e + 2) % 3][w == kJ2Rows - 1 ? w : w + 1][lane];
                R q;
                if (ANY) {
                    if (s2 >= 2 && s2 <= nz - 1) {
                        q = apk.div_in_wave(jac_num(bm, ae, aw, an, as, azz, azz, q1e, q1w, q1n, q1s, q1, q1mm), nd2);
                    } else {
                        const R af = (s2 != nz) ? azz : R(0), ab = (s2 != 1) ? azz : R(0);
                        q = apk.div(jac_num(bm, ae, aw, an, as, af, ab, q1e, q1w, q1n, q1s, q1, q1mm), s2, nz);
                    }
                } else if (e == 2 && kind == kFirst) {  // plane 1
                    q = jac_num(bm, ae, aw, an, as, azz, R(0), q1e, q1w, q1n, q1s, q1, q1mm) / apk.lo;
                } else if (e == KZ + 1 && kind == kLast) {  // plane nz
                    q = jac_num(bm, ae, aw, an, as, R(0), azz, q1e, q1w, q1n, q1s, q1, q1mm) / apk.hi;
                } else {
                    q = apk.div_in_wave(jac_num(bm, ae, aw, an, as, azz, azz, q1e, q1w, q1n, q1s, q1, q1mm), nd2);
                }
                if (st2) pt.st(q, tc, s2 * sy);
            }
            q0m = q0;
            q0 = q0p;
            q1mm = q1m;
            q1m = q1;
            bm = b;
        }
    };
    if (kind != kAny) march(std::false_type{});
    else march(std::true_type{});
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
// roles of the axes exchanged.  Each lane loads at its row clamped into the
// interior, x-ghost lanes take their neighbour lane's value, and the lanes of
// the first/last interior plane also write ghost planes 0 / nz+1.
template <class R>
__device__ __forceinline__ R corr(const Inst<R>& I, R st, R r0, R r1, R p0, R p1, R hi) {
    const R r = R(0.5) * (r0 + r1);
    return st - I.dt / r * (p0 - p1) * hi;
}

// Per-lane constants of the correction's column mapping.
struct CorrLane {
    int i, j, ci, cj, ts, t;
    bool iin, jin;
};

template <class R>
__device__ __forceinline__ CorrLane corr_lane(const Inst<R>& I, const WMap& m) {
    CorrLane L;
    L.i = m.i;
    L.j = m.j;
    L.cj = min(max(m.j, 1), I.ny);
    L.ci = min(max(m.i, 1), I.nx);
    L.iin = (m.i >= 1 && m.i <= I.nx);
    L.jin = (m.j >= 1 && m.j <= I.ny);
    L.ts = L.cj * I.sx + m.i;  // load position: row clamped into the interior
    L.t = m.t;
    return L;
}

// Plane k (offset o) of the correction, from p, rho, u*, v*, w* of the plane at
// the lane's row, p and rho one row below and one plane below; every lane of
// the wave active (shuffles).
template <class R>
__device__ __forceinline__ void correct_plane(const Inst<R>& I, const WMap& m, const CorrLane& L, int k, int o, R pk,
                                              R pjm, R rk, R rjm, R usk, R vsk, R wsk, R p_km, R r_km,
                                              const Buf<R>& u, const Buf<R>& v, const Buf<R>& w) {
    const int i = L.i, j = L.j, nx = I.nx, ny = I.ny, nz = I.nz, sx = I.sx, sy = I.sy;
    const int ci = L.ci, cj = L.cj, ts = L.ts, t = L.t;
    const bool iin = L.iin, jin = L.jin;
    const R pim = sh_up(pk), rim = sh_up(rk);
    const R uc = corr(I, usk, rk, rim, pk, pim, I.dxi);
    const R vc = corr(I, vsk, rk, rjm, pk, pjm, I.dyi);
    const R wc = corr(I, wsk, rk, r_km, pk, p_km, I.dzi);
    const R v_r = sh_dn(vc), v_l = sh_up(vc), w_r = sh_dn(wc), w_l = sh_up(wc);
    if (m.own) {
        // A tangential ghost is the nearest interior value times I.slip (+1 free
        // slip, -1 no-slip) per ghost axis it lies on, in the passes that copy
        // it: ti on the x-ghost columns, tj on the y-ghost rows, and I.slip
        // again on the z-ghost planes.  The wall faces are zeroed by their own
        // pass, so only the later passes' signs reach them (u's by k only, v's
        // by i and k, w's none).  With +1 every product is the plain copy.
        const R ti = iin ? R(1) : I.slip, tj = jin ? R(1) : I.slip;
        // u: walls at i = 1, nx+1; plane i = 0 keeps/copies its own values
        R uval = R(0);
        bool uw = true;
        if (i >= 2 && i <= nx) uval = tj * uc;
        else if (i == 0) { uw = !jin; if (uw) uval = tj * u.ld(ts, o); }
        if (uw) u.st(uval, t, o);
        // v: walls at j = 1, ny+1; row j = 0 keeps/copies; x ghosts copy the neighbour
        R vval = ti * R(0);
        bool vw = true;
        if (j >= 2 && j <= ny) vval = ti * (iin ? vc : (i == 0 ? v_r : v_l));
        else if (j == 0) { vw = !iin; if (vw) vval = ti * v.ld(ci, o); }
        if (vw) v.st(vval, t, o);
        // w: wall at k = 1
        const R wval = (k == 1) ? R(0) : ti * (tj * (iin ? wc : (i == 0 ? w_r : w_l)));
        w.st(wval, t, o);
        if (k == 1) {  // ghost plane 0: u, v copy plane 1; w copies its own plane-0 interior
            u.st(I.slip * ((i == 0) ? tj * u.ld(ts, sy) : uval), t, 0);
            v.st(I.slip * ((j == 0) ? ti * v.ld(ci, sy) : vval), t, 0);
            if (!(iin && jin)) w.st(ti * (tj * w.ld(cj * sx + ci, 0)), t, 0);
        }
        if (k == nz) {  // ghost plane nz+1: u, v copy plane nz; w wall
            u.st(I.slip * ((i == 0) ? tj * u.ld(ts, o) : uval), t, o + sy);
            v.st(I.slip * ((j == 0) ? ti * v.ld(ci, o) : vval), t, o + sy);
            w.st(R(0), t, o + sy);
        }
    }
}

// Any chunk length: operands of plane k fetched one plane ahead.
template <class R>
__global__ __launch_bounds__(256) void k_correct(const EnsArg<R> ea,
                                                 int fsel, int psel, int kz) {
    VOF_PROLOGUE(R)
    WMap m;
    if (!wmap<R, 1>(I, lb, kz, 1, I.nz, m)) return;
    const CorrLane L = corr_lane(I, m);
    const int nz = I.nz, sx = I.sx, sy = I.sy;
    const int ts = L.ts, tsm = ts - sx;
    const Buf<R> F(sel2(I.F, fsel), I.ntot), p(sel2(I.p, psel), I.ntot);
    const Buf<R> us(I.us, I.ntot), vs(I.vs, I.ntot), ws(I.ws, I.ntot);
    const Buf<R> u(I.u, I.ntot), v(I.v, I.ntot), w(I.w, I.ntot);
    int o = m.k0 * sy;
    const int olim = nz * sy;
    R p_km = p.ld(ts, o - sy);
    R r_km = rho_of(I, F.ld(ts, o - sy));
    // operands of plane k, fetched one plane ahead
    R n_p = p.ld(ts, o), n_pjm = p.ld(tsm, o), n_f = F.ld(ts, o), n_fjm = F.ld(tsm, o);
    R n_us = us.ld(ts, o), n_vs = vs.ld(ts, o), n_ws = ws.ld(ts, o);
    for (int k = m.k0; k <= m.k1; ++k, o += sy) {
        const R pk = n_p, pjm = n_pjm;
        const R rk = rho_of(I, n_f), rjm = rho_of(I, n_fjm);
        const R usk = n_us, vsk = n_vs, wsk = n_ws;
        {
            const int o2 = min(o + sy, olim);
            n_p = p.ld(ts, o2); n_pjm = p.ld(tsm, o2);
            n_f = F.ld(ts, o2); n_fjm = F.ld(tsm, o2);
            n_us = us.ld(ts, o2); n_vs = vs.ld(ts, o2); n_ws = ws.ld(ts, o2);
        }
        correct_plane(I, m, L, k, o, pk, pjm, rk, rjm, usk, vsk, wsk, p_km, r_km, u, v, w);
        p_km = pk;
        r_km = rk;
    }
}

// Chunks of KZ planes with every operand of the chunk (p and F on KZ+1 planes
// of the lane's row, the rest on KZ planes) loaded before the first plane.
template <class R, int KZ>
__global__ __launch_bounds__(256) void k_correct_blk(const EnsArg<R> ea, int fsel, int psel) {
    VOF_PROLOGUE(R)
    WMap m;
    if (!wmap<R, 1>(I, lb, KZ, 1, I.nz, m)) return;
    const CorrLane L = corr_lane(I, m);
    const int nz = I.nz, sx = I.sx, sy = I.sy, k0 = m.k0, k1 = m.k1;
    const int ts = L.ts, tsm = ts - sx;
    const Buf<R> F(sel2(I.F, fsel), I.ntot), p(sel2(I.p, psel), I.ntot);
    const Buf<R> us(I.us, I.ntot), vs(I.vs, I.ntot), ws(I.ws, I.ntot);
    const Buf<R> u(I.u, I.ntot), v(I.v, I.ntot), w(I.w, I.ntot);
    R pc[KZ + 1], fc[KZ + 1], pjm[KZ], fjm[KZ], usw[KZ], vsw[KZ], wsw[KZ];  // pc, fc[e]: plane k0-1+e
#pragma unroll
    for (int e = 0; e < KZ + 1; ++e) {
        const int o = min(k0 - 1 + e, nz) * sy;
        pc[e] = p.ld(ts, o);
        fc[e] = F.ld(ts, o);
    }
#pragma unroll
    for (int e = 0; e < KZ; ++e) {
        const int o = min(k0 + e, nz) * sy;
        pjm[e] = p.ld(tsm, o);
        fjm[e] = F.ld(tsm, o);
        usw[e] = us.ld(ts, o);
        vsw[e] = vs.ld(ts, o);
        wsw[e] = ws.ld(ts, o);
    }
#pragma unroll
    for (int e = 0; e < KZ; ++e) {
        const int k = k0 + e;
        if (k > k1) break;
        correct_plane(I, m, L, k, k * sy, pc[e + 1], pjm[e], rho_of(I, fc[e + 1]), rho_of(I, fjm[e]), usw[e], vsw[e],
                      wsw[e], pc[e], rho_of(I, fc[e]), u, v, w);
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
//
// The last sweep of a step (fin != 0) also performs post_process_f (:502-507),
// a second clamp01 (not idempotent: clamp01 maps the tiny negatives it can
// leave, -x, to +x), and apply_bc_F (:1122-1135), whose three passes leave every
// ghost equal to the nearest interior cell: the lane of a boundary cell writes
// its value to every ghost cell that clamps onto it (in the kernels that end a
// plain step, the ghost-column lanes store the x ghosts: store_F_row_fin).

// x is +0 (not -0).  Where the antidiffusive fluxes a of a cell's two faces
// are +0, their limited products a*C are +0 (C is finite and >= 0), so the
// stage-4 correction ((+0 - +0) / s) * vol / dv is +0 for dv > 0 and
// Fnew = clamp01(Ftd - +0) = clamp01(Ftd) bit for bit.
__device__ __forceinline__ bool pos_zero(float x) { return __float_as_uint(x) == 0u; }
__device__ __forceinline__ bool pos_zero(double x) { return __double_as_longlong(x) == 0; }

template <class R, int DIR>
__device__ __forceinline__ R fct_stage1_scale(const Inst<R>& I, R d) {
    if (DIR == 0) return d * I.dy * I.dz / I.vol;
    if (DIR == 1) return d * I.dy / I.dxdy;
    return d * I.dy * I.dx / I.vol;
}

// Ghost copies of interior cell (i,j,k): every ghost whose nearest interior
// cell it is.  The whole offset is per lane (i, j, k may all differ across
// the wave), so the buffer instruction's scalar offset stays 0.
template <class R>
__device__ __forceinline__ void store_F_ghosts(const Inst<R>& I, const Buf<R>& Fo, int i, int j, int k, R val) {
    const int sx = I.sx, sy = I.sy;
    if (!(i == 1 || i == I.nx || j == 1 || j == I.ny || k == 1 || k == I.nz)) return;
    int ii[3], jj[3], kk[3], ni = 0, nj = 0, nk = 0;
    ii[ni++] = i; if (i == 1) ii[ni++] = 0; if (i == I.nx) ii[ni++] = I.nx + 1;
    jj[nj++] = j; if (j == 1) jj[nj++] = 0; if (j == I.ny) jj[nj++] = I.ny + 1;
    kk[nk++] = k; if (k == 1) kk[nk++] = 0; if (k == I.nz) kk[nk++] = I.nz + 1;
    for (int a = 0; a < ni; ++a)
        for (int b = 0; b < nj; ++b)
            for (int c = 0; c < nk; ++c)
                if (a | b | c) Fo.st(val, kk[c] * sy + jj[b] * sx + ii[a], 0);
}

// Ghost copies in column iw of the value of interior cell (i, j, k), i = iw
// or its interior neighbour: (iw, jj, kk) for jj in {j, j's ghost row} and kk
// in {k, k's ghost plane}, except (iw, j, k) itself.  With iw = i these are the
// cell's j and k ghosts; with iw = 0 or nx+1 the edges and corners of its i
// ghost.  Together with the i ghost (iw, j, k) they are store_F_ghosts.
template <class R>
__device__ __forceinline__ void store_F_ghosts_jk(const Inst<R>& I, const Buf<R>& Fo, int iw, int j, int k, R val) {
    const int sx = I.sx, sy = I.sy;
    if (!(j == 1 || j == I.ny || k == 1 || k == I.nz)) return;
    int jj[3], kk[3], nj = 0, nk = 0;
    jj[nj++] = j; if (j == 1) jj[nj++] = 0; if (j == I.ny) jj[nj++] = I.ny + 1;
    kk[nk++] = k; if (k == 1) kk[nk++] = 0; if (k == I.nz) kk[nk++] = I.nz + 1;
    for (int b = 0; b < nj; ++b)
        for (int c = 0; c < nk; ++c)
            if (b | c) Fo.st(val, kk[c] * sy + jj[b] * sx + iw, 0);
}

// Last-sweep store of one row segment of a wave whose lanes run along x
// (every lane active, so that each shuffle reads a live lane); vo, so: the
// lane's offset and the scalar offset of the row.  Interior lanes (store)
// write their value and its j/k ghosts.  A ghost-column lane (i = 0 or nx+1)
// with nb set takes the value of its interior neighbour lane (lane+1 at i = 0,
// lane-1 at i = nx+1), which must hold a final value, and writes it as the i
// ghost in the same store instruction as the row, so the ghost shares the
// row's cache line instead of costing a lone 4-byte write later, then that
// value's edge and corner ghosts.  Where the ghost lane cannot (its neighbour
// is in another wave or is a halo lane), the interior lane writes that column
// itself: wlo at i = 1 (column 0), whi at i = nx (column nx+1).
template <class R>
__device__ __forceinline__ void store_F_row_fin(const Inst<R>& I, const Buf<R>& Fo, R val, bool store, bool nb,
                                                bool wlo, bool whi, int i, int j, int k, int vo, int so) {
    const R dn = sh_dn(val), up = sh_up(val);
    const R v = nb ? ((i == 0) ? dn : up) : val;
    if (store || nb) {
        Fo.st(v, vo, so);
        store_F_ghosts_jk(I, Fo, i, j, k, v);
    }
    if (store && wlo) {
        Fo.st(val, vo - 1, so);
        store_F_ghosts_jk(I, Fo, 0, j, k, val);
    }
    if (store && whi) {
        Fo.st(val, vo + 1, so);
        store_F_ghosts_jk(I, Fo, I.nx + 1, j, k, val);
    }
}

// Store of cell (i,j,k), k wave-uniform; plus its ghost copies on the last sweep.
template <class R>
__device__ __forceinline__ void store_F(const Inst<R>& I, const Buf<R>& Fo, int i, int j, int k, R val,
                                        bool ghosts) {
    Fo.st(val, j * I.sx + i, k * I.sy);
    if (ghosts) store_F_ghosts(I, Fo, i, j, k, val);
}

// Gas cells stay exactly zero: if F vanishes on the whole 7-cell stencil of a
// cell, every flux is a signed zero, Ftd = +0, rp = rm = 0 and Fnew = clamp01(+-0)
// = +0, so the divisions are skipped (wave-uniformly) with identical results.

// x sweep: lanes run along x over the flattened (x,y) plane (overlapped waves,
// halo 3: a wave stores its middle 58 lanes, whose whole +-3 stencil is inside
// the wave); each wave walks kz planes.  Row ends need no care: ghost lanes
// carry the reference's zero conventions, which decouple consecutive rows.
//
// x-sweep value of the lane's cell from F and u of its row (every lane
// active, so that each shuffle reads a live lane); before the last sweep's
// second clamp, and +0 where the lane is not an interior cell.
template <class R>
__device__ __forceinline__ R fct_x_value(const Inst<R>& I, const R f0, const R u0, bool cell) {
    if (__all(f0 == R(0))) return R(0);
    const R fm = sh_up(f0);
    const R up = sh_dn(u0);
    const R vdt = u0 * I.dt;
    const R L0 = (u0 >= R(0)) ? vdt * fm : vdt * f0;
    const R H0 = (u0 <= R(0)) ? vdt * fm : vdt * f0;
    const R a0 = H0 - L0;
    const R L1 = sh_dn(L0);
    const R a1 = sh_dn(a0);
    const R dv = I.vol - I.dtdydz * (up - u0);
    R ftd = R(0);
    if (cell) {
        R oo = (f0 + fct_stage1_scale<R, 0>(I, L0 - L1)) * I.vol / dv;
        if (oo > R(1) || oo < R(0)) oo = clamp01(oo);
        ftd = oo;
    }
    const R ftm = sh_up(ftd);
    const R ftp = sh_dn(ftd);
    R rp = R(0), rm = R(0);
    if (cell) {
        const R fmx = vmax(ftd, vmax(ftm, ftp));
        const R fmn = vmin(ftd, vmin(ftm, ftp));
        const R pp = vmax(R(0), a0) - vmin(R(0), a1);
        if (pp > R(0)) rp = vmin(R(1), (fmx - ftd) * I.dx / pp);
        const R pm = vmax(R(0), a1) - vmin(R(0), a0);
        if (pm > R(0)) rm = vmin(R(1), (ftd - fmn) * I.dx / pm);
    }
    const R rp1 = sh_dn(rp);
    const R rm1 = sh_dn(rm);
    const R c1 = cell ? ((a1 >= R(0)) ? vmin(rp1, rm) : vmin(rp, rm1)) : R(0);
    const R ac1 = a1 * c1;
    const R ac0 = sh_up(ac1);
    // No limited antidiffusion through either face of any cell of the wave
    // (liquid and gas interiors): the correction is +0 and Fnew = clamp01(Ftd).
    if (__all(!cell || (pos_zero(ac1) && pos_zero(ac0) && dv > R(0)))) return cell ? clamp01(ftd) : R(0);
    return cell ? clamp01(ftd - ((ac1 - ac0) / I.dy) * I.vol / dv) : R(0);
}

// One plane k of a wave: the x sweep of the plane and its store.
template <class R>
__device__ __forceinline__ void fct_x_plane(const Inst<R>& I, const Buf<R>& Fo, const R f0, const R u0, bool cell,
                                            bool store, int i, int j, int k, int fin) {
    const R fn = fct_x_value(I, f0, u0, cell);
    if (store) store_F(I, Fo, i, j, k, fin ? clamp01(fn) : fn, fin != 0);
}

// Any chunk length: planes streamed with F and u fetched one plane ahead.
template <class R>
__global__ __launch_bounds__(256) void k_fct_x(const EnsArg<R> ea,
                                               int src, int kz, int fin) {
    VOF_PROLOGUE_RR(R)
    WMap m;
    if (!wmap<R, 3>(I, lb, kz, 1, I.nz, m)) return;
    const int i = m.i, j = m.j;
    const bool cell = (m.t == m.tc) && i >= 1 && i <= I.nx && j >= 1 && j <= I.ny;
    const bool store = cell && m.own;
    const Buf<R> Fi(sel2(I.F, src), I.ntot), Fo(sel2(I.F, src ^ 1), I.ntot), u(I.u, I.ntot);
    R nf = Fi.ld(m.tc, m.k0 * I.sy), nu = u.ld(m.tc, m.k0 * I.sy);  // plane k, fetched one plane ahead
    for (int k = m.k0; k <= m.k1; ++k) {
        const R f0 = nf, u0 = nu;
        const int on = min(k + 1, m.k1) * I.sy;
        nf = Fi.ld(m.tc, on);
        nu = u.ld(m.tc, on);
        fct_x_plane(I, Fo, f0, u0, cell, store, i, j, k, fin);
    }
}

// Chunks of KZ planes: F of the whole chunk is loaded before the first plane,
// and u only when F is nonzero somewhere on the wave's chunk (an all-gas chunk
// stores zeros, which is what every one of its planes would store).
template <class R, int KZ>
__global__ __launch_bounds__(256) void k_fct_xb(const EnsArg<R> ea,
                                                int src, int fin) {
    VOF_PROLOGUE_RR(R)
    WMap m;
    if (!wmap<R, 3>(I, lb, KZ, 1, I.nz, m)) return;
    const int i = m.i, j = m.j;
    const bool cell = (m.t == m.tc) && i >= 1 && i <= I.nx && j >= 1 && j <= I.ny;
    const bool store = cell && m.own;
    const Buf<R> Fi(sel2(I.F, src), I.ntot), Fo(sel2(I.F, src ^ 1), I.ntot), u(I.u, I.ntot);
    R fw[KZ], uw[KZ];
#pragma unroll
    for (int e = 0; e < KZ; ++e) fw[e] = Fi.ld(m.tc, min(m.k0 + e, m.k1) * I.sy);
    bool gas = true;
#pragma unroll
    for (int e = 0; e < KZ; ++e) gas = gas && (fw[e] == R(0));
    if (__all(gas)) {
        if (store) {
#pragma unroll
            for (int e = 0; e < KZ; ++e)
                if (m.k0 + e <= m.k1) store_F(I, Fo, i, j, m.k0 + e, R(0), fin != 0);
        }
        return;
    }
#pragma unroll
    for (int e = 0; e < KZ; ++e) uw[e] = u.ld(m.tc, min(m.k0 + e, m.k1) * I.sy);
#pragma unroll
    for (int e = 0; e < KZ; ++e)
        if (m.k0 + e <= m.k1) fct_x_plane(I, Fo, fw[e], uw[e], cell, store, i, j, m.k0 + e, fin);
}

// y and z sweeps: lanes run along x (coalesced), each lane marches along the
// sweep line through a register window, producing one Fnew per step two cells
// behind the Ftd front.  Lines are cut into chunks of ch cells with a 5-cell
// warm-up, so that there are enough waves to cover memory latency.
//
// The window of the march, positioned on cell c of the line: F and V of c,
// the low-face flux L and antidiffusive flux a of faces c, c-1, c-2, Ftd and
// dv of cells c-1, c-2, the Zalesak ratios of c-1, c-2 and the limiter of
// face c-2.  zr counts the trailing run of zero F values.
template <class R>
struct FctWin {
    R Fc, Vc, Lc, ac, am1, am2, ftm1, ftm2, dvm1, dvm2, rpm1, rmm1, rpm2, rmm2, cm2;
    int zr;
};

// Window on cell c = m0 - 2 from F(c-1), F(c), V(c).
template <class R>
__device__ __forceinline__ void fct_win_init(const Inst<R>& I, FctWin<R>& w, R Fcm, R Fc, R Vc) {
    w.Fc = Fc;
    w.Vc = Vc;
    const R vdt = Vc * I.dt;
    w.Lc = (Vc >= R(0)) ? vdt * Fcm : vdt * Fc;
    const R Hc = (Vc <= R(0)) ? vdt * Fcm : vdt * Fc;
    w.ac = Hc - w.Lc;
    w.zr = (Fc == R(0)) ? ((Fcm == R(0)) ? 2 : 1) : 0;
    w.am1 = R(0); w.am2 = R(0); w.ftm1 = R(0); w.ftm2 = R(0); w.dvm1 = R(1); w.dvm2 = R(1);
    w.rpm1 = R(0); w.rmm1 = R(0); w.rpm2 = R(0); w.rmm2 = R(0); w.cm2 = R(0);
}

// One march step on cell c of a line of nd cells, given F and V of cell c+1:
// forms Ftd(c), the ratios of c-1 and the limiter of face c-1; when emit,
// returns the corrected value of cell c-2 (before the final-sweep clamp).
// Shifts the window to c+1.
template <class R, int DIR>
__device__ __forceinline__ R fct_win_step(const Inst<R>& I, FctWin<R>& w, int c, int nd, R Fcp, R Vcp,
                                          bool emit) {
    const R kdv = (DIR == 1) ? I.dtdxdz : I.dtdxdy;
    const R qf = (DIR == 1) ? I.dx : I.dz;
    const R s4 = (DIR == 1) ? I.dy : I.dz;
    const R dt = I.dt, vol = I.vol;
    w.zr = (Fcp == R(0)) ? w.zr + 1 : 0;
    const R dvc = vol - kdv * (Vcp - w.Vc);
    R Lcp, acp, ftc, cm1, out = R(0);
    if (w.zr >= 7) {
        // F(c-5..c+1) all zero: every value formed this step is (signed) zero.
        Lcp = R(0); acp = R(0); ftc = R(0);
        w.rpm1 = R(0); w.rmm1 = R(0); cm1 = R(0);
    } else {
        const R vdt = Vcp * dt;
        Lcp = (Vcp >= R(0)) ? vdt * w.Fc : vdt * Fcp;
        const R Hcp = (Vcp <= R(0)) ? vdt * w.Fc : vdt * Fcp;
        acp = Hcp - Lcp;
        ftc = R(0);
        if (c >= 1 && c <= nd) {
            R oo = (w.Fc + fct_stage1_scale<R, DIR>(I, w.Lc - Lcp)) * vol / dvc;
            if (oo > R(1) || oo < R(0)) oo = clamp01(oo);
            ftc = oo;
        }
        // Zalesak ratios of cell c-1
        w.rpm1 = R(0); w.rmm1 = R(0);
        if (c - 1 >= 1 && c - 1 <= nd) {
            const R fmx = vmax(w.ftm1, vmax(w.ftm2, ftc));
            const R fmn = vmin(w.ftm1, vmin(w.ftm2, ftc));
            const R pp = vmax(R(0), w.am1) - vmin(R(0), w.ac);
            if (pp > R(0)) w.rpm1 = vmin(R(1), (fmx - w.ftm1) * qf / pp);
            const R pm = vmax(R(0), w.ac) - vmin(R(0), w.am1);
            if (pm > R(0)) w.rmm1 = vmin(R(1), (w.ftm1 - fmn) * qf / pm);
        }
        // limiter on face c-1
        cm1 = R(0);
        if (c - 1 >= 2 && c - 1 <= nd + 1) cm1 = (w.am1 >= R(0)) ? vmin(w.rpm1, w.rmm2) : vmin(w.rpm2, w.rmm1);
        // corrected value of cell c-2; with no antidiffusion through either of
        // its faces on any lane of the wave, the correction is +0
        if (emit) {
            if (__all(pos_zero(w.am1) && pos_zero(w.am2) && w.dvm2 > R(0))) out = clamp01(w.ftm2);
            else out = clamp01(w.ftm2 - ((w.am1 * cm1 - w.am2 * w.cm2) / s4) * vol / w.dvm2);
        }
    }
    // shift the window
    w.Fc = Fcp; w.Vc = Vcp; w.Lc = Lcp;
    w.am2 = w.am1; w.am1 = w.ac; w.ac = acp;
    w.ftm2 = w.ftm1; w.ftm1 = ftc;
    w.dvm2 = w.dvm1; w.dvm1 = dvc;
    w.rpm2 = w.rpm1; w.rmm2 = w.rmm1;
    w.cm2 = cm1;
    return out;
}

// Lane -> line mapping shared by the two march kernels: lanes along x over the
// flattened (x, other) plane.  The wave-uniform parts (line length nd, stride
// S between line elements, chunk index chn) are formed before any lane exits,
// so that they stay in scalar registers; false for lanes without an interior
// line.
struct MarchLine {
    int nd, S, chn, nbt;  // wave-uniform
    int i, o, base;       // per lane
};

template <class R, int DIR>
__device__ __forceinline__ MarchLine march_uniform(const Inst<R>& I, int lb) {
    MarchLine L;
    L.nd = (DIR == 1) ? I.ny : I.nz;
    L.S = (DIR == 1) ? I.sx : I.sy;
    const int no = (DIR == 1) ? I.nz : I.ny;
    L.nbt = (((no * I.sx + 255) >> 8) + 7) & ~7;  // multiple of 8: chunk neighbours share an XCD
    L.chn = lb / L.nbt;
    return L;
}

// As march_lane, keeping the lanes of the ghost columns i = 0 and nx+1 (g set;
// they march on their own column's values, which are never stored).
template <class R, int DIR>
__device__ __forceinline__ bool march_lane_g(const Inst<R>& I, int lb, MarchLine& L, bool& g) {
    const int no = (DIR == 1) ? I.nz : I.ny;
    const int t = (lb - L.chn * L.nbt) * 256 + (int)threadIdx.x;
    if (t >= no * I.sx) return false;
    L.o = 1 + t / I.sx;
    L.i = t - (L.o - 1) * I.sx;
    g = L.i < 1 || L.i > I.nx;
    L.base = (DIR == 1) ? (L.o * I.sy + L.i) : (L.o * I.sx + L.i);
    return true;
}

template <class R, int DIR>
__device__ __forceinline__ bool march_lane(const Inst<R>& I, int lb, MarchLine& L) {
    const int no = (DIR == 1) ? I.nz : I.ny;
    const int t = (lb - L.chn * L.nbt) * 256 + (int)threadIdx.x;
    if (t >= no * I.sx) return false;
    L.o = 1 + t / I.sx;
    L.i = t - (L.o - 1) * I.sx;
    if (L.i < 1 || L.i > I.nx) return false;
    L.base = (DIR == 1) ? (L.o * I.sy + L.i) : (L.o * I.sx + L.i);
    return true;
}

// Store of line element mm (wave-uniform): per-lane line base + scalar mm*S,
// as the loads.
template <class R, int DIR>
__device__ __forceinline__ void march_put(const Inst<R>& I, const Buf<R>& Fo, const MarchLine& L, int S, int mm,
                                          R val, bool gh) {
    Fo.st(val, L.base, mm * S);
    if (!gh) return;
    if (DIR == 1) store_F_ghosts(I, Fo, L.i, mm, L.o, val);
    else store_F_ghosts(I, Fo, L.i, L.o, mm, val);
}

// Last-sweep store of line element mm (wave-uniform) in k_fct_line, whose
// ghost-column lanes stay active and whose lanes all hold final values (each
// marches its own line): a ghost lane stores its neighbour's value where that
// neighbour is in the same wave; where it is not (i = 0 on lane 63, i = nx+1
// on lane 0), the interior lane writes that column itself.
template <class R, int DIR>
__device__ __forceinline__ void march_put_fin(const Inst<R>& I, const Buf<R>& Fo, const MarchLine& L, bool g,
                                              int S, int mm, R val) {
    const int lane = (int)threadIdx.x & 63;
    const int j = (DIR == 1) ? mm : L.o, k = (DIR == 1) ? L.o : mm;
    const bool nb = g && ((L.i == 0) ? (lane != 63) : (lane != 0));
    store_F_row_fin(I, Fo, val, !g, nb, L.i == 1 && lane == 0, L.i == I.nx && lane == 63, L.i, j, k, L.base,
                    mm * S);
}

// Any chunk length: the line is streamed through the window with F and V
// fetched three cells ahead.
template <class R, int DIR>
__global__ __launch_bounds__(256) void k_fct_march(const EnsArg<R> ea,
                                                   int src, int ch, int fin) {
    VOF_PROLOGUE_RR(R)
    MarchLine L = march_uniform<R, DIR>(I, lb);
    const int S = L.S, nd = L.nd, chn = L.chn;
    if (!march_lane<R, DIR>(I, lb, L)) return;
    const Buf<R> Fb(sel2(I.F, src), I.ntot), Fo(sel2(I.F, src ^ 1), I.ntot), Vb((DIR == 1) ? I.v : I.w, I.ntot);
    auto cl = [nd](int mm) { return mm < 0 ? 0 : (mm > nd + 1 ? nd + 1 : mm); };
    auto F = [&](int mm) { return Fb.ld(L.base, cl(mm) * S); };
    auto V = [&](int mm) { return Vb.ld(L.base, cl(mm) * S); };
    const bool gh = fin != 0;
    const int m0 = 1 + chn * ch;
    const int m1 = min(nd, m0 + ch - 1);
    int c = m0 - 2;
    FctWin<R> w;
    fct_win_init(I, w, F(c - 1), F(c), V(c));
    R Fn = F(c + 1), Vn = V(c + 1), Fn2 = F(c + 2), Vn2 = V(c + 2);
    for (; c <= m1 + 2; ++c) {
        const R Fcp = Fn, Vcp = Vn;
        Fn = Fn2;
        Vn = Vn2;
        Fn2 = F(c + 3);
        Vn2 = V(c + 3);
        const bool emit = c - 2 >= m0;
        const R fn = fct_win_step<R, DIR>(I, w, c, nd, Fcp, Vcp, emit);
        if (emit) march_put<R, DIR>(I, Fo, L, S, c - 2, gh ? clamp01(fn) : fn, gh);
    }
}

// Chunks of CH cells: the whole window of a chunk (F on CH+6 cells, V on CH+5)
// is loaded into registers before the march, so every load of a wave is in
// flight at once.  A wave whose F window is zero on every lane stores zeros
// without loading V (all values of the chunk are +0, as in the march).  The
// lanes of the ghost columns stay active so that the last sweep of a step can
// store the i ghosts with the rows (march_put_fin).
template <class R, int DIR, int CH>
__global__ __launch_bounds__(256) void k_fct_line(const EnsArg<R> ea,
                                                  int src, int fin) {
    VOF_PROLOGUE_RR(R)
    MarchLine L = march_uniform<R, DIR>(I, lb);
    const int S = L.S, nd = L.nd, chn = L.chn;
    bool g = false;  // ghost-column lane: marches, stores only i ghosts of the last sweep
    if (!march_lane_g<R, DIR>(I, lb, L, g)) return;
    const Buf<R> Fb(sel2(I.F, src), I.ntot), Fo(sel2(I.F, src ^ 1), I.ntot), Vb((DIR == 1) ? I.v : I.w, I.ntot);
    auto cl = [nd](int mm) { return mm < 0 ? 0 : (mm > nd + 1 ? nd + 1 : mm); };
    const bool gh = fin != 0;
    const int m0 = 1 + chn * CH;
    const int m1 = min(nd, m0 + CH - 1);
    constexpr int NF = CH + 6, NV = CH + 5;
    R fw[NF], vw[NV];  // F(m0-3+e), V(m0-2+e)
#pragma unroll
    for (int e = 0; e < NF; ++e) fw[e] = Fb.ld(L.base, cl(m0 - 3 + e) * S);
    bool gas = true;
#pragma unroll
    for (int e = 0; e < NF; ++e) gas = gas && (g || fw[e] == R(0));
    if (__all(gas)) {
#pragma unroll
        for (int e = 0; e < CH; ++e) {
            if (m0 + e > m1) continue;
            if (gh) march_put_fin<R, DIR>(I, Fo, L, g, S, m0 + e, R(0));
            else if (!g) Fo.st(R(0), L.base, (m0 + e) * S);
        }
        return;
    }
#pragma unroll
    for (int e = 0; e < NV; ++e) vw[e] = Vb.ld(L.base, cl(m0 - 2 + e) * S);
    FctWin<R> w;
    fct_win_init(I, w, fw[0], fw[1], vw[0]);
#pragma unroll
    for (int s = 0; s < CH + 4; ++s) {  // cell c = m0 - 2 + s
        const int c = m0 - 2 + s;
        const R fn = fct_win_step<R, DIR>(I, w, c, nd, fw[s + 2], vw[s + 1], s >= 4);
        if (s >= 4 && c - 2 <= m1) {
            if (gh) march_put_fin<R, DIR>(I, Fo, L, g, S, c - 2, clamp01(fn));
            else if (!g) Fo.st(fn, L.base, (c - 2) * S);
        }
    }
}

// z sweep directly followed by the x sweep, in one launch (the sweep orders
// (y, z, x) and (z, x, y) of :1384-1396).  Lanes run along x in the
// overlapped waves of the x sweep (halo 3, 58 stored lanes) and march along z
// through a chunk of CH planes as in k_fct_line<R, 2, CH>; each z-sweep value
// the march emits stays in a register, and the x sweep of that plane is formed
// from it at once by lane shifts (every lane of the wave emits the same plane
// in the same step).  The intermediate F never goes to memory.
//
// Lanes outside the interior (ghost columns and rows) carry F' = 0 into the x
// sweep.  The two-launch form reads the intermediate buffer's ghost cells
// there, whose values it does not depend on: at a wall face u = +0, so both
// fluxes through it are signed zeros whose sign cancels in L - L' and H - L,
// and ghost cells' own Ftd, ratios and limiters are zero.  The z part is never
// the last sweep of a step, so only the x part clamps twice and writes ghosts.
template <class R, int CH>
__global__ __launch_bounds__(256) void k_fct_zx(const EnsArg<R> ea,
                                                int src, int fin) {
    VOF_PROLOGUE_RR(R)
    WMap m;
    if (!wmap<R, 3, true>(I, lb, CH, 1, I.nz, m)) return;
    const int i = m.i, j = m.j, nd = I.nz, S = I.sy;
    const bool line = (m.t == m.tc) && i >= 1 && i <= I.nx && j >= 1 && j <= I.ny;
    const bool store = line && m.own;
    // last sweep: a stored ghost-column lane writes its neighbour's value as the i ghost when
    // that neighbour is a stored lane too (lanes 3..60 hold final values; 0-2 and 61-63 are
    // halo); otherwise the neighbour, stored by the adjacent wave on its lane 3 or 60, writes it
    const bool gcol = m.own && (m.t == m.tc) && j >= 1 && j <= I.ny &&
                      ((i == 0 && m.lane < 63 - 3) || (i == I.nx + 1 && m.lane > 3));
    const bool wlo = i == 1 && m.lane == 3, whi = i == I.nx && m.lane == 63 - 3;
    const Buf<R> Fb(sel2(I.F, src), I.ntot), Fo(sel2(I.F, src ^ 1), I.ntot), Wb(I.w, I.ntot), Ub(I.u, I.ntot);
    auto cl = [nd](int mm) { return mm < 0 ? 0 : (mm > nd + 1 ? nd + 1 : mm); };
    const int m0 = m.k0, m1 = m.k1;
    constexpr int NF = CH + 6, NV = CH + 5;
    R fw[NF], vw[NV], uw[CH];  // F(m0-3+e), w(m0-2+e), u(m0+e)
#pragma unroll
    for (int e = 0; e < NF; ++e) fw[e] = Fb.ld(m.tc, cl(m0 - 3 + e) * S);
    bool gas = true;
#pragma unroll
    for (int e = 0; e < NF; ++e) gas = gas && (!line || fw[e] == R(0));
    if (__all(gas)) {  // F' = 0 on every lane and plane, so the x sweep stores +0
#pragma unroll
        for (int e = 0; e < CH; ++e) {
            if (m0 + e > m1) continue;
            if (fin) store_F_row_fin(I, Fo, R(0), store, gcol, wlo, whi, i, j, m0 + e, m.tc, (m0 + e) * S);
            else if (store) Fo.st(R(0), m.tc, (m0 + e) * S);
        }
        return;
    }
#pragma unroll
    for (int e = 0; e < NV; ++e) vw[e] = Wb.ld(m.tc, cl(m0 - 2 + e) * S);
#pragma unroll
    for (int e = 0; e < CH; ++e) uw[e] = Ub.ld(m.tc, min(m0 + e, m1) * S);
    FctWin<R> w;
    fct_win_init(I, w, fw[0], fw[1], vw[0]);
#pragma unroll
    for (int s = 0; s < CH + 4; ++s) {  // cell c = m0 - 2 + s
        const int c = m0 - 2 + s;
        const R fz = fct_win_step<R, 2>(I, w, c, nd, fw[s + 2], vw[s + 1], s >= 4);
        if (s >= 4 && c - 2 <= m1) {
            const R fx = fct_x_value(I, line ? fz : R(0), uw[s - 4], line);
            if (fin) store_F_row_fin(I, Fo, clamp01(fx), store, gcol, wlo, whi, i, j, c - 2, m.tc, (c - 2) * S);
            else if (store) Fo.st(fx, m.tc, (c - 2) * S);
        }
    }
}

// x sweep directly followed by the y sweep, in one launch (the sweep order
// x-y-z of :1384-1396; neither is the last sweep of the step).  Lanes run
// along x over the flattened (x, z) plane of interior planes in overlapped
// waves (halo 3, 58 stored lanes) and march along y through a chunk of CH
// rows as in k_fct_line<R, 1, CH>.  Each row that enters the march window is
// first swept along x by lane shifts; the intermediate F never goes to
// memory.  Rows outside the interior (ghost rows, clamped window rows) and
// ghost columns enter the y march as +0: the two-launch form reads the
// intermediate buffer's ghost cells there, whose values it does not depend on
// (at a wall face v = +0, so both fluxes through it are signed zeros whose
// sign cancels, and ghost cells' own Ftd, ratios and limiters are zero).
template <class R, int CH>
__global__ __launch_bounds__(256) void k_fct_xy(const EnsArg<R> ea,
                                                int src) {
    VOF_PROLOGUE_RR(R)
    constexpr int H = 3, SW = 64 - 2 * H;
    const int P = I.sx * I.nz;
    const int wpp = (P + SW - 1) / SW;
    const int bpp = (((wpp + 3) >> 2) + 7) & ~7;  // multiple of 8: a line's chunks share an XCD
    const int chn = lb / bpp;
    const int wv = (lb - chn * bpp) * 4 + ((int)threadIdx.x >> 6);
    if (wv >= wpp) return;  // wave-uniform
    const int lane = (int)threadIdx.x & 63;
    const int t = wv * SW - H + lane;
    const int tc = min(max(t, 0), P - 1);
    const int k = 1 + tc / I.sx;
    const int i = tc - (k - 1) * I.sx;
    const bool col = (t == tc) && i >= 1 && i <= I.nx;  // an interior line
    const bool store = col && lane >= H && lane < 64 - H;
    const int nd = I.ny, S = I.sx, base = k * I.sy + i;
    const Buf<R> Fb(sel2(I.F, src), I.ntot), Fo(sel2(I.F, src ^ 1), I.ntot), Ub(I.u, I.ntot), Vb(I.v, I.ntot);
    auto cl = [nd](int mm) { return mm < 0 ? 0 : (mm > nd + 1 ? nd + 1 : mm); };
    const int m0 = 1 + chn * CH;
    const int m1 = min(nd, m0 + CH - 1);
    constexpr int NF = CH + 6, NV = CH + 5;
    R fw[NF], uw[NF], vw[NV];  // F and u of rows m0-3+e, v of rows m0-2+e
#pragma unroll
    for (int e = 0; e < NF; ++e) fw[e] = Fb.ld(base, cl(m0 - 3 + e) * S);
    bool gas = true;
#pragma unroll
    for (int e = 0; e < NF; ++e) gas = gas && (fw[e] == R(0));
    if (__all(gas)) {  // every x-swept row is +0, so the y sweep stores +0
        if (store) {
#pragma unroll
            for (int e = 0; e < CH; ++e)
                if (m0 + e <= m1) Fo.st(R(0), base, (m0 + e) * S);
        }
        return;
    }
#pragma unroll
    for (int e = 0; e < NF; ++e) uw[e] = Ub.ld(base, cl(m0 - 3 + e) * S);
#pragma unroll
    for (int e = 0; e < NV; ++e) vw[e] = Vb.ld(base, cl(m0 - 2 + e) * S);
    // x-swept F of window row e (wave-uniform row index)
    auto xrow = [&](int e) __attribute__((always_inline)) -> R {
        const int mm = m0 - 3 + e;
        if (mm < 1 || mm > nd) return R(0);
        return fct_x_value(I, fw[e], uw[e], col);
    };
    FctWin<R> w;
    {
        const R f0 = xrow(0), f1 = xrow(1);
        fct_win_init(I, w, f0, f1, vw[0]);
    }
#pragma unroll
    for (int s = 0; s < CH + 4; ++s) {  // cell c = m0 - 2 + s
        const int c = m0 - 2 + s;
        const R fcp = xrow(s + 2);
        const R fn = fct_win_step<R, 1>(I, w, c, nd, fcp, vw[s + 1], s >= 4);
        if (s >= 4 && c - 2 <= m1 && store) Fo.st(fn, base, (c - 2) * S);
    }
}

// ---------------------------------------------------------------------------
// F ghost layer on its own (after an IC upload or a single test stage): every
// ghost equals the nearest interior cell, F(clamp(i), clamp(j), clamp(k)),
// written once per ghost: k faces (full planes), then j faces and i faces of
// the interior slabs.
template <class R>
__global__ __launch_bounds__(256) void k_ghost_F(const EnsArg<R> ea,
                                                 int fsel) {
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
    R* __restrict__ F = sel2(I.F, fsel);
    const int ci = min(max(i, 1), nx), cj = min(max(j, 1), ny), ck = min(max(k, 1), nz);
    F[k * sy + j * sx + i] = F[ck * sy + cj * sx + ci];
}

// post_process_f (:502-507) on interior cells, for stage-by-stage tests; in a
// step it is folded into the store of the last FCT sweep.
template <class R>
__global__ __launch_bounds__(256) void k_clamp_F(const EnsArg<R> ea,
                                                 int fsel, int kz) {
    VOF_PROLOGUE(R)
    WMap m;
    if (!wmap<R, 1>(I, lb, kz, 1, I.nz, m)) return;
    if (!(m.own && m.i >= 1 && m.i <= I.nx && m.j >= 1 && m.j <= I.ny)) return;
    R* __restrict__ F = sel2(I.F, fsel);
    for (int k = m.k0; k <= m.k1; ++k) {
        const int c = k * I.sy + m.tc;
        F[c] = clamp01(F[c]);
    }
}

// ---------------------------------------------------------------------------
// CSF surface tension (extension, VOF_FLAG_CSF): F smoothed on interior cells,
// the curvature of every interior cell from it, then the surface-tension
// acceleration added to the predicted face velocities (dt a on faces i >= 2,
// j >= 2, k >= 2, as the predictor writes them).  They run after the
// predictor and before the rhs.
template <class R>
__global__ __launch_bounds__(256) void k_csf_smooth(const EnsArg<R> ea,
                                                    int fsel, int kz) {
    VOF_PROLOGUE(R)
    WMap m;
    if (!wmap<R, 1>(I, lb, kz, 1, I.nz, m)) return;
    if (!(m.own && m.i >= 1 && m.i <= I.nx && m.j >= 1 && m.j <= I.ny)) return;
    const R* __restrict__ F = sel2(I.F, fsel);
    for (int k = m.k0; k <= m.k1; ++k) {
        const int c = k * I.sy + m.tc;
        I.fsm[c] = csf_smooth(I, F, c);
    }
}

template <class R>
__global__ __launch_bounds__(256) void k_curvature(const EnsArg<R> ea,
                                                   int kz) {
    VOF_PROLOGUE(R)
    WMap m;
    if (!wmap<R, 1>(I, lb, kz, 1, I.nz, m)) return;
    if (!(m.own && m.i >= 1 && m.i <= I.nx && m.j >= 1 && m.j <= I.ny)) return;
    for (int k = m.k0; k <= m.k1; ++k) I.kappa[k * I.sy + m.tc] = csf_curvature(I, I.fsm, m.i, m.j, k);
}

template <class R>
__global__ __launch_bounds__(256) void k_csf(const EnsArg<R> ea,
                                             int fsel, int kz) {
    VOF_PROLOGUE(R)
    WMap m;
    if (!wmap<R, 1>(I, lb, kz, 1, I.nz, m)) return;
    const int i = m.i, j = m.j;
    if (!(m.own && i >= 1 && i <= I.nx && j >= 1 && j <= I.ny)) return;
    const R* __restrict__ F = sel2(I.F, fsel);
    for (int k = m.k0; k <= m.k1; ++k) {
        const int c = k * I.sy + m.tc;
        if (i >= 2) I.us[c] = I.us[c] + I.dt * csf_accel(I, F, I.kappa, c, 1, I.dxi);
        if (j >= 2) I.vs[c] = I.vs[c] + I.dt * csf_accel(I, F, I.kappa, c, I.sx, I.dyi);
        if (k >= 2) I.ws[c] = I.ws[c] + I.dt * csf_accel(I, F, I.kappa, c, I.sy, I.dzi);
    }
}

// ---------------------------------------------------------------------------
// Mean-pressure removal (extension; no reference counterpart): subtract the
// FP64 interior mean, stored by the reduction, from interior p.
template <class R>
__global__ __launch_bounds__(256) void k_sub_mean(const EnsArg<R> ea,
                                                  int psel, const double* __restrict__ means, int kz) {
    VOF_PROLOGUE(R)
    WMap m;
    if (!wmap<R, 1>(I, lb, kz, 1, I.nz, m)) return;
    if (!(m.own && m.i >= 1 && m.i <= I.nx && m.j >= 1 && m.j <= I.ny)) return;
    const R mean = R(means[inst_]);
    R* __restrict__ p = sel2(I.p, psel);
    for (int k = m.k0; k <= m.k1; ++k) {
        const int c = k * I.sy + m.tc;
        p[c] = p[c] - mean;
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
__global__ __launch_bounds__(256) void k_reduce1(const EnsArg<R> ea,
                                                 int fsel, int psel, int kz, int mode,
                                                 double* __restrict__ part) {
    const int b_ = blockIdx.x;
    int lb = b_;
    Inst<R> I = ea.one;
    if (ea.ninst != 1) {
        const int inst_ = find_inst(ea.boff, ea.ninst, b_);
        lb = b_ - ea.boff[inst_];
        I = ea.tab[inst_];
    }
    double s[NRED] = {0.0, 0.0, 0.0, 0.0, 1e300, -1e300, 0.0};
    WMap m;
    if (wmap<R, 0>(I, lb, kz, 0, I.nz + 1, m) && m.own) {
        const int i = m.i, j = m.j;
        const R* __restrict__ F = sel2(I.F, fsel);
        const R* __restrict__ p = sel2(I.p, psel);
        const bool ij_in = (i >= 1 && i <= I.nx && j >= 1 && j <= I.ny);
        const double ax = double(I.axx), ay = double(I.ayy), az = double(I.azz);
        for (int k = m.k0; k <= m.k1; ++k) {
            const int c = k * I.sy + m.tc;
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

// Pass 2 body: the 256 threads of one work-group sum the partials of
// work-groups [b0, b1) in a fixed order; thread q < NRED returns quantity q.
__device__ __forceinline__ double reduce_parts(const double* __restrict__ part, int b0, int b1) {
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
    double r = 0.0;
    if (threadIdx.x < NRED) {
        const int q = threadIdx.x;
        r = sh[0][q];
        for (int w2 = 1; w2 < 4; ++w2) r = (q == 4) ? fmin(r, sh[w2][q]) : (q == 5) ? fmax(r, sh[w2][q]) : r + sh[w2][q];
    }
    return r;
}

// Pass 2: one work-group per instance.  out[inst*NRED + q].  If means is not
// null, also stores the p mean (sum p / n^3) for k_sub_mean.
__global__ __launch_bounds__(256) void k_reduce2(const int* __restrict__ boff, const double* __restrict__ part,
                                                 const long long* __restrict__ ncell, double* __restrict__ out,
                                                 double* __restrict__ means) {
    const int inst = blockIdx.x;
    const double r = reduce_parts(part, boff[inst], boff[inst + 1]);
    if (threadIdx.x < NRED) {
        const int q = threadIdx.x;
        out[(size_t)inst * NRED + q] = r;
        if (q == 1 && means) means[inst] = r / double(ncell[inst]);
    }
}

// Pass 2 of the allocation-free diagnostics: the partials of nblk
// work-groups summed as k_reduce2 sums them, then turned into the caller's
// vof_diag record in device memory (the same arithmetic the host applies to
// k_reduce2's sums in vof_diagnostics).
__global__ __launch_bounds__(256) void k_reduce_diag(const double* __restrict__ part, int nblk, double vol,
                                                     double ncl, vof_diag* __restrict__ out) {
    const double r = reduce_parts(part, 0, nblk);
    switch (threadIdx.x) {
        case 0: out->liquid_volume = vol * r; break;
        case 1: out->p_mean = r / ncl; break;
        case 2: out->residual_l2 = sqrt(r); break;
        case 3: out->rhs_l2 = sqrt(r); break;
        case 4: out->f_min = r; break;
        case 5: out->f_max = r; break;
        case 6: out->nonfinite = (int64_t)r; break;
        default: break;
    }
}

// ---------------------------------------------------------------------------
// Allocation-free diagnostics (for callers that cannot allocate device
// memory: code under stream capture, serving loops).  The per-work-group
// partials of k_reduce1 live in `scratch`, a file-scope device array of
// kNoAllocMaxBlocks x NRED doubles (3.7 MB) that each GPU holds until the
// process ends.  Host code never uses the array's name as an address: the
// first call on a device, with that device current, looks its address up
// with hipGetSymbolAddress and keeps it per device.  Every call on one GPU
// shares the array; k_reduce1 writes every partial that k_reduce_diag reads,
// so nothing needs clearing, but calls that overlap on one GPU (different
// streams, or threads) race on it.
static __device__ __align__(256) double scratch[kNoAllocMaxBlocks * NRED];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

// Work-groups of k_reduce1 for a chunk of kz planes over the stored planes
// 0..nz+1 (as vof_diagnostics uses with kz = 8).
int noalloc_blocks(int nx, int ny, int nz, int kz) { return blocks_wcol(nx, ny, nz + 2, kz, 0); }

template <class R>
hipError_t launch_diag_noalloc(const Inst<R>& I, int kz, double* part, vof_diag* out, hipStream_t st) {
    const int nblk = noalloc_blocks(I.nx, I.ny, I.nz, kz);
    const EnsArg<R> ea{nullptr, nullptr, 1, I};
    hipLaunchKernelGGL(k_reduce1<R>, dim3(nblk), dim3(256), 0, st, ea, 0, 0, kz, 0, part);
    hipLaunchKernelGGL(k_reduce_diag, dim3(1), dim3(256), 0, st, (const double*)part, nblk, double(I.vol),
                       double(I.nx) * I.ny * I.nz, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Plain streaming copy (16 B per lane), used to measure the attainable HBM
// bandwidth and to calibrate the byte counters.  One work-group per 4 KiB, so
// every access of the copy is in flight from the start (a grid-stride loop
// over a few thousand work-groups reaches only ~70% of the rate at 1 GiB).
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ a, float4* __restrict__ b, long n) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) b[t] = a[t];
}

void launch_copy(const void* a, void* b, long n16, hipStream_t st) {
    const long nb = std::min<long>((n16 + 255) / 256, 1L << 30);
    hipLaunchKernelGGL(k_copy, dim3(unsigned(std::max<long>(nb, 1))), dim3(256), 0, st, (const float4*)a, (float4*)b,
                       n16);
}

// Stream mix: NR arrays read and NW arrays written, n16 16-byte elements
// each, consecutive arrays `pitch` elements apart; every written element
// depends on the elements read at its index, so nothing is elided.  The
// attainable time of a step kernel's compulsory traffic (its fields read and
// written once each, with no arithmetic to hide).
template <int NR, int NW>
__global__ __launch_bounds__(256) void k_stream(const float4* __restrict__ a, float4* __restrict__ b, long n16,
                                                long pitch) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n16; t += stride) {
        float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const float4 x = a[r * pitch + t];
            s.x += x.x;
            s.y += x.y;
            s.z += x.z;
            s.w += x.w;
        }
#pragma unroll
        for (int q = 0; q < NW; ++q) b[q * pitch + t] = s;
    }
}

bool launch_stream(int nread, int nwrite, const void* a, void* b, long n16, long pitch, hipStream_t st) {
    const dim3 grid(kStreamBlocks), blk(256);
    const auto* x = (const float4*)a;
    auto* y = (float4*)b;
#define VOF_STREAM_CASE(R, W) \
    case (R) * 8 + (W): hipLaunchKernelGGL((k_stream<R, W>), grid, blk, 0, st, x, y, n16, pitch); return true;
    switch (nread * 8 + nwrite) {
        VOF_STREAM_CASE(1, 1) VOF_STREAM_CASE(2, 1) VOF_STREAM_CASE(3, 1) VOF_STREAM_CASE(4, 1)
        VOF_STREAM_CASE(5, 1) VOF_STREAM_CASE(6, 1) VOF_STREAM_CASE(1, 2) VOF_STREAM_CASE(2, 2)
        VOF_STREAM_CASE(3, 2) VOF_STREAM_CASE(4, 2) VOF_STREAM_CASE(5, 2) VOF_STREAM_CASE(6, 2)
        VOF_STREAM_CASE(1, 3) VOF_STREAM_CASE(2, 3) VOF_STREAM_CASE(3, 3) VOF_STREAM_CASE(4, 3)
        VOF_STREAM_CASE(5, 3) VOF_STREAM_CASE(6, 3)
        default: return false;
    }
#undef VOF_STREAM_CASE
}

// ---------------------------------------------------------------------------
// Host launchers.
template <class R>
void launch_predict(const Ens<R>& e, const Geo& g, int fsel, hipStream_t st) {
    const dim3 grid(g.nblk), blk(256);
    if (sizeof(R) == 4 && g.rb == kPredictBandRows) {  // bands of rows per lane (the launch shape was counted for them)
        if (g.kz == 4) {
            hipLaunchKernelGGL((k_predict_band<R, kPredictBandRows, 4>), grid, blk, 0, st, ens_arg(e, g), fsel);
            return;
        }
        if (g.kz == 8) {
            hipLaunchKernelGGL((k_predict_band<R, kPredictBandRows, 8>), grid, blk, 0, st, ens_arg(e, g), fsel);
            return;
        }
    }
    switch (g.kz) {  // chunks of 4, 8 or 16 planes: unrolled plane loop
        case 4: hipLaunchKernelGGL((k_predict<R, 4>), grid, blk, 0, st, ens_arg(e, g), fsel, g.kz); break;
        case 8: hipLaunchKernelGGL((k_predict<R, 8>), grid, blk, 0, st, ens_arg(e, g), fsel, g.kz); break;
        case 16: hipLaunchKernelGGL((k_predict<R, 16>), grid, blk, 0, st, ens_arg(e, g), fsel, g.kz); break;
        default: hipLaunchKernelGGL((k_predict<R, 0>), grid, blk, 0, st, ens_arg(e, g), fsel, g.kz);
    }
}
template <class R>
void launch_predict_rho(const Ens<R>& e, const Geo& g, int fsel, hipStream_t st) {
    hipLaunchKernelGGL(k_predict_rho<R>, dim3(g.nblk), dim3(256), 0, st, ens_arg(e, g), fsel);
}
template <class R>
void launch_rhs(const Ens<R>& e, const Geo& g, int fsel, hipStream_t st) {
    const dim3 grid(g.nblk), blk(256);
    switch (g.kz) {  // chunks of 4 or 8 planes: operands preloaded
        case 4: hipLaunchKernelGGL((k_rhs_blk<R, 4>), grid, blk, 0, st, ens_arg(e, g), fsel); break;
        case 8: hipLaunchKernelGGL((k_rhs_blk<R, 8>), grid, blk, 0, st, ens_arg(e, g), fsel); break;
        default: hipLaunchKernelGGL(k_rhs<R>, grid, blk, 0, st, ens_arg(e, g), fsel, g.kz);
    }
}
template <class R>
void launch_jacobi(const Ens<R>& e, const Geo& g, int psrc, hipStream_t st) {
    const dim3 grid(g.nblk), blk(256);
    switch (g.kz) {  // chunks of 4, 8 or 16 planes: operands preloaded
        case 4: hipLaunchKernelGGL((k_jacobi_blk<R, 4>), grid, blk, 0, st, ens_arg(e, g), psrc); break;
        case 8: hipLaunchKernelGGL((k_jacobi_blk<R, 8>), grid, blk, 0, st, ens_arg(e, g), psrc); break;
        case 16: hipLaunchKernelGGL((k_jacobi_blk<R, 16>), grid, blk, 0, st, ens_arg(e, g), psrc); break;
        default: hipLaunchKernelGGL(k_jacobi<R>, grid, blk, 0, st, ens_arg(e, g), psrc, g.kz);
    }
}
template <class R>
void launch_jacobi2(const Ens<R>& e, const Geo& g, int psrc, hipStream_t st) {
    const dim3 grid(g.nblk), blk(1024);
    // FP32 chunks of 4, 8 or 16 planes: operands preloaded.  FP64 streams the
    // planes: its preloaded windows need 76-128 VGPRs, and at occupancy 4-6
    // the streaming kernel is faster (26.5 against 40 us per sweep at 200^3).
    if constexpr (sizeof(R) == 4) {
        switch (g.kz) {
            case 4: hipLaunchKernelGGL((k_jacobi2_blk<R, 4>), grid, blk, 0, st, ens_arg(e, g), psrc); return;
            case 8: hipLaunchKernelGGL((k_jacobi2_blk<R, 8>), grid, blk, 0, st, ens_arg(e, g), psrc); return;
            case 16: hipLaunchKernelGGL((k_jacobi2_blk<R, 16>), grid, blk, 0, st, ens_arg(e, g), psrc); return;
            default: break;
        }
    }
    // One plane of prefetch: a deeper ring (2 planes in FP64, 4 in FP32, tried
    // at 200^3) costs the registers of the second work-group per CU and ran
    // 20-35% slower.
    if constexpr (sizeof(R) == 4) {
        if (jacobi2_streams_lds(g.kz, g.min_nz, 4)) {  // streamed through LDS by the buffer unit
            if (g.kz == kJ2DmaKz) hipLaunchKernelGGL((k_jacobi2_dma<kJ2DmaKz>), grid, blk, 0, st, ens_arg(e, g), psrc);
            else hipLaunchKernelGGL((k_jacobi2_dma<kJ2DmaKzShort>), grid, blk, 0, st, ens_arg(e, g), psrc);
            return;
        }
    }
    hipLaunchKernelGGL((k_jacobi2<R, 1>), grid, blk, 0, st, ens_arg(e, g), psrc, g.kz);
}
template <class R>
bool launch_jacobiT(const Ens<R>& e, const Geo& g, int sweeps, int psrc, hipStream_t st) {
    const dim3 grid(g.nblk), blk(1024);
    switch (sweeps) {
        case 3: hipLaunchKernelGGL((k_jacobiT<R, 3>), grid, blk, 0, st, ens_arg(e, g), psrc, g.kz); return true;
        case 4: hipLaunchKernelGGL((k_jacobiT<R, 4>), grid, blk, 0, st, ens_arg(e, g), psrc, g.kz); return true;
        case 5: hipLaunchKernelGGL((k_jacobiT<R, 5>), grid, blk, 0, st, ens_arg(e, g), psrc, g.kz); return true;
        default: return false;
    }
}
template <class R>
void launch_correct(const Ens<R>& e, const Geo& g, int fsel, int psel, hipStream_t st) {
    const dim3 grid(g.nblk), blk(256);
    switch (g.kz) {  // chunks of 4 or 8 planes: operands preloaded
        case 4: hipLaunchKernelGGL((k_correct_blk<R, 4>), grid, blk, 0, st, ens_arg(e, g), fsel, psel); break;
        case 8: hipLaunchKernelGGL((k_correct_blk<R, 8>), grid, blk, 0, st, ens_arg(e, g), fsel, psel); break;
        default: hipLaunchKernelGGL(k_correct<R>, grid, blk, 0, st, ens_arg(e, g), fsel, psel, g.kz);
    }
}
// Chunks of 8, 16 or 32 cells use the register-blocked kernel; any other
// length the streaming one.
template <class R, int DIR>
void launch_march(const Ens<R>& e, const Geo& g, int src, int fin, hipStream_t st) {
    const dim3 grid(g.nblk), blk(256);
    switch (g.kz) {
        case 8: hipLaunchKernelGGL((k_fct_line<R, DIR, 8>), grid, blk, 0, st, ens_arg(e, g), src, fin); break;
        case 16: hipLaunchKernelGGL((k_fct_line<R, DIR, 16>), grid, blk, 0, st, ens_arg(e, g), src, fin); break;
        case 32: hipLaunchKernelGGL((k_fct_line<R, DIR, 32>), grid, blk, 0, st, ens_arg(e, g), src, fin); break;
        default:
            hipLaunchKernelGGL((k_fct_march<R, DIR>), grid, blk, 0, st, ens_arg(e, g), src, g.kz, fin);
    }
}
template <class R>
void launch_fct(const Ens<R>& e, const Geo& g, int dir, int src, int fin, hipStream_t st) {
    if (dir == 0) {
        const dim3 grid(g.nblk), blk(256);
        switch (g.kz) {  // chunks of 4, 8 or 16 planes: F (and u) preloaded
            case 4: hipLaunchKernelGGL((k_fct_xb<R, 4>), grid, blk, 0, st, ens_arg(e, g), src, fin); break;
            case 8: hipLaunchKernelGGL((k_fct_xb<R, 8>), grid, blk, 0, st, ens_arg(e, g), src, fin); break;
            case 16: hipLaunchKernelGGL((k_fct_xb<R, 16>), grid, blk, 0, st, ens_arg(e, g), src, fin); break;
            default: hipLaunchKernelGGL(k_fct_x<R>, grid, blk, 0, st, ens_arg(e, g), src, g.kz, fin);
        }
    } else if (dir == 1)
        launch_march<R, 1>(e, g, src, fin, st);
    else
        launch_march<R, 2>(e, g, src, fin, st);
}
// x sweep followed by the y sweep (k_fct_xy): reads F[src], writes F[src^1].
template <class R>
bool launch_fct_xy(const Ens<R>& e, const Geo& g, int src, hipStream_t st) {
    const dim3 grid(g.nblk), blk(256);
    switch (g.kz) {
        case 4: hipLaunchKernelGGL((k_fct_xy<R, 4>), grid, blk, 0, st, ens_arg(e, g), src); return true;
        case 8: hipLaunchKernelGGL((k_fct_xy<R, 8>), grid, blk, 0, st, ens_arg(e, g), src); return true;
        case 16: hipLaunchKernelGGL((k_fct_xy<R, 16>), grid, blk, 0, st, ens_arg(e, g), src); return true;
        default: return false;
    }
}
// z sweep followed by the x sweep (k_fct_zx): reads F[src], writes F[src^1].
template <class R>
bool launch_fct_zx(const Ens<R>& e, const Geo& g, int src, int fin, hipStream_t st) {
    const dim3 grid(g.nblk), blk(256);
    switch (g.kz) {
        case 4: hipLaunchKernelGGL((k_fct_zx<R, 4>), grid, blk, 0, st, ens_arg(e, g), src, fin); return true;
        case 8: hipLaunchKernelGGL((k_fct_zx<R, 8>), grid, blk, 0, st, ens_arg(e, g), src, fin); return true;
        case 16: hipLaunchKernelGGL((k_fct_zx<R, 16>), grid, blk, 0, st, ens_arg(e, g), src, fin); return true;
        default: return false;
    }
}
template <class R>
void launch_clamp_F(const Ens<R>& e, const Geo& g, int fsel, hipStream_t st) {
    hipLaunchKernelGGL(k_clamp_F<R>, dim3(g.nblk), dim3(256), 0, st, ens_arg(e, g), fsel, g.kz);
}
template <class R>
void launch_ghost_F(const Ens<R>& e, const Geo& g, int fsel, hipStream_t st) {
    hipLaunchKernelGGL(k_ghost_F<R>, dim3(g.nblk), dim3(256), 0, st, ens_arg(e, g), fsel);
}
template <class R>
void launch_csf(const Ens<R>& e, const Geo& g, int fsel, hipStream_t st) {
    hipLaunchKernelGGL(k_csf_smooth<R>, dim3(g.nblk), dim3(256), 0, st, ens_arg(e, g), fsel, g.kz);
    hipLaunchKernelGGL(k_curvature<R>, dim3(g.nblk), dim3(256), 0, st, ens_arg(e, g), g.kz);
    hipLaunchKernelGGL(k_csf<R>, dim3(g.nblk), dim3(256), 0, st, ens_arg(e, g), fsel, g.kz);
}
template <class R>
void launch_sub_mean(const Ens<R>& e, const Geo& g, int psel, const double* means, hipStream_t st) {
    hipLaunchKernelGGL(k_sub_mean<R>, dim3(g.nblk), dim3(256), 0, st, ens_arg(e, g), psel, means, g.kz);
}
template <class R>
void launch_reduce(const Ens<R>& e, const Geo& g, int fsel, int psel, int mode, double* part,
                   const long long* ncell, double* out, double* means, hipStream_t st) {
    hipLaunchKernelGGL(k_reduce1<R>, dim3(g.nblk), dim3(256), 0, st, ens_arg(e, g), fsel, psel, g.kz, mode,
                       part);
    hipLaunchKernelGGL(k_reduce2, dim3(e.ninst), dim3(256), 0, st, g.boff, (const double*)part, ncell, out, means);
}

#define VOF_INSTANTIATE(R)                                                                               \
    template void launch_predict<R>(const Ens<R>&, const Geo&, int, hipStream_t);                        \
    template void launch_predict_rho<R>(const Ens<R>&, const Geo&, int, hipStream_t);                    \
    template void launch_rhs<R>(const Ens<R>&, const Geo&, int, hipStream_t);                            \
    template void launch_jacobi<R>(const Ens<R>&, const Geo&, int, hipStream_t);                         \
    template void launch_jacobi2<R>(const Ens<R>&, const Geo&, int, hipStream_t);                        \
    template bool launch_jacobiT<R>(const Ens<R>&, const Geo&, int, int, hipStream_t);                   \
    template void launch_correct<R>(const Ens<R>&, const Geo&, int, int, hipStream_t);                   \
    template void launch_fct<R>(const Ens<R>&, const Geo&, int, int, int, hipStream_t);                  \
    template bool launch_fct_zx<R>(const Ens<R>&, const Geo&, int, int, hipStream_t);                    \
    template bool launch_fct_xy<R>(const Ens<R>&, const Geo&, int, hipStream_t);                         \
    template void launch_clamp_F<R>(const Ens<R>&, const Geo&, int, hipStream_t);                        \
    template void launch_ghost_F<R>(const Ens<R>&, const Geo&, int, hipStream_t);                        \
    template void launch_sub_mean<R>(const Ens<R>&, const Geo&, int, const double*, hipStream_t);        \
    template void launch_csf<R>(const Ens<R>&, const Geo&, int, hipStream_t);                            \
    template void launch_reduce<R>(const Ens<R>&, const Geo&, int, int, int, double*, const long long*, \
                                   double*, double*, hipStream_t);                                     \
    template hipError_t launch_diag_noalloc<R>(const Inst<R>&, int, double*, vof_diag*, hipStream_t);

VOF_INSTANTIATE(float)
VOF_INSTANTIATE(double)

}  // namespace vof
