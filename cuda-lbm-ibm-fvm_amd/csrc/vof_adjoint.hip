// vof_adjoint.hip — hand-written reverse-mode kernels of the VOF step.
//
// Each forward stage has a kernel that maps the adjoint of its outputs to the
// adjoint of its inputs (a vector-Jacobian product).  Adjoints are FP64; the
// stage is linearised in FP64 at its forward inputs (the fields of the step
// being reversed, FP32 or FP64, recomputed by the forward kernels), i.e. the
// kernels differentiate the FP64 form of each stage along the trajectory the
// solver actually took.  In FP64 mode that is exactly the derivative of the
// computed step.  The switches of the scheme (upwind side, limiter min/max,
// clips) are differentiated along the branch taken; where a min/max has two
// equal operands the derivative is split evenly between them, which is what a
// central difference sees at such a kink (clamp01 at exactly 0 or 1: 1/2).
//
// Layout: every adjoint array is one field, (nx+2)(ny+2)(nz+2) doubles, x
// fastest.  Column kernels use a 2-D grid: x covers the interior (x,y) plane
// (or the stored plane), y the z index.
#include "vof_adjoint.h"

namespace vof {

namespace {

__device__ __forceinline__ double tie_w(bool win, bool tie) { return win ? 1.0 : (tie ? 0.5 : 0.0); }

__device__ __forceinline__ double clamp01d(double v) { return ((1.0 + v) - fmax(1.0, v)) - fmin(0.0, v); }

// d clamp01(v)/dv: 1 inside (0,1), 1/2 at exactly 0 or 1, 0 outside.
__device__ __forceinline__ double dclamp(double v) {
    if (v > 0.0 && v < 1.0) return 1.0;
    if (v == 0.0 || v == 1.0) return 0.5;
    return 0.0;
}

__device__ __forceinline__ void aadd(double* p, double x) {
    if (x != 0.0) atomicAdd(p, x);
}

template <class R>
__device__ __forceinline__ double rho_d(const Inst<R>& I, R F) {
    const double f = clamp01d(double(F));
    return double(I.rho_g) * (1.0 - f) + double(I.rho_l) * f;
}
template <class R>
__device__ __forceinline__ double nu_d(const Inst<R>& I, R F) {
    const double f = clamp01d(double(F));
    return double(I.nu_l) * f + double(I.nu_g) * (1.0 - f);
}

// Interior cell of a 2-D launch (x: interior plane, y: k-1); false past the end.
template <class R>
__device__ __forceinline__ bool interior_cell(const Inst<R>& I, int& i, int& j, int& k, int& c) {
    const int t = int(blockIdx.x) * 256 + int(threadIdx.x);
    if (t >= I.nx * I.ny) return false;
    j = 1 + t / I.nx;
    i = 1 + (t - (j - 1) * I.nx);
    k = 1 + int(blockIdx.y);
    c = k * I.sy + j * I.sx + i;
    return true;
}

// Stored cell of a 2-D launch (x: stored plane, y: k).
template <class R>
__device__ __forceinline__ bool stored_cell(const Inst<R>& I, int& i, int& j, int& k, int& c) {
    const int t = int(blockIdx.x) * 256 + int(threadIdx.x);
    if (t >= I.sy) return false;
    j = t / I.sx;
    i = t - j * I.sx;
    k = int(blockIdx.y);
    c = k * I.sy + t;
    return true;
}

// Sum NQ per-lane values over the 256 lanes of a work-group into
// part[blk*NQ + q] (blk = linear work-group index), or add them to it when
// ACC.  Every lane must call it.
template <int NQ, bool ACC = false>
__device__ __forceinline__ void block_parts(const double (&v)[NQ], double* part) {
    __shared__ double sh[NQ][256];
    for (int q = 0; q < NQ; ++q) sh[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int q = 0; q < NQ; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const size_t blk = size_t(blockIdx.y) * gridDim.x + blockIdx.x;
        for (int q = 0; q < NQ; ++q) {
            if (ACC) part[blk * NQ + q] += sh[q][0];
            else part[blk * NQ + q] = sh[q][0];
        }
    }
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
    int i, j, k, c;
    double v[2] = {0.0, 0.0};
    if (interior_cell(I, i, j, k, c)) {
        const double f = double(F[c]);
        v[0] = f;
        v[1] = f * ((double(k) - 0.5) * double(I.dz));
    }
    block_parts<2>(v, part);
}

// Seed of the final state's F: aF = gJ * dJ/dF + (seed folded through the
// ghost fill).  The stored ghosts of F copy the nearest interior cell, so the
// adjoint of a ghost entry is added to that cell; ghost entries of aF are 0.
template <class R>
__global__ __launch_bounds__(256) void k_centroid_adj(const Inst<R>* __restrict__ tab, double J, double sumF,
                                                      double gJ, const double* __restrict__ seed,
                                                      double* __restrict__ aF) {
    const Inst<R>& I = tab[0];
    int i, j, k, c;
    if (!stored_cell(I, i, j, k, c)) return;
    const bool in = i >= 1 && i <= I.nx && j >= 1 && j <= I.ny && k >= 1 && k <= I.nz;
    if (!in) { aF[c] = 0.0; return; }
    double g = gJ * (((double(k) - 0.5) * double(I.dz)) - J) / sumF;
    if (seed) {
        int ii[3], jj[3], kk[3], ni = 0, nj = 0, nk = 0;
        ii[ni++] = i; if (i == 1) ii[ni++] = 0; if (i == I.nx) ii[ni++] = I.nx + 1;
        jj[nj++] = j; if (j == 1) jj[nj++] = 0; if (j == I.ny) jj[nj++] = I.ny + 1;
        kk[nk++] = k; if (k == 1) kk[nk++] = 0; if (k == I.nz) kk[nk++] = I.nz + 1;
        for (int a = 0; a < ni; ++a)
            for (int b = 0; b < nj; ++b)
                for (int d = 0; d < nk; ++d) g += seed[kk[d] * I.sy + jj[b] * I.sx + ii[a]];
    }
    aF[c] = g;
}

// ---------------------------------------------------------------------------
// FCT sweep reverse as a march along the line.
//
// Forward quantities of a line (positions 0..n+1; face f lies between cells
// f-1 and f; everything outside faces 1..n+1 / cells 1..n is zero, as in the
// forward sweep):
//   face f   L = vdt F(donor), H = vdt F(downwind), a = H - L     <- F(f-1), F(f), V(f)
//   cell c   dv = vol - kdv (V(c+1) - V(c)),
//            y1 = (F(c) + s1(L(c) - L(c+1))) vol / dv,  Ftd = clip(y1)
//   cell c   rp, rm (Zalesak ratios)                            <- Ftd(c-1..c+1), a(c), a(c+1)
//   face f   C = a(f) >= 0 ? min(rp(f), rm(f-1)) : min(rp(f-1), rm(f))   (f >= 2)
//   cell c   yv = Ftd - ((aC)(c+1) - (aC)(c)) / s4 * vol / dv,  Fn = clamp01(yv)
// For the y and z sweeps a lane walks a chunk of its line (lanes along x, so
// every load is coalesced) with the front at position j.  It forms the
// forward at j (face), j-1 (cell), j-2 (ratios, limiter) and j-3 (yv), and the
// reverse follows: yv and the aC face adjoint at j-3, the ratios at j-4, Ftd,
// dv and the face fluxes at j-5, F at j-6.  Every quantity lives in a register
// window, X<d> = value at position j-d.  A chunk of cells m0..m1 starts the
// walk at m0-6 so that every adjoint it writes is complete.  (The x sweep,
// whose lines run along the contiguous axis, is k_fct_adj_x.)
//
// aFo (adjoint of the sweep output, interior cells) -> aFi (adjoint of the
// sweep input, interior cells), aV += face-velocity adjoints.  Ghost entries
// are neither read nor written: a ghost F only enters a flux through a wall
// face, whose velocity is zero, so its adjoint is zero.  fin: the output went
// through clamp01 a second time (last sweep of a step).
//
// The chunk length CH is a compile-time constant: F and V on the CH+13
// positions the walk reads and aFo on its CH+12 output positions are loaded
// into registers before the walk, so that every load of a wave is in flight at
// once, and the walk is unrolled over the register windows.
template <class R, int DIR, int CH>
__global__ __launch_bounds__(256) void k_fct_adj(const Inst<R>* __restrict__ tab, const R* __restrict__ F,
                                                 const R* __restrict__ V, const double* __restrict__ aFo,
                                                 double* __restrict__ aFi, double* __restrict__ aV, int fin) {
    static_assert(DIR == 1 || DIR == 2, "marches serve the y and z sweeps");
    constexpr int ch = CH;
    const Inst<R>& I = tab[0];
    const int n = (DIR == 1) ? I.ny : I.nz;
    const int na = I.nx;  // lanes along x
    const int nb = (DIR == 2) ? I.ny : I.nz;
    const int nlines = na * nb;
    const int nbt = (nlines + 255) >> 8;
    const int chn = int(blockIdx.x) / nbt;
    const int t = (int(blockIdx.x) - chn * nbt) * 256 + int(threadIdx.x);
    if (t >= nlines) return;
    const int m0 = 1 + chn * ch;
    if (m0 > n) return;
    const int m1 = min(n, m0 + ch - 1);
    const int tb = 1 + t / na, ta = 1 + (t - (tb - 1) * na);
    // line of x = ta at z = tb (y sweep) or y = tb (z sweep)
    const int base = (DIR == 1) ? tb * I.sy + ta : tb * I.sx + ta;
    const int S = (DIR == 1) ? I.sx : I.sy;
    const double dt = I.dt, vol = I.vol, dx = I.dx, dy = I.dy, dz = I.dz;
    const double kdv = (DIR == 1) ? double(I.dtdxdz) : double(I.dtdxdy);
    const double qf = (DIR == 2) ? dz : dx;
    const double s4 = (DIR == 2) ? dz : dy;
    const double dxdy = I.dxdy;
    // stage-1 scaling with the forward's association, and its derivative
    auto s1 = [&](double d) {
        if (DIR == 1) return d * dy / dxdy;
        return d * dy * dx / vol;
    };
    const double ds1 = (DIR == 1) ? dy / dxdy : dy * dx / vol;
    const int j0 = m0 - 6, j1 = m1 + 6;
    // F, V at positions j0-1+e (zero outside 0..n+1); aFo at cell j0-3+e (zero
    // outside 1..n).  Loads use positions clamped into range.
    R fw[CH + 13], vw[CH + 13];
    double gw[CH + 12];
#pragma unroll
    for (int e = 0; e < CH + 13; ++e) {
        const int mm = j0 - 1 + e;
        const int mc = min(max(mm, 0), n + 1);
        const R fv = F[base + mc * S], vv = V[base + mc * S];
        const bool in = (mm >= 0 && mm <= n + 1);
        fw[e] = in ? fv : R(0);
        vw[e] = in ? vv : R(0);
    }
#pragma unroll
    for (int e = 0; e < CH + 12; ++e) {
        const int c = j0 - 3 + e;
        const double g = aFo[base + min(max(c, 1), n) * S];
        gw[e] = (c >= 1 && c <= n) ? g : 0.0;
    }
    // forward windows
    double F0 = double(fw[0]), F1 = 0, F2 = 0, F3 = 0, F4 = 0, F5 = 0, F6 = 0;
    double V0 = double(vw[0]), V1 = 0, V2 = 0, V3 = 0, V4 = 0, V5 = 0;
    double L0 = 0, L1 = 0;
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
    double dv1 = 1, dv2 = 1, dv3 = 1, dv4 = 1, dv5 = 1;
    double Y1 = 0, Y2 = 0, Y3 = 0, Y4 = 0, Y5 = 0;
    double T1 = 0, T2 = 0, T3 = 0, T4 = 0, T5 = 0;
    double P2 = 0, P3 = 0, P4 = 0, M2 = 0, M3 = 0, M4 = 0;
    double C2 = 0, C3 = 0;
    // adjoint windows
    double GA2 = 0, GA3 = 0;            // aC on faces
    double GP3 = 0, GP4 = 0, GM3 = 0, GM4 = 0;  // rp, rm
    double GT3 = 0, GT4 = 0, GT5 = 0;   // Ftd
    double Ga3 = 0, Ga4 = 0, Ga5 = 0;   // a
    double GD3 = 0, GD4 = 0, GD5 = 0;   // dv
    double GL4 = 0, GL5 = 0;            // L (stage-1 part)
    double GF5 = 0, GF6 = 0;            // F
    double GV4 = 0, GV5 = 0;            // V
#pragma unroll
    for (int e = 0; e < CH + 12; ++e) {
        const int j = j0 + e;
        if (j > j1) continue;  // short last chunk (wave-uniform)
        // ---- advance the windows by one position
        F6 = F5; F5 = F4; F4 = F3; F3 = F2; F2 = F1; F1 = F0; F0 = double(fw[e + 1]);
        V5 = V4; V4 = V3; V3 = V2; V2 = V1; V1 = V0; V0 = double(vw[e + 1]);
        L1 = L0;
        a4 = a3; a3 = a2; a2 = a1; a1 = a0;
        dv5 = dv4; dv4 = dv3; dv3 = dv2; dv2 = dv1;
        Y5 = Y4; Y4 = Y3; Y3 = Y2; Y2 = Y1;
        T5 = T4; T4 = T3; T3 = T2; T2 = T1;
        P4 = P3; P3 = P2; M4 = M3; M3 = M2;
        C3 = C2;
        GA3 = GA2; GA2 = 0;
        GP4 = GP3; GP3 = 0; GM4 = GM3; GM3 = 0;
        GT5 = GT4; GT4 = GT3; GT3 = 0;
        Ga5 = Ga4; Ga4 = Ga3; Ga3 = 0;
        GD5 = GD4; GD4 = GD3; GD3 = 0;
        GL5 = GL4; GL4 = 0;
        GF6 = GF5; GF5 = 0;
        GV5 = GV4; GV4 = 0;
        // ---- forward: face j
        L0 = 0; a0 = 0;
        if (j >= 1 && j <= n + 1) {
            const double vdt = V0 * dt;
            const double Lf = (V0 >= 0.0) ? vdt * F1 : vdt * F0;
            const double Hf = (V0 <= 0.0) ? vdt * F1 : vdt * F0;
            L0 = Lf;
            a0 = Hf - Lf;
        }
        // cell j-1: stage 1
        dv1 = 1; Y1 = 0; T1 = 0;
        if (j - 1 >= 1 && j - 1 <= n) {
            dv1 = vol - kdv * (V0 - V1);
            Y1 = (F1 + s1(L1 - L0)) * vol / dv1;
            T1 = (Y1 > 1.0 || Y1 < 0.0) ? clamp01d(Y1) : Y1;
        }
        // cell j-2: Zalesak ratios
        P2 = 0; M2 = 0;
        if (j - 2 >= 1 && j - 2 <= n) {
            const double fmx = fmax(T2, fmax(T3, T1)), fmn = fmin(T2, fmin(T3, T1));
            const double pp = fmax(0.0, a2) - fmin(0.0, a1);
            const double pm = fmax(0.0, a1) - fmin(0.0, a2);
            if (pp > 0.0) P2 = fmin(1.0, (fmx - T2) * qf / pp);
            if (pm > 0.0) M2 = fmin(1.0, (T2 - fmn) * qf / pm);
        }
        // face j-2: limiter
        C2 = 0;
        if (j - 2 >= 2 && j - 2 <= n + 1) C2 = (a2 >= 0.0) ? fmin(P2, M3) : fmin(P3, M2);
        // ---- reverse, cell j-3: yv -> Ftd, aC(j-2), aC(j-3), dv
        if (j - 3 >= 1 && j - 3 <= n) {
            const double num = (a2 * C2 - a3 * C3) / s4;
            const double yv = T3 - num * vol / dv3;
            double g = gw[e] * dclamp(yv);
            if (fin) g *= dclamp(clamp01d(yv));
            if (g != 0.0) {
                const double k = vol / (s4 * dv3);
                GT3 += g;
                GA2 -= g * k;
                GA3 += g * k;
                GD3 += g * num * vol / (dv3 * dv3);
            }
        }
        // face j-3: aC = a C -> a, limiter operands
        Ga3 += GA3 * C3;
        if (j - 3 >= 2 && j - 3 <= n + 1 && GA3 != 0.0) {
            const double gC = GA3 * a3;
            if (a3 >= 0.0) {  // C = min(rp(f), rm(f-1))
                GP3 += gC * tie_w(P3 < M4, P3 == M4);
                GM4 += gC * tie_w(M4 < P3, P3 == M4);
            } else {          // C = min(rp(f-1), rm(f))
                GP4 += gC * tie_w(P4 < M3, P4 == M3);
                GM3 += gC * tie_w(M3 < P4, P4 == M3);
            }
        }
        // cell j-4: ratios -> Ftd(j-5..j-3), a(j-4), a(j-3)
        if (j - 4 >= 1 && j - 4 <= n) {
            const double fc = T4, fl = T5, fr = T3, alo = a4, ahi = a3;
            for (int s = 0; s < 2; ++s) {
                const double gr = s ? GM4 : GP4;
                if (gr == 0.0) continue;
                const double x = s ? ahi : alo, y = s ? alo : ahi;  // P = max(0,x) - min(0,y)
                const double P = fmax(0.0, x) - fmin(0.0, y);
                if (!(P > 0.0)) continue;
                const double m1v = s ? fmin(fl, fr) : fmax(fl, fr);
                const double ext = s ? fmin(fc, m1v) : fmax(fc, m1v);
                const double Q = (s ? (fc - ext) : (ext - fc)) * qf;
                const double ratio = Q / P;
                const double wr = tie_w(ratio < 1.0, ratio == 1.0);
                if (wr == 0.0) continue;
                const double gratio = gr * wr;
                const double gQ = gratio / P;
                const double gP = -gratio * Q / (P * P);
                const double gext = (s ? -gQ : gQ) * qf;
                GT4 += (s ? gQ : -gQ) * qf;
                const double wc = tie_w(s ? (fc < m1v) : (fc > m1v), fc == m1v);
                const double wl = tie_w(s ? (fl < fr) : (fl > fr), fl == fr);
                GT4 += gext * wc;
                GT5 += gext * (1.0 - wc) * wl;
                GT3 += gext * (1.0 - wc) * (1.0 - wl);
                const double gx = gP * tie_w(x > 0.0, x == 0.0), gy = -gP * tie_w(y < 0.0, y == 0.0);
                if (s) { Ga3 += gx; Ga4 += gy; } else { Ga4 += gx; Ga3 += gy; }
            }
        }
        // cell j-5: stage 1 (with its clip) -> F, L(j-5), L(j-4), dv; dv -> V
        if (j - 5 >= 1 && j - 5 <= n) {
            const double o = Y5;
            const double go = GT5 * ((o > 1.0 || o < 0.0) ? dclamp(o) : 1.0);
            if (go != 0.0) {
                const double k = vol / dv5;
                GF5 += go * k;
                GL5 += go * k * ds1;
                GL4 -= go * k * ds1;
                GD5 -= go * o / dv5;
            }
            GV4 -= GD5 * kdv;
            GV5 += GD5 * kdv;
        }
        // face j-5: L, H -> F(j-6), F(j-5), V(j-5)
        if (j - 5 >= 1 && j - 5 <= n + 1) {
            const double gL = GL5 - Ga5, gH = Ga5;
            const double vdt = V5 * dt;
            const bool loL = V5 >= 0.0, loH = V5 <= 0.0;
            if (loL) GF6 += gL * vdt; else GF5 += gL * vdt;
            if (loH) GF6 += gH * vdt; else GF5 += gH * vdt;
            GV5 += gL * dt * (loL ? F6 : F5) + gH * dt * (loH ? F6 : F5);
            const int f = j - 5;
            if ((f >= m0 && f <= m1) || (f == n + 1 && m1 == n)) aV[base + f * S] += GV5;
        }
        // cell j-6: complete
        if (j - 6 >= m0 && j - 6 <= m1) aFi[base + (j - 6) * S] = GF6;
    }
}

// x sweep reverse with lanes along x (coalesced): overlapped waves over the
// flattened (x,y) plane, every lane one position (cell p and face p), each
// stage of the forward and of the reverse formed per lane with the
// neighbours' values taken by whole-wave lane shifts.  Along any chain from an
// input to an output there are at most six shifts each way, so a wave stores
// its middle 64 - 2*6 lanes.  Quantities of invalid positions (outside cells
// 1..nx, faces 1..nx+1, ghost rows) are zero, which also keeps consecutive rows
// of the flattened plane apart.  Each wave walks kz planes.
constexpr int kAdjXHalo = 6;

template <class R>
__global__ __launch_bounds__(256) void k_fct_adj_x(const Inst<R>* __restrict__ tab, const R* __restrict__ F,
                                                   const R* __restrict__ V, const double* __restrict__ aFo,
                                                   double* __restrict__ aFi, double* __restrict__ aV, int kz,
                                                   int fin) {
    const Inst<R>& I = tab[0];
    constexpr int H = kAdjXHalo, S = 64 - 2 * H;
    const int nx = I.nx, ny = I.ny, sx = I.sx, sy = I.sy;
    const int P = sx * (ny + 2);
    const int wpp = (P + S - 1) / S, bpp = (wpp + 3) >> 2;
    const int zb = int(blockIdx.x) / bpp;
    const int wv = (int(blockIdx.x) - zb * bpp) * 4 + (int(threadIdx.x) >> 6);
    if (wv >= wpp) return;  // wave-uniform
    const int lane = int(threadIdx.x) & 63;
    const int t = wv * S - H + lane;
    const int tc = min(max(t, 0), P - 1);
    const int j = tc / sx, i = tc - j * sx;
    const bool row = (t == tc) && j >= 1 && j <= ny;
    const bool own = lane >= H && lane < 64 - H && t == tc;
    const bool fc = row && i >= 1 && i <= nx + 1;   // face i
    const bool cc = row && i >= 1 && i <= nx;       // cell i
    const bool cf2 = row && i >= 2 && i <= nx + 1;  // face i with a limiter
    const int k0 = 1 + zb * kz, k1 = min(I.nz, k0 + kz - 1);
    const double dt = I.dt, vol = I.vol, kdv = I.dtdydz, qf = I.dx, s4 = I.dy;
    const double dy = I.dy, dz = I.dz;
    const double ds1 = dy * dz / vol;
    auto up = [](double x) { return lane_up(x); };  // value of lane-1
    auto dn = [](double x) { return lane_dn(x); };  // value of lane+1
    for (int k = k0; k <= k1; ++k) {
        const int o = k * sy + tc;
        const double Fv = double(F[o]), Vv = double(V[o]);
        const double aFov = cc ? aFo[o] : 0.0;
        // ---- forward
        const double F1 = up(Fv);
        double L = 0.0, a = 0.0;
        if (fc) {
            const double vdt = Vv * dt;
            const double Lf = (Vv >= 0.0) ? vdt * F1 : vdt * Fv;
            const double Hf = (Vv <= 0.0) ? vdt * F1 : vdt * Fv;
            L = Lf;
            a = Hf - Lf;
        }
        const double Vn = dn(Vv), Ln = dn(L);
        double dv = 1.0, y1 = 0.0, T = 0.0;
        if (cc) {
            dv = vol - kdv * (Vn - Vv);
            y1 = (Fv + (L - Ln) * dy * dz / vol) * vol / dv;
            T = (y1 > 1.0 || y1 < 0.0) ? clamp01d(y1) : y1;
        }
        const double Tl = up(T), Tr = dn(T), an = dn(a);
        double rp = 0.0, rm = 0.0;
        if (cc) {
            const double fmx = fmax(T, fmax(Tl, Tr)), fmn = fmin(T, fmin(Tl, Tr));
            const double pp = fmax(0.0, a) - fmin(0.0, an);
            const double pm = fmax(0.0, an) - fmin(0.0, a);
            if (pp > 0.0) rp = fmin(1.0, (fmx - T) * qf / pp);
            if (pm > 0.0) rm = fmin(1.0, (T - fmn) * qf / pm);
        }
        const double rpl = up(rp), rml = up(rm);
        double C = 0.0;
        if (cf2) C = (a >= 0.0) ? fmin(rp, rml) : fmin(rpl, rm);
        const double aC = a * C, aCn = dn(aC);
        // ---- reverse: yv (cell) -> Ftd, aC faces, dv
        double g = 0.0, gk = 0.0, gdv = 0.0;
        if (cc) {
            const double num = (aCn - aC) / s4;
            const double yv = T - num * vol / dv;
            g = aFov * dclamp(yv);
            if (fin) g *= dclamp(clamp01d(yv));
            gk = g * vol / (s4 * dv);
            gdv = g * num * vol / (dv * dv);
        }
        double gT = g;
        const double gaC = gk - up(gk);  // face i: +gk of cell i, -gk of cell i-1
        // aC = a C (face) -> a, limiter operands
        double ga = gaC * C;
        double rp_own = 0.0, rm_own = 0.0, rp_prev = 0.0, rm_prev = 0.0;
        if (cf2 && gaC != 0.0) {
            const double gC = gaC * a;
            if (a >= 0.0) {
                rp_own = gC * tie_w(rp < rml, rp == rml);
                rm_prev = gC * tie_w(rml < rp, rp == rml);
            } else {
                rp_prev = gC * tie_w(rpl < rm, rpl == rm);
                rm_own = gC * tie_w(rm < rpl, rpl == rm);
            }
        }
        const double grp = rp_own + dn(rp_prev), grm = rm_own + dn(rm_prev);
        // ratios (cell) -> Ftd(i-1..i+1), a(i), a(i+1)
        double tL = 0.0, tR = 0.0, gaLo = 0.0, gaHi = 0.0;
        if (cc) {
            for (int s = 0; s < 2; ++s) {
                const double gr = s ? grm : grp;
                if (gr == 0.0) continue;
                const double x = s ? an : a, y = s ? a : an;  // P = max(0,x) - min(0,y)
                const double Pq = fmax(0.0, x) - fmin(0.0, y);
                if (!(Pq > 0.0)) continue;
                const double m1v = s ? fmin(Tl, Tr) : fmax(Tl, Tr);
                const double ext = s ? fmin(T, m1v) : fmax(T, m1v);
                const double Q = (s ? (T - ext) : (ext - T)) * qf;
                const double ratio = Q / Pq;
                const double wr = tie_w(ratio < 1.0, ratio == 1.0);
                if (wr == 0.0) continue;
                const double gratio = gr * wr;
                const double gQ = gratio / Pq;
                const double gP = -gratio * Q / (Pq * Pq);
                const double gext = (s ? -gQ : gQ) * qf;
                gT += (s ? gQ : -gQ) * qf;
                const double wc = tie_w(s ? (T < m1v) : (T > m1v), T == m1v);
                const double wl = tie_w(s ? (Tl < Tr) : (Tl > Tr), Tl == Tr);
                gT += gext * wc;
                tL += gext * (1.0 - wc) * wl;
                tR += gext * (1.0 - wc) * (1.0 - wl);
                const double gx = gP * tie_w(x > 0.0, x == 0.0), gy = -gP * tie_w(y < 0.0, y == 0.0);
                if (s) { gaHi += gx; gaLo += gy; } else { gaLo += gx; gaHi += gy; }
            }
        }
        gT += dn(tL) + up(tR);
        ga += gaLo + up(gaHi);
        // stage 1 (cell) -> F, L(i), L(i+1), dv; dv -> V(i), V(i+1)
        double gF = 0.0, gLc = 0.0, gLn = 0.0, gVo = 0.0, gVn = 0.0;
        if (cc) {
            const double go = gT * ((y1 > 1.0 || y1 < 0.0) ? dclamp(y1) : 1.0);
            const double kk = vol / dv;
            gF = go * kk;
            gLc = go * kk * ds1;
            gLn = -go * kk * ds1;
            gdv -= go * y1 / dv;
            gVo = gdv * kdv;
            gVn = -gdv * kdv;
        }
        // fluxes (face) -> F(i-1), F(i), V(i)
        const double gL = gLc + up(gLn) - ga, gH = ga;
        double gV = gVo + up(gVn), toPrev = 0.0;
        if (fc) {
            const double vdt = Vv * dt;
            const bool loL = Vv >= 0.0, loH = Vv <= 0.0;
            if (loL) toPrev += gL * vdt; else gF += gL * vdt;
            if (loH) toPrev += gH * vdt; else gF += gH * vdt;
            gV += gL * dt * (loL ? F1 : Fv) + gH * dt * (loH ? F1 : Fv);
        }
        gF += dn(toPrev);
        if (own && cc) aFi[o] = gF;
        if (own && fc) aV[o] += gV;
    }
}

// ---------------------------------------------------------------------------
// Transpose of the velocity correction and its boundary closed form.  The
// stored u(i,j,k) is u_in(i, cj, ck) for 2 <= i <= nx, zero on the walls
// i = 1, nx+1 and u_old(0, cj, ck) on plane 0 (cj, ck: index clamped to the
// interior; likewise v, w).  Per interior cell c, gathered (no atomics):
//   G(face) = sum of the adjoints of the stored copies of that face,
//   au*(c) = G(c) (faces i >= 2), ap(c) += the pressure-gradient terms of
//   faces c and c+1, arho(c) = the density terms of faces c and c+1, and on
//   the first interior plane of each axis the plane-0 copies -> au0.
template <class R>
__global__ __launch_bounds__(256) void k_correct_adj(const Inst<R>* __restrict__ tab, const R* __restrict__ F,
                                                     const R* __restrict__ p, const double* __restrict__ au,
                                                     const double* __restrict__ av, const double* __restrict__ aw,
                                                     double* __restrict__ aus, double* __restrict__ avs,
                                                     double* __restrict__ aws, double* __restrict__ ap,
                                                     double* __restrict__ arho, double* __restrict__ au0,
                                                     double* __restrict__ av0, double* __restrict__ aw0) {
    const Inst<R>& I = tab[0];
    int i, j, k, c;
    if (!interior_cell(I, i, j, k, c)) return;
    const int nx = I.nx, ny = I.ny, nz = I.nz, sx = I.sx, sy = I.sy;
    // index sets of the stored copies along each axis
    int I3[3], J3[3], K3[3], ni = 0, nj = 0, nk = 0;
    I3[ni++] = i; if (i == 1) I3[ni++] = 0; if (i == nx) I3[ni++] = nx + 1;
    J3[nj++] = j; if (j == 1) J3[nj++] = 0; if (j == ny) J3[nj++] = ny + 1;
    K3[nk++] = k; if (k == 1) K3[nk++] = 0; if (k == nz) K3[nk++] = nz + 1;
    auto Gu = [&](int ii) {  // u faces at x index ii, copies over (j, k)
        double s = 0.0;
        for (int b = 0; b < nj; ++b)
            for (int d = 0; d < nk; ++d) s += au[K3[d] * sy + J3[b] * sx + ii];
        return s;
    };
    auto Gv = [&](int jj) {
        double s = 0.0;
        for (int a = 0; a < ni; ++a)
            for (int d = 0; d < nk; ++d) s += av[K3[d] * sy + jj * sx + I3[a]];
        return s;
    };
    auto Gw = [&](int kk) {
        double s = 0.0;
        for (int a = 0; a < ni; ++a)
            for (int b = 0; b < nj; ++b) s += aw[kk * sy + J3[b] * sx + I3[a]];
        return s;
    };
    const double dt = I.dt;
    const double rc = rho_d(I, F[c]), pc = double(p[c]);
    double gp = 0.0, gr = 0.0;
    // face between cell c-off (m) and c (hi side); returns its G
    auto face = [&](double G, int m, bool c_is_hi, double h) {
        const double rm = rho_d(I, F[m]), pm = double(p[m]);
        const double rr = 1.0 / (c_is_hi ? 0.5 * (rc + rm) : 0.5 * (rm + rc));
        const double dp = c_is_hi ? (pc - pm) : (pm - pc);
        const double kk = dt * rr * h;  // d u / d p, and d u / d r = kk dp / r
        gp += c_is_hi ? -G * kk : G * kk;
        gr += 0.5 * G * kk * dp * rr;
    };
    const double hx = I.dxi, hy = I.dyi, hz = I.dzi;
    if (i >= 2) { const double G = Gu(i); aus[c] = G; face(G, c - 1, true, hx); }
    if (i + 1 <= nx) face(Gu(i + 1), c + 1, false, hx);
    if (j >= 2) { const double G = Gv(j); avs[c] = G; face(G, c - sx, true, hy); }
    if (j + 1 <= ny) face(Gv(j + 1), c + sx, false, hy);
    if (k >= 2) { const double G = Gw(k); aws[c] = G; face(G, c - sy, true, hz); }
    if (k + 1 <= nz) face(Gw(k + 1), c + sy, false, hz);
    ap[c] += gp;
    arho[c] = gr;
    if (i == 1) au0[k * sy + j * sx] = Gu(0);
    if (j == 1) av0[k * sy + i] = Gv(0);
    if (k == 1) aw0[j * sx + i] = Gw(0);
}

// Transposed Jacobi sweeps.  A forward sweep is p' = D^-1 (rhs - N p) on
// interior cells, with N the off-diagonal part (coefficient a_d towards every
// interior neighbour, zero towards a ghost) and D = diag(ap), ap = -sum of the
// coefficients (summed in the forward's order).  N is symmetric, so with
// h = D^-1 gbar the reverse of one sweep is
//   rhs_bar += h,   gbar_prev = -N h,   h_prev = D^-1 gbar_prev,
// a zero-rhs Jacobi sweep on h.  A launch does one such sweep; in_g: the input
// is gbar (divided by ap on load), out_g: write gbar_prev instead of h_prev
// (last launch), first: assign rhs_bar instead of adding.  Only interior
// entries are read or written.
template <class R>
__device__ __forceinline__ double jac_ap(const Inst<R>& I, int i, int j, int k) {
    const R ae = (i != I.nx) ? I.axx : R(0), aw = (i != 1) ? I.axx : R(0);
    const R an = (j != I.ny) ? I.ayy : R(0), as = (j != 1) ? I.ayy : R(0);
    const R af = (k != I.nz) ? I.azz : R(0), ab = (k != 1) ? I.azz : R(0);
    return double(-(ae + aw + an + as + af + ab));
}

// The same reverse sweep as a column march.  Lanes cover the flattened
// (x,y) plane in overlapped waves that store their middle 62 lanes; z
// neighbours are carried in registers, x neighbours come from the adjacent
// lanes (the neighbour terms are summed in the order above), and the divisions
// by ap are multiplications by 1/ap, formed once per lane for the three kinds
// of plane (k = 1, interior, k = nz) of the rows j-1, j, j+1.  in_g: every
// loaded value is gbar and becomes h on load.  first: rhs_bar is assigned.
// acc: 0 leaves rhs_bar alone, 1 adds the input h, 2 adds the input and the
// output h (the next launch's input, which then adds nothing), so rhs_bar is
// read and written once per pair of sweeps.  The adjoint is compared with its
// references to a tolerance, not bit for bit, so the reciprocals are allowed
// here.
// ri: 1/ap of a cell whose row and column are interior, for k = 1, interior
// planes and k = nz (formed on the host with the same sums); the other cells
// divide on the device.
template <class R>
__global__ __launch_bounds__(256) void k_jacobi_adj_col(const Inst<R>* __restrict__ tab,
                                                        const double* __restrict__ in, double* __restrict__ out,
                                                        double* __restrict__ arhs, int in_g, int out_g, int first,
                                                        int acc, int kz, double ri_lo, double ri_in, double ri_hi) {
    const Inst<R> I = tab[0];
    constexpr int H = 1, S = 64 - 2 * H;
    const int nx = I.nx, ny = I.ny, nz = I.nz, sx = I.sx, sy = I.sy;
    const int P = sx * (ny + 2);
    const int wpp = (P + S - 1) / S, bpp = (wpp + 3) >> 2;
    const int zb = int(blockIdx.x) / bpp;
    const int wv = (int(blockIdx.x) - zb * bpp) * 4 + (int(threadIdx.x) >> 6);
    if (wv >= wpp) return;  // wave-uniform
    const int lane = int(threadIdx.x) & 63;
    const int t = wv * S - H + lane;
    const int tc = min(max(t, 0), P - 1);
    const int j = tc / sx, i = tc - j * sx;
    const bool cell = lane >= H && lane < 64 - H && t == tc && i >= 1 && i <= nx && j >= 1 && j <= ny;
    const int tjm = max(tc - sx, 0), tjp = min(tc + sx, P - 1);
    const int k0 = 1 + zb * kz, k1 = min(nz, k0 + kz - 1);
    const double ax = I.axx, ay = I.ayy, az = I.azz;
    // 1/ap of rows j-1, j, j+1 (index 0, 1, 2) for k = 1, interior planes, k = nz;
    // the side rows only when their loads are gbar
    double rlo[3] = {ri_lo, ri_lo, ri_lo}, rin[3] = {ri_in, ri_in, ri_in}, rhi[3] = {ri_hi, ri_hi, ri_hi};
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int jr = min(max(j - 1 + r, 0), ny + 1);
        if ((r != 1 && !in_g) || (i >= 2 && i <= nx - 1 && jr >= 2 && jr <= ny - 1)) continue;
        rlo[r] = 1.0 / jac_ap(I, i, jr, 1);
        rhi[r] = 1.0 / jac_ap(I, i, jr, nz);
        rin[r] = (nz > 2) ? 1.0 / jac_ap(I, i, jr, 2) : 0.0;
    }
    auto rk = [&](int r, int k) { return (k == 1) ? rlo[r] : ((k == nz) ? rhi[r] : rin[r]); };
    auto as_h = [&](double x, int r, int k) { return in_g ? x * rk(r, k) : x; };
    const Buf<double> hb(in, I.ntot);
    int o = k0 * sy;
    double hkm = as_h(hb.ld(tc, o - sy), 1, k0 - 1), hk = as_h(hb.ld(tc, o), 1, k0);
    double n_kp = as_h(hb.ld(tc, o + sy), 1, k0 + 1);
    double n_jm = as_h(hb.ld(tjm, o), 0, k0), n_jp = as_h(hb.ld(tjp, o), 2, k0);
    for (int k = k0; k <= k1; ++k, o += sy) {
        const double hkp = n_kp, hjm = n_jm, hjp = n_jp;
        const int k2 = min(k + 1, nz);  // plane whose rows are fetched next
        const int o2 = k2 * sy;
        n_kp = as_h(hb.ld(tc, o2 + sy), 1, k2 + 1);
        n_jm = as_h(hb.ld(tjm, o2), 0, k2);
        n_jp = as_h(hb.ld(tjp, o2), 2, k2);
        const double hw = lane_up(hk), he = lane_dn(hk);
        if (cell) {
            double sm = 0.0;
            if (i < nx) sm -= ax * he;
            if (i > 1) sm -= ax * hw;
            if (j < ny) sm -= ay * hjp;
            if (j > 1) sm -= ay * hjm;
            if (k < nz) sm -= az * hkp;
            if (k > 1) sm -= az * hkm;
            const int c = o + tc;
            const double ho = sm * rk(1, k);
            out[c] = out_g ? sm : ho;
            const double add = (acc == 2) ? hk + ho : hk;
            if (first) arhs[c] = add;
            else if (acc) arhs[c] += add;
        }
        hkm = hk;
        hk = hkp;
    }
}

// Two of those reverse sweeps per launch (temporal blocking, as the forward's
// two-sweep kernels): sweep a takes the input h (gbar divided by ap on load
// when in_g) to h1 = D^-1 (-N h), sweep b takes h1 to the output (-N h1, or
// its D^-1 multiple unless out_g).  rhs_bar takes h + h1 of every cell (the
// two sweep inputs, added in that order, assigned when first), so a pair
// launch leaves out and rhs_bar bit for bit as two single-sweep launches of
// k_jacobi_adj_col (acc 2 then 0) do: every sum keeps that kernel's order.
// Tile and march as the forward's k_jacobi2: a work-group of 16 waves covers
// 16 rows x 64 lanes, lane l at position 60*xtile - 2 + l of the flattened
// plane of (i, row band) positions, wave w at row j = 14*band + w of the
// lane's band; sweep a is formed on every lane of every row and
// shared across rows through a 3-plane LDS ring, sweep b is stored on lanes
// 2..61 of rows 1..14, one plane behind sweep a; z neighbours of both sweeps
// in registers, one plane of loads in flight.  Every wave reaches every
// barrier: the chunk bounds are uniform over the work-group.
template <class R>
__global__ __launch_bounds__(1024) void k_jacobi2_adj(const Inst<R>* __restrict__ tab,
                                                      const double* __restrict__ in, double* __restrict__ out,
                                                      double* __restrict__ arhs, int in_g, int out_g, int first,
                                                      int kz, double ri_lo, double ri_in, double ri_hi) {
    const Inst<R> I = tab[0];
    __shared__ double ring[3][16][64];
    const int nx = I.nx, ny = I.ny, nz = I.nz, sx = I.sx, sy = I.sy;
    // tiles over the flattened plane of (i, row band) positions, as the
    // forward two-sweep kernels: lane l of tile xt holds position 60 xt - 2 + l
    const int nbd = (ny + 13) / 14, npos = sx * nbd;
    const int per_z = (npos + 59) / 60;
    const int b = int(blockIdx.x);
    const int zc = b / per_z, xt = b - zc * per_z;
    const int lane = int(threadIdx.x) & 63, w = int(threadIdx.x) >> 6;
    const int t = min(max(60 * xt - 2 + lane, 0), npos - 1);
    const int bd = t / sx;
    const int i = t - bd * sx, j = 14 * bd + w;
    const int k0 = 1 + zc * kz, k1 = min(nz, k0 + kz - 1);
    const bool in1 = (i >= 1 && i <= nx && j >= 1 && j <= ny);
    const bool st2 = in1 && lane >= 2 && lane <= 61 && w >= 1 && w <= 14;
    const int ic = min(max(i, 0), nx + 1), jc = min(max(j, 0), ny + 1);
    const int tc = jc * sx + ic;
    const int tjm = max(jc - 1, 0) * sx + ic, tjp = min(jc + 1, ny + 1) * sx + ic;
    const double ax = I.axx, ay = I.ayy, az = I.azz;
    // 1/ap of rows j-1, j, j+1 (index 0, 1, 2) for k = 1, interior planes, k = nz,
    // as k_jacobi_adj_col forms them (the side rows matter only when in_g)
    double rlo[3] = {ri_lo, ri_lo, ri_lo}, rin[3] = {ri_in, ri_in, ri_in}, rhi[3] = {ri_hi, ri_hi, ri_hi};
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int jr = min(max(j - 1 + r, 0), ny + 1);
        if ((r != 1 && !in_g) || (ic >= 2 && ic <= nx - 1 && jr >= 2 && jr <= ny - 1)) continue;
        rlo[r] = 1.0 / jac_ap(I, ic, jr, 1);
        rhi[r] = 1.0 / jac_ap(I, ic, jr, nz);
        rin[r] = (nz > 2) ? 1.0 / jac_ap(I, ic, jr, 2) : 0.0;
    }
    auto rk = [&](int r, int k) { return (k == 1) ? rlo[r] : ((k == nz) ? rhi[r] : rin[r]); };
    auto as_h = [&](double x, int r, int k) { return in_g ? x * rk(r, k) : x; };
    const int olast = (nz + 1) * sy;
    auto po = [&](int s) { return min(max(s, 0) * sy, olast); };  // clamped plane offset
    const Buf<double> hb(in, I.ntot);
    // march: sweep a at plane s, sweep b at plane s-1, for s = k0-1 .. k1+1
    const int s_beg = k0 - 1, s_end = k1 + 1;
    double h0m = as_h(hb.ld(tc, po(s_beg - 1)), 1, s_beg - 1), h0 = as_h(hb.ld(tc, po(s_beg)), 1, s_beg);
    double h1mm = 0.0, h1m = 0.0;  // sweep a at s-2, s-1
    double nq = hb.ld(tc, po(s_beg + 1)), njm = hb.ld(tjm, po(s_beg)), njp = hb.ld(tjp, po(s_beg));
    for (int s = s_beg; s <= s_end; ++s) {
        const double h0p = as_h(nq, 1, s + 1), h0jm = as_h(njm, 0, s), h0jp = as_h(njp, 2, s);
        nq = hb.ld(tc, po(s + 2));
        njm = hb.ld(tjm, po(s + 1));
        njp = hb.ld(tjp, po(s + 1));
        // sweep a at plane s
        const double hw = lane_up(h0), he = lane_dn(h0);
        double h1 = 0.0;
        if (in1 && s >= 1 && s <= nz) {
            double sm = 0.0;
            if (i < nx) sm -= ax * he;
            if (i > 1) sm -= ax * hw;
            if (j < ny) sm -= ay * h0jp;
            if (j > 1) sm -= ay * h0jm;
            if (s < nz) sm -= az * h0p;
            if (s > 1) sm -= az * h0m;
            h1 = sm * rk(1, s);
        }
        ring[s % 3][w][lane] = h1;
        __syncthreads();
        // sweep b at plane s-1
        const int s2 = s - 1;
        const double h1w = lane_up(h1m), h1e = lane_dn(h1m);
        if (st2 && s2 >= k0) {
            const int ps = (s + 2) % 3;
            const double h1s = ring[ps][w - 1][lane], h1n = ring[ps][w + 1][lane];
            double sm = 0.0;
            if (i < nx) sm -= ax * h1e;
            if (i > 1) sm -= ax * h1w;
            if (j < ny) sm -= ay * h1n;
            if (j > 1) sm -= ay * h1s;
            if (s2 < nz) sm -= az * h1;
            if (s2 > 1) sm -= az * h1mm;
            const int c = s2 * sy + tc;
            out[c] = out_g ? sm : sm * rk(1, s2);
            const double add = h0m + h1m;  // the two sweep inputs at plane s-1
            if (first) arhs[c] = add;
            else arhs[c] += add;
        }
        // the slot written at step s+3 was last read at step s+1, before the
        // barrier of step s+2 that every wave passes first
        h0m = h0;
        h0 = h0p;
        h1mm = h1m;
        h1m = h1;
    }
}

// Transpose of rhs = rho/dt * div(u*) on interior cells: each interior face
// i >= 2 (j >= 2, k >= 2) gathers from its two cells; arho += arhs div / dt.
template <class R>
__global__ __launch_bounds__(256) void k_rhs_adj(const Inst<R>* __restrict__ tab, const R* __restrict__ F,
                                                 const R* __restrict__ us, const R* __restrict__ vs,
                                                 const R* __restrict__ ws, const double* __restrict__ arhs,
                                                 double* __restrict__ aus, double* __restrict__ avs,
                                                 double* __restrict__ aws, double* __restrict__ arho) {
    const Inst<R>& I = tab[0];
    int i, j, k, c;
    if (!interior_cell(I, i, j, k, c)) return;
    const int sx = I.sx, sy = I.sy;
    const double rdt = 1.0 / double(I.dt);
    const double hx = I.dxi, hy = I.dyi, hz = I.dzi;
    const double gc = arhs[c] * rho_d(I, F[c]) * rdt;
    if (i >= 2) aus[c] += (arhs[c - 1] * rho_d(I, F[c - 1]) * rdt - gc) * hx;
    if (j >= 2) avs[c] += (arhs[c - sx] * rho_d(I, F[c - sx]) * rdt - gc) * hy;
    if (k >= 2) aws[c] += (arhs[c - sy] * rho_d(I, F[c - sy]) * rdt - gc) * hz;
    const double div = (double(us[c + 1]) - double(us[c])) * hx + (double(vs[c + sx]) - double(vs[c])) * hy +
                       (double(ws[c + sy]) - double(ws[c])) * hz;
    arho[c] += arhs[c] * div * rdt;
}

// Partial derivatives of one predictor face value
//   out = c + dt * (visc + adv + g),  visc = nu sum_d (m_d - 2c + p_d) / h_d^2,
//   adv = -sum_d a_d * (a_d > 0 ? (c - m_d) : (p_d - c)) / h_d
// times g_out, accumulated into the operand adjoints.  own: the axis whose
// advecting velocity is the face value c itself.
struct MomAdj {
    double c = 0, xm = 0, xp = 0, ym = 0, yp = 0, zm = 0, zp = 0, ax = 0, ay = 0, az = 0;
};

__device__ __forceinline__ void mom_adj(double hx, double hy, double hz, double dt, double g, double c, double xm,
                                        double xp, double ym, double yp, double zm, double zp, double ax, double ay,
                                        double az, double n, int own, MomAdj& d, double& dnu) {
    const double gd = g * dt;
    dnu += gd * ((xm - 2.0 * c + xp) * hx * hx + (ym - 2.0 * c + yp) * hy * hy + (zm - 2.0 * c + zp) * hz * hz);
    d.c += g - 2.0 * gd * n * (hx * hx + hy * hy + hz * hz);
    d.xm += gd * n * hx * hx; d.xp += gd * n * hx * hx;
    d.ym += gd * n * hy * hy; d.yp += gd * n * hy * hy;
    d.zm += gd * n * hz * hz; d.zp += gd * n * hz * hz;
    double ddx, ddy, ddz;
    if (ax > 0.0) { ddx = (c - xm) * hx; d.c -= gd * ax * hx; d.xm += gd * ax * hx; }
    else { ddx = (xp - c) * hx; d.c += gd * ax * hx; d.xp -= gd * ax * hx; }
    if (ay > 0.0) { ddy = (c - ym) * hy; d.c -= gd * ay * hy; d.ym += gd * ay * hy; }
    else { ddy = (yp - c) * hy; d.c += gd * ay * hy; d.yp -= gd * ay * hy; }
    if (az > 0.0) { ddz = (c - zm) * hz; d.c -= gd * az * hz; d.zm += gd * az * hz; }
    else { ddz = (zp - c) * hz; d.c += gd * az * hz; d.zp -= gd * az * hz; }
    if (own == 0) d.c -= gd * ddx; else d.ax -= gd * ddx;
    if (own == 1) d.c -= gd * ddy; else d.ay -= gd * ddy;
    if (own == 2) d.c -= gd * ddz; else d.az -= gd * ddz;
}

// Transpose of the momentum predictor, fused with that of rho/nu.
//
// A work-group owns a column of PTX x PTY interior cells and walks a chunk of
// kz planes.  Per plane every lane forms the partial derivatives of the three
// faces of its cell (u* on face i >= 2, v* on j >= 2, w* on k >= 2) and adds
// them, times the face adjoints, into an LDS window of (PTX+2) x (PTY+2) cells
// x 3 planes x 3 velocity components.  Once a plane can receive nothing more
// from this work-group it is written out: a cell whose whole stencil lies in
// the column and the chunk gets a plain store, the others (shared with a
// neighbouring column or chunk) a global atomic add onto arrays the host
// zeroed.  The lane then adds to aF(c) the F adjoint of rho(c) and nu(c) (nu(c)
// enters only the three faces of cell c).  Work-group sums for
// dJ/d(rho_l, rho_g, nu_l, nu_g, gx, gy, gz).
constexpr int PTX = 32, PTY = 8;
constexpr int PWX = PTX + 2, PWY = PTY + 2, PWC = PWX * PWY;

__device__ __forceinline__ int pslot(int p) { return p % 3; }

template <class R>
__global__ __launch_bounds__(256) void k_predict_adj(const Inst<R>* __restrict__ tab, const R* __restrict__ F,
                                                     const R* __restrict__ u, const R* __restrict__ v,
                                                     const R* __restrict__ w, const double* __restrict__ aus,
                                                     const double* __restrict__ avs, const double* __restrict__ aws,
                                                     const double* __restrict__ arho, double* __restrict__ au,
                                                     double* __restrict__ av, double* __restrict__ aw,
                                                     double* __restrict__ aF, double* __restrict__ part, int kz) {
    const Inst<R>& I = tab[0];
    __shared__ double win[3][3][PWC];  // [plane slot][component][window cell]
    const int nx = I.nx, ny = I.ny, nz = I.nz, sx = I.sx, sy = I.sy;
    const int nbx = (nx + PTX - 1) / PTX, nby = (ny + PTY - 1) / PTY;
    const int b = int(blockIdx.x);
    const int bz = b / (nbx * nby), bxy = b - bz * nbx * nby;
    const int by = bxy / nbx, bx = bxy - by * nbx;
    const int x0 = 1 + bx * PTX, y0 = 1 + by * PTY, k0 = 1 + bz * kz, k1 = min(nz, k0 + kz - 1);
    const int tid = int(threadIdx.x), tx = tid % PTX, ty = tid / PTX;
    const int i = x0 + tx, j = y0 + ty;
    const bool valid = i <= nx && j <= ny;
    for (int q = tid; q < 9 * PWC; q += 256) (&win[0][0][0])[q] = 0.0;
    __syncthreads();
    const double hx = I.dxi, hy = I.dyi, hz = I.dzi, dt = I.dt;
    double sums[7] = {0, 0, 0, 0, 0, 0, 0};
    double* const outp[3] = {au, av, aw};
    // write plane p of the window (inner: cells with a complete stencil may be stored)
    auto flush = [&](int p, bool inner_plane) {
        if (p < 0 || p > nz + 1) return;
        const int s = pslot(p);
        for (int q = tid; q < 3 * PWC; q += 256) {
            const int comp = q / PWC, wc = q - comp * PWC;
            const int ly = wc / PWX, lx = wc - ly * PWX;
            const int gx = x0 - 1 + lx, gy = y0 - 1 + ly;
            if (gx > nx + 1 || gy > ny + 1) continue;
            const double val = win[s][comp][wc];
            double* dst = outp[comp] + (size_t(p) * sy + size_t(gy) * sx + gx);
            const bool inner = inner_plane && lx >= 2 && lx <= PTX - 1 && ly >= 2 && ly <= PTY - 1;
            if (inner) *dst = val;
            else if (val != 0.0) atomicAdd(dst, val);
        }
    };
    for (int k = k0; k <= k1; ++k) {
        if (valid) {
            const int c = k * sy + j * sx + i;
            const double n0 = nu_d(I, F[c]);
            double gn = 0.0;
            auto U = [&](int q) { return double(u[q]); };
            auto Vv = [&](int q) { return double(v[q]); };
            auto W = [&](int q) { return double(w[q]); };
            // add g to component comp of the cell at offset (di, dj, dk)
            auto put = [&](int comp, int di, int dj, int dk, double g) {
                if (g != 0.0) atomicAdd(&win[pslot(k + dk)][comp][(ty + 1 + dj) * PWX + (tx + 1 + di)], g);
            };
            auto put7 = [&](int comp, const MomAdj& d) {
                put(comp, 0, 0, 0, d.c); put(comp, -1, 0, 0, d.xm); put(comp, 1, 0, 0, d.xp);
                put(comp, 0, -1, 0, d.ym); put(comp, 0, 1, 0, d.yp); put(comp, 0, 0, -1, d.zm);
                put(comp, 0, 0, 1, d.zp);
            };
            if (i >= 2 && aus[c] != 0.0) {  // u*: advecting (u, vh, wh)
                const double vh = 0.25 * (Vv(c - 1) + Vv(c - 1 + sx) + Vv(c) + Vv(c + sx));
                const double wh = 0.25 * (W(c - 1) + W(c - 1 + sy) + W(c) + W(c + sy));
                MomAdj d;
                mom_adj(hx, hy, hz, dt, aus[c], U(c), U(c - 1), U(c + 1), U(c - sx), U(c + sx), U(c - sy),
                        U(c + sy), U(c), vh, wh, n0, 0, d, gn);
                put7(0, d);
                const double qy = 0.25 * d.ay, qz = 0.25 * d.az;  // v at (i-1|i, j|j+1), w at (i-1|i, k|k+1)
                put(1, -1, 0, 0, qy); put(1, -1, 1, 0, qy); put(1, 0, 0, 0, qy); put(1, 0, 1, 0, qy);
                put(2, -1, 0, 0, qz); put(2, -1, 0, 1, qz); put(2, 0, 0, 0, qz); put(2, 0, 0, 1, qz);
                sums[4] += aus[c] * dt;
            }
            if (j >= 2 && avs[c] != 0.0) {  // v*: advecting (uh, v, wh)
                const double uh = 0.25 * (U(c - sx) + U(c) + U(c + 1 - sx) + U(c + 1));
                const double wh = 0.25 * (W(c - sx + sy) + W(c - sx) + W(c) + W(c + sy));
                MomAdj d;
                mom_adj(hx, hy, hz, dt, avs[c], Vv(c), Vv(c - 1), Vv(c + 1), Vv(c - sx), Vv(c + sx), Vv(c - sy),
                        Vv(c + sy), uh, Vv(c), wh, n0, 1, d, gn);
                put7(1, d);
                const double qx = 0.25 * d.ax, qz = 0.25 * d.az;  // u at (i|i+1, j-1|j), w at (j-1|j, k|k+1)
                put(0, 0, -1, 0, qx); put(0, 0, 0, 0, qx); put(0, 1, -1, 0, qx); put(0, 1, 0, 0, qx);
                put(2, 0, -1, 1, qz); put(2, 0, -1, 0, qz); put(2, 0, 0, 0, qz); put(2, 0, 0, 1, qz);
                sums[5] += avs[c] * dt;
            }
            if (k >= 2 && aws[c] != 0.0) {  // w*: advecting (uh, vh, w)
                const double uh = 0.25 * (U(c + 1 - sy) + U(c - sy) + U(c + 1) + U(c));
                const double vh = 0.25 * (Vv(c + sx - sy) + Vv(c - sy) + Vv(c) + Vv(c + sx));
                MomAdj d;
                mom_adj(hx, hy, hz, dt, aws[c], W(c), W(c - 1), W(c + 1), W(c - sx), W(c + sx), W(c - sy),
                        W(c + sy), uh, vh, W(c), n0, 2, d, gn);
                put7(2, d);
                const double qx = 0.25 * d.ax, qy = 0.25 * d.ay;  // u at (i|i+1, k-1|k), v at (j|j+1, k-1|k)
                put(0, 1, 0, -1, qx); put(0, 0, 0, -1, qx); put(0, 1, 0, 0, qx); put(0, 0, 0, 0, qx);
                put(1, 0, 1, -1, qy); put(1, 0, 0, -1, qy); put(1, 0, 0, 0, qy); put(1, 0, 1, 0, qy);
                sums[6] += aws[c] * dt;
            }
            // rho(c), nu(c) <- f = clamp01(F(c))
            const double gr = arho[c];
            const double f = clamp01d(double(F[c]));
            const double gf = gr * (double(I.rho_l) - double(I.rho_g)) + gn * (double(I.nu_l) - double(I.nu_g));
            aF[c] += gf * dclamp(double(F[c]));
            sums[0] += gr * f;
            sums[1] += gr * (1.0 - f);
            sums[2] += gn * f;
            sums[3] += gn * (1.0 - f);
        }
        __syncthreads();
        // plane k-1 is complete for this work-group (faces of planes k-2..k done)
        flush(k - 1, k - 1 >= k0 + 1);
        __syncthreads();
        const int s = pslot(k - 1);  // reused for plane k+2
        for (int q = tid; q < 3 * PWC; q += 256) win[s][q / PWC][q % PWC] = 0.0;
        __syncthreads();
    }
    flush(k1, false);
    flush(k1 + 1, false);
    block_parts<7, true>(sums, part);  // accumulated over the steps; summed once by adj_predict_sums
}

// ---------------------------------------------------------------------------
// CSF surface tension in reverse (VOF_FLAG_CSF).  Forward, per face between
// cells m (low) and q (high): u* += dt a, a = sigma kf gF / rf with
// kf = (kappa(m)+kappa(q))/2, gF = (F(q)-F(m)) h, rf = (rho(m)+rho(q))/2;
// kappa from the unit vertex normals of Fs, Fs the smoothed F.  The step's Fs
// and kappa are first recomputed from its F with the forward's functions.
template <class R>
__global__ __launch_bounds__(256) void k_adj_smooth(const Inst<R>* __restrict__ tab, const R* __restrict__ F,
                                                    R* __restrict__ fsm) {
    const Inst<R>& I = tab[0];
    int i, j, k, c;
    if (interior_cell(I, i, j, k, c)) fsm[c] = csf_smooth(I, F, c);
}
template <class R>
__global__ __launch_bounds__(256) void k_adj_curv(const Inst<R>* __restrict__ tab, const R* __restrict__ fsm,
                                                  R* __restrict__ kap) {
    const Inst<R>& I = tab[0];
    int i, j, k, c;
    if (interior_cell(I, i, j, k, c)) kap[c] = csf_curvature(I, fsm, i, j, k);
}

// Face-force VJP, gathered per interior cell c over the (up to) six faces it
// touches: akap(c) = (assigned), aF(c) += (the F difference), arho(c) += (the
// face density); work-group sums of dJ/dsigma (each face counted once, at its
// high cell).
template <class R>
__global__ __launch_bounds__(256) void k_csf_adj(const Inst<R>* __restrict__ tab, const R* __restrict__ F,
                                                 const R* __restrict__ kap, const double* __restrict__ aus,
                                                 const double* __restrict__ avs, const double* __restrict__ aws,
                                                 double* __restrict__ akap, double* __restrict__ aF,
                                                 double* __restrict__ arho, double* __restrict__ part) {
    const Inst<R>& I = tab[0];
    int i, j, k, c;
    double sg[1] = {0.0};
    if (interior_cell(I, i, j, k, c)) {
        const double sigma = I.sigma, dt = I.dt;
        double gk = 0.0, gf = 0.0, gr = 0.0;
        // face with high cell q, low cell q - off; c_hi: c is the high cell
        auto face = [&](const double* ag, int q, int off, double h, bool c_hi) {
            const double g = ag[q] * dt;
            if (g == 0.0) return;
            const int m = q - off;
            const double kf = 0.5 * (double(kap[m]) + double(kap[q]));
            const double gF = (double(F[q]) - double(F[m])) * h;
            const double rr = 1.0 / (0.5 * (rho_d(I, F[m]) + rho_d(I, F[q])));
            const double a = sigma * kf * gF * rr;
            if (c_hi) sg[0] += g * kf * gF * rr;
            gk += 0.5 * g * sigma * gF * rr;
            gf += (c_hi ? g : -g) * sigma * kf * h * rr;
            gr -= 0.5 * g * a * rr;
        };
        const int sx = I.sx, sy = I.sy;
        const double hx = I.dxi, hy = I.dyi, hz = I.dzi;
        if (i >= 2) face(aus, c, 1, hx, true);
        if (i + 1 <= I.nx) face(aus, c + 1, 1, hx, false);
        if (j >= 2) face(avs, c, sx, hy, true);
        if (j + 1 <= I.ny) face(avs, c + sx, sx, hy, false);
        if (k >= 2) face(aws, c, sy, hz, true);
        if (k + 1 <= I.nz) face(aws, c + sy, sy, hz, false);
        akap[c] = gk;
        aF[c] += gf;
        arho[c] += gr;
    }
    block_parts<1>(sg, part);
}

// Curvature VJP per vertex (low corner (i,j,k), 0 <= i <= nx ...): gather the
// unit-normal adjoint from the curvature adjoints of the (interior) cells
// around it, through the normalisation to grad Fs, and scatter onto the 8
// (index-clamped) cells that formed it.  2-D launch over (nx+1)(ny+1) x (nz+1).
template <class R>
__global__ __launch_bounds__(256) void k_curv_adj(const Inst<R>* __restrict__ tab, const R* __restrict__ fsm,
                                                  const double* __restrict__ akap, double* __restrict__ afs) {
    const Inst<R>& I = tab[0];
    const int nx = I.nx, ny = I.ny, nz = I.nz, sx = I.sx, sy = I.sy;
    const int t = int(blockIdx.x) * 256 + int(threadIdx.x);
    if (t >= (nx + 1) * (ny + 1)) return;
    const int j = t / (nx + 1), i = t - j * (nx + 1), k = int(blockIdx.y);
    const double qx = 0.25 * double(I.dxi), qy = 0.25 * double(I.dyi), qz = 0.25 * double(I.dzi);
    // kappa(cell) = -(kx+ky+kz): the vertex is on the high side of cell i (di=0), low side of i+1
    double amx = 0.0, amy = 0.0, amz = 0.0;
    for (int dk = 0; dk < 2; ++dk)
        for (int dj = 0; dj < 2; ++dj)
            for (int di = 0; di < 2; ++di) {
                const int ci = i + di, cj = j + dj, ck = k + dk;
                if (ci < 1 || ci > nx || cj < 1 || cj > ny || ck < 1 || ck > nz) continue;
                const double g = akap[ck * sy + cj * sx + ci];
                amx -= (di ? -qx : qx) * g;
                amy -= (dj ? -qy : qy) * g;
                amz -= (dk ? -qz : qz) * g;
            }
    if (amx == 0.0 && amy == 0.0 && amz == 0.0) return;
    const int i0 = max(i, 1), i1 = min(i + 1, nx), j0 = max(j, 1), j1 = min(j + 1, ny);
    const int k0 = max(k, 1), k1 = min(k + 1, nz);
    const int cell[8] = {k0 * sy + j0 * sx + i0, k0 * sy + j0 * sx + i1, k0 * sy + j1 * sx + i0,
                         k0 * sy + j1 * sx + i1, k1 * sy + j0 * sx + i0, k1 * sy + j0 * sx + i1,
                         k1 * sy + j1 * sx + i0, k1 * sy + j1 * sx + i1};  // (di,dj,dk) = bits 0,1,2
    double f[8];
    for (int b = 0; b < 8; ++b) f[b] = fsm[cell[b]];
    const double nxv = (((f[1] + f[3]) + (f[5] + f[7])) - ((f[0] + f[2]) + (f[4] + f[6]))) * qx;
    const double nyv = (((f[2] + f[3]) + (f[6] + f[7])) - ((f[0] + f[1]) + (f[4] + f[5]))) * qy;
    const double nzv = (((f[4] + f[5]) + (f[6] + f[7])) - ((f[0] + f[1]) + (f[2] + f[3]))) * qz;
    const double mag = sqrt((nxv * nxv + nyv * nyv) + nzv * nzv);
    const double inv = 1.0 / (mag + 1e-8 * double(I.dxi));
    double anx = amx * inv, any = amy * inv, anz = amz * inv;
    if (mag > 0.0) {
        const double s = (nxv * amx + nyv * amy + nzv * amz) * inv * inv / mag;
        anx -= nxv * s;
        any -= nyv * s;
        anz -= nzv * s;
    }
    for (int b = 0; b < 8; ++b) {
        const double g = ((b & 1) ? anx : -anx) * qx + ((b & 2) ? any : -any) * qy + ((b & 4) ? anz : -anz) * qz;
        aadd(&afs[cell[b]], g);
    }
}

// Smoothing transpose: every interior cell spreads its Fs adjoint back over
// the 27 cells it averaged, a ghost cell standing for its nearest interior cell.
template <class R>
__global__ __launch_bounds__(256) void k_smooth_adj(const Inst<R>* __restrict__ tab, const double* __restrict__ afs,
                                                    double* __restrict__ aF) {
    const Inst<R>& I = tab[0];
    int i, j, k, c;
    if (!interior_cell(I, i, j, k, c)) return;
    const double g = afs[c];
    if (g == 0.0) return;
    const double w[3] = {0.25, 0.5, 0.25};
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int ti = min(max(i + dx, 1), I.nx), tj = min(max(j + dy, 1), I.ny);
                const int tk = min(max(k + dz, 1), I.nz);
                aadd(&aF[tk * I.sy + tj * I.sx + ti], w[dx + 1] * w[dy + 1] * w[dz + 1] * g);
            }
}

__global__ __launch_bounds__(256) void k_zero3(double* __restrict__ a, double* __restrict__ b,
                                               double* __restrict__ c, long n) {
    const long t = long(blockIdx.x) * 256 + threadIdx.x;
    if (t < n) { a[t] = 0.0; b[t] = 0.0; c[t] = 0.0; }
}

// Up to five whole-field copies in one launch.
template <class R>
__global__ __launch_bounds__(256) void k_copy_fields(FieldCopy<R> fc, long n) {
    const long t = long(blockIdx.x) * 256 + threadIdx.x;
    if (t >= n) return;
    for (int q = 0; q < fc.nf; ++q) fc.dst[q][t] = fc.src[q][t];
}

// ---------------------------------------------------------------------------
// Host launchers.
namespace {
template <class R> dim3 grid_interior(const Inst<R>& h) { return dim3(adj_blocks(long(h.nx) * h.ny), h.nz); }
template <class R> dim3 grid_stored(const Inst<R>& h) { return dim3(adj_blocks(long(h.sy)), h.nz + 2); }
}  // namespace

template <class R>
int adj_centroid_parts(const Inst<R>* tab, const Inst<R>& h, const R* F, double* part, hipStream_t st) {
    const dim3 g = grid_interior(h);
    hipLaunchKernelGGL(k_centroid_parts<R>, g, dim3(256), 0, st, tab, F, part);
    return int(g.x * g.y);
}
template <class R>
void adj_centroid(const Inst<R>* tab, const Inst<R>& h, double J, double sumF, double gJ, const double* seed,
                  double* aF, hipStream_t st) {
    hipLaunchKernelGGL(k_centroid_adj<R>, grid_stored(h), dim3(256), 0, st, tab, J, sumF, gJ, seed, aF);
}
template <class R>
void adj_fct(const Inst<R>* tab, const Inst<R>& h, int dir, const R* F, const R* V, const double* aFo,
             double* aFi, double* aV, int fin, hipStream_t st) {
    if (dir == 0) {  // lanes along x; z chunk of 8 planes, fewer while under 2048 work-groups
        constexpr int S = 64 - 2 * kAdjXHalo;
        const int P = h.sx * (h.ny + 2);
        const int bpp = ((P + S - 1) / S + 3) / 4;
        int kz = 8;
        while (kz > 1 && bpp * ((h.nz + kz - 1) / kz) < 2048) kz /= 2;
        hipLaunchKernelGGL(k_fct_adj_x<R>, dim3(bpp * ((h.nz + kz - 1) / kz)), dim3(256), 0, st, tab, F, V, aFo, aFi,
                           aV, kz, fin);
        return;
    }
    // y and z: marches with lanes along x in chunks of 8 cells (a chunk's
    // windows, 21 positions of F and V and 20 output adjoints, stay in
    // registers; 16-cell chunks would spill them)
    const int n = (dir == 1) ? h.ny : h.nz;
    const long lines = (dir == 1) ? long(h.nx) * h.nz : long(h.nx) * h.ny;
    const int nbt = adj_blocks(lines);
    constexpr int ch = 8;
    const dim3 grid(nbt * ((n + ch - 1) / ch));
    if (dir == 1) hipLaunchKernelGGL((k_fct_adj<R, 1, ch>), grid, dim3(256), 0, st, tab, F, V, aFo, aFi, aV, fin);
    else hipLaunchKernelGGL((k_fct_adj<R, 2, ch>), grid, dim3(256), 0, st, tab, F, V, aFo, aFi, aV, fin);
}
template <class R>
void adj_correct(const Inst<R>* tab, const Inst<R>& h, const R* F, const R* p, const double* au, const double* av,
                 const double* aw, double* aus, double* avs, double* aws, double* ap, double* arho, double* au0,
                 double* av0, double* aw0, hipStream_t st) {
    hipLaunchKernelGGL(k_correct_adj<R>, grid_interior(h), dim3(256), 0, st, tab, F, p, au, av, aw, aus, avs, aws,
                       ap, arho, au0, av0, aw0);
}
template <class R>
void adj_jacobi(const Inst<R>* tab, const Inst<R>& h, const double* in, double* out, double* arhs, int in_g,
                int out_g, int first, int acc, hipStream_t st) {
    // column march; z chunk of 8 planes, fewer while under 900 work-groups
    // (at 96^3: 4-plane chunks, 936 work-groups; the whole backward sweep ran
    // 4% faster than with the 1-plane chunks of 3744 work-groups)
    const int bpp = ((h.sx * (h.ny + 2) + 61) / 62 + 3) / 4;
    int kz = 8;
    while (kz > 1 && bpp * ((h.nz + kz - 1) / kz) < 900) kz /= 2;
    // 1/ap of interior rows and columns, summed as jac_ap sums (field precision)
    const R ax = h.axx, ay = h.ayy, az = h.azz, zf = (h.nz != 1) ? h.azz : R(0);
    const double ri_lo = 1.0 / double(-(ax + ax + ay + ay + zf + R(0)));
    const double ri_hi = 1.0 / double(-(ax + ax + ay + ay + R(0) + zf));
    const double ri_in = 1.0 / double(-(ax + ax + ay + ay + az + az));
    hipLaunchKernelGGL(k_jacobi_adj_col<R>, dim3(bpp * ((h.nz + kz - 1) / kz)), dim3(256), 0, st, tab, in, out, arhs,
                       in_g, out_g, first, acc, kz, ri_lo, ri_in, ri_hi);
}
template <class R>
void adj_jacobi2(const Inst<R>* tab, const Inst<R>& h, const double* in, double* out, double* arhs, int in_g,
                 int out_g, int first, hipStream_t st) {
    // z chunk: the longest of 16, 8, 4, 2 planes that still gives a work-group
    // per CU (256); a chunk re-forms the first sweep on its two end planes, so
    // shorter chunks cost more (at 96^3: 4-plane chunks and 336 work-groups,
    // backward 132 ms; 2-plane chunks, 672 work-groups, 138 ms; 8-plane, 134)
    const int per_z = ((h.nx + 2) * ((h.ny + 13) / 14) + 59) / 60;
    int kz = 16;
    while (kz > 2 && per_z * ((h.nz + kz - 1) / kz) < 256) kz /= 2;
    const R ax = h.axx, ay = h.ayy, az = h.azz, zf = (h.nz != 1) ? h.azz : R(0);
    const double ri_lo = 1.0 / double(-(ax + ax + ay + ay + zf + R(0)));
    const double ri_hi = 1.0 / double(-(ax + ax + ay + ay + R(0) + zf));
    const double ri_in = 1.0 / double(-(ax + ax + ay + ay + az + az));
    hipLaunchKernelGGL(k_jacobi2_adj<R>, dim3(per_z * ((h.nz + kz - 1) / kz)), dim3(1024), 0, st, tab, in, out, arhs,
                       in_g, out_g, first, kz, ri_lo, ri_in, ri_hi);
}
template <class R>
void adj_rhs(const Inst<R>* tab, const Inst<R>& h, const R* F, const R* us, const R* vs, const R* ws,
             const double* arhs, double* aus, double* avs, double* aws, double* arho, hipStream_t st) {
    hipLaunchKernelGGL(k_rhs_adj<R>, grid_interior(h), dim3(256), 0, st, tab, F, us, vs, ws, arhs, aus, avs, aws,
                       arho);
}
int adj_predict_blocks(int nx, int ny, int nz, int* kz_out);
template <class R>
void adj_predict(const Inst<R>* tab, const Inst<R>& h, const R* F, const R* u, const R* v, const R* w,
                 const double* aus, const double* avs, const double* aws, const double* arho, double* au, double* av,
                 double* aw, double* aF, double* part, double* acc, hipStream_t st) {
    int kz = 0;
    const int nb = adj_predict_blocks(h.nx, h.ny, h.nz, &kz);
    hipLaunchKernelGGL(k_predict_adj<R>, dim3(nb), dim3(256), 0, st, tab, F, u, v, w, aus, avs, aws, arho, au, av,
                       aw, aF, part, kz);
}
// Work-groups of the predictor transpose (and its z chunk): 16 planes, fewer
// while that leaves under 1024 work-groups.
int adj_predict_blocks(int nx, int ny, int nz, int* kz_out) {
    const int nxy = ((nx + PTX - 1) / PTX) * ((ny + PTY - 1) / PTY);
    int kz = 16;
    while (kz > 2 && nxy * ((nz + kz - 1) / kz) < 1024) kz /= 2;
    if (kz_out) *kz_out = kz;
    return nxy * ((nz + kz - 1) / kz);
}
void adj_predict_sums(int nx, int ny, int nz, const double* part, double* acc, hipStream_t st) {
    hipLaunchKernelGGL(k_sum_parts<7>, dim3(1), dim3(256), 0, st, part, adj_predict_blocks(nx, ny, nz, nullptr),
                       acc);
}
template <class R>
void adj_csf(const Inst<R>* tab, const Inst<R>& h, const R* F, R* fsm, R* kap, const double* aus, const double* avs,
             const double* aws, double* akap, double* afs, double* aF, double* arho, double* part, double* acc,
             hipStream_t st) {
    const dim3 g = grid_interior(h);
    hipLaunchKernelGGL(k_adj_smooth<R>, g, dim3(256), 0, st, tab, F, fsm);
    hipLaunchKernelGGL(k_adj_curv<R>, g, dim3(256), 0, st, tab, fsm, kap);
    hipLaunchKernelGGL(k_csf_adj<R>, g, dim3(256), 0, st, tab, F, kap, aus, avs, aws, akap, aF, arho, part);
    hipLaunchKernelGGL(k_sum_parts<1>, dim3(1), dim3(256), 0, st, part, int(g.x * g.y), acc + ADJ_SIGMA);
    (void)hipMemsetAsync(afs, 0, size_t(h.ntot) * sizeof(double), st);
    hipLaunchKernelGGL(k_curv_adj<R>, dim3(adj_blocks(long(h.nx + 1) * (h.ny + 1)), h.nz + 1), dim3(256), 0, st, tab,
                       fsm, akap, afs);
    hipLaunchKernelGGL(k_smooth_adj<R>, g, dim3(256), 0, st, tab, afs, aF);
}
template void adj_csf<float>(const Inst<float>*, const Inst<float>&, const float*, float*, float*, const double*,
                             const double*, const double*, double*, double*, double*, double*, double*, double*,
                             hipStream_t);
template void adj_csf<double>(const Inst<double>*, const Inst<double>&, const double*, double*, double*,
                              const double*, const double*, const double*, double*, double*, double*, double*,
                              double*, double*, hipStream_t);

void adj_zero3(double* a, double* b, double* c, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero3, dim3(adj_blocks(n)), dim3(256), 0, st, a, b, c, n);
}
template <class R>
void copy_fields(const FieldCopy<R>& fc, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_copy_fields<R>, dim3(adj_blocks(n)), dim3(256), 0, st, fc, n);
}
template void copy_fields<float>(const FieldCopy<float>&, long, hipStream_t);
template void copy_fields<double>(const FieldCopy<double>&, long, hipStream_t);

#define VOF_ADJ_INSTANTIATE(R)                                                                                     \
    template int adj_centroid_parts<R>(const Inst<R>*, const Inst<R>&, const R*, double*, hipStream_t);           \
    template void adj_centroid<R>(const Inst<R>*, const Inst<R>&, double, double, double, const double*, double*, \
                                  hipStream_t);                                                                    \
    template void adj_fct<R>(const Inst<R>*, const Inst<R>&, int, const R*, const R*, const double*, double*,     \
                             double*, int, hipStream_t);                                                           \
    template void adj_correct<R>(const Inst<R>*, const Inst<R>&, const R*, const R*, const double*, const double*, \
                                 const double*, double*, double*, double*, double*, double*, double*, double*,     \
                                 double*, hipStream_t);                                                            \
    template void adj_jacobi<R>(const Inst<R>*, const Inst<R>&, const double*, double*, double*, int, int, int,   \
                                int, hipStream_t);                                                                 \
    template void adj_jacobi2<R>(const Inst<R>*, const Inst<R>&, const double*, double*, double*, int, int, int,  \
                                 hipStream_t);                                                                     \
    template void adj_rhs<R>(const Inst<R>*, const Inst<R>&, const R*, const R*, const R*, const R*,              \
                             const double*, double*, double*, double*, double*, hipStream_t);                      \
    template void adj_predict<R>(const Inst<R>*, const Inst<R>&, const R*, const R*, const R*, const R*,          \
                                 const double*, const double*, const double*, const double*, double*, double*,     \
                                 double*, double*, double*, double*, hipStream_t);

VOF_ADJ_INSTANTIATE(float)
VOF_ADJ_INSTANTIATE(double)

}  // namespace vof
