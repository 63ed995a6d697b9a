// vof_device.h — device-side types and helpers shared by every gfx950 kernel.
//
// One `Inst` describes one dam-break instance: its grid, the constants every
// stage needs (pre-multiplied exactly the way the reference's expressions
// associate them, so results stay bit-identical), and the device pointers of
// its fields.  The plain solver is an ensemble of one instance; the batched
// solver is an ensemble of many, found per work-group through a prefix table of
// work-group counts.  Field layout is the reference's: (nx+2)(ny+2)(nz+2),
// one ghost layer, x fastest (voftrial1.cu:66-70).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vof {

template <class R>
struct Inst {
    int nx, ny, nz;  // interior cells
    int sx, sy;      // row pitch (nx+2), plane pitch (nx+2)(ny+2), in elements
    int ntot;        // (nx+2)(ny+2)(nz+2)
    R dt, dx, dy, dz, dxi, dyi, dzi;
    R vol;                     // (dx*dy)*dz                        (voftrial1.cu:518)
    R dxdy;                    // dx*dy, y-sweep stage-1 divisor     (:672)
    R axx, ayy, azz;           // dxi*dxi, dyi*dyi, dzi*dzi          (:424-429)
    R dtdydz, dtdxdz, dtdxdy;  // (dt*dy)*dz etc., dilation factors  (:519, :661, :800)
    R rho_l, rho_g, nu_l, nu_g;
    R gx, gy, gz;
    R sigma;
    R slip;   // sign of a tangential velocity ghost: +1 copies the interior (free slip,
              // the reference's walls), -1 negates it (no-slip, VOF_FLAG_NO_SLIP)
    R* F[2];  // volume fraction, ping-pong across directional sweeps
    R* u;
    R* v;
    R* w;     // face velocities (MAC: u on the low-x face of cell i)
    R* us;
    R* vs;
    R* ws;    // predicted velocities u*, v*, w*
    R* p[2];  // pressure, ping-pong across Jacobi launches
    R* rhs;   // rho/dt * div(u*), formed once per step
    R* fsm;   // smoothed F for the interface normals (CSF extension only; null otherwise)
    R* kappa; // interface curvature (CSF extension only; null otherwise)
};

// Pick one of a ping-pong pair with constant indices only, so that a copied
// Inst stays in registers (a runtime array index would force it to memory).
template <class T>
__device__ __forceinline__ T sel2(const T (&a)[2], int s) { return s ? a[1] : a[0]; }

// Work-group -> instance lookup.  boff has ninst+1 entries: boff[i] is the
// first work-group of instance i for this launch shape, boff[ninst] the total.
__device__ __forceinline__ int find_inst(const int* __restrict__ boff, int ninst, int b) {
    int lo = 0, hi = ninst;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (boff[mid] <= b) lo = mid; else hi = mid;
    }
    return lo;
}

// Work-groups are dealt round-robin to the 8 XCDs (each with its own L2).
// Renumber so that every XCD works through one contiguous range of logical
// work-groups; neighbouring tiles then share an L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// ---------------------------------------------------------------------------
// Raw buffer access.  A field is addressed through a 128-bit buffer resource
// (base, size in bytes): the per-lane offset is a 32-bit VGPR that stays
// fixed while a lane marches, and the plane offset is a scalar (SGPR)
// operand, so a load or store costs no per-lane 64-bit address arithmetic.
// Accesses outside [0, size) return 0 / are dropped by the hardware.  Sizes
// and byte offsets are 32-bit unsigned quantities (a field of up to 4 GiB;
// the host rejects larger grids), formed in unsigned arithmetic.
constexpr int kRsrcWord3 = 0x00020000;  // raw, untyped 32-bit records (gfx9 family)

__device__ __forceinline__ int byte_off(int elems, unsigned esz) { return (int)((unsigned)elems * esz); }

template <class R>
struct Buf {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ Buf(const R* p, int n)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<R*>(p), (short)0, byte_off(n, sizeof(R)), kRsrcWord3)) {}
    // vo: per-lane element offset, so: wave-uniform element offset
    __device__ __forceinline__ R ld(int vo, int so) const;
    __device__ __forceinline__ void st(R v, int vo, int so) const;
};

template <>
__device__ __forceinline__ float Buf<float>::ld(int vo, int so) const {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, byte_off(vo, 4u), byte_off(so, 4u), 0));
}
template <>
__device__ __forceinline__ void Buf<float>::st(float v, int vo, int so) const {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, byte_off(vo, 4u), byte_off(so, 4u), 0);
}
template <>
__device__ __forceinline__ double Buf<double>::ld(int vo, int so) const {
    const auto x = __builtin_amdgcn_raw_buffer_load_b64(r, byte_off(vo, 8u), byte_off(so, 8u), 0);
    return __hiloint2double((int)x[1], (int)x[0]);
}
template <>
__device__ __forceinline__ void Buf<double>::st(double v, int vo, int so) const {
    decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)) x;
    x[0] = (unsigned)__double2loint(v);
    x[1] = (unsigned)__double2hiint(v);
    __builtin_amdgcn_raw_buffer_store_b64(x, r, byte_off(vo, 8u), byte_off(so, 8u), 0);
}

// ---------------------------------------------------------------------------
// Whole-wave lane shifts by one lane (DPP wave_shr:1 / wave_shl:1; the value
// entering at the wave's edge is 0).  The zero comes from bound_ctrl (a lane
// whose source lane does not exist reads 0), not from a zeroed old value, so
// the shift needs no initialising move and can be folded into the DPP
// operand of the instruction that consumes it.
__device__ __forceinline__ int dpp_up_i(int x) { return __builtin_amdgcn_mov_dpp(x, 0x138, 0xf, 0xf, true); }
__device__ __forceinline__ int dpp_dn_i(int x) { return __builtin_amdgcn_mov_dpp(x, 0x130, 0xf, 0xf, true); }

__device__ __forceinline__ float lane_up(float x) { return __int_as_float(dpp_up_i(__float_as_int(x))); }  // lane-1
__device__ __forceinline__ float lane_dn(float x) { return __int_as_float(dpp_dn_i(__float_as_int(x))); }  // lane+1
__device__ __forceinline__ double lane_up(double x) {
    return __hiloint2double(dpp_up_i(__double2hiint(x)), dpp_up_i(__double2loint(x)));
}
__device__ __forceinline__ double lane_dn(double x) {
    return __hiloint2double(dpp_dn_i(__double2hiint(x)), dpp_dn_i(__double2loint(x)));
}

// ---------------------------------------------------------------------------
// Division by a negative divisor whose reciprocal y = RN(1/b) is formed once
// (by an IEEE division) and reused: Markstein's correction.  q0 = RN(a y) is
// within two ulps of a/b; one correction with the residual r0 = a - b q0
// (formed by an FMA) gives a faithful q1; the residual of a faithful quotient
// is exact, and a second correction then returns RN(a/b), the IEEE quotient
// (Markstein, IBM J. Res. Dev. 34 (1990); Muller et al., Handbook of
// Floating-Point Arithmetic, Thm. "Markstein").  The proof needs every
// intermediate in the normal range: with -b in [2^-40, 2^40] (ok_b, checked
// once per divisor) that holds for |a| in [2^-60, 2^80].  A zero a also gives
// the IEEE result when b < 0: q0 = a y is the signed zero a/b, both residuals
// are +0, and the corrections keep q0's sign (with b > 0 and a = -0 they would
// not).  Any other a (tiny, huge, inf, NaN) takes the IEEE division.  So the
// result is the IEEE quotient's bits for every a; the common case costs one
// multiply and four FMAs instead of the ~10-instruction scaled division with
// its quarter-rate reciprocal.
__device__ __forceinline__ bool div_b_ok(float b) {
    const unsigned u = __float_as_uint(b);
    return u - 0xab800000u <= 0xd3800000u - 0xab800000u;  // 2^-40 <= -b <= 2^40
}
__device__ __forceinline__ bool div_b_ok(double b) {
    const unsigned u = unsigned(__double2hiint(b));
    return u - 0xbd700000u <= 0xc2700000u - 0xbd700000u;  // 2^-40 <= -b < 2^40 (1 + 2^-20)
}
__device__ __forceinline__ bool div_a_ok(float a) {
    const unsigned u = __float_as_uint(a) & 0x7fffffffu;
    return u - 0x21800000u <= 0x67800000u - 0x21800000u || u == 0u;  // 2^-60 <= |a| <= 2^80, or a = +-0
}
__device__ __forceinline__ bool div_a_ok(double a) {
    const unsigned u = unsigned(__double2hiint(a)) & 0x7fffffffu;
    // 2^-60 <= |a| < 2^80 (1 + 2^-20), or a = +-0 (high word 0 and low word 0)
    return u - 0x3c300000u <= 0x44f00000u - 0x3c300000u || (u | unsigned(__double2loint(a))) == 0u;
}
__device__ __forceinline__ float fma_r(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_r(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <class R>
__device__ __forceinline__ R div_rn(R a, R b, R y, bool ok_b) {
    R q;
    if (ok_b && div_a_ok(a)) {
        const R q0 = a * y;
        const R r0 = fma_r(-b, q0, a);
        const R q1 = fma_r(r0, y, q0);
        const R r1 = fma_r(-b, q1, a);
        q = fma_r(r1, y, q1);
    } else {
        q = a / b;
    }
    return q;
}

// The same quotient without divergent branches on the common path: the
// corrected quotient is formed on every lane, and only when some lane of the
// wave needs it (a wave-uniform test) do those lanes redo it as the IEEE
// division.  The lane sets are wave masks formed once by the caller: need,
// the lanes whose result is kept; must, those of them whose divisor fails
// ok_b.  The range test is one compare per point whose lane mask is used as
// is; zeros (exact with b < 0) are sorted out only when the range test fails.
// a outside the range of div_a_ok (zero, NaN and inf included): above it
// (or unordered) or below it, each one compare of |a| (a source modifier)
// whose lane mask the ballot takes as is
__device__ __forceinline__ bool div_a_above(float a) { return !(fabsf(a) <= 0x1p80f); }
__device__ __forceinline__ bool div_a_below(float a) { return fabsf(a) < 0x1p-60f; }
__device__ __forceinline__ bool div_a_above(double a) { return !(fabs(a) < 0x1.00001p80); }
__device__ __forceinline__ bool div_a_below(double a) { return fabs(a) < 0x1p-60; }

template <class R>
__device__ __forceinline__ R div_rn_wave(R a, R b, R y, uint64_t must, uint64_t need) {
    const R q0 = a * y;
    const R r0 = fma_r(-b, q0, a);
    const R q1 = fma_r(r0, y, q0);
    const R r1 = fma_r(-b, q1, a);
    R q = fma_r(r1, y, q1);
    uint64_t bad =
        ((__builtin_amdgcn_ballot_w64(div_a_above(a)) | __builtin_amdgcn_ballot_w64(div_a_below(a))) & need) | must;
    if (bad != 0) {
        bad &= ~__builtin_amdgcn_ballot_w64(a == R(0)) | must;
        if (bad != 0) {
            // The empty volatile asm keeps the division inside this branch (the
            // compiler would otherwise compute it on every lane and select).
            R aa = a;
            asm volatile("" : "+v"(aa));
            const R qs = aa / b;
            const int lane = (int)__lane_id();
            q = ((bad >> lane) & 1u) ? qs : q;
        }
    }
    return q;
}

template <class R> __device__ __forceinline__ R vmax(R a, R b);
template <class R> __device__ __forceinline__ R vmin(R a, R b);
template <> __device__ __forceinline__ float vmax<float>(float a, float b) { return fmaxf(a, b); }
template <> __device__ __forceinline__ float vmin<float>(float a, float b) { return fminf(a, b); }
template <> __device__ __forceinline__ double vmax<double>(double a, double b) { return fmax(a, b); }
template <> __device__ __forceinline__ double vmin<double>(double a, double b) { return fmin(a, b); }

// median(0, 1, v) as ((0 + 1) + v) - max(max(0,1), v) - min(min(0,1), v); the
// constant sub-terms fold exactly, leaving the form below.  Not a plain clamp:
// values in (0,1) are rounded to the 2^-23 grid of 1+v (voftrial1.cu:52-60).
template <class R>
__device__ __forceinline__ R clamp01(R v) {
    return ((R(1) + v) - vmax(R(1), v)) - vmin(R(0), v);
}

template <class R>
__device__ __forceinline__ R rho_of(const Inst<R>& I, R F) {  // :100
    const R f = clamp01(F);
    return I.rho_g * (R(1) - f) + I.rho_l * f;
}

template <class R>
__device__ __forceinline__ R nu_of(const Inst<R>& I, R F) {  // :101
    const R f = clamp01(F);
    return I.nu_l * f + I.nu_g * (R(1) - f);
}

// ---------------------------------------------------------------------------
// CSF surface tension (extension; voftrial1.cu stores sigma, :35, :1279, and
// never uses it).  Continuum surface force a = sigma kappa grad F / rho:
//  * the interface normals come from Fs, F smoothed once by the separable
//    (1/4, 1/2, 1/4) filter on interior cells; Fs outside the interior is the
//    nearest interior value (index clamping);
//  * the normal n = grad Fs at a vertex comes from the 2x2x2 cells around it
//    (the four on the high side of each axis minus the four on the low side,
//    times 1/(4h)), the unit normal is m = n / (|n| + eps), eps = 1e-8 / dx;
//  * the curvature of a cell is kappa = -div m from its 8 vertices (the four
//    on the high side of each axis minus the four on the low side, times 1/(4h));
//  * on a face between cells c-1 and c the acceleration is
//    sigma (kappa(c-1)+kappa(c))/2 (F(c)-F(c-1))/h / ((rho(c-1)+rho(c))/2).
// With the smoothing, the face curvature of a resolved sphere is 2/R to 1%.
// The association written here is the contract the CPU oracle repeats.
template <class R>
__device__ __forceinline__ R csf_smooth(const Inst<R>& I, const R* __restrict__ F, int c) {
    const int sx = I.sx, sy = I.sy;
    R ry[3];
    for (int dz = -1; dz <= 1; ++dz) {
        R rx[3];
        for (int dy = -1; dy <= 1; ++dy) {
            const int q = c + dz * sy + dy * sx;
            rx[dy + 1] = (R(0.25) * F[q - 1] + R(0.5) * F[q]) + R(0.25) * F[q + 1];
        }
        ry[dz + 1] = (R(0.25) * rx[0] + R(0.5) * rx[1]) + R(0.25) * rx[2];
    }
    return (R(0.25) * ry[0] + R(0.5) * ry[1]) + R(0.25) * ry[2];
}

template <class R>
struct VNormal {
    R x, y, z;
};

// Unit normal at the vertex whose low corner is cell (i,j,k), 0 <= i <= nx etc.
template <class R>
__device__ __forceinline__ VNormal<R> csf_vertex_normal(const Inst<R>& I, const R* __restrict__ Fs, int i, int j,
                                                        int k) {
    const int sx = I.sx, sy = I.sy;
    const int i0 = max(i, 1), i1 = min(i + 1, I.nx), j0 = max(j, 1), j1 = min(j + 1, I.ny);
    const int k0 = max(k, 1), k1 = min(k + 1, I.nz);
    const R f000 = Fs[k0 * sy + j0 * sx + i0], f100 = Fs[k0 * sy + j0 * sx + i1];
    const R f010 = Fs[k0 * sy + j1 * sx + i0], f110 = Fs[k0 * sy + j1 * sx + i1];
    const R f001 = Fs[k1 * sy + j0 * sx + i0], f101 = Fs[k1 * sy + j0 * sx + i1];
    const R f011 = Fs[k1 * sy + j1 * sx + i0], f111 = Fs[k1 * sy + j1 * sx + i1];
    VNormal<R> n;
    n.x = (((f100 + f110) + (f101 + f111)) - ((f000 + f010) + (f001 + f011))) * (R(0.25) * I.dxi);
    n.y = (((f010 + f110) + (f011 + f111)) - ((f000 + f100) + (f001 + f101))) * (R(0.25) * I.dyi);
    n.z = (((f001 + f101) + (f011 + f111)) - ((f000 + f100) + (f010 + f110))) * (R(0.25) * I.dzi);
    const R mag = sqrt((n.x * n.x + n.y * n.y) + n.z * n.z);
    const R inv = R(1) / (mag + R(1e-8) * I.dxi);
    n.x = n.x * inv;
    n.y = n.y * inv;
    n.z = n.z * inv;
    return n;
}

// Curvature of interior cell (i,j,k) from the smoothed field Fs.
template <class R>
__device__ __forceinline__ R csf_curvature(const Inst<R>& I, const R* __restrict__ Fs, int i, int j, int k) {
    const VNormal<R> v000 = csf_vertex_normal(I, Fs, i - 1, j - 1, k - 1), v100 = csf_vertex_normal(I, Fs, i, j - 1, k - 1);
    const VNormal<R> v010 = csf_vertex_normal(I, Fs, i - 1, j, k - 1), v110 = csf_vertex_normal(I, Fs, i, j, k - 1);
    const VNormal<R> v001 = csf_vertex_normal(I, Fs, i - 1, j - 1, k), v101 = csf_vertex_normal(I, Fs, i, j - 1, k);
    const VNormal<R> v011 = csf_vertex_normal(I, Fs, i - 1, j, k), v111 = csf_vertex_normal(I, Fs, i, j, k);
    const R kx = (((v100.x + v110.x) + (v101.x + v111.x)) - ((v000.x + v010.x) + (v001.x + v011.x))) * (R(0.25) * I.dxi);
    const R ky = (((v010.y + v110.y) + (v011.y + v111.y)) - ((v000.y + v100.y) + (v001.y + v101.y))) * (R(0.25) * I.dyi);
    const R kz = (((v001.z + v101.z) + (v011.z + v111.z)) - ((v000.z + v100.z) + (v010.z + v110.z))) * (R(0.25) * I.dzi);
    return -((kx + ky) + kz);
}

// CSF acceleration on the face between cells c-off (low) and c (high), h = 1/spacing.
template <class R>
__device__ __forceinline__ R csf_accel(const Inst<R>& I, const R* __restrict__ F, const R* __restrict__ kap, int c,
                                       int off, R h) {
    const R kf = R(0.5) * (kap[c - off] + kap[c]);
    const R gF = (F[c] - F[c - off]) * h;
    const R rf = R(0.5) * (rho_of(I, F[c - off]) + rho_of(I, F[c]));
    return I.sigma * kf * gF / rf;
}

}  // namespace vof
