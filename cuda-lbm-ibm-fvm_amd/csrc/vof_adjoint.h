// vof_adjoint.h — host launchers of the reverse-mode kernels (vof_adjoint.hip).
//
// Every launcher works on one instance (table entry 0) and takes FP64 adjoint
// arrays laid out like the fields ((nx+2)(ny+2)(nz+2), x fastest).
#pragma once
#include "vof_device.h"

namespace vof {

// Scalar gradients accumulated on the device, in this order.
enum AdjScalar { ADJ_RHO_L = 0, ADJ_RHO_G, ADJ_NU_L, ADJ_NU_G, ADJ_GX, ADJ_GY, ADJ_GZ, ADJ_SIGMA, ADJ_NSCALAR };

inline int adj_blocks(long n) { return int((n + 255) / 256); }
// doubles needed by the `part` scratch of the launchers below
inline size_t adj_part_size(int nx, int ny, int nz) {
    const size_t a = size_t(adj_blocks(long(nx) * ny)), b = size_t((nx + 31) / 32) * size_t((ny + 7) / 8);
    return size_t(ADJ_NSCALAR) * (a > b ? a : b) * size_t(nz);
}

// Work-group partial sums (sum F, sum F z) of the centroid objective; returns
// the number of work-groups (2 doubles each in part).
template <class R> int adj_centroid_parts(const Inst<R>* tab, const Inst<R>& h, const R* F, double* part, hipStream_t st);
// aF = gJ dJ/dF + seed folded through the ghost fill (seed may be null).
template <class R> void adj_centroid(const Inst<R>* tab, const Inst<R>& h, double J, double sumF, double gJ,
                                     const double* seed, double* aF, hipStream_t st);
// One FCT sweep in reverse: aFo (interior) -> aFi (interior), aV += velocity adjoint.
template <class R> void adj_fct(const Inst<R>* tab, const Inst<R>& h, int dir, const R* F, const R* V,
                                const double* aFo, double* aFi, double* aV, int fin, hipStream_t st);
// Correction + velocity BC in reverse: au* = (assigned), ap += , arho = , plane-0 copies -> au0 (assigned).
template <class R> void adj_correct(const Inst<R>* tab, const Inst<R>& h, const R* F, const R* p, const double* au,
                                    const double* av, const double* aw, double* aus, double* avs, double* aws,
                                    double* ap, double* arho, double* au0, double* av0, double* aw0, hipStream_t st);
// One Jacobi sweep in reverse, on h = D^-1 gbar: in_g: the input is gbar,
// out_g: write gbar instead of h, first: assign arhs instead of adding; acc:
// 0 arhs untouched, 1 arhs (+)= input h, 2 arhs (+)= input h + output h.
template <class R> void adj_jacobi(const Inst<R>* tab, const Inst<R>& h, const double* in, double* out,
                                   double* arhs, int in_g, int out_g, int first, int acc, hipStream_t st);
// Two such sweeps in one launch: arhs (+)= input h + the first sweep's h (the
// two sweep inputs), out = the second sweep's h (gbar when out_g); the same
// bits as two adj_jacobi launches with acc 2 and 0.
template <class R> void adj_jacobi2(const Inst<R>* tab, const Inst<R>& h, const double* in, double* out,
                                    double* arhs, int in_g, int out_g, int first, hipStream_t st);
template <class R> void adj_rhs(const Inst<R>* tab, const Inst<R>& h, const R* F, const R* us, const R* vs,
                                const R* ws, const double* arhs, double* aus, double* avs, double* aws, double* arho,
                                hipStream_t st);
// Predictor + rho/nu in reverse: au += (atomics), aF += ; adds its work-group
// partial sums of dJ/d(rho_l, rho_g, nu_l, nu_g, gx, gy, gz) to part (7 per
// work-group; zero it before the first step), which adj_predict_sums reduces
// into acc[0..7) once after the last step.
template <class R> void adj_predict(const Inst<R>* tab, const Inst<R>& h, const R* F, const R* u, const R* v,
                                    const R* w, const double* aus, const double* avs, const double* aws,
                                    const double* arho, double* au, double* av, double* aw, double* aF, double* part,
                                    double* acc, hipStream_t st);
void adj_predict_sums(int nx, int ny, int nz, const double* part, double* acc, hipStream_t st);
// doubles of the predictor's partial-sum region (7 per work-group)
inline size_t adj_predict_part_size(int nx, int ny, int nz) { return adj_part_size(nx, ny, nz); }
// CSF in reverse (VOF_FLAG_CSF): recompute Fs and kappa of the step into fsm,
// kap; then akap = (scratch), aF += and arho += the face-force VJP, afs =
// (scratch) the curvature VJP, aF += its smoothing transpose (ghosts folded to
// the nearest interior cell); adds dJ/dsigma to acc[ADJ_SIGMA].
template <class R> void adj_csf(const Inst<R>* tab, const Inst<R>& h, const R* F, R* fsm, R* kap, const double* aus,
                                const double* avs, const double* aws, double* akap, double* afs, double* aF,
                                double* arho, double* part, double* acc, hipStream_t st);
// a = b = c = 0 over n doubles.
void adj_zero3(double* a, double* b, double* c, long n, hipStream_t st);

// Up to five copies of n elements in one launch.
template <class R>
struct FieldCopy {
    const R* src[5];
    R* dst[5];
    int nf;
};
template <class R> void copy_fields(const FieldCopy<R>& fc, long n, hipStream_t st);

}  // namespace vof
