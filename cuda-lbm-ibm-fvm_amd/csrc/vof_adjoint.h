// vof_adjoint.h — host launchers of the reverse-mode kernels (vof_adjoint.hip).
//
// Every launcher works on one instance (table entry 0) and takes FP64 adjoint
// arrays laid out like the fields ((nx+2)(ny+2)(nz+2), x fastest).
#pragma once
#include "vof_device.h"

namespace vof {

// Per-line scratch of the FCT reverse sweep: forward intermediates (R) and
// line-private adjoint accumulators (FP64), each one field in size.
template <class R>
struct FctTape {
    R *ftd, *a, *L, *rp, *rm, *C, *dv, *y;
    double *gftd, *ga, *gL, *grp, *grm, *gdv;
};

// Scalar gradients accumulated on the device, in this order.
enum AdjScalar { ADJ_RHO_L = 0, ADJ_RHO_G, ADJ_NU_L, ADJ_NU_G, ADJ_GX, ADJ_GY, ADJ_GZ, ADJ_NSCALAR };

inline int adj_blocks(long n) { return int((n + 255) / 256); }

template <class R> void adj_centroid_parts(const Inst<R>* tab, const Inst<R>& h, const R* F, double* part, hipStream_t st);
template <class R> void adj_centroid(const Inst<R>* tab, const Inst<R>& h, double J, double sumF, double gJ,
                                     const double* seed, double* aF, hipStream_t st);
template <class R> void adj_ghostF(const Inst<R>* tab, const Inst<R>& h, double* aF, hipStream_t st);
template <class R> void adj_fct(const Inst<R>* tab, const Inst<R>& h, int dir, const R* F, const R* V, double* aF,
                                double* aV, const FctTape<R>& tp, int fin, hipStream_t st);
template <class R> void adj_correct(const Inst<R>* tab, const Inst<R>& h, const R* F, const R* p, const double* au,
                                    const double* av, const double* aw, double* aus, double* avs, double* aws,
                                    double* ap, double* arho, double* au0, double* av0, double* aw0, hipStream_t st);
template <class R> void adj_jacobi(const Inst<R>* tab, const Inst<R>& h, const double* gin, double* gout,
                                   double* arhs, hipStream_t st);
template <class R> void adj_rhs(const Inst<R>* tab, const Inst<R>& h, const R* F, const R* us, const R* vs,
                                const R* ws, const double* arhs, double* aus, double* avs, double* aws, double* arho,
                                hipStream_t st);
// Adds dJ/dg (3 values) to acc[ADJ_GX..ADJ_GZ]; part needs 3 * adj_blocks(nx*ny*nz) doubles.
template <class R> void adj_predict(const Inst<R>* tab, const Inst<R>& h, const R* F, const R* u, const R* v,
                                    const R* w, const double* aus, const double* avs, const double* aws, double* au,
                                    double* av, double* aw, double* anu, double* part, double* acc, hipStream_t st);
// Adds the F adjoint of rho and nu to aF and dJ/d(rho_l, rho_g, nu_l, nu_g) to
// acc[ADJ_RHO_L..ADJ_NU_G]; part needs 4 * adj_blocks(ntot) doubles.
template <class R> void adj_rhonu(const Inst<R>* tab, const Inst<R>& h, const R* F, const double* arho,
                                  const double* anu, double* aF, double* part, double* acc, hipStream_t st);
// dst += src over n doubles.
void adj_axpy(const double* src, double* dst, long n, hipStream_t st);

}  // namespace vof
