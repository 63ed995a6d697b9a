// vof_kernels.h — host-side launchers of the gfx950 kernels (vof_kernels.hip).
#pragma once
#include "../../include/vof_capi.h"
#include "vof_device.h"

namespace vof {

// Launch shape of one kernel over an ensemble: prefix table of work-group
// counts per instance (device memory, ninst+1 entries), the total, and the
// number of cells each lane walks along its march direction.
struct Geo {
    int* boff = nullptr;
    int nblk = 0;
    int kz = 1;
    int min_nz = 0;  // fewest planes over the instances
    int rb = 0;      // predictor: rows per lane of the band mapping (0: one row per lane)
};

template <class R>
struct Ens {
    Inst<R>* tab = nullptr;  // device copy of the instance table
    int ninst = 0;
    Inst<R> one{};           // host copy of instance 0 (the whole table when ninst == 1)
};

// Kernel argument block of every step kernel: the instance table and the
// launch shape's work-group prefix table; for a single instance also the
// record itself, by value, so that a work-group has it with its kernel
// arguments instead of after two dependent scalar loads (prefix, record).
template <class R>
struct EnsArg {
    const Inst<R>* tab;
    const int* boff;
    int ninst;
    Inst<R> one;  // valid when ninst == 1
};

template <class R>
inline EnsArg<R> ens_arg(const Ens<R>& e, const Geo& g) {
    return EnsArg<R>{e.tab, g.boff, e.ninst, e.one};
}

constexpr int kNumReduce = 7;
// Chunk lengths of the FP32 two-sweep Jacobi streamed through LDS requests
// (k_jacobi2_dma): 25 planes, or 20 (at 200^3 two work-groups on nearly every
// CU in one round); any other length uses the preloaded or register-streamed
// kernels.
constexpr int kJ2DmaKz = 25;
constexpr int kJ2DmaKzShort = 20;
// Most Jacobi sweeps one launch performs (jacobi_tblock 1 .. kJacobiMaxT).
constexpr int kJacobiMaxT = 5;
// Largest k_reduce1 launch the allocation-free diagnostics serve: its
// partials fill a static device array of this many work-groups x kNumReduce.
constexpr int kNoAllocMaxBlocks = 65536;

void launch_copy(const void* a, void* b, long n16, hipStream_t st);
// Work-groups of the stream-mix kernel (grid-stride over the arrays).
constexpr int kStreamBlocks = 8192;
// nread in 1..6 arrays read, nwrite in 1..3 written, n16 16-byte elements each,
// consecutive arrays pitch elements apart in a and in b; false for another mix.
bool launch_stream(int nread, int nwrite, const void* a, void* b, long n16, long pitch, hipStream_t st);
template <class R> void launch_predict(const Ens<R>& e, const Geo& g, int fsel, hipStream_t st);
template <class R> void launch_predict_rho(const Ens<R>& e, const Geo& g, int fsel, hipStream_t st);
template <class R> void launch_rhs(const Ens<R>& e, const Geo& g, int fsel, hipStream_t st);
template <class R> void launch_jacobi(const Ens<R>& e, const Geo& g, int psrc, hipStream_t st);
template <class R> void launch_jacobi2(const Ens<R>& e, const Geo& g, int psrc, hipStream_t st);
// sweeps in 3 .. kJacobiMaxT sweeps in one launch; false for another count.
template <class R> bool launch_jacobiT(const Ens<R>& e, const Geo& g, int sweeps, int psrc, hipStream_t st);
template <class R> void launch_correct(const Ens<R>& e, const Geo& g, int fsel, int psel, hipStream_t st);
template <class R> void launch_fct(const Ens<R>& e, const Geo& g, int dir, int src, int fin, hipStream_t st);
// z sweep followed by the x sweep in one launch, chunks of g.kz in {4, 8, 16}
// planes (false, and nothing launched, for any other chunk).
template <class R> bool launch_fct_zx(const Ens<R>& e, const Geo& g, int src, int fin, hipStream_t st);
// x sweep followed by the y sweep in one launch, chunks of g.kz in {4, 8, 16}
// rows (false, and nothing launched, for any other chunk).
template <class R> bool launch_fct_xy(const Ens<R>& e, const Geo& g, int src, hipStream_t st);
template <class R> void launch_clamp_F(const Ens<R>& e, const Geo& g, int fsel, hipStream_t st);
template <class R> void launch_ghost_F(const Ens<R>& e, const Geo& g, int fsel, hipStream_t st);
template <class R> void launch_sub_mean(const Ens<R>& e, const Geo& g, int psel, const double* means, hipStream_t st);
// CSF extension: curvature, then the surface-tension acceleration on u*, v*, w*.
template <class R> void launch_csf(const Ens<R>& e, const Geo& g, int fsel, hipStream_t st);
template <class R>
void launch_reduce(const Ens<R>& e, const Geo& g, int fsel, int psel, int mode, double* part,
                   const long long* ncell, double* out, double* means, hipStream_t st);

// Work-groups per instance for each launch shape (host mirror of the device
// index arithmetic).  Overlapped-wave columns: waves of 64 lanes storing
// 64-2h positions of the (nx+2)(ny+2) plane, 4 waves per work-group, one
// work-group row per chunk of kz planes out of nzrange.
inline int blocks_wcol(int nx, int ny, int nzrange, int kz, int h) {
    const int plane = (nx + 2) * (ny + 2);
    const int s = 64 - 2 * h;
    const int wpp = (plane + s - 1) / s;
    return ((wpp + 3) / 4) * ((nzrange + kz - 1) / kz);
}
// The same columns with each chunk's work-groups padded to a multiple of 8
// (k_fct_zx: the chunks of one column share an XCD and its L2).
inline int blocks_wcol_pad8(int nx, int ny, int nzrange, int kz, int h) {
    const int plane = (nx + 2) * (ny + 2);
    const int s = 64 - 2 * h;
    const int wpp = (plane + s - 1) / s;
    return ((((wpp + 3) / 4) + 7) & ~7) * ((nzrange + kz - 1) / kz);
}
// k_fct_xy: overlapped waves (halo 3) over the flattened (x, z) plane of
// interior planes, padded to a multiple of 8 work-groups, one such row of
// work-groups per chunk of ch rows along y.
inline int blocks_xy(int nx, int ny, int nz, int ch) {
    const int lines = (nx + 2) * nz;
    const int wpp = (lines + 57) / 58;
    return ((((wpp + 3) / 4) + 7) & ~7) * ((ny + ch - 1) / ch);
}
// Predictor with bands of rb rows per lane (k_predict_band): overlapped waves
// of 62 stored lanes over the (nx+2) x ny/rb positions of band rows, per chunk
// of kz planes.  Used for chunks of 4 or 8 planes when every instance's ny is a
// multiple of rb (the bands then end at ny, and their upper neighbour row is
// the ghost row).
constexpr int kPredictBandRows = 4;
inline int blocks_band(int nx, int ny, int nz, int kz, int rb) {
    const int pos = (nx + 2) * ((ny + rb - 1) / rb);
    const int wpp = (pos + 61) / 62;
    return ((wpp + 3) / 4) * ((nz + kz - 1) / kz);
}
// k_predict_rho: one lane per interior cell, 256 per work-group.
inline int blocks_cells(int nx, int ny, int nz) { return int((int64_t(nx) * ny * nz + 255) / 256); }
// The LDS-streamed two-sweep kernel: FP32, chunks of kJ2DmaKz or
// kJ2DmaKzShort planes on grids of at least that many planes.
inline bool jacobi2_streams_lds(int kz, int min_nz, int esz) {
    return esz == 4 && (kz == kJ2DmaKz || kz == kJ2DmaKzShort) && min_nz >= kz;
}
// Two-sweep Jacobi tiles (every two-sweep kernel): 16 rows x 64 lanes per
// work-group of 1024 lanes over the flattened plane of (i, row band)
// positions, 60 stored positions x 14 stored rows each, so x tiles need not
// end at the row; one work-group row per chunk of kz planes.
inline int blocks_jacobi2_flat(int nx, int ny, int nz, int kz) {
    const int pos = (nx + 2) * ((ny + 13) / 14);
    return ((pos + 59) / 60) * ((nz + kz - 1) / kz);
}
// T-sweep Jacobi tiles (3 <= T <= kJacobiMaxT): 64 - 2T stored lanes x
// 18 - 2T stored rows per work-group of 1024 lanes.
inline int blocks_jacobiT(int nx, int ny, int nz, int kz, int T) {
    const int ol = 64 - 2 * T, orow = 18 - 2 * T;
    return ((nx + ol - 1) / ol) * ((ny + orow - 1) / orow) * ((nz + kz - 1) / kz);
}
// y/z line kernels: 256 lanes per work-group over the no*(nx+2) lines, the
// count rounded up to a multiple of 8 so that under the round-robin XCD
// dispatch the chunks of one line all run on the same XCD (their overlapping
// windows hit its L2); one such row of work-groups per chunk of ch cells.
inline int blocks_march(int nx, int nd, int no, int ch) {
    const int ntr = no * (nx + 2);
    return ((((ntr + 255) / 256) + 7) & ~7) * ((nd + ch - 1) / ch);
}
inline int blocks_ghost(int nx, int ny, int nz) {
    const int sx = nx + 2;
    const long n = 2L * sx * (ny + 2) + 2L * sx * nz + 2L * ny * nz;
    return int((n + 255) / 256);
}

// Allocation-free diagnostics: per-device address of the static partials
// array (looked up on first use), the work-groups a chunk length gives, and
// the two-kernel launch writing a vof_diag record to device memory.
hipError_t noalloc_scratch(double** out);
int noalloc_blocks(int nx, int ny, int nz, int kz);
template <class R>
hipError_t launch_diag_noalloc(const Inst<R>& I, int kz, double* part, vof_diag* out, hipStream_t st);

}  // namespace vof
