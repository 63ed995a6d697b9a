// vof_host.hip — C ABI (include/vof_capi.h) over the gfx950 kernels.
//
// One Ensemble<R> owns the device arena, the instance table, the launch
// shapes and one HIP stream.  The plain solver is an ensemble of one
// instance; the batched solver an ensemble of many.  A step is 15 launches
// for the reference configuration (predictor, rhs, 10 Jacobi sweeps,
// correction+BC, 3 FCT sweeps, F ghosts) against the reference's 36, and can
// be replayed as one hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/vof_capi.h"
#include "vof_adjoint.h"
#include "vof_kernels.h"
#include "vof_vtk.h"

namespace {

thread_local std::string g_err;

void set_err(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

#define VOF_HIP(call)                                                                             \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            set_err("HIP failure at %s:%d : %s", __FILE__, __LINE__, hipGetErrorString(e_));     \
            return VOF_ERR_HIP;                                                                   \
        }                                                                                         \
    } while (0)

#define VOF_TRY(call)                       \
    do {                                    \
        vof_status s_ = (call);             \
        if (s_ != VOF_OK) return s_;        \
    } while (0)

// Liquid extents of the reference IC: number of interior cells whose lower
// face coordinate float(i-1)*dx satisfies 0 <= x <= L*frac, evaluated in float
// exactly as set_init_F_kernel does (voftrial1.cu:80-91).
int ref_ic_extent(int n, double L, float divisor) {
    const float fL = static_cast<float>(L);
    const float d = fL / static_cast<float>(n);
    const float lim = fL / divisor;
    int cnt = 0;
    for (int i = 1; i <= n; ++i) {
        const float x = static_cast<float>(i - 1) * d;
        if (x >= 0.0f && x <= lim) cnt = i;
    }
    return cnt;
}

// Extension box IC: count of cells with lower face (i-1)*L/n < b, i.e.
// i-1 < b*n/L, decided with a tolerance of 1e-9 cells so that exact ties
// (b*n/L integral) are not counted.
int box_extent(int n, double L, double b) {
    const double q = b * double(n) / L;
    const int c = int(std::ceil(q - 1e-9));
    return std::max(0, std::min(n, c));
}

// Tuning cache: a JSON object, one entry per line,
//   "gfx950|200x200x200|p4|n1|c8000000|s10": {"kz_predict": 4, ..., "fct_xy_fused": 2},
constexpr int kNumTuneFields = 9;
const char* const kTuneFields[kNumTuneFields] = {"kz_predict", "kz_rhs",    "kz_jacobi",     "kz_correct",
                                                 "fct_chunk",  "jacobi_tblock", "kz_fct_x", "fct_zx_fused",
                                                 "fct_xy_fused"};

int32_t* tune_slot(vof_tuning* t, int f) {
    int32_t* s[kNumTuneFields] = {&t->kz_predict, &t->kz_rhs, &t->kz_jacobi, &t->kz_correct,
                                  &t->fct_chunk, &t->jacobi_tblock, &t->kz_fct_x, &t->fct_zx_fused,
                                  &t->fct_xy_fused};
    return s[f];
}

std::vector<std::string> read_lines(const char* path) {
    std::vector<std::string> lines;
    std::FILE* fp = std::fopen(path, "r");
    if (!fp) return lines;
    char buf[1024];
    while (std::fgets(buf, sizeof(buf), fp)) lines.emplace_back(buf);
    std::fclose(fp);
    return lines;
}

bool load_tuning(const char* path, const std::string& key, vof_tuning* t) {
    const std::string tag = "\"" + key + "\":";
    for (const auto& l : read_lines(path)) {
        if (l.find(tag) == std::string::npos) continue;
        vof_tuning r = *t;
        for (int f = 0; f < kNumTuneFields; ++f) {
            const std::string name = std::string("\"") + kTuneFields[f] + "\":";
            const size_t at = l.find(name);
            if (at == std::string::npos) return false;
            const int v = std::atoi(l.c_str() + at + name.size());
            if (v <= 0) return false;
            *tune_slot(&r, f) = v;
        }
        *t = r;
        return true;
    }
    return false;
}

void store_tuning(const char* path, const std::string& key, vof_tuning t) {
    std::vector<std::string> keep;
    const std::string tag = "\"" + key + "\":";
    for (auto l : read_lines(path)) {
        if (l.find("\": {") == std::string::npos || l.find(tag) != std::string::npos) continue;
        while (!l.empty() && (l.back() == '\n' || l.back() == ',' || l.back() == ' ')) l.pop_back();
        keep.push_back(l);
    }
    std::string entry = " " + tag + " {";
    for (int f = 0; f < kNumTuneFields; ++f) {
        char b[64];
        std::snprintf(b, sizeof(b), "%s\"%s\": %d", f ? ", " : "", kTuneFields[f], int(*tune_slot(&t, f)));
        entry += b;
    }
    entry += "}";
    keep.push_back(entry);
    const std::string tmp = std::string(path) + ".tmp";
    std::FILE* fp = std::fopen(tmp.c_str(), "w");
    if (!fp) return;
    std::fputs("{\n", fp);
    for (size_t i = 0; i < keep.size(); ++i) {
        std::fputs(keep[i].c_str(), fp);
        std::fputs(i + 1 < keep.size() ? ",\n" : "\n", fp);
    }
    std::fputs("}\n", fp);
    std::fclose(fp);
    std::rename(tmp.c_str(), path);
}

struct EnsBase {
    virtual ~EnsBase() {}
    int prec = 4;
    virtual vof_status init_extents(const std::vector<int>& ex, const std::vector<int>& ey,
                                    const std::vector<int>& ez) = 0;
    virtual vof_status step(int64_t first, int n, bool graph) = 0;
    virtual vof_status download(int inst, vof_field f, void* host, size_t count) = 0;
    virtual vof_status upload(int inst, vof_field f, const void* host, size_t count) = 0;
    virtual vof_status state_io(bool save, std::FILE* fp, int64_t* istep) = 0;  // restart file, instance 0
    virtual vof_status diagnostics(int inst, vof_diag* d) = 0;
    virtual vof_status sync() = 0;
    virtual vof_status time_steps(int64_t first, int warmup, int steps, bool graph, double* ms,
                                  double* stage_ms) = 0;
    virtual vof_status set_tuning(const vof_tuning* t) = 0;
    virtual void get_tuning(vof_tuning* t) const = 0;
    virtual vof_status debug_stage(int stage) = 0;
    virtual int64_t interior_cells() const = 0;
    virtual int64_t total_cells(int inst) const = 0;
    virtual int ninst() const = 0;
    virtual vof_status autotune(int64_t first, int steps, const char* cache, vof_tuning* out) = 0;
    virtual const vof_params& params(int inst) const = 0;
    virtual const void* field_dev(int inst, vof_field f) = 0;  // current device buffer of a stored field
    virtual hipStream_t stream() const = 0;
    virtual vof_status adj_forward(int64_t first, int n, int every, double* J) = 0;
    virtual vof_status adj_seed(vof_field f, const double* host, size_t count) = 0;
    virtual vof_status adj_backward(double gJ, vof_adj_scalars* out) = 0;
    virtual vof_status adj_download(vof_field f, double* host, size_t count) = 0;
    virtual void adj_times(double* fwd, double* bwd) const = 0;
    virtual void adj_release() = 0;
};

template <class R>
struct Ensemble : EnsBase {
    std::vector<vof_params> prm;
    std::vector<vof::Inst<R>> ih;
    vof::Ens<R> e;
    vof::Geo g_pred, g_rhs, g_jac, g_jac2, g_jacT, g_corr, g_fx, g_fy, g_fz, g_fzx, g_fxy, g_ghost, g_red, g_sub, g_cell;
    vof_tuning tun{4, 4, 4, 4, 8, 1, 4, 2, 2};
    int fsel = 0, psel = 0;
    int sweeps = 10;
    bool launch_failed = false;  // a launcher refused its launch shape (set while enqueueing)
    uint32_t flags = 0;
    hipStream_t st = nullptr;
    void* arena = nullptr;
    size_t arena_bytes = 0;
    std::vector<int*> boffs;
    double* part = nullptr;
    double* red_out = nullptr;
    double* means = nullptr;
    long long* ncell = nullptr;
    size_t part_cap = 0;
    std::map<int, hipGraphExec_t> graphs;
    std::vector<hipGraph_t> graph_defs;

    // ---- reverse mode (single instance).  State = F, u, v, w, p, each ntot R,
    // stored back to back.  store: nck checkpoints, then the K states of the
    // interval being reversed, then the final state of the forward.
    struct Adj {
        int64_t first = 0;
        int n = 0, K = 0, nck = 0;
        size_t nt = 0;
        R* store = nullptr;
        size_t store_states = 0;
        int f_fin = 0, p_fin = 0;
        double J = 0.0, sumF = 0.0;
        double* d = nullptr;  // FP64 arrays, ntot each, see adj_alloc
        double *aF[2] = {}, *au[2] = {}, *av[2] = {}, *aw[2] = {}, *ap[2] = {}, *gt = nullptr;
        double *aus = nullptr, *avs = nullptr, *aws = nullptr, *arhs = nullptr, *arho = nullptr;
        double *akap = nullptr, *afs = nullptr;  // CSF only
        double* seed[5] = {};  // F, u, v, w, p seeds of the final state (nullptr: zero)
        hipStream_t st2 = nullptr;  // taped forward passes, overlapped with the reverse
        hipEvent_t ev_start = nullptr, ev_tape[2] = {}, ev_rev[2] = {};
        R* csf_scratch = nullptr;  // smoothed F and curvature of the reverse (VOF_FLAG_CSF)
        double* seed_mem = nullptr;
        double* part = nullptr;
        double* acc = nullptr;
        int cur = 0, fcur = 0;  // which of the pairs holds the live adjoint state
        bool fwd_done = false, bwd_done = false;
        double fwd_ms = 0.0, bwd_ms = 0.0;
    } adj;

    void adj_release() override {
        Adj& a = adj;
        for (void* p : {(void*)a.store, (void*)a.d, (void*)a.seed_mem, (void*)a.part, (void*)a.acc,
                        (void*)a.csf_scratch})
            if (p) (void)hipFree(p);
        for (hipEvent_t ev : {a.ev_start, a.ev_tape[0], a.ev_tape[1], a.ev_rev[0], a.ev_rev[1]})
            if (ev) (void)hipEventDestroy(ev);
        if (a.st2) (void)hipStreamDestroy(a.st2);
        a = Adj();
    }

    ~Ensemble() override {
        adj_release();
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        for (auto g : graph_defs) (void)hipGraphDestroy(g);
        for (int* p : boffs) (void)hipFree(p);
        if (part) (void)hipFree(part);
        if (red_out) (void)hipFree(red_out);
        if (means) (void)hipFree(means);
        if (ncell) (void)hipFree(ncell);
        if (e.tab) (void)hipFree(e.tab);
        if (arena) (void)hipFree(arena);
        if (st) (void)hipStreamDestroy(st);
    }

    int ninst() const override { return int(ih.size()); }
    const vof_params& params(int i) const override { return prm[size_t(i)]; }
    int64_t interior_cells() const override {
        int64_t s = 0;
        for (auto& I : ih) s += int64_t(I.nx) * I.ny * I.nz;
        return s;
    }
    int64_t total_cells(int i) const override { return ih[size_t(i)].ntot; }

    static void derive(const vof_params& p, vof::Inst<R>& I) {
        I.nx = p.nx; I.ny = p.ny; I.nz = p.nz;
        I.sx = p.nx + 2;
        I.sy = (p.nx + 2) * (p.ny + 2);
        I.ntot = I.sy * (p.nz + 2);
        // Inputs rounded to R first; derived values formed in R (voftrial1.cu:1272-1290).
        const R Lx = R(p.Lx), Ly = R(p.Ly), Lz = R(p.Lz);
        I.dt = R(p.dt);
        I.dx = Lx / static_cast<R>(p.nx);
        I.dy = Ly / static_cast<R>(p.ny);
        I.dz = Lz / static_cast<R>(p.nz);
        I.dxi = R(1) / I.dx; I.dyi = R(1) / I.dy; I.dzi = R(1) / I.dz;
        I.vol = I.dx * I.dy * I.dz;
        I.dxdy = I.dx * I.dy;
        I.axx = I.dxi * I.dxi; I.ayy = I.dyi * I.dyi; I.azz = I.dzi * I.dzi;
        I.dtdydz = I.dt * I.dy * I.dz;
        I.dtdxdz = I.dt * I.dx * I.dz;
        I.dtdxdy = I.dt * I.dx * I.dy;
        I.rho_l = R(p.rho_l); I.rho_g = R(p.rho_g);
        I.nu_l = R(p.nu_l); I.nu_g = R(p.nu_g);
        I.gx = R(p.gx); I.gy = R(p.gy); I.gz = R(p.gz);
        I.sigma = R(p.sigma);
        I.slip = (p.flags & VOF_FLAG_NO_SLIP) ? R(-1) : R(1);
    }

    vof_status make_geo(vof::Geo& g, int kz, int (*count)(const vof::Inst<R>&, int)) {
        std::vector<int> off(ih.size() + 1, 0);
        for (size_t i = 0; i < ih.size(); ++i) off[i + 1] = off[i] + count(ih[i], kz);
        if (!g.boff) {
            VOF_HIP(hipMalloc(&g.boff, off.size() * sizeof(int)));
            boffs.push_back(g.boff);
        }
        VOF_HIP(hipMemcpy(g.boff, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
        g.nblk = off.back();
        g.kz = kz;
        g.min_nz = ih.empty() ? 0 : ih[0].nz;
        for (const auto& I : ih) g.min_nz = std::min(g.min_nz, I.nz);
        return VOF_OK;
    }

    vof_status build_geos() {
        auto col1 = [](const vof::Inst<R>& I, int kz) { return vof::blocks_wcol(I.nx, I.ny, I.nz, kz, 1); };
        // the predictor's band mapping for chunks of 4 or 8 planes when every
        // instance's ny is a multiple of the band height
        bool band = sizeof(R) == 4 && (tun.kz_predict == 4 || tun.kz_predict == 8);
        for (const auto& I : ih) band = band && (I.ny % vof::kPredictBandRows == 0);
        if (band) {
            VOF_TRY(make_geo(g_pred, tun.kz_predict, [](const vof::Inst<R>& I, int kz) {
                return vof::blocks_band(I.nx, I.ny, I.nz, kz, vof::kPredictBandRows); }));
        } else {
            VOF_TRY(make_geo(g_pred, tun.kz_predict, col1));
        }
        g_pred.rb = band ? vof::kPredictBandRows : 0;
        VOF_TRY(make_geo(g_rhs, tun.kz_rhs, col1));
        VOF_TRY(make_geo(g_jac, tun.kz_jacobi, col1));
        VOF_TRY(make_geo(g_jac2, tun.kz_jacobi, [](const vof::Inst<R>& I, int kz) {
            return vof::blocks_jacobi2_flat(I.nx, I.ny, I.nz, kz); }));
        if (tun.jacobi_tblock < 1 || tun.jacobi_tblock > vof::kJacobiMaxT) {
            set_err("jacobi_tblock %d: sweeps per Jacobi launch must be 1 .. %d", tun.jacobi_tblock,
                    vof::kJacobiMaxT);
            return VOF_ERR_INVALID;
        }
        // the T-sweep tiles depend on T; built for T = 3 when fewer are fused (unused then)
        switch (std::max(tun.jacobi_tblock, 3)) {
            case 3: VOF_TRY(make_geo(g_jacT, tun.kz_jacobi, [](const vof::Inst<R>& I, int kz) {
                        return vof::blocks_jacobiT(I.nx, I.ny, I.nz, kz, 3); })); break;
            case 4: VOF_TRY(make_geo(g_jacT, tun.kz_jacobi, [](const vof::Inst<R>& I, int kz) {
                        return vof::blocks_jacobiT(I.nx, I.ny, I.nz, kz, 4); })); break;
            default: VOF_TRY(make_geo(g_jacT, tun.kz_jacobi, [](const vof::Inst<R>& I, int kz) {
                        return vof::blocks_jacobiT(I.nx, I.ny, I.nz, kz, 5); })); break;
        }
        VOF_TRY(make_geo(g_sub, 8, col1));
        VOF_TRY(make_geo(g_corr, tun.kz_correct, col1));
        VOF_TRY(make_geo(g_red, 8, [](const vof::Inst<R>& I, int kz) {
            return vof::blocks_wcol(I.nx, I.ny, I.nz + 2, kz, 0); }));
        VOF_TRY(make_geo(g_fx, tun.kz_fct_x, [](const vof::Inst<R>& I, int kz) {
            return vof::blocks_wcol(I.nx, I.ny, I.nz, kz, 3); }));
        VOF_TRY(make_geo(g_fzx, tun.kz_fct_x, [](const vof::Inst<R>& I, int kz) {
            return vof::blocks_wcol_pad8(I.nx, I.ny, I.nz, kz, 3); }));
        VOF_TRY(make_geo(g_fxy, tun.fct_chunk, [](const vof::Inst<R>& I, int ch) {
            return vof::blocks_xy(I.nx, I.ny, I.nz, ch); }));
        VOF_TRY(make_geo(g_fy, tun.fct_chunk, [](const vof::Inst<R>& I, int ch) {
            return vof::blocks_march(I.nx, I.ny, I.nz, ch); }));
        VOF_TRY(make_geo(g_fz, tun.fct_chunk, [](const vof::Inst<R>& I, int ch) {
            return vof::blocks_march(I.nx, I.nz, I.ny, ch); }));
        VOF_TRY(make_geo(g_ghost, 1, [](const vof::Inst<R>& I, int) {
            return vof::blocks_ghost(I.nx, I.ny, I.nz); }));
        VOF_TRY(make_geo(g_cell, 1, [](const vof::Inst<R>& I, int) {
            return vof::blocks_cells(I.nx, I.ny, I.nz); }));
        const size_t need = size_t(g_red.nblk) * vof::kNumReduce;
        if (need > part_cap) {
            if (part) (void)hipFree(part);
            part = nullptr;
            VOF_HIP(hipMalloc(&part, need * sizeof(double)));
            part_cap = need;
        }
        return VOF_OK;
    }

    vof_status create(const vof_params* p, int n) {
        prec = sizeof(R);
        prm.assign(p, p + n);
        sweeps = p[0].jacobi_sweeps;
        flags = p[0].flags;
        if (p[0].jacobi_tblock > 0) tun.jacobi_tblock = p[0].jacobi_tblock;
        ih.resize(size_t(n));
        const bool csf = (flags & VOF_FLAG_CSF) != 0;
        const size_t nfield = csf ? 13 : 11;  // + smoothed F and curvature with CSF
        size_t total = 0;
        for (int i = 0; i < n; ++i) {
            derive(p[i], ih[size_t(i)]);
            total += size_t(ih[size_t(i)].ntot) * nfield;
        }
        VOF_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        const size_t bytes = total * sizeof(R);
        arena_bytes = bytes;
        if (hipMalloc(&arena, bytes) != hipSuccess) {
            arena = nullptr;
            set_err("device allocation of %zu bytes failed", bytes);
            return VOF_ERR_ALLOC;
        }
        // Every field starts at zero: never-written entries (ghost p, wall faces
        // of u*, ...) must read zero, as after the reference's cudaMemset (:1202).
        VOF_HIP(hipMemset(arena, 0, bytes));
        R* cur = static_cast<R*>(arena);
        for (auto& I : ih) {
            R** slots[13] = {&I.F[0], &I.F[1], &I.u, &I.v, &I.w, &I.us, &I.vs, &I.ws, &I.p[0], &I.p[1], &I.rhs,
                             &I.fsm, &I.kappa};
            I.fsm = nullptr;
            I.kappa = nullptr;
            for (size_t q = 0; q < nfield; ++q) {
                *slots[q] = cur;
                cur += I.ntot;
            }
        }
        VOF_HIP(hipMalloc(&e.tab, ih.size() * sizeof(vof::Inst<R>)));
        VOF_HIP(hipMemcpy(e.tab, ih.data(), ih.size() * sizeof(vof::Inst<R>), hipMemcpyHostToDevice));
        e.ninst = n;
        e.one = ih[0];
        VOF_HIP(hipMalloc(&red_out, size_t(n) * vof::kNumReduce * sizeof(double)));
        VOF_HIP(hipMalloc(&means, size_t(n) * sizeof(double)));
        VOF_HIP(hipMalloc(&ncell, size_t(n) * sizeof(long long)));
        std::vector<long long> nc(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) nc[size_t(i)] = (long long)p[i].nx * p[i].ny * p[i].nz;
        VOF_HIP(hipMemcpy(ncell, nc.data(), nc.size() * sizeof(long long), hipMemcpyHostToDevice));
        VOF_TRY(build_geos());
        VOF_HIP(hipDeviceSynchronize());
        return VOF_OK;
    }

    void drop_graphs() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
        for (auto g : graph_defs) (void)hipGraphDestroy(g);
        graph_defs.clear();
    }

    vof_status set_tuning(const vof_tuning* t) override {
        if (!t) { set_err("null tuning"); return VOF_ERR_INVALID; }
        const vof_tuning old = tun;
        if (t->kz_predict > 0) tun.kz_predict = t->kz_predict;
        if (t->kz_rhs > 0) tun.kz_rhs = t->kz_rhs;
        if (t->kz_jacobi > 0) tun.kz_jacobi = t->kz_jacobi;
        if (t->kz_correct > 0) tun.kz_correct = t->kz_correct;
        if (t->fct_chunk > 0) tun.fct_chunk = t->fct_chunk;
        if (t->jacobi_tblock > 0) tun.jacobi_tblock = t->jacobi_tblock;
        if (t->kz_fct_x > 0) tun.kz_fct_x = t->kz_fct_x;
        if (t->fct_zx_fused > 0) tun.fct_zx_fused = t->fct_zx_fused;
        if (t->fct_xy_fused > 0) tun.fct_xy_fused = t->fct_xy_fused;
        VOF_HIP(hipStreamSynchronize(st));
        drop_graphs();
        const vof_status rc = build_geos();
        if (rc != VOF_OK) {  // keep the launch tables and the tuning they were built for together
            const std::string why = g_err;
            tun = old;
            (void)build_geos();
            set_err("%s", why.c_str());
        }
        return rc;
    }
    void get_tuning(vof_tuning* t) const override { *t = tun; }

    // ---- initial condition: F = [ci <= ex][cj <= ey][ck <= ez] over every
    // stored cell (c* = index clamped to the interior), i.e. the liquid box and
    // its zero-gradient ghosts, as IC + apply_bc_F leave it.
    vof_status init_extents(const std::vector<int>& ex, const std::vector<int>& ey,
                            const std::vector<int>& ez) override {
        VOF_HIP(hipStreamSynchronize(st));
        for (size_t n = 0; n < ih.size(); ++n) {
            const auto& I = ih[n];
            std::vector<R> h(size_t(I.ntot), R(0));
            for (int k = 0; k < I.nz + 2; ++k) {
                const int ck = std::min(std::max(k, 1), I.nz);
                if (ck > ez[n]) continue;
                for (int j = 0; j < I.ny + 2; ++j) {
                    const int cj = std::min(std::max(j, 1), I.ny);
                    if (cj > ey[n]) continue;
                    for (int i = 0; i < I.nx + 2; ++i) {
                        const int ci = std::min(std::max(i, 1), I.nx);
                        if (ci <= ex[n]) h[size_t(k) * I.sy + size_t(j) * I.sx + i] = R(1);
                    }
                }
            }
            VOF_HIP(hipMemcpy(I.F[fsel], h.data(), h.size() * sizeof(R), hipMemcpyHostToDevice));
        }
        return VOF_OK;
    }

    // ---- the Jacobi sweeps: T = jacobi_tblock sweeps per launch while at
    // least T are left (T >= 3: the T-sweep tiles), then pairs (T >= 2), then
    // single sweeps; psel follows the launch ping-pong.
    int jacobi_launches() const {
        const int T = tun.jacobi_tblock;
        const int deep = (T >= 3) ? sweeps / T : 0, left = sweeps - deep * T;
        return deep + ((T >= 2) ? (left / 2 + (left & 1)) : left);
    }
    void enqueue_jacobi() {
        int left = sweeps;
        if (tun.jacobi_tblock >= 3) {
            for (; left >= tun.jacobi_tblock; left -= tun.jacobi_tblock) {
                if (!vof::launch_jacobiT(e, g_jacT, tun.jacobi_tblock, psel, st)) launch_failed = true;
                psel ^= 1;
            }
        }
        if (tun.jacobi_tblock >= 2) {
            for (; left >= 2; left -= 2) {
                vof::launch_jacobi2(e, g_jac2, psel, st);
                psel ^= 1;
            }
        }
        for (; left > 0; --left) {
            vof::launch_jacobi(e, g_jac, psel, st);
            psel ^= 1;
        }
    }

    // ---- one step, the reference loop body (voftrial1.cu:1362-1401).
    void enqueue_predict() {
        if (flags & VOF_FLAG_RHO_PREDICTOR)
            vof::launch_predict_rho(e, g_cell, fsel, st);      // density-weighted viscous term
        else
            vof::launch_predict(e, g_pred, fsel, st);          // :1364-1368
        if (flags & VOF_FLAG_CSF) vof::launch_csf(e, g_sub, fsel, st);
    }

    // ---- the FCT sweeps of step istep, in the order of :1384-1396; the last
    // one also does :1398-1401.  Where the z sweep is directly followed by the
    // x sweep (orders y-z-x and z-x-y) the pair is one k_fct_zx launch when
    // fct_zx_fused == 2 and the x chunk is one it is built for; likewise the x
    // and y sweeps of the order x-y-z (k_fct_xy, fct_xy_fused, the march
    // chunk).  mark(q) is called after each launch with the stage-time slot q
    // of that launch.
    bool zx_fused(int r) const {
        const int c = tun.kz_fct_x;
        return r != 0 && tun.fct_zx_fused >= 2 && (c == 4 || c == 8 || c == 16);
    }
    bool xy_fused(int r) const {
        const int c = tun.fct_chunk;
        return r == 0 && tun.fct_xy_fused >= 2 && (c == 4 || c == 8 || c == 16);
    }
    int fct_launches(int64_t istep) const {
        const int r = int(istep % 3);
        return (zx_fused(r) || xy_fused(r)) ? 2 : 3;
    }
    template <class Mark>
    void enqueue_fct(int64_t istep, Mark&& mark) {
        static const int order[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};
        const int r = int(istep % 3);
        for (int d = 0; d < 3; ++d) {
            const int dir = order[r][d];
            if (dir == 2 && zx_fused(r)) {  // z, then x as sweep d + 1
                if (!vof::launch_fct_zx(e, g_fzx, fsel, d + 1 == 2, st)) launch_failed = true;
                fsel ^= 1;
                ++d;
                mark(9);
                continue;
            }
            if (dir == 0 && xy_fused(r)) {  // x, then y as sweep d + 1 (never the last)
                if (!vof::launch_fct_xy(e, g_fxy, fsel, st)) launch_failed = true;
                fsel ^= 1;
                ++d;
                mark(10);
                continue;
            }
            vof::Geo& g = (dir == 0) ? g_fx : (dir == 1) ? g_fy : g_fz;
            vof::launch_fct(e, g, dir, fsel, d == 2, st);
            fsel ^= 1;
            mark(5 + dir);
        }
    }

    void enqueue_step(int64_t istep) {
        enqueue_predict();
        vof::launch_rhs(e, g_rhs, fsel, st);                  // rhs of :418-422
        enqueue_jacobi();                                     // :1371-1375
        if (flags & VOF_FLAG_REMOVE_P_MEAN) {
            vof::launch_reduce(e, g_red, fsel, psel, 1, part, ncell, red_out, means, st);
            vof::launch_sub_mean(e, g_sub, psel, means, st);
        }
        vof::launch_correct(e, g_corr, fsel, psel, st);       // :1377-1382
        enqueue_fct(istep, [](int) {});                       // :1384-1401
    }

    // A launcher that refused its launch shape (a fused FCT pair whose chunk it
    // has no kernel for) sets launch_failed while the step is enqueued; the
    // sweep pair did not run, so the step is reported failed.
    vof_status launch_status() {
        if (!launch_failed) return VOF_OK;
        launch_failed = false;
        set_err("no kernel for the fused FCT launch shape (kz_fct_x %d, fct_chunk %d)", tun.kz_fct_x, tun.fct_chunk);
        return VOF_ERR_INVALID;
    }

    vof_status replay(int64_t istep) {
        const int key = int(istep % 3) * 4 + fsel * 2 + psel;
        auto it = graphs.find(key);
        if (it == graphs.end()) {
            hipGraph_t g = nullptr;
            const int f0 = fsel, p0 = psel;
            VOF_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            enqueue_step(istep);
            const hipError_t ce = hipStreamEndCapture(st, &g);
            fsel = f0; psel = p0;  // capture recorded the step without running it
            if (ce != hipSuccess) {
                set_err("HIP failure at %s:%d : %s", __FILE__, __LINE__, hipGetErrorString(ce));
                return VOF_ERR_HIP;
            }
            if (launch_failed) {
                (void)hipGraphDestroy(g);
                return launch_status();
            }
            hipGraphExec_t ex = nullptr;
            VOF_HIP(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
            graph_defs.push_back(g);
            it = graphs.emplace(key, ex).first;
        }
        VOF_HIP(hipGraphLaunch(it->second, st));
        // advance the buffer parities exactly as enqueue_step does
        psel ^= (jacobi_launches() & 1);
        fsel ^= (fct_launches(istep) & 1);
        return VOF_OK;
    }

    vof_status step(int64_t first, int n, bool graph) override {
        if (n < 0) { set_err("negative step count"); return VOF_ERR_INVALID; }
        for (int s = 0; s < n; ++s) {
            if (graph) {
                VOF_TRY(replay(first + s));
            } else {
                enqueue_step(first + s);
                VOF_TRY(launch_status());
            }
        }
        VOF_HIP(hipGetLastError());
        if (flags & VOF_FLAG_WATCHDOG) {
            // clamp01 returns the rounding error of 1+v, so it can leave [0,1]
            // by half an ulp of 1: [-2^-25, 1+2^-23] in float, [-2^-54, 1+2^-52] in double.
            const double lo = (sizeof(R) == 4) ? -std::ldexp(1.0, -25) : -std::ldexp(1.0, -54);
            const double hi = 1.0 + ((sizeof(R) == 4) ? std::ldexp(1.0, -23) : std::ldexp(1.0, -52));
            for (int i = 0; i < ninst(); ++i) {
                vof_diag d;
                VOF_TRY(diagnostics(i, &d));
                if (d.nonfinite != 0 || !(d.f_min >= lo) || !(d.f_max <= hi)) {
                    set_err("watchdog: instance %d has %lld non-finite values, F in [%g, %g]", i,
                            (long long)d.nonfinite, d.f_min, d.f_max);
                    return VOF_ERR_NUMERICS;
                }
            }
        }
        return VOF_OK;
    }

    vof_status debug_stage(int stage) override {
        switch (stage) {
            case 0:
                if (flags & VOF_FLAG_RHO_PREDICTOR) vof::launch_predict_rho(e, g_cell, fsel, st);
                else vof::launch_predict(e, g_pred, fsel, st);
                break;
            case 1: vof::launch_rhs(e, g_rhs, fsel, st); break;
            case 2: vof::launch_jacobi(e, g_jac, psel, st); psel ^= 1; break;
            case 8: vof::launch_jacobi2(e, g_jac2, psel, st); psel ^= 1; break;
            case 3: vof::launch_correct(e, g_corr, fsel, psel, st); break;
            case 4: vof::launch_fct(e, g_fx, 0, fsel, 0, st); fsel ^= 1; break;
            case 5: vof::launch_fct(e, g_fy, 1, fsel, 0, st); fsel ^= 1; break;
            case 6: vof::launch_fct(e, g_fz, 2, fsel, 0, st); fsel ^= 1; break;
            case 7: vof::launch_clamp_F(e, g_sub, fsel, st); vof::launch_ghost_F(e, g_ghost, fsel, st); break;
            case 9:
                if (!(flags & VOF_FLAG_CSF)) { set_err("stage 9 (CSF) needs VOF_FLAG_CSF"); return VOF_ERR_INVALID; }
                vof::launch_csf(e, g_sub, fsel, st);
                break;
            case 10:
                if (!vof::launch_jacobiT(e, g_jacT, tun.jacobi_tblock, psel, st)) {
                    set_err("stage 10 needs jacobi_tblock 3 .. %d (is %d)", vof::kJacobiMaxT, tun.jacobi_tblock);
                    return VOF_ERR_INVALID;
                }
                psel ^= 1;
                break;
            default: set_err("unknown stage %d", stage); return VOF_ERR_INVALID;
        }
        VOF_HIP(hipGetLastError());
        VOF_HIP(hipStreamSynchronize(st));
        return VOF_OK;
    }

    R* field_ptr(int inst, vof_field f) {
        auto& I = ih[size_t(inst)];
        switch (f) {
            case VOF_FIELD_F: return I.F[fsel];
            case VOF_FIELD_U: return I.u;
            case VOF_FIELD_V: return I.v;
            case VOF_FIELD_W: return I.w;
            case VOF_FIELD_P: return I.p[psel];
            case VOF_FIELD_USTAR: return I.us;
            case VOF_FIELD_VSTAR: return I.vs;
            case VOF_FIELD_WSTAR: return I.ws;
            case VOF_FIELD_RHS: return I.rhs;
            case VOF_FIELD_KAPPA: return I.kappa;
            default: return nullptr;
        }
    }

    const void* field_dev(int inst, vof_field f) override {
        return (inst >= 0 && inst < ninst()) ? field_ptr(inst, f) : nullptr;
    }
    hipStream_t stream() const override { return st; }

    vof_status download(int inst, vof_field f, void* host, size_t count) override {
        if (inst < 0 || inst >= ninst()) { set_err("instance %d out of range", inst); return VOF_ERR_INVALID; }
        const auto& I = ih[size_t(inst)];
        if (!host || count != size_t(I.ntot)) {
            set_err("download: count %zu != %d stored cells", count, I.ntot);
            return VOF_ERR_INVALID;
        }
        if (f == VOF_FIELD_RHO || f == VOF_FIELD_NU) {
            // Derived from F on the host with the same float operations
            // (cal_nu_rho_kernel, :95-104).
            std::vector<R> F(count);
            VOF_TRY(download(inst, VOF_FIELD_F, F.data(), count));
            R* out = static_cast<R*>(host);
            for (size_t t = 0; t < count; ++t) {
                const R v = F[t];
                const R fc = ((R(1) + v) - std::fmax(R(1), v)) - std::fmin(R(0), v);
                out[t] = (f == VOF_FIELD_RHO) ? I.rho_g * (R(1) - fc) + I.rho_l * fc
                                              : I.nu_l * fc + I.nu_g * (R(1) - fc);
            }
            return VOF_OK;
        }
        R* src = field_ptr(inst, f);
        if (!src) { set_err("download: unknown field %d", int(f)); return VOF_ERR_INVALID; }
        VOF_HIP(hipMemcpyAsync(host, src, count * sizeof(R), hipMemcpyDeviceToHost, st));
        VOF_HIP(hipStreamSynchronize(st));
        return VOF_OK;
    }

    // Restart file.  A step reads F (both buffers: the intermediate FCT sweeps
    // read ghost cells of the second one, written in an earlier step), u, v, w
    // and p of the current buffer; u*, v*, w*, rhs and the other p buffer are
    // rewritten before they are read, their never-written entries are zero.
    // Layout: 8-byte magic, then uint32 version, precision, nx, ny, nz, nfields,
    // int64 istep, then nfields arrays of (nx+2)(ny+2)(nz+2) values each:
    // F (current), F (other), u, v, w, p.  Native byte order (little-endian).
    vof_status state_io(bool save, std::FILE* fp, int64_t* istep) override {
        static const char kMagic[8] = {'V', 'O', 'F', 'S', 'T', 'A', 'T', 'E'};
        const auto& I = ih[0];
        const size_t n = size_t(I.ntot);
        uint32_t hdr[6] = {1u, uint32_t(sizeof(R)), uint32_t(I.nx), uint32_t(I.ny), uint32_t(I.nz), 6u};
        R* fields[6] = {I.F[fsel], I.F[fsel ^ 1], I.u, I.v, I.w, I.p[psel]};
        std::vector<R> buf(n);
        if (save) {
            if (std::fwrite(kMagic, 1, 8, fp) != 8 || std::fwrite(hdr, sizeof(hdr), 1, fp) != 1 ||
                std::fwrite(istep, sizeof(int64_t), 1, fp) != 1) {
                set_err("state file: write failed");
                return VOF_ERR_INVALID;
            }
            for (R* f : fields) {
                VOF_HIP(hipMemcpyAsync(buf.data(), f, n * sizeof(R), hipMemcpyDeviceToHost, st));
                VOF_HIP(hipStreamSynchronize(st));
                if (std::fwrite(buf.data(), sizeof(R), n, fp) != n) { set_err("state file: write failed"); return VOF_ERR_INVALID; }
            }
            return VOF_OK;
        }
        char magic[8];
        uint32_t h[6];
        int64_t is = 0;
        if (std::fread(magic, 1, 8, fp) != 8 || std::memcmp(magic, kMagic, 8) != 0 ||
            std::fread(h, sizeof(h), 1, fp) != 1 || std::fread(&is, sizeof(is), 1, fp) != 1) {
            set_err("state file: not a VOF state file");
            return VOF_ERR_INVALID;
        }
        if (h[0] != hdr[0] || h[1] != hdr[1] || h[2] != hdr[2] || h[3] != hdr[3] || h[4] != hdr[4] || h[5] != hdr[5]) {
            set_err("state file: version %u, %u-byte values, grid %ux%ux%u, %u fields; this solver needs version %u, "
                    "%u-byte values, grid %ux%ux%u, %u fields", h[0], h[1], h[2], h[3], h[4], h[5], hdr[0], hdr[1],
                    hdr[2], hdr[3], hdr[4], hdr[5]);
            return VOF_ERR_INVALID;
        }
        if (is < 0) {
            set_err("state file: step index %lld is negative", (long long)is);
            return VOF_ERR_INVALID;
        }
        // the whole file must be there before the state is touched; then one
        // field at a time through one host buffer
        const long data_at = std::ftell(fp);
        if (data_at < 0 || std::fseek(fp, 0, SEEK_END) != 0) { set_err("state file: cannot seek"); return VOF_ERR_INVALID; }
        const long end = std::ftell(fp);
        if (end < 0 || uint64_t(end - data_at) != uint64_t(6) * n * sizeof(R)) {
            set_err("state file: %ld data bytes, expected %llu", end - data_at,
                    (unsigned long long)(uint64_t(6) * n * sizeof(R)));
            return VOF_ERR_INVALID;
        }
        if (std::fseek(fp, data_at, SEEK_SET) != 0) { set_err("state file: cannot seek"); return VOF_ERR_INVALID; }
        for (R* f : fields) {
            if (std::fread(buf.data(), sizeof(R), n, fp) != n) { set_err("state file: read failed"); return VOF_ERR_INVALID; }
            VOF_HIP(hipMemcpyAsync(f, buf.data(), n * sizeof(R), hipMemcpyHostToDevice, st));
            VOF_HIP(hipStreamSynchronize(st));
        }
        *istep = is;
        return VOF_OK;
    }

    vof_status upload(int inst, vof_field f, const void* host, size_t count) override {
        if (inst < 0 || inst >= ninst()) { set_err("instance %d out of range", inst); return VOF_ERR_INVALID; }
        const auto& I = ih[size_t(inst)];
        if (!host || count != size_t(I.ntot)) {
            set_err("upload: count %zu != %d stored cells", count, I.ntot);
            return VOF_ERR_INVALID;
        }
        if (f == VOF_FIELD_RHO || f == VOF_FIELD_NU || f == VOF_FIELD_KAPPA) {
            set_err("rho, nu and kappa are derived from F");
            return VOF_ERR_INVALID;
        }
        R* dst = field_ptr(inst, f);
        if (!dst) { set_err("upload: unknown field %d", int(f)); return VOF_ERR_INVALID; }
        VOF_HIP(hipMemcpyAsync(dst, host, count * sizeof(R), hipMemcpyHostToDevice, st));
        VOF_HIP(hipStreamSynchronize(st));
        return VOF_OK;
    }

    vof_status diagnostics(int inst, vof_diag* d) override {
        if (!d || inst < 0 || inst >= ninst()) { set_err("diagnostics: bad argument"); return VOF_ERR_INVALID; }
        vof::launch_reduce(e, g_red, fsel, psel, 0, part, ncell, red_out, nullptr, st);
        VOF_HIP(hipGetLastError());
        double h[vof::kNumReduce];
        VOF_HIP(hipMemcpyAsync(h, red_out + size_t(inst) * vof::kNumReduce, sizeof(h), hipMemcpyDeviceToHost, st));
        VOF_HIP(hipStreamSynchronize(st));
        const auto& I = ih[size_t(inst)];
        const double ncl = double(I.nx) * I.ny * I.nz;
        d->liquid_volume = double(I.vol) * h[0];
        d->p_mean = h[1] / ncl;
        d->residual_l2 = std::sqrt(h[2]);
        d->rhs_l2 = std::sqrt(h[3]);
        d->f_min = h[4];
        d->f_max = h[5];
        d->nonfinite = (int64_t)h[6];
        return VOF_OK;
    }

    // ---- autotuner: time candidate launch shapes stage by stage on the live
    // state (snapshotted and restored), keep the fastest per stage, cache it.
    std::string tune_key() const {
        hipDeviceProp_t prop;
        int dev = 0;
        (void)hipGetDevice(&dev);
        std::string arch = (hipGetDeviceProperties(&prop, dev) == hipSuccess) ? prop.gcnArchName : "unknown";
        const size_t colon = arch.find(':');
        if (colon != std::string::npos) arch = arch.substr(0, colon);
        char buf[160];
        int64_t cells = interior_cells();
        std::snprintf(buf, sizeof(buf), "%s|%dx%dx%d|p%d|n%d|c%lld|s%d", arch.c_str(), ih[0].nx, ih[0].ny, ih[0].nz,
                      int(sizeof(R)), ninst(), (long long)cells, sweeps);
        return buf;
    }

    vof_status autotune(int64_t first, int steps, const char* cache, vof_tuning* out) override {
        const std::string key = tune_key();
        if (cache && load_tuning(cache, key, &tun)) {
            VOF_TRY(set_tuning(&tun));
            if (out) *out = tun;
            return VOF_OK;
        }
        // Whole rotations of the sweep order, every trial at the same phase of
        // it, so that each trial times the same mix of FCT launches.
        if (steps < 1) steps = 1;
        steps = 3 * ((steps + 2) / 3);
        VOF_HIP(hipStreamSynchronize(st));
        void* snap = nullptr;
        if (hipMalloc(&snap, arena_bytes) != hipSuccess) { set_err("autotune: no memory for the state snapshot"); return VOF_ERR_ALLOC; }
        hipError_t ce = hipMemcpyAsync(snap, arena, arena_bytes, hipMemcpyDeviceToDevice, st);
        const int f0 = fsel, p0 = psel;
        vof_status rc = (ce == hipSuccess) ? VOF_OK : VOF_ERR_HIP;
        vof_tuning best = tun;
        int64_t is = first;
        auto trial = [&](const vof_tuning& t, double* stage) -> vof_status {
            vof_tuning tt = t;
            vof_status s = set_tuning(&tt);
            if (s != VOF_OK) return s;
            double ms = 0.0;
            s = time_steps(is, 1, steps, false, &ms, stage);
            is += steps + 3;  // one warm-up step, rounded up to a whole rotation
            return s;
        };
        // column kernels and the x sweep: one z-chunk length for all, best per stage
        const int kzs[] = {2, 4, 8, 16};
        double bestt[VOF_NSTAGE_TIMES];
        for (double& x : bestt) x = 1e300;
        for (int c : kzs) {
            if (rc != VOF_OK) break;
            vof_tuning t = tun;
            t.kz_predict = t.kz_rhs = t.kz_jacobi = t.kz_correct = t.kz_fct_x = c;
            double stage[VOF_NSTAGE_TIMES];
            rc = trial(t, stage);
            if (rc != VOF_OK) break;
            if (stage[0] < bestt[0]) { bestt[0] = stage[0]; best.kz_predict = c; }
            if (stage[1] < bestt[1]) { bestt[1] = stage[1]; best.kz_rhs = c; }
            if (stage[2] < bestt[2]) { bestt[2] = stage[2]; best.kz_jacobi = c; best.jacobi_tblock = tun.jacobi_tblock; }
            if (stage[4] < bestt[4]) { bestt[4] = stage[4]; best.kz_correct = c; }
            // the x chunk also sets the fused z+x launches' chunk: x, z and z+x together
            const double fx = stage[5] + stage[7] + stage[9];  // x, z, z+x
            if (fx < bestt[5]) { bestt[5] = fx; best.kz_fct_x = c; }
        }
        // Jacobi temporal-block depth with its own chunk lengths
        const int other_tb = (tun.jacobi_tblock >= 2) ? 1 : 2;
        const int kzj[] = {2, 4, 8, 16, vof::kJ2DmaKzShort, vof::kJ2DmaKz};
        for (int c : kzj) {
            if (rc != VOF_OK) break;
            vof_tuning t = best;
            t.jacobi_tblock = other_tb;
            t.kz_jacobi = c;
            double stage[VOF_NSTAGE_TIMES];
            rc = trial(t, stage);
            if (rc == VOF_OK && stage[2] < bestt[2]) { bestt[2] = stage[2]; best.kz_jacobi = c; best.jacobi_tblock = other_tb; }
        }
        // deeper blocks: T sweeps per launch (the remainder in pairs)
        for (int T = 3; T <= vof::kJacobiMaxT && T <= sweeps; ++T) {
            for (int c : {16, 25, 50}) {
                if (rc != VOF_OK) break;
                vof_tuning t = best;
                t.jacobi_tblock = T;
                t.kz_jacobi = c;
                double stage[VOF_NSTAGE_TIMES];
                rc = trial(t, stage);
                if (rc == VOF_OK && stage[2] < bestt[2]) { bestt[2] = stage[2]; best.kz_jacobi = c; best.jacobi_tblock = T; }
            }
        }
        // y/z march chunk
        const int chs[] = {4, 8, 16, 32};
        double bestm = 1e300;
        for (int c : chs) {
            if (rc != VOF_OK) break;
            vof_tuning t = best;
            t.fct_chunk = c;
            double stage[VOF_NSTAGE_TIMES];
            rc = trial(t, stage);
            const double fm = stage[6] + stage[7] + stage[10];  // y, z, x+y
            if (rc == VOF_OK && fm < bestm) { bestm = fm; best.fct_chunk = c; }
        }
        // sweep pairs in one launch or two, with the chosen chunks
        for (int pair = 0; pair < 2; ++pair) {
            double fz[2] = {1e300, 1e300};
            for (int k = 0; k < 2; ++k) {
                if (rc != VOF_OK) break;
                vof_tuning t = best;
                (pair ? t.fct_xy_fused : t.fct_zx_fused) = k + 1;
                double stage[VOF_NSTAGE_TIMES];
                rc = trial(t, stage);
                if (rc == VOF_OK) fz[k] = stage[5] + stage[6] + stage[7] + stage[9] + stage[10];
            }
            if (rc == VOF_OK) (pair ? best.fct_xy_fused : best.fct_zx_fused) = (fz[1] <= fz[0]) ? 2 : 1;
        }
        // restore the state and apply the winner
        const hipError_t re = hipMemcpyAsync(arena, snap, arena_bytes, hipMemcpyDeviceToDevice, st);
        (void)hipStreamSynchronize(st);
        (void)hipFree(snap);
        fsel = f0;
        psel = p0;
        if (re != hipSuccess && rc == VOF_OK) { set_err("autotune: state restore failed"); rc = VOF_ERR_HIP; }
        if (rc != VOF_OK) return rc;
        VOF_TRY(set_tuning(&best));
        if (cache) store_tuning(cache, key, best);
        if (out) *out = best;
        return VOF_OK;
    }

    vof_status sync() override {
        VOF_HIP(hipStreamSynchronize(st));
        return VOF_OK;
    }

    vof_status time_steps(int64_t first, int warmup, int steps, bool graph, double* ms,
                          double* stage_ms) override {
        int64_t is = first;
        VOF_TRY(step(is, warmup, graph));
        is += warmup;
        VOF_HIP(hipStreamSynchronize(st));
        hipEvent_t a, b;
        VOF_HIP(hipEventCreate(&a));
        VOF_HIP(hipEventCreate(&b));
        if (stage_ms) {
            // Per-stage device time: an event after every stage of eager steps.
            // Slot q of ev brackets stage q-1; stages follow the order of the
            // header (predictor, rhs, jacobi, p-mean, correction, fct x, y, z,
            // ghosts, fused fct z+x, fused fct x+y); slots 6.. take the FCT
            // launches of a step.
            constexpr int NS = VOF_NSTAGE_TIMES;
            for (int q = 0; q < NS; ++q) stage_ms[q] = 0.0;
            hipEvent_t ev[NS + 1];
            for (auto& x : ev) VOF_HIP(hipEventCreate(&x));
            for (int s = 0; s < steps; ++s, ++is) {
                VOF_HIP(hipEventRecord(ev[0], st));
                enqueue_predict();
                VOF_HIP(hipEventRecord(ev[1], st));
                vof::launch_rhs(e, g_rhs, fsel, st);
                VOF_HIP(hipEventRecord(ev[2], st));
                enqueue_jacobi();
                VOF_HIP(hipEventRecord(ev[3], st));
                if (flags & VOF_FLAG_REMOVE_P_MEAN) {
                    vof::launch_reduce(e, g_red, fsel, psel, 1, part, ncell, red_out, means, st);
                    vof::launch_sub_mean(e, g_sub, psel, means, st);
                }
                VOF_HIP(hipEventRecord(ev[4], st));
                vof::launch_correct(e, g_corr, fsel, psel, st);
                VOF_HIP(hipEventRecord(ev[5], st));
                int nl = 0, slot[3];
                hipError_t me = hipSuccess;
                enqueue_fct(is, [&](int q) {
                    slot[nl] = q;
                    const hipError_t x = hipEventRecord(ev[6 + nl], st);
                    if (me == hipSuccess) me = x;
                    ++nl;
                });
                VOF_HIP(me);
                VOF_TRY(launch_status());
                // F ghosts are written by the last sweep: the last slot stays ~0
                VOF_HIP(hipEventRecord(ev[NS], st));
                VOF_HIP(hipEventSynchronize(ev[NS]));
                float t = 0.f;
                for (int q = 0; q < 5; ++q) {
                    VOF_HIP(hipEventElapsedTime(&t, ev[q], ev[q + 1]));
                    stage_ms[q] += t;
                }
                hipEvent_t prev = ev[5];
                for (int d = 0; d < nl; ++d) {
                    VOF_HIP(hipEventElapsedTime(&t, prev, ev[6 + d]));
                    stage_ms[slot[d]] += t;
                    prev = ev[6 + d];
                }
                VOF_HIP(hipEventElapsedTime(&t, prev, ev[NS]));
                stage_ms[8] += t;
            }
            for (auto& x : ev) (void)hipEventDestroy(x);
            *ms = 0.0;
            for (int q = 0; q < NS; ++q) *ms += stage_ms[q];
        } else {
            VOF_HIP(hipEventRecord(a, st));
            VOF_TRY(step(is, steps, graph));
            VOF_HIP(hipEventRecord(b, st));
            VOF_HIP(hipEventSynchronize(b));
            float t = 0.f;
            VOF_HIP(hipEventElapsedTime(&t, a, b));
            *ms = t;
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return VOF_OK;
    }

    // ---------------------------------------------------------------- reverse mode
    // Store layout, in units of 5 fields: checkpoints [0, nck), the K+1 states
    // of the interval being reversed [nck, nck+K+1), the final state nck+K+1,
    // and the K step tapes (u*, v*, w*, F after the first and the second sweep)
    // [nck+K+2, nck+2K+2).
    // Store layout (states of 5 fields): checkpoints [0, nck); two interval
    // buffers of K+1 states each (the states entering the steps of an interval
    // and its end state); the final state of the forward; two tape buffers of
    // K tapes.  Buffer b = interval & 1, so that one interval can be recomputed
    // while the next one up is reversed.
    R* adj_state(size_t idx) const { return adj.store + idx * 5 * adj.nt; }
    R* adj_ival(int b, size_t m) const { return adj_state(size_t(adj.nck) + size_t(b) * size_t(adj.K + 1) + m); }
    size_t adj_final() const { return size_t(adj.nck) + 2 * size_t(adj.K + 1); }
    R* adj_tape(int b, size_t m) const { return adj_state(adj_final() + 1 + size_t(b) * size_t(adj.K) + m); }
    static size_t adj_store_states(int nck, int K) { return size_t(nck) + 4 * size_t(K) + 3; }

    vof_status save_state(R* dst) {
        const auto& I = ih[0];
        const size_t nt = adj.nt;
        vof::FieldCopy<R> fc{{I.F[fsel], I.u, I.v, I.w, I.p[psel]}, {dst, dst + nt, dst + 2 * nt, dst + 3 * nt, dst + 4 * nt}, 5};
        vof::copy_fields(fc, long(nt), st);
        return VOF_OK;
    }
    // Load a stored state into F[fs], u, v, w, p[ps] and select those buffers.
    vof_status load_state(const R* s, int fs, int ps) {
        const auto& I = ih[0];
        const size_t nt = adj.nt;
        vof::FieldCopy<R> fc{{s, s + nt, s + 2 * nt, s + 3 * nt, s + 4 * nt}, {I.F[fs], I.u, I.v, I.w, I.p[ps]}, 5};
        vof::copy_fields(fc, long(nt), st);
        fsel = fs;
        psel = ps;
        return VOF_OK;
    }

    // One step (as enqueue_step) that also tapes u*, v*, w* and F after the
    // first and the second FCT sweep into tape (5 fields).
    void enqueue_step_taped(int64_t istep, R* tape) {
        const auto& I = ih[0];
        const size_t nt = adj.nt;
        enqueue_predict();
        vof::launch_rhs(e, g_rhs, fsel, st);
        enqueue_jacobi();
        vof::launch_correct(e, g_corr, fsel, psel, st);
        static const int order[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};
        const int* od = order[int(istep % 3)];
        vof::Geo* gd[3] = {&g_fx, &g_fy, &g_fz};
        for (int d = 0; d < 3; ++d) {
            if (d == 2) {  // F[fsel^1] = after the first sweep, F[fsel] = after the second
                vof::FieldCopy<R> fc{{I.F[fsel ^ 1], I.F[fsel]}, {tape + 3 * nt, tape + 4 * nt}, 2};
                vof::copy_fields(fc, long(nt), st);
            }
            vof::launch_fct(e, *gd[od[d]], od[d], fsel, d == 2, st);
            fsel ^= 1;
        }
        vof::FieldCopy<R> fc{{I.us, I.vs, I.ws}, {tape, tape + nt, tape + 2 * nt}, 3};
        vof::copy_fields(fc, long(nt), st);
    }

    vof_status adj_forward(int64_t first, int n, int every, double* J) override {
        if (ninst() != 1) { set_err("reverse mode runs on a single instance"); return VOF_ERR_UNSUPPORTED; }
        if (flags & ~uint32_t(VOF_FLAG_WATCHDOG | VOF_FLAG_CSF)) {
            set_err("reverse mode supports flags VOF_FLAG_CSF and VOF_FLAG_WATCHDOG only (got 0x%x)", flags);
            return VOF_ERR_UNSUPPORTED;
        }
        if (n < 1 || every < 1 || !J) { set_err("adj_forward: need nsteps >= 1, checkpoint_every >= 1, J"); return VOF_ERR_INVALID; }
        const int K = std::min(every, n);
        const int nck = (n + K - 1) / K;
        const size_t nt = size_t(ih[0].ntot);
        const size_t states = adj_store_states(nck, K);
        if (adj.store && (adj.nt != nt || adj.store_states < states)) {
            (void)hipFree(adj.store);
            adj.store = nullptr;
        }
        adj.nt = nt;
        if (!adj.store) {
            const size_t bytes = states * 5 * nt * sizeof(R);
            if (hipMalloc(&adj.store, bytes) != hipSuccess) {
                adj.store = nullptr;
                set_err("reverse mode: checkpoint store of %zu bytes failed", bytes);
                return VOF_ERR_ALLOC;
            }
            adj.store_states = states;
        }
        if (!adj.part) {
            // two regions: the predictor's partial sums, accumulated over a
            // whole backward sweep, then scratch for the other reductions
            const size_t np = vof::adj_part_size(ih[0].nx, ih[0].ny, ih[0].nz);
            VOF_HIP(hipMalloc(&adj.part, 2 * np * sizeof(double)));
            VOF_HIP(hipMalloc(&adj.acc, vof::ADJ_NSCALAR * sizeof(double)));
        }
        for (double*& s : adj.seed) s = nullptr;
        adj.first = first;
        adj.n = n;
        adj.K = K;
        adj.nck = nck;
        adj.bwd_done = false;
        hipEvent_t e0, e1;
        VOF_HIP(hipEventCreate(&e0));
        VOF_HIP(hipEventCreate(&e1));
        VOF_HIP(hipEventRecord(e0, st));
        for (int s = 0; s < n; ++s) {
            if (s % K == 0) VOF_TRY(save_state(adj_state(size_t(s / K))));
            VOF_TRY(step(first + s, 1, true));
        }
        VOF_TRY(save_state(adj_state(adj_final())));
        adj.f_fin = fsel;
        adj.p_fin = psel;
        const auto& I = ih[0];
        double* scratch = adj.part + vof::adj_part_size(I.nx, I.ny, I.nz);
        const int nb = vof::adj_centroid_parts(e.tab, I, I.F[fsel], scratch, st);
        VOF_HIP(hipGetLastError());
        VOF_HIP(hipEventRecord(e1, st));
        std::vector<double> h(2 * size_t(nb));
        VOF_HIP(hipMemcpyAsync(h.data(), scratch, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        VOF_HIP(hipStreamSynchronize(st));
        float ms = 0.f;
        VOF_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        double s0 = 0.0, s1 = 0.0;
        for (int b = 0; b < nb; ++b) {
            s0 += h[2 * size_t(b)];
            s1 += h[2 * size_t(b) + 1];
        }
        if (!(s0 > 0.0)) { set_err("reverse mode: no liquid left, the centroid is undefined"); return VOF_ERR_NUMERICS; }
        adj.sumF = s0;
        adj.J = s1 / s0;
        adj.fwd_ms = ms;
        adj.fwd_done = true;
        *J = adj.J;
        return VOF_OK;
    }

    static int field_slot(vof_field f) {
        switch (f) {
            case VOF_FIELD_F: return 0;
            case VOF_FIELD_U: return 1;
            case VOF_FIELD_V: return 2;
            case VOF_FIELD_W: return 3;
            case VOF_FIELD_P: return 4;
            default: return -1;
        }
    }

    vof_status adj_seed(vof_field f, const double* host, size_t count) override {
        if (!adj.fwd_done) { set_err("adj_seed: call vof_adj_forward first"); return VOF_ERR_INVALID; }
        const int q = field_slot(f);
        if (q < 0 || !host || count != adj.nt) { set_err("adj_seed: field must be F/U/V/W/P with %zu values", adj.nt); return VOF_ERR_INVALID; }
        if (!adj.seed_mem) VOF_HIP(hipMalloc(&adj.seed_mem, 5 * adj.nt * sizeof(double)));
        adj.seed[q] = adj.seed_mem + size_t(q) * adj.nt;
        VOF_HIP(hipMemcpyAsync(adj.seed[q], host, count * sizeof(double), hipMemcpyHostToDevice, st));
        VOF_HIP(hipStreamSynchronize(st));
        return VOF_OK;
    }

    // FP64 adjoint arrays, one field each, allocated zeroed once: entries a
    // kernel never writes (the F ghosts, the unused u* wall faces) stay zero.
    vof_status adj_alloc() {
        const size_t nt = adj.nt;
        if (!adj.st2) {
            VOF_HIP(hipStreamCreateWithFlags(&adj.st2, hipStreamNonBlocking));
            for (hipEvent_t* ev : {&adj.ev_start, &adj.ev_tape[0], &adj.ev_tape[1], &adj.ev_rev[0], &adj.ev_rev[1]})
                VOF_HIP(hipEventCreateWithFlags(ev, hipEventDisableTiming));
        }
        if ((flags & VOF_FLAG_CSF) && !adj.csf_scratch) VOF_HIP(hipMalloc(&adj.csf_scratch, 2 * nt * sizeof(R)));
        const size_t kArrays = (flags & VOF_FLAG_CSF) ? 18 : 16;
        if (!adj.d) {
            VOF_HIP(hipMalloc(&adj.d, kArrays * nt * sizeof(double)));
            VOF_HIP(hipMemsetAsync(adj.d, 0, kArrays * nt * sizeof(double), st));
        }
        double* x = adj.d;
        auto next = [&]() { double* r = x; x += nt; return r; };
        for (int s = 0; s < 2; ++s) {
            adj.aF[s] = next(); adj.au[s] = next(); adj.av[s] = next(); adj.aw[s] = next(); adj.ap[s] = next();
        }
        adj.gt = next(); adj.aus = next(); adj.avs = next(); adj.aws = next();
        adj.arhs = next(); adj.arho = next();
        if (flags & VOF_FLAG_CSF) { adj.akap = next(); adj.afs = next(); }
        return VOF_OK;
    }

    // Reverse of step `istep`: S is its input state, S1 its output state and
    // T its tape (u*, v*, w*, F after the first and the second sweep), all
    // recorded by the forward kernels.  Applies the stage adjoints in reverse
    // order.  On entry aF[fcur] and au/av/aw/ap[cur] hold the adjoint of the
    // step's output; on exit those of its input.
    vof_status reverse_step(int64_t istep, const R* S, const R* S1, const R* T) {
        auto& I = ih[0];
        const size_t nt = adj.nt;
        const R *SF = S, *Su = S + nt, *Sv = S + 2 * nt, *Sw = S + 3 * nt;
        const R *us = T, *vs = T + nt, *ws = T + 2 * nt, *p1 = S1 + 4 * nt;
        static const int order[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};
        const int* od = order[int(istep % 3)];
        const R* Fin[3] = {SF, T + 3 * nt, T + 4 * nt};
        const R* vel[3] = {S1 + nt, S1 + 2 * nt, S1 + 3 * nt};
        const int c = adj.cur, x = c ^ 1;
        double* avel[3] = {adj.au[c], adj.av[c], adj.aw[c]};
        int f = adj.fcur;
        for (int d = 2; d >= 0; --d, f ^= 1)
            vof::adj_fct(e.tab, I, od[d], Fin[d], vel[od[d]], adj.aF[f], adj.aF[f ^ 1], avel[od[d]], d == 2, st);
        adj.fcur = f;
        vof::adj_zero3(adj.au[x], adj.av[x], adj.aw[x], long(nt), st);
        vof::adj_correct(e.tab, I, SF, p1, adj.au[c], adj.av[c], adj.aw[c], adj.aus, adj.avs, adj.aws,
                         adj.ap[c], adj.arho, adj.au[x], adj.av[x], adj.aw[x], st);
        if (sweeps == 0) {
            VOF_HIP(hipMemcpyAsync(adj.ap[x], adj.ap[c], nt * sizeof(double), hipMemcpyDeviceToDevice, st));
        } else {
            // two sweeps per launch (an odd count ends with one single
            // sweep); the outputs alternate so that the last lands in ap[x]
            const int nl = (sweeps + 1) / 2;
            const double* in = adj.ap[c];
            for (int l = 0; l < nl; ++l) {
                double* out = ((nl - 1 - l) % 2 == 0) ? adj.ap[x] : adj.gt;
                // rhs_bar = the sum of every sweep's input h: a pair launch adds
                // its input and its first sweep's output (launch 0 assigns), a
                // last single sweep its input only
                if (2 * l + 1 < sweeps)
                    vof::adj_jacobi2(e.tab, I, in, out, adj.arhs, l == 0, l == nl - 1, l == 0, st);
                else
                    vof::adj_jacobi(e.tab, I, in, out, adj.arhs, l == 0, 1, l == 0, 1, st);
                in = out;
            }
            vof::adj_rhs(e.tab, I, SF, us, vs, ws, adj.arhs, adj.aus, adj.avs, adj.aws, adj.arho, st);
        }
        if (flags & VOF_FLAG_CSF)
            vof::adj_csf(e.tab, I, SF, adj.csf_scratch, adj.csf_scratch + nt, adj.aus, adj.avs, adj.aws, adj.akap,
                         adj.afs, adj.aF[f], adj.arho, adj.part + vof::adj_part_size(I.nx, I.ny, I.nz), adj.acc, st);
        vof::adj_predict(e.tab, I, SF, Su, Sv, Sw, adj.aus, adj.avs, adj.aws, adj.arho, adj.au[x], adj.av[x],
                         adj.aw[x], adj.aF[f], adj.part, adj.acc, st);
        adj.cur = x;
        VOF_HIP(hipGetLastError());
        return VOF_OK;
    }

    vof_status adj_backward(double gJ, vof_adj_scalars* out) override {
        if (!adj.fwd_done) { set_err("adj_backward: call vof_adj_forward first"); return VOF_ERR_INVALID; }
        VOF_TRY(adj_alloc());
        auto& I = ih[0];
        const size_t nt = adj.nt;
        hipEvent_t e0, e1;
        VOF_HIP(hipEventCreate(&e0));
        VOF_HIP(hipEventCreate(&e1));
        VOF_HIP(hipEventRecord(e0, st));
        VOF_HIP(hipMemsetAsync(adj.acc, 0, vof::ADJ_NSCALAR * sizeof(double), st));
        // the predictor transposes of all steps add their work-group sums here
        VOF_HIP(hipMemsetAsync(adj.part, 0, vof::adj_part_size(I.nx, I.ny, I.nz) * sizeof(double), st));
        adj.cur = 0;
        adj.fcur = 0;
        vof::adj_centroid(e.tab, I, adj.J, adj.sumF, gJ, adj.seed[0], adj.aF[0], st);
        double* init[4] = {adj.au[0], adj.av[0], adj.aw[0], adj.ap[0]};
        for (int q = 0; q < 4; ++q) {
            if (adj.seed[q + 1])
                VOF_HIP(hipMemcpyAsync(init[q], adj.seed[q + 1], nt * sizeof(double), hipMemcpyDeviceToDevice, st));
            else
                VOF_HIP(hipMemsetAsync(init[q], 0, nt * sizeof(double), st));
        }
        // Per interval, last first: a taped forward pass from its checkpoint
        // (the state entering every step, and its tapes) into buffer k & 1,
        // then its steps in reverse.  The taped pass of interval k-1 runs on
        // the second stream while interval k is reversed on the first: it
        // writes only the solver's fields and buffer (k-1) & 1, which the
        // reverse of interval k+1 (waited for) was the last to read; the
        // reverse reads only the store and its own arrays.
        auto retape = [&](int k) -> vof_status {
            const int b = k & 1, s0 = k * adj.K, len = std::min(adj.n, s0 + adj.K) - s0;
            VOF_TRY(load_state(adj_state(size_t(k)), 0, 0));
            for (int m = 0; m < len; ++m) {
                VOF_TRY(save_state(adj_ival(b, size_t(m))));
                enqueue_step_taped(adj.first + s0 + m, adj_tape(b, size_t(m)));
            }
            return save_state(adj_ival(b, size_t(len)));
        };
        const hipStream_t st_rev = st;
        VOF_HIP(hipEventRecord(adj.ev_start, st_rev));  // solver fields as the caller left them
        VOF_HIP(hipStreamWaitEvent(adj.st2, adj.ev_start, 0));
        auto retape_on_st2 = [&](int k) -> vof_status {
            st = adj.st2;  // every enqueue of the taped pass goes to the second stream
            const vof_status rc = retape(k);
            st = st_rev;
            if (rc != VOF_OK) return rc;
            VOF_HIP(hipEventRecord(adj.ev_tape[k & 1], adj.st2));
            return VOF_OK;
        };
        VOF_TRY(retape_on_st2(adj.nck - 1));
        for (int k = adj.nck - 1; k >= 0; --k) {
            if (k >= 1) {
                if (k + 1 <= adj.nck - 1) VOF_HIP(hipStreamWaitEvent(adj.st2, adj.ev_rev[(k + 1) & 1], 0));
                VOF_TRY(retape_on_st2(k - 1));
            }
            VOF_HIP(hipStreamWaitEvent(st_rev, adj.ev_tape[k & 1], 0));
            const int b = k & 1, s0 = k * adj.K, len = std::min(adj.n, s0 + adj.K) - s0;
            for (int m = len - 1; m >= 0; --m)
                VOF_TRY(reverse_step(adj.first + s0 + m, adj_ival(b, size_t(m)), adj_ival(b, size_t(m) + 1),
                                     adj_tape(b, size_t(m))));
            VOF_HIP(hipEventRecord(adj.ev_rev[b], st_rev));
        }
        vof::adj_predict_sums(I.nx, I.ny, I.nz, adj.part, adj.acc, st);
        VOF_TRY(load_state(adj_state(adj_final()), adj.f_fin, adj.p_fin));
        VOF_HIP(hipEventRecord(e1, st));
        double acc[vof::ADJ_NSCALAR];
        VOF_HIP(hipMemcpyAsync(acc, adj.acc, sizeof(acc), hipMemcpyDeviceToHost, st));
        VOF_HIP(hipStreamSynchronize(st));
        float ms = 0.f;
        VOF_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        adj.bwd_ms = ms;
        adj.bwd_done = true;
        if (out) {
            out->d_rho_l = acc[vof::ADJ_RHO_L];
            out->d_rho_g = acc[vof::ADJ_RHO_G];
            out->d_nu_l = acc[vof::ADJ_NU_L];
            out->d_nu_g = acc[vof::ADJ_NU_G];
            out->d_gx = acc[vof::ADJ_GX];
            out->d_gy = acc[vof::ADJ_GY];
            out->d_gz = acc[vof::ADJ_GZ];
            out->d_sigma = acc[vof::ADJ_SIGMA];
        }
        return VOF_OK;
    }

    vof_status adj_download(vof_field f, double* host, size_t count) override {
        if (!adj.bwd_done) { set_err("adj_download: call vof_adj_backward first"); return VOF_ERR_INVALID; }
        const int q = field_slot(f);
        if (q < 0 || !host || count != adj.nt) { set_err("adj_download: field must be F/U/V/W/P with %zu values", adj.nt); return VOF_ERR_INVALID; }
        const int c = adj.cur;
        const double* src[5] = {adj.aF[adj.fcur], adj.au[c], adj.av[c], adj.aw[c], adj.ap[c]};
        VOF_HIP(hipMemcpyAsync(host, src[q], count * sizeof(double), hipMemcpyDeviceToHost, st));
        VOF_HIP(hipStreamSynchronize(st));
        if (f == VOF_FIELD_P) {
            // ghost pressures are multiplied by zero coefficients and never
            // read otherwise: they carry no sensitivity
            const auto& I = ih[0];
            double* h = host;
            for (int k = 0; k < I.nz + 2; ++k)
                for (int j = 0; j < I.ny + 2; ++j)
                    for (int i = 0; i < I.nx + 2; ++i)
                        if (i == 0 || j == 0 || k == 0 || i == I.nx + 1 || j == I.ny + 1 || k == I.nz + 1)
                            h[size_t(k) * I.sy + size_t(j) * I.sx + i] = 0.0;
        }
        return VOF_OK;
    }

    void adj_times(double* fwd, double* bwd) const override {
        if (fwd) *fwd = adj.fwd_ms;
        if (bwd) *bwd = adj.bwd_ms;
    }
};

vof_status validate(const vof_params& p) {
    if (p.nx < 1 || p.ny < 1 || p.nz < 1) { set_err("grid must have at least one cell per axis"); return VOF_ERR_INVALID; }
    const double ntot = double(p.nx + 2) * (p.ny + 2) * (p.nz + 2);
    if (ntot >= 2147483647.0) { set_err("grid too large: %.0f stored cells exceed 32-bit indexing", ntot); return VOF_ERR_INVALID; }
    if (p.precision != 4 && p.precision != 8) { set_err("precision must be 4 or 8, got %d", p.precision); return VOF_ERR_INVALID; }
    if (ntot * p.precision > 4294967295.0) {  // kernels address each field through one 32-bit buffer resource
        set_err("grid too large: a field of %.0f bytes exceeds the 4 GiB one buffer resource addresses", ntot * p.precision);
        return VOF_ERR_INVALID;
    }
    if (p.jacobi_sweeps < 0) { set_err("jacobi_sweeps must be >= 0"); return VOF_ERR_INVALID; }
    if (!(p.dt > 0.0) || !(p.Lx > 0.0) || !(p.Ly > 0.0) || !(p.Lz > 0.0)) { set_err("dt and L must be positive"); return VOF_ERR_INVALID; }
    const uint32_t known = VOF_FLAG_REMOVE_P_MEAN | VOF_FLAG_CSF | VOF_FLAG_RHO_PREDICTOR | VOF_FLAG_NO_SLIP |
                           VOF_FLAG_WATCHDOG;
    if (p.flags & ~known) {
        set_err("flags 0x%x are not defined", p.flags & ~known);
        return VOF_ERR_UNSUPPORTED;
    }
    return VOF_OK;
}

EnsBase* make_ensemble(const vof_params* p, int n, vof_status* st) {
    for (int i = 0; i < n; ++i) {
        *st = validate(p[i]);
        if (*st != VOF_OK) return nullptr;
        if (p[i].precision != p[0].precision) { set_err("all instances must share one precision"); *st = VOF_ERR_INVALID; return nullptr; }
        // one launch per stage serves every instance: the stage sequence
        // (flags, sweep count, sweeps per launch) is the ensemble's, not per instance
        if (p[i].flags != p[0].flags || p[i].jacobi_sweeps != p[0].jacobi_sweeps ||
            p[i].jacobi_tblock != p[0].jacobi_tblock) {
            set_err("instance %d: flags, jacobi_sweeps and jacobi_tblock must equal instance 0's", i);
            *st = VOF_ERR_INVALID;
            return nullptr;
        }
    }
    EnsBase* e = nullptr;
    try {
        if (p[0].precision == 8) {
            auto* x = new Ensemble<double>();
            *st = x->create(p, n);
            e = x;
        } else {
            auto* x = new Ensemble<float>();
            *st = x->create(p, n);
            e = x;
        }
    } catch (const std::bad_alloc&) {
        set_err("host allocation failed");
        *st = VOF_ERR_ALLOC;
        return nullptr;
    }
    if (*st != VOF_OK) {
        delete e;
        return nullptr;
    }
    return e;
}

vof_status init_ref_ic(EnsBase* e, int ic) {
    const int n = e->ninst();
    std::vector<int> ex(static_cast<size_t>(n)), ey(static_cast<size_t>(n)), ez(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        const vof_params& p = e->params(i);
        if (ic == 1) {
            ex[size_t(i)] = ref_ic_extent(p.nx, p.Lx, 3.0f);
            ey[size_t(i)] = ref_ic_extent(p.ny, p.Ly, 2.0f);
            ez[size_t(i)] = ref_ic_extent(p.nz, p.Lz, 3.0f);
        } else {
            ex[size_t(i)] = ey[size_t(i)] = ez[size_t(i)] = 0;
        }
    }
    return e->init_extents(ex, ey, ez);
}

}  // namespace

// Asynchronous VTK output of F.  F is copied on the solver's stream into a
// device staging slot (in step order; a device copy of one field), then into
// pinned host memory on a second stream while later steps run, and a host
// thread writes the file from there.  Two slots: a slot is reused once its
// previous file is written.  Failures of a writer thread are kept here and
// reported by the next call or by vof_output_wait.
struct AsyncVtk {
    size_t bytes = 0;
    void* dev[2] = {};
    void* host[2] = {};
    hipEvent_t staged[2] = {}, copied[2] = {};
    hipStream_t cst = nullptr;
    std::thread th[2];
    std::mutex mu;
    std::string err;
    int next = 0;

    void join_all() {
        for (auto& t : th)
            if (t.joinable()) t.join();
    }
    bool take_error(std::string* out) {
        std::lock_guard<std::mutex> g(mu);
        if (err.empty()) return false;
        *out = err;
        err.clear();
        return true;
    }
    ~AsyncVtk() {
        join_all();
        for (int i = 0; i < 2; ++i) {
            if (staged[i]) (void)hipEventDestroy(staged[i]);
            if (copied[i]) (void)hipEventDestroy(copied[i]);
            if (dev[i]) (void)hipFree(dev[i]);
            if (host[i]) (void)hipHostFree(host[i]);
        }
        if (cst) (void)hipStreamDestroy(cst);
    }
};

struct vof_solver {
    EnsBase* e;
    AsyncVtk* out = nullptr;
};
struct vof_batch {
    EnsBase* e;
};

extern "C" {

void vof_params_default(vof_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->nx = p->ny = p->nz = 200;
    p->precision = 4;
    p->Lx = p->Ly = p->Lz = 0.1;
    p->rho_l = 1000.0; p->rho_g = 50.0;
    p->nu_l = 1.0e-6; p->nu_g = 1.5e-5;
    p->sigma = 0.007;
    p->gx = 0.0; p->gy = -5.0; p->gz = 0.0;
    p->dt = 4.0e-6;
    p->jacobi_sweeps = 10;
    p->flags = 0;
    p->jacobi_tblock = 0;
}

vof_status vof_create(const vof_params* p, vof_solver** out) {
    if (!p || !out) { set_err("null argument"); return VOF_ERR_INVALID; }
    *out = nullptr;
    vof_status st = VOF_OK;
    EnsBase* e = make_ensemble(p, 1, &st);
    if (!e) return st;
    *out = new (std::nothrow) vof_solver{e};
    if (!*out) { delete e; set_err("host allocation failed"); return VOF_ERR_ALLOC; }
    return VOF_OK;
}

void vof_destroy(vof_solver* s) {
    if (!s) return;
    delete s->out;  // waits for pending files before the solver's stream goes away
    delete s->e;
    delete s;
}

#define VOF_NEED(s) do { if (!(s) || !(s)->e) { set_err("null handle"); return VOF_ERR_INVALID; } } while (0)
// Step indices are 1-based (the reference's ++istep, voftrial1.cu:1362) and
// pick the sweep rotation istep % 3: a negative one is refused.
#define VOF_NEED_STEP(i) do { if ((i) < 0) { set_err("step index %lld is negative", (long long)(i)); return VOF_ERR_INVALID; } } while (0)

vof_status vof_init_ic(vof_solver* s, int ic) {
    VOF_NEED(s);
    if (ic < 1 || ic > 3) { set_err("Invalid initial condition. Allowed values: 1, 2, 3"); return VOF_ERR_INVALID; }
    return init_ref_ic(s->e, ic);
}

vof_status vof_init_box(vof_solver* s, double bx, double by, double bz) {
    VOF_NEED(s);
    const vof_params& p = s->e->params(0);
    return s->e->init_extents({box_extent(p.nx, p.Lx, bx)}, {box_extent(p.ny, p.Ly, by)},
                              {box_extent(p.nz, p.Lz, bz)});
}

vof_status vof_step(vof_solver* s, int64_t first_istep, int nsteps) {
    VOF_NEED(s);
    VOF_NEED_STEP(first_istep);
    return s->e->step(first_istep, nsteps, false);
}
vof_status vof_step_graph(vof_solver* s, int64_t first_istep, int nsteps) {
    VOF_NEED(s);
    VOF_NEED_STEP(first_istep);
    return s->e->step(first_istep, nsteps, true);
}
vof_status vof_download(vof_solver* s, vof_field f, void* host, size_t count) {
    VOF_NEED(s);
    return s->e->download(0, f, host, count);
}
vof_status vof_upload(vof_solver* s, vof_field f, const void* host, size_t count) {
    VOF_NEED(s);
    return s->e->upload(0, f, host, count);
}
vof_status vof_diagnostics(vof_solver* s, vof_diag* d) {
    VOF_NEED(s);
    return s->e->diagnostics(0, d);
}
vof_status vof_sync(vof_solver* s) {
    VOF_NEED(s);
    return s->e->sync();
}
vof_status vof_time_steps(vof_solver* s, int64_t first_istep, int warmup, int steps, int use_graph,
                          double* ms_total, double* stage_ms) {
    VOF_NEED(s);
    VOF_NEED_STEP(first_istep);
    if (!ms_total || warmup < 0 || steps < 1) { set_err("time_steps: bad argument"); return VOF_ERR_INVALID; }
    return s->e->time_steps(first_istep, warmup, steps, use_graph != 0, ms_total, stage_ms);
}
vof_status vof_set_tuning(vof_solver* s, const vof_tuning* t) {
    VOF_NEED(s);
    return s->e->set_tuning(t);
}
vof_status vof_get_tuning(const vof_solver* s, vof_tuning* t) {
    if (!s || !s->e || !t) { set_err("null argument"); return VOF_ERR_INVALID; }
    s->e->get_tuning(t);
    return VOF_OK;
}
vof_status vof_autotune(vof_solver* s, int64_t first_istep, int steps_per_trial, const char* cache_path,
                        vof_tuning* out) {
    VOF_NEED(s);
    VOF_NEED_STEP(first_istep);
    return s->e->autotune(first_istep, steps_per_trial, cache_path, out);
}
vof_status vof_batch_autotune(vof_batch* b, int64_t first_istep, int steps_per_trial, const char* cache_path,
                              vof_tuning* out) {
    VOF_NEED(b);
    VOF_NEED_STEP(first_istep);
    return b->e->autotune(first_istep, steps_per_trial, cache_path, out);
}
vof_status vof_debug_stage(vof_solver* s, int stage) {
    VOF_NEED(s);
    return s->e->debug_stage(stage);
}

extern "C++" {
namespace {
template <class R>
vof_status diag_noalloc(const vof_params& p, const void* F, const void* u, const void* v, const void* w,
                        const void* pres, const void* rhs, vof_diag* out, hipStream_t st) {
    vof::Inst<R> I{};
    Ensemble<R>::derive(p, I);
    I.F[0] = I.F[1] = static_cast<R*>(const_cast<void*>(F));
    I.u = static_cast<R*>(const_cast<void*>(u));
    I.v = static_cast<R*>(const_cast<void*>(v));
    I.w = static_cast<R*>(const_cast<void*>(w));
    I.p[0] = I.p[1] = static_cast<R*>(const_cast<void*>(pres));
    I.rhs = static_cast<R*>(const_cast<void*>(rhs));
    // the chunk length vof_diagnostics uses, longer only where its launch
    // would need more partials than the static array holds
    int kz = 8;
    while (vof::noalloc_blocks(p.nx, p.ny, p.nz, kz) > vof::kNoAllocMaxBlocks && kz < p.nz + 2) kz *= 2;
    const int nblk = vof::noalloc_blocks(p.nx, p.ny, p.nz, kz);
    if (nblk > vof::kNoAllocMaxBlocks) {
        set_err("diagnostics_noalloc: %dx%d stored cells per plane need %d partials, more than the %d the static "
                "scratch holds", p.nx + 2, p.ny + 2, nblk, vof::kNoAllocMaxBlocks);
        return VOF_ERR_INVALID;
    }
    double* part = nullptr;
    hipError_t e = vof::noalloc_scratch(&part);
    if (e != hipSuccess) {
        set_err("diagnostics_noalloc: scratch address lookup failed: %s", hipGetErrorString(e));
        return VOF_ERR_HIP;
    }
    e = vof::launch_diag_noalloc<R>(I, kz, part, out, st);
    if (e != hipSuccess) {
        set_err("HIP failure at %s:%d : %s", __FILE__, __LINE__, hipGetErrorString(e));
        return VOF_ERR_HIP;
    }
    return VOF_OK;
}
}  // namespace
}  // extern "C++"

vof_status vof_diagnostics_noalloc(const vof_params* p, const void* F, const void* u, const void* v, const void* w,
                                   const void* pres, const void* rhs, vof_diag* out, void* stream) {
    if (!p || !F || !u || !v || !w || !pres || !rhs || !out) { set_err("diagnostics_noalloc: null argument"); return VOF_ERR_INVALID; }
    const vof_status st = validate(*p);
    if (st != VOF_OK) return st;
    hipStream_t hs = static_cast<hipStream_t>(stream);
    return (p->precision == 8) ? diag_noalloc<double>(*p, F, u, v, w, pres, rhs, out, hs)
                               : diag_noalloc<float>(*p, F, u, v, w, pres, rhs, out, hs);
}

vof_status vof_write_vtk(vof_solver* s, const char* path) {
    VOF_NEED(s);
    if (!path) { set_err("null path"); return VOF_ERR_INVALID; }
    const vof_params& p = s->e->params(0);
    const size_t n = size_t(s->e->total_cells(0));
    int rc;
    if (p.precision == 8) {
        std::vector<double> h(n);
        VOF_TRY(s->e->download(0, VOF_FIELD_F, h.data(), n));
        rc = vof::write_legacy_vtk_f64(path, h.data(), p.nx + 2, p.ny + 2, p.nz + 2);
    } else {
        std::vector<float> h(n);
        VOF_TRY(s->e->download(0, VOF_FIELD_F, h.data(), n));
        rc = vof::write_legacy_vtk_f32(path, h.data(), p.nx + 2, p.ny + 2, p.nz + 2);
    }
    if (rc != 0) { set_err("Failed to open output file: %s", path); return VOF_ERR_INVALID; }
    return VOF_OK;
}

vof_status vof_write_vtk_async(vof_solver* s, const char* path) {
    VOF_NEED(s);
    if (!path) { set_err("null path"); return VOF_ERR_INVALID; }
    const vof_params& p = s->e->params(0);
    const size_t esz = (p.precision == 8) ? 8 : 4;
    if (!s->out) {
        auto* a = new (std::nothrow) AsyncVtk();
        if (!a) { set_err("async output: out of host memory"); return VOF_ERR_ALLOC; }
        a->bytes = size_t(s->e->total_cells(0)) * esz;
        hipError_t e = hipStreamCreateWithFlags(&a->cst, hipStreamNonBlocking);
        for (int i = 0; i < 2 && e == hipSuccess; ++i) {
            if ((e = hipMalloc(&a->dev[i], a->bytes)) != hipSuccess) break;
            if ((e = hipHostMalloc(&a->host[i], a->bytes, hipHostMallocDefault)) != hipSuccess) break;
            if ((e = hipEventCreateWithFlags(&a->staged[i], hipEventDisableTiming)) != hipSuccess) break;
            e = hipEventCreateWithFlags(&a->copied[i], hipEventDisableTiming);
        }
        if (e != hipSuccess) {
            delete a;
            set_err("async output: allocation failed: %s", hipGetErrorString(e));
            return VOF_ERR_ALLOC;
        }
        s->out = a;
    }
    AsyncVtk& a = *s->out;
    const int k = a.next;
    if (a.th[k].joinable()) a.th[k].join();  // the slot's previous file is written
    std::string err;
    if (a.take_error(&err)) { set_err("%s", err.c_str()); return VOF_ERR_INVALID; }
    const void* F = s->e->field_dev(0, VOF_FIELD_F);
    hipStream_t st = s->e->stream();
    VOF_HIP(hipMemcpyAsync(a.dev[k], F, a.bytes, hipMemcpyDeviceToDevice, st));
    VOF_HIP(hipEventRecord(a.staged[k], st));
    VOF_HIP(hipStreamWaitEvent(a.cst, a.staged[k], 0));
    VOF_HIP(hipMemcpyAsync(a.host[k], a.dev[k], a.bytes, hipMemcpyDeviceToHost, a.cst));
    VOF_HIP(hipEventRecord(a.copied[k], a.cst));
    AsyncVtk* ap = s->out;
    const std::string file(path);
    const int nxT = p.nx + 2, nyT = p.ny + 2, nzT = p.nz + 2;
    const bool dbl = (p.precision == 8);
    auto write = [ap, k, file, nxT, nyT, nzT, dbl]() {
        const hipError_t e = hipEventSynchronize(ap->copied[k]);
        int rc = -1;
        if (e == hipSuccess)
            rc = dbl ? vof::write_legacy_vtk_f64(file.c_str(), static_cast<const double*>(ap->host[k]), nxT, nyT, nzT)
                     : vof::write_legacy_vtk_f32(file.c_str(), static_cast<const float*>(ap->host[k]), nxT, nyT, nzT);
        if (e != hipSuccess || rc != 0) {
            std::lock_guard<std::mutex> g(ap->mu);
            if (ap->err.empty())
                ap->err = (e != hipSuccess) ? std::string("async output: ") + hipGetErrorString(e)
                                            : "Failed to open output file: " + file;
        }
    };
    try {
        a.th[k] = std::thread(write);
    } catch (...) {  // no thread to be had: write the file here instead
        write();
    }
    a.next = k ^ 1;
    return VOF_OK;
}

vof_status vof_save_state(vof_solver* s, const char* path, int64_t istep) {
    VOF_NEED(s);
    VOF_NEED_STEP(istep);
    if (!path) { set_err("null path"); return VOF_ERR_INVALID; }
    if (s->e->ninst() != 1) { set_err("state files hold one instance"); return VOF_ERR_INVALID; }
    VOF_HIP(hipStreamSynchronize(s->e->stream()));
    const std::string tmp = std::string(path) + ".tmp";
    std::FILE* fp = std::fopen(tmp.c_str(), "wb");
    if (!fp) { set_err("Failed to open state file: %s", path); return VOF_ERR_INVALID; }
    vof_status rc = s->e->state_io(true, fp, &istep);
    if (std::fclose(fp) != 0 && rc == VOF_OK) { set_err("state file: write failed"); rc = VOF_ERR_INVALID; }
    if (rc == VOF_OK && std::rename(tmp.c_str(), path) != 0) { set_err("state file: cannot rename to %s", path); rc = VOF_ERR_INVALID; }
    if (rc != VOF_OK) std::remove(tmp.c_str());
    return rc;
}

vof_status vof_load_state(vof_solver* s, const char* path, int64_t* istep) {
    VOF_NEED(s);
    if (!path || !istep) { set_err("null argument"); return VOF_ERR_INVALID; }
    std::FILE* fp = std::fopen(path, "rb");
    if (!fp) { set_err("Failed to open state file: %s", path); return VOF_ERR_INVALID; }
    const vof_status rc = s->e->state_io(false, fp, istep);
    std::fclose(fp);
    return rc;
}

vof_status vof_output_wait(vof_solver* s) {
    VOF_NEED(s);
    if (!s->out) return VOF_OK;
    s->out->join_all();
    std::string err;
    if (s->out->take_error(&err)) { set_err("%s", err.c_str()); return VOF_ERR_INVALID; }
    return VOF_OK;
}

vof_status vof_copy_bandwidth(int64_t bytes, int reps, double* gbs) {
    if (bytes < (1 << 20) || reps < 1 || !gbs) { set_err("copy_bandwidth: bad argument"); return VOF_ERR_INVALID; }
    bytes &= ~int64_t(15);
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    vof_status rc = VOF_OK;
    auto fail = [&](hipError_t e) {
        set_err("HIP failure in copy_bandwidth: %s", hipGetErrorString(e));
        rc = VOF_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipMalloc(&a, size_t(bytes))) != hipSuccess || (e = hipMalloc(&b, size_t(bytes))) != hipSuccess) {
        fail(e);
    } else if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess ||
               (e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess ||
               (e = hipMemsetAsync(a, 0, size_t(bytes), st)) != hipSuccess) {
        fail(e);
    } else {
        const long n16 = long(bytes / 16);
        vof::launch_copy(a, b, n16, st);  // warm-up
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < reps; ++r) vof::launch_copy(a, b, n16, st);
        (void)hipEventRecord(e1, st);
        if ((e = hipEventSynchronize(e1)) != hipSuccess) {
            fail(e);
        } else {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            *gbs = 2.0 * double(bytes) * reps / (double(ms) * 1e-3) / 1e9;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

vof_status vof_stream_bandwidth(int64_t field_bytes, int nread, int nwrite, int resident, int reps,
                                double* us_per_pass) {
    if (field_bytes < (1 << 20) || field_bytes > (int64_t(1) << 32) || nread < 1 || nread > 6 || nwrite < 1 ||
        nwrite > 3 || reps < 1 || !us_per_pass) {
        set_err("stream_bandwidth: bad argument");
        return VOF_ERR_INVALID;
    }
    const int64_t pitch16 = (field_bytes + 255) / 256 * 16;  // 16-byte elements per array, 256-byte aligned
    const long n16 = long(field_bytes / 16);
    const int64_t set_bytes = int64_t(nread + nwrite) * pitch16 * 16;
    // Unless `resident`, passes rotate over enough array sets that a pass's
    // arrays were last touched 1 GiB (four times the MALL) of traffic earlier:
    // its reads come from HBM, as a step kernel's inputs do.
    const int sets = resident ? 1 : int(((int64_t(1) << 30) + set_bytes - 1) / set_bytes) + 1;
    char *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    vof_status rc = VOF_OK;
    hipError_t e = hipSuccess;
    const size_t abytes = size_t(sets) * nread * pitch16 * 16, bbytes = size_t(sets) * nwrite * pitch16 * 16;
    auto pass = [&](int q) {
        return vof::launch_stream(nread, nwrite, a + size_t(q) * nread * pitch16 * 16,
                                  b + size_t(q) * nwrite * pitch16 * 16, n16, long(pitch16), st);
    };
    if ((e = hipMalloc(&a, abytes)) != hipSuccess || (e = hipMalloc(&b, bbytes)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess ||
        (e = hipMemsetAsync(a, 0, abytes, st)) != hipSuccess) {
        set_err("HIP failure in stream_bandwidth: %s", hipGetErrorString(e));
        rc = VOF_ERR_HIP;
    } else {
        for (int q = 0; q < sets; ++q) (void)pass(q);  // warm-up: every set once
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < reps; ++r) (void)pass(r % sets);
        (void)hipEventRecord(e1, st);
        if ((e = hipEventSynchronize(e1)) != hipSuccess || (e = hipGetLastError()) != hipSuccess) {
            set_err("HIP failure in stream_bandwidth: %s", hipGetErrorString(e));
            rc = VOF_ERR_HIP;
        } else {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            *us_per_pass = double(ms) * 1e3 / reps;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

vof_status vof_write_vtk_host(const char* path, const void* data, int precision, int nxTot, int nyTot,
                              int nzTot) {
    if (!path || !data || nxTot < 2 || nyTot < 2 || nzTot < 2 || (precision != 4 && precision != 8)) {
        set_err("write_vtk_host: bad argument");
        return VOF_ERR_INVALID;
    }
    const int rc = (precision == 8)
                       ? vof::write_legacy_vtk_f64(path, static_cast<const double*>(data), nxTot, nyTot, nzTot)
                       : vof::write_legacy_vtk_f32(path, static_cast<const float*>(data), nxTot, nyTot, nzTot);
    if (rc != 0) { set_err("Failed to open output file: %s", path); return VOF_ERR_INVALID; }
    return VOF_OK;
}

int64_t vof_interior_cells(const vof_solver* s) { return (s && s->e) ? s->e->interior_cells() : 0; }
int64_t vof_total_cells(const vof_solver* s) { return (s && s->e) ? s->e->total_cells(0) : 0; }

const char* vof_status_str(vof_status st) {
    switch (st) {
        case VOF_OK: return "ok";
        case VOF_ERR_INVALID: return "invalid argument";
        case VOF_ERR_HIP: return "HIP runtime error";
        case VOF_ERR_ALLOC: return "allocation failed";
        case VOF_ERR_UNSUPPORTED: return "unsupported";
        case VOF_ERR_NUMERICS: return "numerical watchdog tripped";
    }
    return "unknown status";
}
const char* vof_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------- batch
vof_status vof_batch_create(const vof_params* params, int ninst, vof_batch** out) {
    if (!params || !out || ninst < 1) { set_err("batch_create: bad argument"); return VOF_ERR_INVALID; }
    *out = nullptr;
    vof_status st = VOF_OK;
    EnsBase* e = make_ensemble(params, ninst, &st);
    if (!e) return st;
    *out = new (std::nothrow) vof_batch{e};
    if (!*out) { delete e; set_err("host allocation failed"); return VOF_ERR_ALLOC; }
    return VOF_OK;
}
void vof_batch_destroy(vof_batch* b) {
    if (!b) return;
    delete b->e;
    delete b;
}
vof_status vof_batch_init_box(vof_batch* b, const double* bx, const double* by, const double* bz) {
    VOF_NEED(b);
    if (!bx || !by || !bz) { set_err("null box arrays"); return VOF_ERR_INVALID; }
    const int n = b->e->ninst();
    std::vector<int> ex(static_cast<size_t>(n)), ey(static_cast<size_t>(n)), ez(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        const vof_params& p = b->e->params(i);
        ex[size_t(i)] = box_extent(p.nx, p.Lx, bx[i]);
        ey[size_t(i)] = box_extent(p.ny, p.Ly, by[i]);
        ez[size_t(i)] = box_extent(p.nz, p.Lz, bz[i]);
    }
    return b->e->init_extents(ex, ey, ez);
}
vof_status vof_batch_init_ic(vof_batch* b, int ic) {
    VOF_NEED(b);
    if (ic < 1 || ic > 3) { set_err("Invalid initial condition. Allowed values: 1, 2, 3"); return VOF_ERR_INVALID; }
    return init_ref_ic(b->e, ic);
}
vof_status vof_batch_step(vof_batch* b, int64_t first_istep, int nsteps, int use_graph) {
    VOF_NEED(b);
    VOF_NEED_STEP(first_istep);
    return b->e->step(first_istep, nsteps, use_graph != 0);
}
vof_status vof_batch_download(vof_batch* b, int inst, vof_field f, void* host, size_t count) {
    VOF_NEED(b);
    return b->e->download(inst, f, host, count);
}
vof_status vof_batch_upload(vof_batch* b, int inst, vof_field f, const void* host, size_t count) {
    VOF_NEED(b);
    return b->e->upload(inst, f, host, count);
}
vof_status vof_batch_diagnostics(vof_batch* b, int inst, vof_diag* d) {
    VOF_NEED(b);
    return b->e->diagnostics(inst, d);
}
vof_status vof_batch_sync(vof_batch* b) {
    VOF_NEED(b);
    return b->e->sync();
}
vof_status vof_batch_time_steps(vof_batch* b, int64_t first_istep, int warmup, int steps, int use_graph,
                                double* ms_total) {
    VOF_NEED(b);
    VOF_NEED_STEP(first_istep);
    if (!ms_total || warmup < 0 || steps < 1) { set_err("time_steps: bad argument"); return VOF_ERR_INVALID; }
    return b->e->time_steps(first_istep, warmup, steps, use_graph != 0, ms_total, nullptr);
}
vof_status vof_batch_set_tuning(vof_batch* b, const vof_tuning* t) {
    VOF_NEED(b);
    return b->e->set_tuning(t);
}
int64_t vof_batch_interior_cells(const vof_batch* b) { return (b && b->e) ? b->e->interior_cells() : 0; }
int vof_batch_size(const vof_batch* b) { return (b && b->e) ? b->e->ninst() : 0; }

// ---------------------------------------------------------------- reverse mode
vof_status vof_adj_forward(vof_solver* s, int64_t first_istep, int nsteps, int checkpoint_every, double* J) {
    VOF_NEED(s);
    VOF_NEED_STEP(first_istep);
    return s->e->adj_forward(first_istep, nsteps, checkpoint_every, J);
}
vof_status vof_adj_seed(vof_solver* s, vof_field f, const double* host, size_t count) {
    VOF_NEED(s);
    return s->e->adj_seed(f, host, count);
}
vof_status vof_adj_backward(vof_solver* s, double gJ, vof_adj_scalars* out) {
    VOF_NEED(s);
    return s->e->adj_backward(gJ, out);
}
vof_status vof_adj_download(vof_solver* s, vof_field f, double* host, size_t count) {
    VOF_NEED(s);
    return s->e->adj_download(f, host, count);
}
vof_status vof_adj_times(const vof_solver* s, double* fwd_ms, double* bwd_ms) {
    if (!s || !s->e) { set_err("null handle"); return VOF_ERR_INVALID; }
    s->e->adj_times(fwd_ms, bwd_ms);
    return VOF_OK;
}
vof_status vof_adj_release(vof_solver* s) {
    VOF_NEED(s);
    VOF_TRY(s->e->sync());
    s->e->adj_release();
    return VOF_OK;
}

}  // extern "C"
