// Sanitizer driver for the CPU oracle (tests/test_sanitizers.py): whole steps
// in FP32 and FP64 with every physics flag, field get/set, stage calls,
// volume accounting, diagnostics and the VTK writer, on a small odd-sized grid.
// Exit status 0 when every call returns success; the sanitizers report the rest.
#include "../../include/vof_capi.h"

#include <cstdio>
#include <cstdlib>
#include <vector>

extern "C" {
void* orc_create(const vof_params* p);
void orc_destroy(void* h);
int orc_init_box(void* h, double bx, double by, double bz);
int orc_step(void* h, long long first_istep, int nsteps);
int orc_stage(void* h, int stage);
int orc_stats_enable(void* h, int on);
int orc_stats(void* h, double* out);
int orc_get(void* h, int f, double* out);
int orc_set(void* h, int f, const double* in);
int orc_diag(void* h, vof_diag* d);
long long orc_total(void* h);
int orc_write_vtk_float(const char* path, const float* field, int nxTot, int nyTot, int nzTot);
}

static int fail(const char* what) {
    std::fprintf(stderr, "driver: %s failed\n", what);
    return 1;
}

int main(int argc, char** argv) {
    const char* vtk = argc > 1 ? argv[1] : "oracle_driver.vtk";
    const unsigned flag_sets[] = {0u, VOF_FLAG_CSF | VOF_FLAG_REMOVE_P_MEAN, VOF_FLAG_RHO_PREDICTOR | VOF_FLAG_NO_SLIP};
    for (int prec : {4, 8}) {
        for (unsigned flags : flag_sets) {
            vof_params p{};
            p.nx = 13; p.ny = 11; p.nz = 9; p.precision = prec;
            p.Lx = 0.1; p.Ly = 0.09; p.Lz = 0.08;
            p.rho_l = 1000; p.rho_g = 50; p.nu_l = 1e-6; p.nu_g = 1.5e-5; p.sigma = 0.07;
            p.gx = 0; p.gy = -5; p.gz = -9.81; p.dt = 2e-4;
            p.jacobi_sweeps = 10; p.flags = flags; p.jacobi_tblock = 0;
            void* h = orc_create(&p);
            if (!h) return fail("orc_create");
            if (orc_init_box(h, 0.05, 0.05, 0.04)) return fail("orc_init_box");
            const long long n = orc_total(h);
            std::vector<double> a(static_cast<size_t>(n));
            if (orc_get(h, VOF_FIELD_U, a.data())) return fail("orc_get");
            for (long long c = 0; c < n; ++c) a[c] = 0.3 * ((c * 7919) % 13 - 6) / 6.0;
            if (orc_set(h, VOF_FIELD_U, a.data())) return fail("orc_set");
            if (orc_stats_enable(h, 1)) return fail("orc_stats_enable");
            if (orc_step(h, 1, 4)) return fail("orc_step");
            double st[6];
            if (orc_stats(h, st)) return fail("orc_stats");
            for (int s = 0; s < 4; ++s)
                if (orc_stage(h, s)) return fail("orc_stage");
            vof_diag d;
            if (orc_diag(h, &d)) return fail("orc_diag");
            if (orc_get(h, VOF_FIELD_F, a.data())) return fail("orc_get F");
            std::vector<float> f(a.begin(), a.end());
            if (orc_write_vtk_float(vtk, f.data(), p.nx + 2, p.ny + 2, p.nz + 2)) return fail("orc_write_vtk_float");
            orc_destroy(h);
        }
    }
    std::printf("oracle driver done\n");
    return 0;
}
