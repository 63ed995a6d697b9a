// Sanitizer driver for the multithreaded VTK writer (tests/test_sanitizers.py):
// FP32 and FP64 fields large enough to be formatted on several threads, with
// the thread-start limit at none, one and unlimited (the writer falls back to
// formatting on the calling thread), and an unwritable path.
#include <cstdio>
#include <vector>

#include "../../cuda-lbm-ibm-fvm_amd/csrc/vof_vtk.h"

extern "C" void vof_debug_vtk_thread_limit(int n);

int main(int argc, char** argv) {
    const char* path = argc > 1 ? argv[1] : "vtk_driver.vtk";
    const int nx = 67, ny = 45, nz = 39;
    std::vector<float> f(static_cast<size_t>(nx) * ny * nz);
    std::vector<double> d(f.size());
    for (size_t i = 0; i < f.size(); ++i) {
        f[i] = static_cast<float>((i * 2654435761u) % 1000003u) / 1000003.0f;
        d[i] = static_cast<double>(f[i]) / 3.0;
    }
    for (int limit : {-1, 0, 1, 3}) {
        vof_debug_vtk_thread_limit(limit);
        if (vof::write_legacy_vtk_f32(path, f.data(), nx, ny, nz) != 0) return 1;
        if (vof::write_legacy_vtk_f64(path, d.data(), nx, ny, nz) != 0) return 1;
    }
    vof_debug_vtk_thread_limit(-1);
    if (vof::write_legacy_vtk_f32("/nonexistent-dir/x.vtk", f.data(), nx, ny, nz) != -1) return 1;
    std::printf("vtk driver done\n");
    return 0;
}
