// vof_vtk.cpp — multithreaded legacy VTK writer (see vof_vtk.h).
//
// std::ostream inserts a float by promoting it to double and formatting it
// with "%.*g" at the stream precision; snprintf with the same format gives the
// same bytes, and lets the 8.24 M values of a 200^3 dump be formatted in
// parallel slabs instead of one ostream call each (voftrial1.cu:934-940).
#include "vof_vtk.h"

#include <algorithm>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

namespace vof {
namespace {

template <class T>
void format_slab(const T* field, size_t begin, size_t end, std::string& out) {
    out.clear();
    out.reserve((end - begin) * 12);
    char buf[40];
    for (size_t t = begin; t < end; ++t) {
        const int n = std::snprintf(buf, sizeof(buf), "%.8g\n", static_cast<double>(field[t]));
        out.append(buf, static_cast<size_t>(n));
    }
}

template <class T>
int write_vtk(const char* path, const T* field, int nxTot, int nyTot, int nzTot) {
    std::FILE* fp = std::fopen(path, "wb");
    if (!fp) return -1;
    char hdr[512];
    const float sx = 1.0f / static_cast<float>(nxTot - 1);
    const float sy = 1.0f / static_cast<float>(nyTot - 1);
    const float sz = 1.0f / static_cast<float>(nzTot - 1);
    const int hn = std::snprintf(hdr, sizeof(hdr),
                                 "# vtk DataFile Version 3.0\n"
                                 "3D VOF field\n"
                                 "ASCII\n"
                                 "DATASET STRUCTURED_POINTS\n"
                                 "DIMENSIONS %d %d %d\n"
                                 "ORIGIN 0 0 0\n"
                                 "SPACING %g %g %g\n"
                                 "POINT_DATA %lld\n"
                                 "SCALARS VOF float 1\n"
                                 "LOOKUP_TABLE default\n",
                                 nxTot, nyTot, nzTot, static_cast<double>(sx), static_cast<double>(sy),
                                 static_cast<double>(sz),
                                 static_cast<long long>(nxTot) * nyTot * nzTot);
    bool ok = std::fwrite(hdr, 1, static_cast<size_t>(hn), fp) == static_cast<size_t>(hn);
    const size_t n = static_cast<size_t>(nxTot) * nyTot * nzTot;
    unsigned hw = std::thread::hardware_concurrency();
    const int nth = static_cast<int>(std::max(1u, std::min(hw == 0 ? 4u : hw, 32u)));
    const size_t chunk = size_t(1) << 20;  // values per slab
    std::vector<std::string> bufs(static_cast<size_t>(nth));
    for (size_t base = 0; base < n && ok; base += chunk * nth) {
        std::vector<std::thread> th;
        for (int t = 0; t < nth; ++t) {
            const size_t b = base + chunk * t;
            const size_t e = std::min(n, b + chunk);
            if (b >= e) { bufs[t].clear(); continue; }
            th.emplace_back([&, b, e, t] { format_slab(field, b, e, bufs[t]); });
        }
        for (auto& x : th) x.join();
        for (int t = 0; t < nth && ok; ++t)
            ok = std::fwrite(bufs[t].data(), 1, bufs[t].size(), fp) == bufs[t].size();
    }
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? 0 : -1;
}

}  // namespace

int write_legacy_vtk_f32(const char* path, const float* field, int nxTot, int nyTot, int nzTot) {
    return write_vtk(path, field, nxTot, nyTot, nzTot);
}
int write_legacy_vtk_f64(const char* path, const double* field, int nxTot, int nyTot, int nzTot) {
    return write_vtk(path, field, nxTot, nyTot, nzTot);
}

}  // namespace vof
