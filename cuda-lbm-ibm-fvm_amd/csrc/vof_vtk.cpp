// vof_vtk.cpp — multithreaded legacy VTK writer (see vof_vtk.h).
//
// std::ostream inserts a float by promoting it to double and formatting it
// with "%.*g" at the stream precision; snprintf with the same format gives the
// same bytes, and lets the 8.24 M values of a 200^3 dump be formatted in
// parallel slabs instead of one ostream call each (voftrial1.cu:934-940).
#include "vof_vtk.h"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <new>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "../../include/vof_capi.h"

namespace vof {
namespace {

// Test hook: when >= 0, at most this many formatting threads are started per
// batch and any further start fails as if the system had no thread to give.
std::atomic<int> g_spawn_limit{-1};

// Start a formatting thread; false when none can be started (the caller then
// formats that slab itself).
template <class F>
bool try_spawn(std::vector<std::thread>& th, F&& fn) {
    const int lim = g_spawn_limit.load(std::memory_order_relaxed);
    if (lim >= 0 && int(th.size()) >= lim) return false;
    try {
        th.emplace_back(std::forward<F>(fn));
        return true;
    } catch (const std::system_error&) {
        return false;
    } catch (const std::bad_alloc&) {
        return false;
    }
}

// Formats values [begin, end) into out; false (out cleared) when the host
// memory for the text cannot be had.  Never throws, so it can run on a
// formatting thread or on the caller's.
template <class T>
bool format_slab(const T* field, size_t begin, size_t end, std::string& out) noexcept {
    try {
        out.clear();
        out.reserve((end - begin) * 12);
        char buf[40];
        for (size_t t = begin; t < end; ++t) {
            const int n = std::snprintf(buf, sizeof(buf), "%.8g\n", static_cast<double>(field[t]));
            out.append(buf, static_cast<size_t>(n));
        }
        return true;
    } catch (...) {
        out.clear();
        return false;
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
    std::vector<std::string> bufs;
    try {
        bufs.resize(static_cast<size_t>(nth));
    } catch (...) {
        std::fclose(fp);
        return -1;
    }
    std::atomic<bool> fmt_ok{true};
    for (size_t base = 0; base < n && ok; base += chunk * nth) {
        // A slab whose thread cannot be started is formatted on this thread
        // at once; nothing here throws, so every started thread is joined.
        std::vector<std::thread> th;
        try {
            th.reserve(static_cast<size_t>(nth));
        } catch (...) {
        }
        for (int t = 0; t < nth; ++t) {
            const size_t b = base + chunk * t;
            const size_t e = std::min(n, b + chunk);
            if (b >= e) { bufs[t].clear(); continue; }
            std::string* out = &bufs[t];
            auto job = [field, b, e, out, &fmt_ok] {
                if (!format_slab(field, b, e, *out)) fmt_ok = false;
            };
            if (!try_spawn(th, job)) job();
        }
        for (auto& x : th) x.join();
        ok = ok && fmt_ok;
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

extern "C" void vof_debug_vtk_thread_limit(int n) { vof::g_spawn_limit.store(n < 0 ? -1 : n); }
