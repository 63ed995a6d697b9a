// vof_vtk.h — legacy ASCII VTK writer of the volume fraction.
#pragma once
#include <cstddef>

namespace vof {

// Byte-identical to write_legacy_vtk (voftrial1.cu:907-941): header with the
// SPACING 1/(n-1) printed at the stream's default precision (6 significant
// digits, "%g"), then one value per line at precision 8 ("%.8g"), x fastest,
// ghost layers included.  Values are formatted on several threads and written
// in order.  Returns 0 on success, -1 if the file cannot be opened or written.
int write_legacy_vtk_f32(const char* path, const float* field, int nxTot, int nyTot, int nzTot);
int write_legacy_vtk_f64(const char* path, const double* field, int nxTot, int nyTot, int nzTot);

}  // namespace vof
