// Text output of the command-line drivers: std::ostream-compatible number formatting and
// the CSV files of hydro2d/main.cxx:166-235 (hydro3d/main.cxx:172-247).
#pragma once

#include <stddef.h>

#include <string>

namespace hydro_io {

// std::ostream default formatting of a double (precision 6, general notation == %g).
std::string format_general(double v);

// std::fixed << std::setprecision(4) (== %.4f), the time label of the CSV file names.
std::string format_fixed4(double v);

// Write out_<it>_<t>.csv: header "x,y,rho,p,u,v" (3D "x,y,z,rho,p,u,v,w"), then one row per
// interior cell in i-major order: 0-based indices, rho, E, rho*u, rho*v[, rho*w], each
// value as std::ostream prints it.  Rows are formatted by `threads` threads (0 = all
// cores, at most 16) in blocks and written in order.  Returns false on an I/O error.
template <typename T>
bool write_csv(const std::string& path, const T* rhoE, const T* mom, int dim, const size_t* nb, int threads);

}  // namespace hydro_io
