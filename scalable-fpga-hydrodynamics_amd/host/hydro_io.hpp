// Text output of the command-line drivers: std::ostream-compatible number formatting and
// the CSV files of hydro2d/main.cxx:166-235 (hydro3d/main.cxx:172-247).
#pragma once

#include <stddef.h>

#include <string>
#include <vector>

namespace hydro_io {

// std::ostream default formatting of a double (precision 6, general notation == %g).
std::string format_general(double v);

// std::fixed << std::setprecision(4) (== %.4f), the time label of the CSV file names.
std::string format_fixed4(double v);

// A box of interior cells: 0-based base and extent per axis (z extent 1 in 2D), and the
// MPI rank of the reference run that holds it.
struct Box {
    size_t base[3];
    size_t n[3];
    size_t rank;
};

// The reference's subdomains in the order its output walks them (hydro2d/main.cxx:166-265,
// 620-657; hydro3d/main.cxx:172-281, 667-713): an sd[0] x sd[1] (x sd[2]) split, subdomain
// s = (a*sd[1] + b)*sd[2] + c of size nb/sd per axis (the last along an axis adding
// nb % sd) at base (nb/sd)*index, held by rank s % ranks; ranks in order, each rank's
// subdomains in s order.  Requires sd[d] >= 1 and ranks >= 1; with sd[d] > nb[d] the
// leading subdomains along that axis are empty, as in the reference.
std::vector<Box> reference_order(int dim, const size_t* nb, const size_t* sd, size_t ranks);

// Write out_<it>_<t>.csv: header "x,y,rho,p,u,v" (3D "x,y,z,rho,p,u,v,w"), then one row per
// interior cell, box after box in `order` (each box in i-major order): 0-based global
// indices, rho, E, rho*u, rho*v[, rho*w], each value as std::ostream prints it.  Rows are
// formatted by `threads` threads (0 = all cores, at most 16) in blocks and written in
// order.  Returns false on an I/O error.
template <typename T>
bool write_csv(const std::string& path, const T* rhoE, const T* mom, int dim, const size_t* nb, int threads,
               const std::vector<Box>& order);

// The same with the whole domain as one box (the reference's order with one subdomain
// along Y and Z).
template <typename T>
bool write_csv(const std::string& path, const T* rhoE, const T* mom, int dim, const size_t* nb, int threads);

}  // namespace hydro_io
