#include "hydro_io.hpp"

#include <stdio.h>

#include <algorithm>
#include <thread>
#include <vector>

namespace hydro_io {

std::string format_general(double v)
{
    char b[64];
    snprintf(b, sizeof(b), "%g", v);
    return b;
}

std::string format_fixed4(double v)
{
    char b[64];
    snprintf(b, sizeof(b), "%.4f", v);
    return b;
}

namespace {

// Append the rows [r0, r1) (row = linear interior index in i-major order).
template <typename T>
void format_rows(std::string& out, const T* rhoE, const T* mom, int dim, const size_t* nb, size_t r0, size_t r1)
{
    char line[256];
    const size_t nz = dim == 3 ? nb[2] : 1;
    const size_t zs = dim == 3 ? nb[2] + 2 : 1;
    const size_t ys = nb[1] + 2;
    for (size_t r = r0; r < r1; ++r) {
        const size_t k = r % nz;
        const size_t q = r / nz;
        const size_t j = q % nb[1];
        const size_t i = q / nb[1];
        int len;
        if (dim == 2) {
            const size_t n = (i + 1) * ys + (j + 1);
            len = snprintf(line, sizeof(line), "%zu,%zu,%g,%g,%g,%g\n", i, j, static_cast<double>(rhoE[2 * n]),
                           static_cast<double>(rhoE[2 * n + 1]), static_cast<double>(mom[2 * n]),
                           static_cast<double>(mom[2 * n + 1]));
        } else {
            const size_t n = ((i + 1) * ys + (j + 1)) * zs + (k + 1);
            len = snprintf(line, sizeof(line), "%zu,%zu,%zu,%g,%g,%g,%g,%g\n", i, j, k,
                           static_cast<double>(rhoE[2 * n]), static_cast<double>(rhoE[2 * n + 1]),
                           static_cast<double>(mom[3 * n]), static_cast<double>(mom[3 * n + 1]),
                           static_cast<double>(mom[3 * n + 2]));
        }
        out.append(line, static_cast<size_t>(len));
    }
}

}  // namespace

template <typename T>
bool write_csv(const std::string& path, const T* rhoE, const T* mom, int dim, const size_t* nb, int threads)
{
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return false;
    const char* header = dim == 2 ? "x,y,rho,p,u,v\n" : "x,y,z,rho,p,u,v,w\n";
    bool ok = fputs(header, f) >= 0;
    size_t rows = nb[0] * nb[1] * (dim == 3 ? nb[2] : 1);
    int nt = threads > 0 ? threads : static_cast<int>(std::thread::hardware_concurrency());
    nt = std::max(1, std::min(nt, 16));
    const size_t block = static_cast<size_t>(nt) * (1u << 16);
    std::vector<std::string> parts(static_cast<size_t>(nt));
    for (size_t b0 = 0; ok && b0 < rows; b0 += block) {
        const size_t b1 = std::min(rows, b0 + block);
        const size_t per = (b1 - b0 + static_cast<size_t>(nt) - 1) / static_cast<size_t>(nt);
        std::vector<std::thread> pool;
        for (int t = 0; t < nt; ++t) {
            const size_t r0 = std::min(b1, b0 + static_cast<size_t>(t) * per);
            const size_t r1 = std::min(b1, r0 + per);
            parts[static_cast<size_t>(t)].clear();
            if (r0 >= r1) continue;
            if (nt == 1) format_rows<T>(parts[0], rhoE, mom, dim, nb, r0, r1);
            else
                pool.emplace_back([&, t, r0, r1] {
                    format_rows<T>(parts[static_cast<size_t>(t)], rhoE, mom, dim, nb, r0, r1);
                });
        }
        for (auto& th : pool) th.join();
        for (auto& p : parts)
            if (!p.empty() && fwrite(p.data(), 1, p.size(), f) != p.size()) ok = false;
    }
    if (fclose(f) != 0) ok = false;
    return ok;
}

template bool write_csv<double>(const std::string&, const double*, const double*, int, const size_t*, int);
template bool write_csv<float>(const std::string&, const float*, const float*, int, const size_t*, int);

}  // namespace hydro_io
