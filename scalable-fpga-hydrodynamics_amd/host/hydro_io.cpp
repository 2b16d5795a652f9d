#include "hydro_io.hpp"

#include <stdio.h>

#include <algorithm>
#include <charconv>
#include <deque>
#include <future>
#include <thread>
#include <vector>

namespace hydro_io {

std::string format_general(double v)
{
    char b[64];
    const auto r = std::to_chars(b, b + sizeof(b), v, std::chars_format::general, 6);
    return std::string(b, r.ptr);
}

std::string format_fixed4(double v)
{
    char b[64];
    snprintf(b, sizeof(b), "%.4f", v);
    return b;
}

namespace {

// Longest row: 3 indices of up to 20 digits, 5 values of up to 13 characters ("-1.23457e-308"
// or "-nan"), separators and the newline.
constexpr size_t kMaxRow = 3 * 20 + 5 * 16 + 8;

// std::to_chars with chars_format::general and precision 6 is specified to produce what
// printf("%g") produces in the C locale, which is what std::ostream's default formatting
// of a double produces (hydro2d/main.cxx:210-229 streams the values with the default
// precision 6); indices are plain decimal as ostream prints size_t.
inline char* put_value(char* p, double v)
{
    return std::to_chars(p, p + 32, v, std::chars_format::general, 6).ptr;
}

inline char* put_index(char* p, size_t v) { return std::to_chars(p, p + 24, v).ptr; }

// Rows [r0, r1) (row = linear interior index in i-major order) as text.
template <typename T>
std::string format_rows(const T* rhoE, const T* mom, int dim, const size_t* nb, size_t r0, size_t r1)
{
    std::string out;
    out.resize((r1 - r0) * kMaxRow);
    char* p = out.data();
    const size_t nz = dim == 3 ? nb[2] : 1;
    const size_t zs = dim == 3 ? nb[2] + 2 : 1;
    const size_t ys = nb[1] + 2;
    for (size_t r = r0; r < r1; ++r) {
        const size_t k = r % nz;
        const size_t q = r / nz;
        const size_t j = q % nb[1];
        const size_t i = q / nb[1];
        p = put_index(p, i);
        *p++ = ',';
        p = put_index(p, j);
        *p++ = ',';
        if (dim == 2) {
            const size_t n = (i + 1) * ys + (j + 1);
            p = put_value(p, static_cast<double>(rhoE[2 * n]));
            *p++ = ',';
            p = put_value(p, static_cast<double>(rhoE[2 * n + 1]));
            *p++ = ',';
            p = put_value(p, static_cast<double>(mom[2 * n]));
            *p++ = ',';
            p = put_value(p, static_cast<double>(mom[2 * n + 1]));
        } else {
            p = put_index(p, k);
            *p++ = ',';
            const size_t n = ((i + 1) * ys + (j + 1)) * zs + (k + 1);
            p = put_value(p, static_cast<double>(rhoE[2 * n]));
            *p++ = ',';
            p = put_value(p, static_cast<double>(rhoE[2 * n + 1]));
            *p++ = ',';
            p = put_value(p, static_cast<double>(mom[3 * n]));
            *p++ = ',';
            p = put_value(p, static_cast<double>(mom[3 * n + 1]));
            *p++ = ',';
            p = put_value(p, static_cast<double>(mom[3 * n + 2]));
        }
        *p++ = '\n';
    }
    out.resize(static_cast<size_t>(p - out.data()));
    return out;
}

}  // namespace

template <typename T>
bool write_csv(const std::string& path, const T* rhoE, const T* mom, int dim, const size_t* nb, int threads)
{
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return false;
    const char* header = dim == 2 ? "x,y,rho,p,u,v\n" : "x,y,z,rho,p,u,v,w\n";
    bool ok = fputs(header, f) >= 0;
    const size_t rows = nb[0] * nb[1] * (dim == 3 ? nb[2] : 1);
    int nt = threads > 0 ? threads : static_cast<int>(std::thread::hardware_concurrency());
    nt = std::max(1, std::min(nt, 16));
    // Blocks of rows are formatted by up to 2*nt concurrent tasks while this thread
    // writes the finished blocks in order.
    constexpr size_t kBlockRows = size_t(1) << 17;
    const size_t nblocks = (rows + kBlockRows - 1) / kBlockRows;
    auto task = [&](size_t b) {
        const size_t r0 = b * kBlockRows;
        return format_rows<T>(rhoE, mom, dim, nb, r0, std::min(rows, r0 + kBlockRows));
    };
    std::deque<std::future<std::string>> inflight;
    size_t next = 0;
    for (size_t b = 0; b < nblocks; ++b) {
        while (next < nblocks && inflight.size() < static_cast<size_t>(2 * nt)) {
            inflight.push_back(nt == 1 ? std::async(std::launch::deferred, task, next)
                                       : std::async(std::launch::async, task, next));
            ++next;
        }
        const std::string s = inflight.front().get();
        inflight.pop_front();
        if (ok && fwrite(s.data(), 1, s.size(), f) != s.size()) ok = false;
    }
    if (fclose(f) != 0) ok = false;
    return ok;
}

template bool write_csv<double>(const std::string&, const double*, const double*, int, const size_t*, int);
template bool write_csv<float>(const std::string&, const float*, const float*, int, const size_t*, int);

}  // namespace hydro_io
