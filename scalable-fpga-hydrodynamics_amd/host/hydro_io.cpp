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

// Rows [r0, r1) of the file (row = position in the box order, each box i-major) as text;
// `first` holds the file row at which each box starts.
template <typename T>
std::string format_rows(const T* rhoE, const T* mom, int dim, const size_t* nb, const std::vector<Box>& order,
                        const std::vector<size_t>& first, size_t r0, size_t r1)
{
    std::string out;
    out.resize((r1 - r0) * kMaxRow);
    char* p = out.data();
    const size_t zs = dim == 3 ? nb[2] + 2 : 1;
    const size_t ys = nb[1] + 2;
    size_t bi = static_cast<size_t>(std::upper_bound(first.begin(), first.end(), r0) - first.begin()) - 1;
    for (size_t r = r0; r < r1; ++r) {
        while (r >= first[bi + 1]) ++bi;
        const Box& b = order[bi];
        const size_t local = r - first[bi];
        const size_t k = b.base[2] + local % b.n[2];
        const size_t q = local / b.n[2];
        const size_t j = b.base[1] + q % b.n[1];
        const size_t i = b.base[0] + q / b.n[1];
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

std::vector<Box> reference_order(int dim, const size_t* nb, const size_t* sd, size_t ranks)
{
    const size_t sdz = dim == 3 ? sd[2] : 1, nsd = sd[0] * sd[1] * sdz;
    const size_t nz = dim == 3 ? nb[2] : 1;
    auto axis = [](size_t n, size_t s, size_t idx, size_t& base, size_t& len) {
        const size_t size = n / s;
        base = size * idx;
        len = size + (idx == s - 1 ? n % s : 0);
    };
    std::vector<Box> order;
    order.reserve(nsd);
    for (size_t r = 0; r < ranks; ++r)
        for (size_t s = r; s < nsd; s += ranks) {
            Box b{};
            axis(nb[0], sd[0], s / (sd[1] * sdz), b.base[0], b.n[0]);
            axis(nb[1], sd[1], (s / sdz) % sd[1], b.base[1], b.n[1]);
            axis(nz, sdz, s % sdz, b.base[2], b.n[2]);
            b.rank = r;
            order.push_back(b);
        }
    return order;
}

template <typename T>
bool write_csv(const std::string& path, const T* rhoE, const T* mom, int dim, const size_t* nb, int threads)
{
    const Box whole{ { 0, 0, 0 }, { nb[0], nb[1], dim == 3 ? nb[2] : 1 }, 0 };
    return write_csv<T>(path, rhoE, mom, dim, nb, threads, std::vector<Box>{ whole });
}

template <typename T>
bool write_csv(const std::string& path, const T* rhoE, const T* mom, int dim, const size_t* nb, int threads,
               const std::vector<Box>& order)
{
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return false;
    const char* header = dim == 2 ? "x,y,rho,p,u,v\n" : "x,y,z,rho,p,u,v,w\n";
    bool ok = fputs(header, f) >= 0;
    std::vector<size_t> first(order.size() + 1, 0);
    for (size_t b = 0; b < order.size(); ++b) first[b + 1] = first[b] + order[b].n[0] * order[b].n[1] * order[b].n[2];
    const size_t rows = first.back();
    int nt = threads > 0 ? threads : static_cast<int>(std::thread::hardware_concurrency());
    nt = std::max(1, std::min(nt, 16));
    // Blocks of rows are formatted by up to 2*nt concurrent tasks while this thread
    // writes the finished blocks in order.
    constexpr size_t kBlockRows = size_t(1) << 17;
    const size_t nblocks = (rows + kBlockRows - 1) / kBlockRows;
    auto task = [&](size_t b) {
        const size_t r0 = b * kBlockRows;
        return format_rows<T>(rhoE, mom, dim, nb, order, first, r0, std::min(rows, r0 + kBlockRows));
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
template bool write_csv<double>(const std::string&, const double*, const double*, int, const size_t*, int,
                                const std::vector<Box>&);
template bool write_csv<float>(const std::string&, const float*, const float*, int, const size_t*, int,
                               const std::vector<Box>&);

}  // namespace hydro_io
