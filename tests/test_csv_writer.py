"""The drivers' CSV writer (host/hydro_io.cpp) against printf's %g, the formatting
std::ostream applies by default (hydro2d/main.cxx:166-235): a harness built with g++
writes the same random states (special values included) through hydro_io::write_csv and
through snprintf, and the files must be byte-identical, for 2D and 3D, several 131072-row
blocks per file, one and many formatting threads."""
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
HOST = ROOT / "scalable-fpga-hydrodynamics_amd" / "host"

HARNESS = r"""
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>
#include "hydro_io.hpp"

int main(int argc, char** argv) {
    const int dim = atoi(argv[1]);
    size_t nb[3] = { (size_t)atoll(argv[2]), (size_t)atoll(argv[3]), dim == 3 ? (size_t)atoll(argv[4]) : 1 };
    const int threads = atoi(argv[5]);
    const size_t nz = dim == 3 ? nb[2] + 2 : 1;
    const size_t cells = (nb[0] + 2) * (nb[1] + 2) * nz;
    std::vector<double> re(2 * cells), mo(dim * cells);
    std::mt19937_64 rng(42);
    const double specials[] = { 0.0, -0.0, 1e-5, 9.999995e-5, 123456.5, 999999.5, 1e300, 5e-324, NAN, -NAN,
                                INFINITY, -INFINITY, 1.2, 0.25, 1e6 };
    auto draw = [&]() {
        const uint64_t u = rng();
        if (u % 17 == 0) return specials[(u >> 8) % (sizeof(specials) / sizeof(double))];
        if (u % 5 == 0) { double v; uint64_t w = rng(); memcpy(&v, &w, 8); return v; }
        return std::ldexp((double)(rng() >> 11) * 0x1p-53, (int)(rng() % 60) - 30) * ((u & 2) ? -1.0 : 1.0);
    };
    for (auto& v : re) v = draw();
    for (auto& v : mo) v = draw();
    // Optional: sdx sdy sdz ranks -> the reference's subdomain/rank row order.
    std::vector<hydro_io::Box> order;
    if (argc > 9) {
        const size_t sd[3] = { (size_t)atoll(argv[6]), (size_t)atoll(argv[7]), (size_t)atoll(argv[8]) };
        order = hydro_io::reference_order(dim, nb, sd, (size_t)atoll(argv[9]));
        if (!hydro_io::write_csv<double>("got.csv", re.data(), mo.data(), dim, nb, threads, order)) return 2;
    } else {
        order.push_back(hydro_io::Box{ { 0, 0, 0 }, { nb[0], nb[1], dim == 3 ? nb[2] : 1 }, 0 });
        if (!hydro_io::write_csv<double>("got.csv", re.data(), mo.data(), dim, nb, threads)) return 2;
    }
    FILE* f = fopen("want.csv", "wb");
    fputs(dim == 2 ? "x,y,rho,p,u,v\n" : "x,y,z,rho,p,u,v,w\n", f);
    const size_t ys = nb[1] + 2;
    for (const hydro_io::Box& bx : order)
    for (size_t i = bx.base[0]; i < bx.base[0] + bx.n[0]; ++i)
        for (size_t j = bx.base[1]; j < bx.base[1] + bx.n[1]; ++j)
            for (size_t k = bx.base[2]; k < bx.base[2] + bx.n[2]; ++k) {
                if (dim == 2) {
                    const size_t n = (i + 1) * ys + (j + 1);
                    fprintf(f, "%zu,%zu,%g,%g,%g,%g\n", i, j, re[2 * n], re[2 * n + 1], mo[2 * n], mo[2 * n + 1]);
                } else {
                    const size_t n = ((i + 1) * ys + (j + 1)) * nz + (k + 1);
                    fprintf(f, "%zu,%zu,%zu,%g,%g,%g,%g,%g\n", i, j, k, re[2 * n], re[2 * n + 1], mo[3 * n],
                            mo[3 * n + 1], mo[3 * n + 2]);
                }
            }
    fclose(f);
    return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("csv")
    src = d / "h.cpp"
    src.write_text(HARNESS)
    exe = d / "h"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", f"-I{HOST}", str(src), str(HOST / "hydro_io.cpp"), "-o",
                    str(exe)], check=True, capture_output=True)
    return exe


@pytest.mark.parametrize("dim,nb,threads", [(2, (300, 500, 1), 4), (2, (7, 3, 1), 1), (3, (40, 30, 120), 16),
                                            (3, (1, 1, 1), 0)])
def test_csv_writer_matches_printf(harness, tmp_path, dim, nb, threads):
    subprocess.run([str(harness), str(dim), *map(str, nb), str(threads)], cwd=tmp_path, check=True, timeout=120)
    got = (tmp_path / "got.csv").read_bytes()
    want = (tmp_path / "want.csv").read_bytes()
    assert got.count(b"\n") == nb[0] * nb[1] * (nb[2] if dim == 3 else 1) + 1
    assert got == want


@pytest.mark.parametrize("dim,nb,threads,split", [(2, (301, 499, 1), 4, (3, 4, 1, 5)), (2, (7, 3, 1), 2, (2, 5, 1, 3)),
                                                  (3, (41, 30, 119), 16, (2, 3, 4, 7)),
                                                  (2, (600, 500, 1), 8, (2, 1, 1, 2))])
def test_csv_writer_in_subdomain_rank_order(harness, tmp_path, dim, nb, threads, split):
    """Rows in the reference's order for a split and a rank count (each rank's
    subdomains in turn), blocks of 131072 rows crossing subdomain boundaries, and
    subdomains left empty where a split exceeds the grid (7 x 3 in 2 x 5): the same bytes
    as printf walking the same boxes."""
    subprocess.run([str(harness), str(dim), *map(str, nb), str(threads), *map(str, split)], cwd=tmp_path, check=True,
                   timeout=120)
    got = (tmp_path / "got.csv").read_bytes()
    assert got.count(b"\n") == nb[0] * nb[1] * (nb[2] if dim == 3 else 1) + 1
    assert got == (tmp_path / "want.csv").read_bytes()


@pytest.fixture(scope="module")
def harness_sanitized(tmp_path_factory):
    d = tmp_path_factory.mktemp("csv_san")
    src = d / "h.cpp"
    src.write_text(HARNESS)
    exe = d / "h"
    subprocess.run(["g++", "-O1", "-g", "-fno-omit-frame-pointer", "-std=c++17", "-pthread",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=all", f"-I{HOST}", str(src),
                    str(HOST / "hydro_io.cpp"), "-o", str(exe)], check=True, capture_output=True)
    return exe


def test_csv_writer_clean_under_asan_and_ubsan(harness_sanitized, tmp_path):
    """Several 131072-row blocks formatted by 8 threads in a 3 x 2 split on 4 ranks, under
    AddressSanitizer and UndefinedBehaviorSanitizer: the same bytes as printf and no report."""
    import os

    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(harness_sanitized), "2", "400", "900", "1", "8", "3", "2", "1", "4"], cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=300)
    report = r.stdout + r.stderr
    assert r.returncode == 0, report
    assert "Sanitizer" not in report and "runtime error" not in report, report
    assert (tmp_path / "got.csv").read_bytes() == (tmp_path / "want.csv").read_bytes()
