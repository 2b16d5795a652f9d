/*
 * fp_model_study.c -- runs the oracle (hydro_oracle.c, FP64) under another
 * floating-point model, to measure how far the reference's published FPGA runs
 * (hydro2d/reference_run/*.res, hydro3d/reference_run/*.res) are from the strict scheme.
 * Test infrastructure only, like the oracle itself; built by `make -C oracle study`.
 *
 * The model is chosen at build time (contraction: -ffp-contract=off or =fast -mfma) and
 * at run time (the x86 rounding mode of every thread: nearest, toward zero, down, up).
 * The run is the reference's blast and step loop (hydro2d/main.cxx:686-815,
 * hydro3d/main.cxx:742-876); the mass is compute_total_mass (hydro2d/main.cxx:237-265)
 * per rank of an X split into n_ranks slabs, the ranks' doubles summed, so the printed
 * change in mass is the reference's |m0 - m1| / cells (hydro2d/main.cxx:267-278).
 *
 * usage: fp_study_<build> NB_X NB_Y NB_Z(0 = 2D) C max_t max_it round(0 nearest,
 *        1 toward zero, 2 down, 3 up) n_ranks out_prefix
 * writes <out_prefix>.txt (summary line, then "it t Dt" per iteration) and
 * <out_prefix>.state (the final rhoE then momentum arrays, raw float64).
 */
#include <fenv.h>

#include "hydro_oracle.c"

static double slab_mass(const double *rhoE, size_t nx, size_t ny, size_t nz, size_t n_ranks)
{
    const size_t plane = (ny + 2) * (nz ? nz + 2 : 1);
    const size_t sub = nx / n_ranks;
    double total = 0.0;
    for (size_t r = 0; r < n_ranks; ++r) {
        const size_t rows = sub + (r == n_ranks - 1 ? nx % n_ranks : 0);
        const double *base = rhoE + 2 * r * sub * plane; /* plane r*sub is the slab's lower neighbour */
        total += nz ? oracle_mass3d_f64(base, rows, ny, nz) : oracle_mass2d_f64(base, rows, ny);
    }
    return total;
}

int main(int argc, char **argv)
{
    if (argc != 10) {
        fprintf(stderr, "usage: %s NB_X NB_Y NB_Z C max_t max_it round n_ranks out_prefix\n", argv[0]);
        return 2;
    }
    const size_t nx = strtoul(argv[1], NULL, 10), ny = strtoul(argv[2], NULL, 10), nz = strtoul(argv[3], NULL, 10);
    const double C = atof(argv[4]), max_t = atof(argv[5]);
    const size_t max_it = strtoul(argv[6], NULL, 10);
    const int rmode = atoi(argv[7]);
    const size_t n_ranks = strtoul(argv[8], NULL, 10);
    const char *out = argv[9];
    const int modes[4] = { FE_TONEAREST, FE_TOWARDZERO, FE_DOWNWARD, FE_UPWARD };
    if (rmode < 0 || rmode > 3 || n_ranks < 1 || nx < n_ranks || ny < 1) return 2;
    const int dim = nz ? 3 : 2;
    const size_t cells = (nx + 2) * (ny + 2) * (nz ? nz + 2 : 1);
    double *rhoE = calloc(2 * cells, sizeof(double)), *mom = calloc((size_t)dim * cells, sizeof(double));
    double *s_r = calloc(2 * cells, sizeof(double)), *s_m = calloc((size_t)dim * cells, sizeof(double));
    double *tt = calloc(max_it ? max_it : 1, sizeof(double)), *td = calloc(max_it ? max_it : 1, sizeof(double));
    if (!rhoE || !mom || !s_r || !s_m || !tt || !td) return 3;

    const double gamma = 1.4;
    double ms = mmin(1.0 / (double)(nx + 1), 1.0 / (double)(ny + 1));
    if (nz) ms = mmin(ms, 1.0 / (double)(nz + 1));
    const double C_min = C * ms;
    double dt0;
    if (nz)
        oracle_ic3d_f64(rhoE, mom, nx, ny, nz, 0, 0, 0, nx, ny, nz, 1.0, 1.2, 10.0, 0.1, C_min, gamma, 0, 0.0, &dt0);
    else
        oracle_ic2d_f64(rhoE, mom, nx, ny, 0, 0, nx, ny, 1.0, 1.2, 10.0, 0.1, C_min, gamma, 0, 0.0, &dt0);
    const double m0 = slab_mass(rhoE, nx, ny, nz, n_ranks);

    /* the rounding mode is per thread: set it in every OpenMP worker and here */
#pragma omp parallel
    fesetround(modes[rmode]);
    fesetround(modes[rmode]);
    double t, dt = dt0;
    const size_t it = nz ? oracle_run3d_f64(rhoE, mom, s_r, s_m, nx, ny, nz, max_t, max_it, C, BC_PERIODIC,
                                            BC_REFLECTIVE, BC_REFLECTIVE, &dt, &t, tt, td, 0)
                         : oracle_run2d_f64(rhoE, mom, s_r, s_m, nx, ny, max_t, max_it, C, BC_PERIODIC, BC_REFLECTIVE,
                                            &dt, &t, tt, td, 0);
#pragma omp parallel
    fesetround(FE_TONEAREST);
    fesetround(FE_TONEAREST);

    const double m1 = slab_mass(rhoE, nx, ny, nz, n_ranks);
    const double change = fabs(m0 - m1) / ((double)nx * (double)ny * (double)(nz ? nz : 1));
    char path[4096];
    snprintf(path, sizeof path, "%s.txt", out);
    FILE *f = fopen(path, "w");
    if (!f) return 4;
    fprintf(f, "it %zu t %.10e dt %.10e m0 %.10e m1 %.10e change %.3e\n", it, t, dt, m0, m1, change);
    for (size_t k = 0; k < it; ++k) fprintf(f, "%zu %.10e %.10e\n", k, tt[k], td[k]);
    fclose(f);
    snprintf(path, sizeof path, "%s.state", out);
    FILE *g = fopen(path, "wb");
    if (!g) return 4;
    fwrite(rhoE, sizeof(double), 2 * cells, g);
    fwrite(mom, sizeof(double), (size_t)dim * cells, g);
    fclose(g);
    printf("it %zu t %.6e change in mass per cell %.3e\n", it, t, change);
    return 0;
}
