/* Body of oracle/sanitize_check.c, included once per precision with `real` and SUF
 * defined: runs every entry point of the oracle on small inputs, in every boundary mode
 * and with one and two threads, so that a build with -fsanitize=address,undefined
 * reports any out-of-bounds access, use of freed memory or undefined behaviour. */

void SUF(oracle_flux2)(const real *L, const real *R, real K, real gamma, real divgamma, real *F);
void SUF(oracle_flux3)(const real *L, const real *R, real K, real gamma, real divgamma, real *F);
void SUF(oracle_launcher2d)(const real *rhoE, const real *uv, real *rhoE_next, real *uv_next, real *Dt_next,
                            real C_min, real gamma, real gmo, real divgamma, real K, size_t nb_x, size_t nb_y,
                            real DtDx, real DtDy, int bc_y, int nthreads);
void SUF(oracle_launcher3d)(const real *rhoE, const real *uvw, real *rhoE_next, real *uvw_next, real *Dt_next,
                            real C_min, real gamma, real gmo, real divgamma, real K, size_t nb_x, size_t nb_y,
                            size_t nb_z, real DtDx, real DtDy, real DtDz, int bc_y, int bc_z, int nthreads);
void SUF(oracle_ghost_fill2d)(real *rhoE, real *uv, size_t nb_x, size_t nb_y, int bc_x, int bc_y);
void SUF(oracle_ghost_fill3d)(real *rhoE, real *uvw, size_t nb_x, size_t nb_y, size_t nb_z, int bc_x, int bc_y,
                              int bc_z);
void SUF(oracle_ic2d)(real *rhoE, real *uv, size_t nb_x, size_t nb_y, size_t base_i, size_t base_j, size_t NB_X,
                      size_t NB_Y, real rho_in, real rho_out, real p_in, real p_out, real C_min, real gamma,
                      int ic_mode, double radius_frac, real *min_dt);
void SUF(oracle_ic3d)(real *rhoE, real *uvw, size_t nb_x, size_t nb_y, size_t nb_z, size_t base_i, size_t base_j,
                      size_t base_k, size_t NB_X, size_t NB_Y, size_t NB_Z, real rho_in, real rho_out, real p_in,
                      real p_out, real C_min, real gamma, int ic_mode, double radius_frac, real *min_dt);
void SUF(oracle_ic_mass_dt)(size_t NB_X, size_t NB_Y, size_t NB_Z, size_t n_sub, real C, double *mass, real *dt0);
double SUF(oracle_mass2d)(const real *rhoE, size_t nb_x, size_t nb_y);
double SUF(oracle_mass3d)(const real *rhoE, size_t nb_x, size_t nb_y, size_t nb_z);
size_t SUF(oracle_run2d)(real *rhoE, real *uv, real *scratch_rhoE, real *scratch_uv, size_t nb_x, size_t nb_y,
                         double max_t, size_t max_it, real C, int bc_x, int bc_y, real *dt, double *t_out,
                         double *traj_t, real *traj_dt, int nthreads);
size_t SUF(oracle_run3d)(real *rhoE, real *uvw, real *scratch_rhoE, real *scratch_uvw, size_t nb_x, size_t nb_y,
                         size_t nb_z, double max_t, size_t max_it, real C, int bc_x, int bc_y, int bc_z, real *dt,
                         double *t_out, double *traj_t, real *traj_dt, int nthreads);
int SUF(oracle_write_csv2d)(const char *path, const real *rhoE, const real *uv, size_t nb_x, size_t nb_y);
int SUF(oracle_write_csv3d)(const char *path, const real *rhoE, const real *uvw, size_t nb_x, size_t nb_y,
                            size_t nb_z);

static int SUF(check)(const char *csv_path)
{
    const real gamma = (real)1.4, gmo = gamma - (real)1, divgamma = (real)1 / gmo, K = (real)1.1;
    int bad = 0;

    /* Face fluxes on one pair of states. */
    const real L2[4] = { 1, (real)0.3, (real)-0.2, 2 }, R2[4] = { (real)1.2, (real)-0.1, (real)0.4, 1 };
    const real L3[5] = { 1, (real)0.3, (real)-0.2, (real)0.1, 2 }, R3[5] = { (real)1.2, (real)-0.1, (real)0.4, 0, 1 };
    real F[5];
    SUF(oracle_flux2)(L2, R2, K, gamma, divgamma, F);
    SUF(oracle_flux3)(L3, R3, K, gamma, divgamma, F);

    /* 2D: IC, ghosts, one step in each mode, a short run; mass is conserved with
     * periodic X and reflective Y. */
    {
        const size_t nx = 37, ny = 29, cells = (nx + 2) * (ny + 2);
        real *a = calloc(cells * 2, sizeof(real)), *b = calloc(cells * 2, sizeof(real));
        real *c = calloc(cells * 2, sizeof(real)), *d = calloc(cells * 2, sizeof(real));
        double *tt = calloc(64, sizeof(double));
        real *td = calloc(64, sizeof(real));
        const real C = (real)0.5, DX = (real)1 / (real)(nx + 1), DY = (real)1 / (real)(ny + 1);
        const real C_min = C * (DX < DY ? DX : DY);
        real dt = 0, dtn = 0;
        double t = 0;
        SUF(oracle_ic2d)(a, b, nx, ny, 0, 0, nx, ny, 1, (real)1.2, 10, (real)0.1, C_min, gamma, 0, 0.0, &dt);
        const double m0 = SUF(oracle_mass2d)(a, nx, ny);
        for (int bx = 0; bx < 2; ++bx)
            for (int by = 0; by < 2; ++by) {
                SUF(oracle_ghost_fill2d)(a, b, nx, ny, bx, by);
                for (int nt = 1; nt <= 2; ++nt)
                    SUF(oracle_launcher2d)(a, b, c, d, &dtn, C_min, gamma, gmo, divgamma, K, nx, ny, dt / DX, dt / DY, by,
                                           nt);
            }
        SUF(oracle_ghost_fill2d)(a, b, nx, ny, 0, 1);
        const size_t it = SUF(oracle_run2d)(a, b, c, d, nx, ny, 2.0, 40, C, 0, 1, &dt, &t, tt, td, 2);
        const double m1 = SUF(oracle_mass2d)(a, nx, ny);
        if (it != 40 || !(dt > 0) || fabs(m1 - m0) > 1e-5 * m0) {
            fprintf(stderr, "2D: it %zu dt %g mass %.17g -> %.17g\n", it, (double)dt, m0, m1);
            bad = 1;
        }
        if (SUF(oracle_write_csv2d)(csv_path, a, b, nx, ny) != 0) bad = 1;
        free(a);
        free(b);
        free(c);
        free(d);
        free(tt);
        free(td);
    }

    /* 3D, the same. */
    {
        const size_t nx = 9, ny = 11, nz = 13, cells = (nx + 2) * (ny + 2) * (nz + 2);
        real *a = calloc(cells * 2, sizeof(real)), *b = calloc(cells * 3, sizeof(real));
        real *c = calloc(cells * 2, sizeof(real)), *d = calloc(cells * 3, sizeof(real));
        double *tt = calloc(64, sizeof(double));
        real *td = calloc(64, sizeof(real));
        const real C = (real)0.4, DX = (real)1 / (real)(nx + 1), DY = (real)1 / (real)(ny + 1),
                   DZ = (real)1 / (real)(nz + 1);
        real ms = DX < DY ? DX : DY;
        ms = ms < DZ ? ms : DZ;
        const real C_min = C * ms;
        real dt = 0, dtn = 0;
        double t = 0;
        SUF(oracle_ic3d)(a, b, nx, ny, nz, 0, 0, 0, nx, ny, nz, 1, (real)1.2, 10, (real)0.1, C_min, gamma, 1, 0.2,
                         &dt);
        const double m0 = SUF(oracle_mass3d)(a, nx, ny, nz);
        for (int bx = 0; bx < 2; ++bx)
            for (int by = 0; by < 2; ++by)
                for (int bz = 0; bz < 2; ++bz) {
                    SUF(oracle_ghost_fill3d)(a, b, nx, ny, nz, bx, by, bz);
                    for (int nt = 1; nt <= 2; ++nt)
                        SUF(oracle_launcher3d)(a, b, c, d, &dtn, C_min, gamma, gmo, divgamma, K, nx, ny, nz, dt / DX,
                                               dt / DY, dt / DZ, by, bz, nt);
                }
        SUF(oracle_ghost_fill3d)(a, b, nx, ny, nz, 0, 1, 1);
        const size_t it = SUF(oracle_run3d)(a, b, c, d, nx, ny, nz, 2.0, 30, C, 0, 1, 1, &dt, &t, tt, td, 2);
        const double m1 = SUF(oracle_mass3d)(a, nx, ny, nz);
        if (it != 30 || !(dt > 0) || fabs(m1 - m0) > 1e-5 * m0) {
            fprintf(stderr, "3D: it %zu dt %g mass %.17g -> %.17g\n", it, (double)dt, m0, m1);
            bad = 1;
        }
        if (SUF(oracle_write_csv3d)(csv_path, a, b, nx, ny, nz) != 0) bad = 1;
        free(a);
        free(b);
        free(c);
        free(d);
        free(tt);
        free(td);
    }

    /* Published-size IC mass and initial Dt by subdomain counting, on a small grid. */
    double mass = 0;
    real dt0 = 0;
    SUF(oracle_ic_mass_dt)(40, 30, 0, 2, (real)0.5, &mass, &dt0);
    SUF(oracle_ic_mass_dt)(12, 10, 8, 2, (real)0.4, &mass, &dt0);
    if (!(mass > 0) || !(dt0 > 0)) bad = 1;
    return bad;
}
