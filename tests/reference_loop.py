"""The reference's host step loop, driven by the CPU oracle, for comparing the command-
line drivers' stdout and CSV files.

Follows hydro2d/main.cxx:659-836 (hydro3d/main.cxx:715-898): initial CSV at it=0 when
w > 0, then per iteration: ghost fill, DtDx = Dt / DX, launcher, swap, Dt <- Dt_next,
write when `w && it && it % w == 0 && it < MAX_IT && t < MAX_T` (labelled with the
current it and t), print `[it=%4lu] ... t=%.4e (Dt=%.3e)`, ++it, t += Dt; finally the
mass line and the last CSV labelled with the final it and t.
"""
from __future__ import annotations

import pathlib

import numpy as np


def run(O, nb, max_t, max_it, w, out_prefix, dtype=np.float64, bc=None, sd=None, ranks=1):
    """sd, ranks: the subdomain split and MPI rank count whose CSV row order and mass sums
    to follow (hydro2d/main.cxx:166-265); None: the whole domain on one rank."""
    dim = len(nb)
    C = 0.5 if dim == 2 else 0.4
    bc = list(bc) if bc is not None else [0] + [1] * (dim - 1)
    if dim == 2:
        rhoE, mom, dt = O.ic2d(*nb, dtype=dtype, C_cfl=C)
    else:
        rhoE, mom, dt = O.ic3d(*nb, dtype=dtype, C_cfl=C)
    k = O.constants(nb, C, dtype)
    if sd is None:
        mass = O.mass2d if dim == 2 else O.mass3d
        write = O.write_csv2d if dim == 2 else O.write_csv3d
    else:
        def mass(r):
            return O.mass_sd(r, sd, ranks)

        def write(path, r, m):
            O.write_csv_sd(path, r, m, sd, ranks)
    m0 = mass(rhoE)
    lines = [f"Initial total mass       : {m0:.7e}", f"Initial Dt : {float(dt):.2e}"]
    files = []

    def emit(it, t):
        name = pathlib.Path(f"{out_prefix}_{it}_{t:.4f}.csv")
        write(name, rhoE, mom)
        files.append(name)

    if w:
        emit(0, 0.0)
    nr, nu = np.zeros_like(rhoE), np.zeros_like(mom)
    t, it = 0.0, 0
    while t < max_t and it < max_it:
        if dim == 2:
            O.ghost_fill2d(rhoE, mom, bc[0], bc[1])
            nr, nu, dtn = O.launcher2d(rhoE, mom, dt / k["D"][0], dt / k["D"][1], k["C_min"], bc_y=bc[1],
                                       out=(nr, nu))
        else:
            O.ghost_fill3d(rhoE, mom, *bc)
            nr, nu, dtn = O.launcher3d(rhoE, mom, dt / k["D"][0], dt / k["D"][1], dt / k["D"][2], k["C_min"],
                                       bc_y=bc[1], bc_z=bc[2], out=(nr, nu))
        rhoE, nr = nr, rhoE
        mom, nu = nu, mom
        dt = dtn
        if w and it and it % w == 0 and it < max_it and t < max_t:
            m = mass(rhoE)
            lines.append(f"[it={it:4d}] t={t:.4e} Total density : {m:.7e}, change in mass: "
                         f"{abs(m0 - m) / np.prod(nb):.1e} (should be close to 0)")
            emit(it, t)
        lines.append(f"[it={it:4d}] hydro_fvm compute time t={t:.4e} (Dt={float(dt):.3e})")
        it += 1
        t += float(dt)
    m = mass(rhoE)
    lines.append(f"[it={it:4d}] t={t:.4e} Total density : {m:.7e}, change in mass: "
                 f"{abs(m0 - m) / np.prod(nb):.1e} (should be close to 0)")
    emit(it, t)
    return lines, files
