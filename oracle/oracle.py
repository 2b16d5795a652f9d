"""numpy/ctypes front end of the CPU oracle (hydro_oracle.c).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker, never by the product.  Builds liboracle_hydro.so
with make on first use if it is missing.

Array layout is the reference's (hydro2d/main.cxx:118-128, hydro3d/main.cxx:122-133):
2D rhoE/uv of shape (nb_x+2, nb_y+2, 2), 3D rhoE (nb_x+2, nb_y+2, nb_z+2, 2) and
uvw (..., 3), C order, ghost width 1.
"""
from __future__ import annotations

import ctypes as C
import json
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle_hydro.so"

BC_PERIODIC = 0
BC_REFLECTIVE = 1

GAMMA = 1.4
K = 1.1


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(LIB_PATH))
        _declare(_lib)
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _declare(L: C.CDLL) -> None:
    vp, sz, i32, f64p = C.c_void_p, C.c_size_t, C.c_int, C.POINTER(C.c_double)
    for sfx, R in (("f64", C.c_double), ("f32", C.c_float)):
        g = lambda n: getattr(L, f"{n}_{sfx}")  # noqa: E731
        g("oracle_flux2").argtypes = [vp, vp, R, R, R, vp]
        g("oracle_flux3").argtypes = [vp, vp, R, R, R, vp]
        g("oracle_launcher2d").argtypes = [vp, vp, vp, vp, vp, R, R, R, R, R, sz, sz, R, R, i32, i32]
        g("oracle_launcher3d").argtypes = [vp, vp, vp, vp, vp, R, R, R, R, R, sz, sz, sz, R, R, R, i32, i32, i32]
        g("oracle_ghost_fill2d").argtypes = [vp, vp, sz, sz, i32, i32]
        g("oracle_ghost_fill3d").argtypes = [vp, vp, sz, sz, sz, i32, i32, i32]
        g("oracle_ic2d").argtypes = [vp, vp, sz, sz, sz, sz, sz, sz, R, R, R, R, R, R, i32, C.c_double, vp]
        g("oracle_ic3d").argtypes = [vp, vp, sz, sz, sz, sz, sz, sz, sz, sz, sz, R, R, R, R, R, R, i32, C.c_double, vp]
        g("oracle_ic_mass_dt").argtypes = [sz, sz, sz, sz, R, f64p, vp]
        g("oracle_mass2d").argtypes = [vp, sz, sz]
        g("oracle_mass2d").restype = C.c_double
        g("oracle_mass3d").argtypes = [vp, sz, sz, sz]
        g("oracle_mass3d").restype = C.c_double
        g("oracle_run2d").argtypes = [vp, vp, vp, vp, sz, sz, C.c_double, sz, R, i32, i32, vp, f64p, vp, vp, i32]
        g("oracle_run2d").restype = sz
        g("oracle_run3d").argtypes = [vp, vp, vp, vp, sz, sz, sz, C.c_double, sz, R, i32, i32, i32, vp, f64p, vp, vp,
                                      i32]
        g("oracle_run3d").restype = sz
        g("oracle_write_csv2d").argtypes = [C.c_char_p, vp, vp, sz, sz]
        g("oracle_write_csv3d").argtypes = [C.c_char_p, vp, vp, sz, sz, sz]
        szp = C.POINTER(C.c_size_t)
        g("oracle_mass_sd").argtypes = [vp, i32, szp, szp, sz]
        g("oracle_mass_sd").restype = C.c_double
        g("oracle_write_csv_sd").argtypes = [C.c_char_p, vp, vp, i32, szp, szp, sz]


def _fn(name: str, dtype) -> C._CFuncPtr:
    return getattr(lib(), f"{name}_{'f64' if np.dtype(dtype) == np.float64 else 'f32'}")


# ----------------------------------------------------------------------------- constants

def constants(nb, C_cfl: float, dtype=np.float64) -> dict:
    """The physics constants of hydro2d/main.cxx:482-484, 595-607 in the state dtype."""
    t = np.dtype(dtype).type
    one = t(1)
    gamma = t(GAMMA)
    gmo = gamma - one
    d = [one / t(n + 1) for n in nb]
    ms = d[0]
    for v in d[1:]:
        ms = v if v < ms else ms
    return {
        "gamma": gamma, "gmo": gmo, "divgamma": one / gmo, "K": t(K), "C_min": t(C_cfl) * ms, "D": d,
    }


# ----------------------------------------------------------------------------- state

def alloc2d(nb_x: int, nb_y: int, dtype=np.float64):
    return (np.zeros((nb_x + 2, nb_y + 2, 2), dtype), np.zeros((nb_x + 2, nb_y + 2, 2), dtype))


def alloc3d(nb_x: int, nb_y: int, nb_z: int, dtype=np.float64):
    return (np.zeros((nb_x + 2, nb_y + 2, nb_z + 2, 2), dtype), np.zeros((nb_x + 2, nb_y + 2, nb_z + 2, 3), dtype))


def ic2d(nb_x, nb_y, dtype=np.float64, C_cfl=0.5, ic_mode=0, radius_frac=0.0, rho_in=1.0, rho_out=1.2, p_in=10.0,
         p_out=0.1):
    """Blast IC (hydro2d/main.cxx:280-338); returns (rhoE, uv, initial Dt)."""
    rhoE, uv = alloc2d(nb_x, nb_y, dtype)
    k = constants((nb_x, nb_y), C_cfl, dtype)
    t = np.dtype(dtype).type
    dt = np.zeros(1, dtype)
    _fn("oracle_ic2d", dtype)(_ptr(rhoE), _ptr(uv), nb_x, nb_y, 0, 0, nb_x, nb_y, t(rho_in), t(rho_out), t(p_in),
                              t(p_out), k["C_min"], k["gamma"], ic_mode, radius_frac, _ptr(dt))
    return rhoE, uv, dt[0]


def ic3d(nb_x, nb_y, nb_z, dtype=np.float64, C_cfl=0.4, ic_mode=0, radius_frac=0.0, rho_in=1.0, rho_out=1.2,
         p_in=10.0, p_out=0.1):
    rhoE, uvw = alloc3d(nb_x, nb_y, nb_z, dtype)
    k = constants((nb_x, nb_y, nb_z), C_cfl, dtype)
    t = np.dtype(dtype).type
    dt = np.zeros(1, dtype)
    _fn("oracle_ic3d", dtype)(_ptr(rhoE), _ptr(uvw), nb_x, nb_y, nb_z, 0, 0, 0, nb_x, nb_y, nb_z, t(rho_in),
                              t(rho_out), t(p_in), t(p_out), k["C_min"], k["gamma"], ic_mode, radius_frac, _ptr(dt))
    return rhoE, uvw, dt[0]


def ic_mass_dt(nb, n_sub=1, C_cfl=None, dtype=np.float64):
    """Initial mass and Dt of the reference blast IC on a global grid, streamed."""
    dim = len(nb)
    if C_cfl is None:
        C_cfl = 0.5 if dim == 2 else 0.4
    m = C.c_double()
    dt = np.zeros(1, dtype)
    nz = nb[2] if dim == 3 else 0
    _fn("oracle_ic_mass_dt", dtype)(nb[0], nb[1], nz, n_sub, np.dtype(dtype).type(C_cfl), C.byref(m), _ptr(dt))
    return m.value, dt[0]


# ----------------------------------------------------------------------------- kernels

def launcher2d(rhoE, uv, DtDx, DtDy, C_min, bc_y=BC_REFLECTIVE, nthreads=0, out=None):
    """One launcher call (hydro2d/kernel.cxx:331-367): returns (rhoE_next, uv_next, Dt_next).
    Ghosts of the outputs are copied from `out` if given, else zero."""
    dtype = rhoE.dtype
    nb_x, nb_y = rhoE.shape[0] - 2, rhoE.shape[1] - 2
    k = constants((nb_x, nb_y), 0.5, dtype)
    rn, un = (np.zeros_like(rhoE), np.zeros_like(uv)) if out is None else out
    dt = np.zeros(1, dtype)
    t = np.dtype(dtype).type
    _fn("oracle_launcher2d", dtype)(_ptr(rhoE), _ptr(uv), _ptr(rn), _ptr(un), _ptr(dt), t(C_min), k["gamma"], k["gmo"],
                                    k["divgamma"], k["K"], nb_x, nb_y, t(DtDx), t(DtDy), bc_y, nthreads)
    return rn, un, dt[0]


def launcher3d(rhoE, uvw, DtDx, DtDy, DtDz, C_min, bc_y=BC_REFLECTIVE, bc_z=BC_REFLECTIVE, nthreads=0, out=None):
    dtype = rhoE.dtype
    nb_x, nb_y, nb_z = (s - 2 for s in rhoE.shape[:3])
    k = constants((nb_x, nb_y, nb_z), 0.4, dtype)
    rn, un = (np.zeros_like(rhoE), np.zeros_like(uvw)) if out is None else out
    dt = np.zeros(1, dtype)
    t = np.dtype(dtype).type
    _fn("oracle_launcher3d", dtype)(_ptr(rhoE), _ptr(uvw), _ptr(rn), _ptr(un), _ptr(dt), t(C_min), k["gamma"],
                                    k["gmo"], k["divgamma"], k["K"], nb_x, nb_y, nb_z, t(DtDx), t(DtDy), t(DtDz), bc_y,
                                    bc_z, nthreads)
    return rn, un, dt[0]


def ghost_fill2d(rhoE, uv, bc_x=BC_PERIODIC, bc_y=BC_REFLECTIVE):
    _fn("oracle_ghost_fill2d", rhoE.dtype)(_ptr(rhoE), _ptr(uv), rhoE.shape[0] - 2, rhoE.shape[1] - 2, bc_x, bc_y)


def ghost_fill3d(rhoE, uvw, bc_x=BC_PERIODIC, bc_y=BC_REFLECTIVE, bc_z=BC_REFLECTIVE):
    nb = [s - 2 for s in rhoE.shape[:3]]
    _fn("oracle_ghost_fill3d", rhoE.dtype)(_ptr(rhoE), _ptr(uvw), *nb, bc_x, bc_y, bc_z)


def flux2(Lq, Rq, dtype=np.float64):
    k = constants((1, 1), 0.5, dtype)
    L = np.asarray(Lq, dtype)
    R = np.asarray(Rq, dtype)
    F = np.zeros(4, dtype)
    _fn("oracle_flux2", dtype)(_ptr(L), _ptr(R), k["K"], k["gamma"], k["divgamma"], _ptr(F))
    return F


def flux3(Lq, Rq, dtype=np.float64):
    k = constants((1, 1, 1), 0.4, dtype)
    L = np.asarray(Lq, dtype)
    R = np.asarray(Rq, dtype)
    F = np.zeros(5, dtype)
    _fn("oracle_flux3", dtype)(_ptr(L), _ptr(R), k["K"], k["gamma"], k["divgamma"], _ptr(F))
    return F


def mass2d(rhoE) -> float:
    return _fn("oracle_mass2d", rhoE.dtype)(_ptr(rhoE), rhoE.shape[0] - 2, rhoE.shape[1] - 2)


def mass3d(rhoE) -> float:
    return _fn("oracle_mass3d", rhoE.dtype)(_ptr(rhoE), *(s - 2 for s in rhoE.shape[:3]))


def _sizes(v):
    return (C.c_size_t * 3)(*(list(v) + [1] * (3 - len(v))))


def mass_sd(rhoE, sd, ranks: int) -> float:
    """compute_total_mass under an sd split and `ranks` MPI ranks (hydro2d/main.cxx:237-265,
    hydro3d/main.cxx:249-281): per rank a compensated long double sum over its subdomains,
    the ranks' doubles added in rank order."""
    dim = rhoE.ndim - 1
    nb = [s - 2 for s in rhoE.shape[:dim]]
    return _fn("oracle_mass_sd", rhoE.dtype)(_ptr(rhoE), dim, _sizes(nb), _sizes(sd), int(ranks))


def write_csv_sd(path, rhoE, mom, sd, ranks: int) -> None:
    """write_results under an sd split and `ranks` MPI ranks (hydro2d/main.cxx:166-235)."""
    dim = rhoE.ndim - 1
    nb = [s - 2 for s in rhoE.shape[:dim]]
    rc = _fn("oracle_write_csv_sd", rhoE.dtype)(str(path).encode(), _ptr(rhoE), _ptr(mom), dim, _sizes(nb),
                                                 _sizes(sd), int(ranks))
    if rc != 0:
        raise OSError(f"oracle csv write failed: {path}")


# ----------------------------------------------------------------------------- runs

def run2d(rhoE, uv, dt0, max_t, max_it, C_cfl=0.5, bc_x=BC_PERIODIC, bc_y=BC_REFLECTIVE, nthreads=0):
    """The reference step loop (hydro2d/main.cxx:686-815) in place on (rhoE, uv).
    Returns dict(it, t, dt, traj_t, traj_dt)."""
    dtype = rhoE.dtype
    nb_x, nb_y = rhoE.shape[0] - 2, rhoE.shape[1] - 2
    sr, su = np.zeros_like(rhoE), np.zeros_like(uv)
    tt = np.zeros(max(max_it, 1), np.float64)
    td = np.zeros(max(max_it, 1), dtype)
    dt = np.array([dt0], dtype)
    t_out = C.c_double()
    n = _fn("oracle_run2d", dtype)(_ptr(rhoE), _ptr(uv), _ptr(sr), _ptr(su), nb_x, nb_y, float(max_t), int(max_it),
                                   np.dtype(dtype).type(C_cfl), bc_x, bc_y, _ptr(dt), C.byref(t_out), _ptr(tt),
                                   _ptr(td), nthreads)
    return {"it": n, "t": t_out.value, "dt": dt[0], "traj_t": tt[:n], "traj_dt": td[:n]}


def run3d(rhoE, uvw, dt0, max_t, max_it, C_cfl=0.4, bc_x=BC_PERIODIC, bc_y=BC_REFLECTIVE, bc_z=BC_REFLECTIVE,
          nthreads=0):
    dtype = rhoE.dtype
    nb = [s - 2 for s in rhoE.shape[:3]]
    sr, su = np.zeros_like(rhoE), np.zeros_like(uvw)
    tt = np.zeros(max(max_it, 1), np.float64)
    td = np.zeros(max(max_it, 1), dtype)
    dt = np.array([dt0], dtype)
    t_out = C.c_double()
    n = _fn("oracle_run3d", dtype)(_ptr(rhoE), _ptr(uvw), _ptr(sr), _ptr(su), *nb, float(max_t), int(max_it),
                                   np.dtype(dtype).type(C_cfl), bc_x, bc_y, bc_z, _ptr(dt), C.byref(t_out), _ptr(tt),
                                   _ptr(td), nthreads)
    return {"it": n, "t": t_out.value, "dt": dt[0], "traj_t": tt[:n], "traj_dt": td[:n]}


def write_csv2d(path, rhoE, uv) -> None:
    rc = _fn("oracle_write_csv2d", rhoE.dtype)(str(path).encode(), _ptr(rhoE), _ptr(uv), rhoE.shape[0] - 2,
                                                rhoE.shape[1] - 2)
    if rc != 0:
        raise OSError(f"oracle csv write failed: {path}")


def write_csv3d(path, rhoE, uvw) -> None:
    rc = _fn("oracle_write_csv3d", rhoE.dtype)(str(path).encode(), _ptr(rhoE), _ptr(uvw),
                                                *(s - 2 for s in rhoE.shape[:3]))
    if rc != 0:
        raise OSError(f"oracle csv write failed: {path}")


# ----------------------------------------------------------------------------- golden logs

def load_res_logs() -> dict:
    p = HERE.parent / "tests" / "golden" / "res_logs.json"
    return json.loads(p.read_text())
