"""MI355X-native first-order finite-volume compressible Euler solver (2D / 3D).

Python front end of libhydro_hip.so, whose C ABI is declared in include/hydro_hip.h.
The compute path is the hand-written HIP code of csrc/; PyTorch only provides device
memory and streams.  There is no CPU fallback: if the library is missing or no GPU is
visible, the calls below raise.

State tensors use the reference layout (hydro2d/main.cxx:118-128,
hydro3d/main.cxx:122-133): 2D rhoE, uv of shape (NB_X+2, NB_Y+2, 2); 3D rhoE of
shape (NB_X+2, NB_Y+2, NB_Z+2, 2) and uvw (..., 3); contiguous, float64 (float32 for
the USE_FLOAT form), one ghost layer.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib
import subprocess

HERE = pathlib.Path(__file__).resolve().parent
# HYDRO_LIB_PATH selects another build of the same library (e.g. a tuning variant).
LIB_PATH = pathlib.Path(os.environ.get("HYDRO_LIB_PATH", HERE / "libhydro_hip.so"))

BC_PERIODIC = 0
BC_REFLECTIVE = 1

# Precisions of the run / batch APIs (include/hydro_hip.h)
FP64 = 64
FP32 = 32
FP32_FLUX = 3264

GAMMA = 1.4
K = 1.1

_lib = None


def build(jobs: int = 8) -> None:
    """Compile the HIP library for gfx950 in-tree (hipcc via make)."""
    subprocess.run(["make", "-s", f"-j{jobs}", "-C", str(HERE)], check=True)


class HydroError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """Load libhydro_hip.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise HydroError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C {HERE}")
        _lib = C.CDLL(str(LIB_PATH))
        _declare(_lib)
    return _lib


def _declare(L: C.CDLL) -> None:
    vp, sz, i32, d, f = C.c_void_p, C.c_size_t, C.c_int, C.c_double, C.c_float
    L.launcher.argtypes = [vp] * 5 + [d] * 5 + [sz, sz, d, d, vp]
    L.launcher.restype = d
    L.launcher3d.argtypes = [vp] * 5 + [d] * 5 + [sz, sz, sz, d, d, d, vp]
    L.launcher3d.restype = d
    L.launcher_f32.argtypes = [vp] * 5 + [f] * 5 + [sz, sz, f, f, vp]
    L.launcher_f32.restype = d
    L.launcher3d_f32.argtypes = [vp] * 5 + [f] * 5 + [sz, sz, sz, f, f, f, vp]
    L.launcher3d_f32.restype = d
    L.hydro_step2d.argtypes = [vp] * 5 + [d] * 5 + [sz, sz, d, d, i32, vp]
    L.hydro_step3d.argtypes = [vp] * 5 + [d] * 5 + [sz, sz, sz, d, d, d, i32, i32, vp]
    L.hydro_ghost_fill2d.argtypes = [vp, vp, sz, sz, i32, i32, vp]
    L.hydro_ghost_fill3d.argtypes = [vp, vp, sz, sz, sz, i32, i32, i32, vp]
    L.hydro_ghost_fill2d_f32.argtypes = [vp, vp, sz, sz, i32, i32, vp]
    L.hydro_ghost_fill3d_f32.argtypes = [vp, vp, sz, sz, sz, i32, i32, i32, vp]
    L.hydro_ic.argtypes = [vp, vp, i32, vp, vp, vp, d, d, d, d, i32, d, d, d, vp, vp]
    L.hydro_ic_f32.argtypes = [vp, vp, i32, vp, vp, vp, f, f, f, f, i32, d, f, f, vp, vp]
    L.hydro_run_create.argtypes = [i32, vp, vp, d, d, d, sz, i32, vp]
    L.hydro_run_create.restype = vp
    L.hydro_run_advance.argtypes = [vp, vp, vp, vp, vp, sz, vp]
    L.hydro_run_status.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.hydro_run_trajectory.argtypes = [vp, sz, sz, vp, vp, vp]
    L.hydro_run_destroy.argtypes = [vp]
    L.hydro_run_destroy.restype = None
    L.hydro_run_set_timing.argtypes = [vp, i32]
    L.hydro_run_kernel_time.argtypes = [vp, vp, vp]
    L.hydro_run_noalloc.argtypes = [i32, vp, vp, d, d, d, sz, i32, vp] + [vp] * 4 + [vp] * 6 + [vp]
    L.hydro_tune_set.argtypes = [C.c_char_p, i32]
    L.hydro_tune_get.argtypes = [C.c_char_p]
    L.hydro_total_mass.argtypes = [vp, sz, sz, sz, i32, vp, vp]
    L.hydro_total_energy.argtypes = [vp, sz, sz, sz, i32, vp, vp]
    L.hydro_stream_copy.argtypes = [vp, vp, sz, vp]
    L.hydro_stream_copy_nt.argtypes = [vp, vp, sz, vp]
    L.hydro_stream_release.argtypes = [vp]
    L.hydro_stream_states.argtypes = [vp, vp]
    L.hydro_stream_read.argtypes = [vp, sz, vp, vp]
    L.hydro_stream_fill.argtypes = [vp, sz, C.c_ulonglong, vp]
    L.hydro_stream_pattern2d.argtypes = [vp, vp, vp, vp, sz, sz, vp]
    L.hydro_batch_create.argtypes = [i32, i32, sz]
    L.hydro_batch_create.restype = vp
    L.hydro_batch_destroy.argtypes = [vp]
    L.hydro_batch_destroy.restype = None
    L.hydro_batch_stream.argtypes = [vp]
    L.hydro_batch_stream.restype = vp
    L.hydro_batch_add.argtypes = [vp, vp, vp, d, d, d, sz, vp, vp, vp, vp]
    L.hydro_batch_add_blast.argtypes = [vp, vp, d, d, d, d, d, d, d, sz, vp, vp, vp, vp]
    L.hydro_batch_remove.argtypes = [vp, i32]
    L.hydro_batch_advance.argtypes = [vp, sz]
    L.hydro_batch_sync.argtypes = [vp]
    L.hydro_batch_status.argtypes = [vp, i32, vp, vp, vp, vp, vp]
    L.hydro_batch_done_flags.argtypes = [vp, vp, sz]
    L.hydro_batch_set_tiling.argtypes = [vp, i32, i32, i32]
    L.hydro_batch_get_tiling.argtypes = [vp, vp, vp, vp]
    L.hydro_batch_set_graph.argtypes = [vp, sz]
    L.hydro_batch_autotune.argtypes = [vp, sz, i32, vp]
    L.hydro_server_create.argtypes = [sz, sz, sz]
    L.hydro_server_create.restype = vp
    L.hydro_server_destroy.argtypes = [vp]
    L.hydro_server_destroy.restype = None
    L.hydro_server_now.argtypes = [vp]
    L.hydro_server_now.restype = d
    L.hydro_server_submit.argtypes = [vp, i32, vp, i32, d, d, d, d, d, d, d, sz, i32, d]
    L.hydro_server_tick.argtypes = [vp, vp]
    L.hydro_server_pending.argtypes = [vp, vp, vp, vp]
    L.hydro_server_request.argtypes = [vp, i32, vp, vp, vp, vp, vp, vp]
    L.hydro_server_result.argtypes = [vp, i32, vp, vp]
    L.hydro_server_release.argtypes = [vp, i32]
    L.hydro_server_stats.argtypes = [vp, vp, vp]
    L.hydro_server_set_limits.argtypes = [vp, d, sz]
    L.hydro_server_set_policy.argtypes = [vp, d]
    L.hydro_server_set_backlog.argtypes = [vp, C.c_size_t]
    L.hydro_last_error.restype = C.c_char_p
    L.hydro_version.restype = i32


def last_error() -> str:
    return lib().hydro_last_error().decode()


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise HydroError(f"{what} failed ({rc}): {last_error()}")


def _stream(stream) -> C.c_void_p:
    if stream is None:
        import torch

        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return C.c_void_p(stream)
    return C.c_void_p(stream.cuda_stream)


def _p(t) -> C.c_void_p:
    if not t.is_cuda:
        raise HydroError("device tensor expected (the solver has no CPU path)")
    if not t.is_contiguous():
        raise HydroError("contiguous tensor expected")
    return C.c_void_p(t.data_ptr())


def _is64(t) -> bool:
    import torch

    if t.dtype == torch.float64:
        return True
    if t.dtype == torch.float32:
        return False
    raise HydroError(f"float64 or float32 state expected, got {t.dtype}")


# The C ABI trusts the sizes it is given, so every tensor is checked here before its
# pointer is passed on: a wrong shape, dtype or device would make a kernel or a copy
# read or write past an allocation.  The checks run in the order dtype, shape, device,
# contiguity, so that shape and dtype errors are reported the same with or without a GPU.

def _state(rhoE, mom, what: str, dim: int | None = None, dtype=None):
    """Check a (rhoE, momentum) pair in the reference layout: rhoE (NB+2..., 2) and mom
    (NB+2..., dim) of one float dtype on one device.  Returns (dim, nb, is64)."""
    is64 = _is64(rhoE)
    if dtype is not None and rhoE.dtype != dtype:
        raise HydroError(f"{what}: {dtype} state expected, got {rhoE.dtype}")
    if mom.dtype != rhoE.dtype:
        raise HydroError(f"{what}: momentum dtype {mom.dtype} differs from rhoE dtype {rhoE.dtype}")
    d = rhoE.dim() - 1
    if d not in (2, 3) or (dim is not None and d != dim) or rhoE.shape[-1] != 2:
        want = f"{dim}D" if dim is not None else "2D or 3D"
        raise HydroError(f"{what}: rhoE of shape (NB+2, ..., 2) for {want} expected, got {tuple(rhoE.shape)}")
    nb = [int(s) - 2 for s in rhoE.shape[:d]]
    if min(nb) < 1:
        raise HydroError(f"{what}: every interior size must be >= 1, got {nb}")
    want_mom = tuple(rhoE.shape[:d]) + (d,)
    if tuple(mom.shape) != want_mom:
        raise HydroError(f"{what}: momentum of shape {want_mom} expected, got {tuple(mom.shape)}")
    if rhoE.device != mom.device:
        raise HydroError(f"{what}: rhoE on {rhoE.device} and momentum on {mom.device}")
    return d, nb, is64


def _like(ref, what: str, *others) -> None:
    """Each tensor in `others` has the shape, dtype and device of `ref`."""
    for t in others:
        if t.dtype != ref.dtype:
            raise HydroError(f"{what}: dtype {t.dtype} differs from {ref.dtype}")
        if tuple(t.shape) != tuple(ref.shape):
            raise HydroError(f"{what}: shape {tuple(t.shape)} differs from {tuple(ref.shape)}")
        if t.device != ref.device:
            raise HydroError(f"{what}: tensor on {t.device}, expected {ref.device}")


def _scalar(dt, ref, what: str) -> None:
    """The Dt_next output: at least one element of the state's dtype on its device."""
    if dt.dtype != ref.dtype:
        raise HydroError(f"{what}: Dt_next dtype {dt.dtype} differs from the state's {ref.dtype}")
    if dt.numel() < 1:
        raise HydroError(f"{what}: Dt_next must hold one element")
    if dt.device != ref.device:
        raise HydroError(f"{what}: Dt_next on {dt.device}, expected {ref.device}")


def _nbytes(x) -> int:
    """Size in bytes of a contiguous torch tensor or numpy array."""
    if hasattr(x, "data_ptr"):
        if not x.is_contiguous():
            raise HydroError("contiguous destination expected")
        return x.numel() * x.element_size()
    if not x.flags["C_CONTIGUOUS"]:
        raise HydroError("contiguous destination expected")
    return x.nbytes


def _sizes(a) -> C.Array:
    return (C.c_size_t * len(a))(*[int(x) for x in a])


# ----------------------------------------------------------------------------- physics

def constants(nb, C_cfl: float, dtype: str = "f64") -> dict:
    """gamma, gamma-1, 1/(gamma-1), K, DX.., C_min as hydro2d/main.cxx:482-484, 595-607."""
    import numpy as np

    t = np.float64 if dtype == "f64" else np.float32
    one = t(1)
    gamma = t(GAMMA)
    gmo = gamma - one
    D = [one / t(n + 1) for n in nb]
    ms = D[0]
    for v in D[1:]:
        ms = v if v < ms else ms
    return {"gamma": gamma, "gmo": gmo, "divgamma": one / gmo, "K": t(K), "D": D, "C_min": t(C_cfl) * ms}


# ----------------------------------------------------------------------------- kernels

def launcher(rhoE, uv, rhoE_next, uv_next, Dt_next, C_min, gamma, gmo, divgamma, Kc, DtDx, DtDy, stream=None):
    """The drop-in 2D launcher (hydro2d/kernel.hpp:63-68): one synchronous step."""
    _, (nb_x, nb_y), is64 = _state(rhoE, uv, "launcher", dim=2)
    _like(rhoE, "launcher rhoE_next", rhoE_next)
    _like(uv, "launcher uv_next", uv_next)
    _scalar(Dt_next, rhoE, "launcher")
    fn = lib().launcher if is64 else lib().launcher_f32
    us = fn(_p(rhoE), _p(uv), _p(rhoE_next), _p(uv_next), _p(Dt_next), C_min, gamma, gmo, divgamma, Kc, nb_x, nb_y,
            DtDx, DtDy, _stream(stream))
    if us < 0:
        raise HydroError(f"launcher failed: {last_error()}")
    return us


def launcher3d(rhoE, uvw, rhoE_next, uvw_next, Dt_next, C_min, gamma, gmo, divgamma, Kc, DtDx, DtDy, DtDz,
               stream=None):
    """The drop-in 3D launcher (hydro3d/kernel.hpp:64-69)."""
    _, nb, is64 = _state(rhoE, uvw, "launcher3d", dim=3)
    _like(rhoE, "launcher3d rhoE_next", rhoE_next)
    _like(uvw, "launcher3d uvw_next", uvw_next)
    _scalar(Dt_next, rhoE, "launcher3d")
    fn = lib().launcher3d if is64 else lib().launcher3d_f32
    us = fn(_p(rhoE), _p(uvw), _p(rhoE_next), _p(uvw_next), _p(Dt_next), C_min, gamma, gmo, divgamma, Kc, *nb, DtDx,
            DtDy, DtDz, _stream(stream))
    if us < 0:
        raise HydroError(f"launcher3d failed: {last_error()}")
    return us


def step2d(rhoE, uv, rhoE_next, uv_next, Dt_next, C_min, gamma, gmo, divgamma, Kc, DtDx, DtDy, bc_y=BC_REFLECTIVE,
           stream=None):
    """One asynchronous FP64 step (hydro_step2d)."""
    import torch

    _, (nb_x, nb_y), _ = _state(rhoE, uv, "step2d", dim=2, dtype=torch.float64)
    _like(rhoE, "step2d rhoE_next", rhoE_next)
    _like(uv, "step2d uv_next", uv_next)
    _scalar(Dt_next, rhoE, "step2d")
    _check(lib().hydro_step2d(_p(rhoE), _p(uv), _p(rhoE_next), _p(uv_next), _p(Dt_next), C_min, gamma, gmo, divgamma,
                              Kc, nb_x, nb_y, DtDx, DtDy, bc_y, _stream(stream)), "hydro_step2d")


def step3d(rhoE, uvw, rhoE_next, uvw_next, Dt_next, C_min, gamma, gmo, divgamma, Kc, DtDx, DtDy, DtDz,
           bc_y=BC_REFLECTIVE, bc_z=BC_REFLECTIVE, stream=None):
    """One asynchronous FP64 step (hydro_step3d)."""
    import torch

    _, nb, _ = _state(rhoE, uvw, "step3d", dim=3, dtype=torch.float64)
    _like(rhoE, "step3d rhoE_next", rhoE_next)
    _like(uvw, "step3d uvw_next", uvw_next)
    _scalar(Dt_next, rhoE, "step3d")
    _check(lib().hydro_step3d(_p(rhoE), _p(uvw), _p(rhoE_next), _p(uvw_next), _p(Dt_next), C_min, gamma, gmo, divgamma,
                              Kc, *nb, DtDx, DtDy, DtDz, bc_y, bc_z, _stream(stream)), "hydro_step3d")


def ghost_fill2d(rhoE, uv, bc_x=BC_PERIODIC, bc_y=BC_REFLECTIVE, stream=None):
    _, (nb_x, nb_y), is64 = _state(rhoE, uv, "ghost_fill2d", dim=2)
    fn = lib().hydro_ghost_fill2d if is64 else lib().hydro_ghost_fill2d_f32
    _check(fn(_p(rhoE), _p(uv), nb_x, nb_y, bc_x, bc_y, _stream(stream)), "hydro_ghost_fill2d")


def ghost_fill3d(rhoE, uvw, bc_x=BC_PERIODIC, bc_y=BC_REFLECTIVE, bc_z=BC_REFLECTIVE, stream=None):
    _, nb, is64 = _state(rhoE, uvw, "ghost_fill3d", dim=3)
    fn = lib().hydro_ghost_fill3d if is64 else lib().hydro_ghost_fill3d_f32
    _check(fn(_p(rhoE), _p(uvw), *nb, bc_x, bc_y, bc_z, _stream(stream)), "hydro_ghost_fill3d")


def ic(rhoE, mom, C_min, NB=None, base=None, ic_mode=0, radius_frac=0.0, rho_in=1.0, rho_out=1.2, p_in=10.0, p_out=0.1,
       stream=None):
    """Blast IC on the device (hydro2d/main.cxx:280-338); returns the initial Dt."""
    import numpy as np

    dim, nb, is64 = _state(rhoE, mom, "ic")
    NB = nb if NB is None else NB
    if len(NB) != dim or (base is not None and len(base) != dim):
        raise HydroError(f"ic: NB and base must have {dim} entries")
    out = (C.c_double if is64 else C.c_float)()
    fn = lib().hydro_ic if is64 else lib().hydro_ic_f32
    t = np.float64 if is64 else np.float32
    _check(fn(_p(rhoE), _p(mom), dim, _sizes(nb), _sizes(base) if base is not None else None, _sizes(NB),
              t(rho_in), t(rho_out), t(p_in), t(p_out), ic_mode, radius_frac, t(C_min), t(GAMMA), C.byref(out),
              _stream(stream)), "hydro_ic")
    return t(out.value)


def _rhoE_only(rhoE, what: str):
    """Check rhoE alone (the mass and energy sums read no momentum)."""
    is64 = _is64(rhoE)
    dim = rhoE.dim() - 1
    if dim not in (2, 3) or rhoE.shape[-1] != 2 or min(rhoE.shape[:dim]) < 3:
        raise HydroError(f"{what}: rhoE of shape (NB+2, ..., 2) expected, got {tuple(rhoE.shape)}")
    return dim, [int(s) - 2 for s in rhoE.shape[:dim]], is64


def total_mass(rhoE, stream=None) -> float:
    dim, nb, _ = _rhoE_only(rhoE, "total_mass")
    m = C.c_double()
    _check(lib().hydro_total_mass(_p(rhoE), nb[0], nb[1], nb[2] if dim == 3 else 0, 64 if _is64(rhoE) else 32,
                                  C.byref(m), _stream(stream)), "hydro_total_mass")
    return m.value


def total_energy(rhoE, stream=None) -> float:
    """Sum of interior E (compensated, on the device)."""
    dim, nb, _ = _rhoE_only(rhoE, "total_energy")
    m = C.c_double()
    _check(lib().hydro_total_energy(_p(rhoE), nb[0], nb[1], nb[2] if dim == 3 else 0, 64 if _is64(rhoE) else 32,
                                    C.byref(m), _stream(stream)), "hydro_total_energy")
    return m.value


# ----------------------------------------------------------------------------- resident run

class Run:
    """The step loop of hydro2d/main.cxx:686-815 with Dt/t/it resident on the device."""

    def __init__(self, nb, D, C_min, dt0, max_t, max_it, precision=64, bc=None):
        self.dim = len(nb)
        self.nb = [int(n) for n in nb]
        self.precision = int(precision)
        if len(D) != self.dim:
            raise HydroError(f"Run: {self.dim} spacings expected, got {len(D)}")
        bcs = list(bc) if bc is not None else [BC_PERIODIC] + [BC_REFLECTIVE] * (self.dim - 1)
        self._bc = (C.c_int * 3)(*(bcs + [0] * (3 - len(bcs))))
        self.max_it = int(max_it)
        h = lib().hydro_run_create(self.dim, _sizes(nb), (C.c_double * 3)(*[float(x) for x in D]), float(C_min),
                                   float(dt0), float(max_t), self.max_it, precision, self._bc)
        if not h:
            raise HydroError(f"hydro_run_create failed: {last_error()}")
        self._h = C.c_void_p(h)

    def _check_buffers(self, a, b) -> None:
        """Both buffer pairs match the run: its dimension, interior sizes and precision
        (float32 state for FP32, float64 for FP64 and FP32_FLUX)."""
        import torch

        dtype = torch.float32 if self.precision == FP32 else torch.float64
        for name, pair in (("a", a), ("b", b)):
            _, nb, _ = _state(pair[0], pair[1], f"Run.advance buffer {name}", dim=self.dim, dtype=dtype)
            if nb != self.nb:
                raise HydroError(f"Run.advance buffer {name}: interior {nb}, the run has {self.nb}")
        _like(a[0], "Run.advance buffer b rhoE", b[0])
        _like(a[1], "Run.advance buffer b momentum", b[1])

    def advance(self, a, b, n_steps, stream=None):
        self._check_buffers(a, b)
        _check(lib().hydro_run_advance(self._h, _p(a[0]), _p(a[1]), _p(b[0]), _p(b[1]), int(n_steps), _stream(stream)),
               "hydro_run_advance")

    def status(self, stream=None) -> dict:
        it, t, dt, done, cur = C.c_size_t(), C.c_double(), C.c_double(), C.c_int(), C.c_int()
        rc = lib().hydro_run_status(self._h, C.byref(it), C.byref(t), C.byref(dt), C.byref(done), C.byref(cur),
                                    _stream(stream))
        st = {"it": it.value, "t": t.value, "dt": dt.value, "done": bool(done.value), "current": cur.value, "rc": rc}
        if rc != 0:
            st["error"] = last_error()
        return st

    def trajectory(self, n, stream=None):
        import numpy as np

        tt = np.zeros(max(n, 1), np.float64)
        td = np.zeros(max(n, 1), np.float64)
        _check(lib().hydro_run_trajectory(self._h, 0, int(n), tt.ctypes.data_as(C.c_void_p),
                                          td.ctypes.data_as(C.c_void_p), _stream(stream)), "hydro_run_trajectory")
        return tt[:n], td[:n]

    def set_timing(self, enable: bool = True) -> None:
        _check(lib().hydro_run_set_timing(self._h, 1 if enable else 0), "hydro_run_set_timing")

    def kernel_time(self):
        """(total ms, launches) of the step kernels recorded since set_timing / the last call."""
        ms, n = C.c_double(), C.c_size_t()
        _check(lib().hydro_run_kernel_time(self._h, C.byref(ms), C.byref(n)), "hydro_run_kernel_time")
        return ms.value, n.value

    def close(self):
        if getattr(self, "_h", None):
            lib().hydro_run_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many independent blasts advanced together by one kernel launch per step
    (include/hydro_hip.h, hydro_batch_*).  Each blast is a full run of the reference step
    loop with X periodic and Y/Z reflective boundaries; its buffers (a, b) are pairs
    (rhoE, mom) of contiguous float64 tensors in the reference layout."""

    def __init__(self, dim: int, precision: int = FP64, capacity: int = 256):
        h = lib().hydro_batch_create(int(dim), int(precision), int(capacity))
        if not h:
            raise HydroError(f"hydro_batch_create failed: {last_error()}")
        self._h = C.c_void_p(h)
        self.dim = dim
        self.precision = int(precision)
        self.capacity = capacity
        self._bufs = {}

    @property
    def stream(self) -> int:
        return lib().hydro_batch_stream(self._h) or 0

    def _after_current_stream(self) -> None:
        """The batch works on its own non-blocking stream: order it after the work already
        queued on torch's current stream (the fills and copies of the buffers being added)."""
        import torch

        torch.cuda.ExternalStream(self.stream).wait_stream(torch.cuda.current_stream())

    def _check_buffers(self, a, b, what: str):
        """Both buffer pairs hold float64 state (FP64 and FP32_FLUX keep FP64 state) of
        the batch's dimension and of equal sizes; returns the interior sizes."""
        import torch

        _, nb, _ = _state(a[0], a[1], f"{what} buffer a", dim=self.dim, dtype=torch.float64)
        _state(b[0], b[1], f"{what} buffer b", dim=self.dim, dtype=torch.float64)
        _like(a[0], f"{what} buffer b rhoE", b[0])
        _like(a[1], f"{what} buffer b momentum", b[1])
        return tuple(nb)

    def add(self, a, b, D, C_min, dt0, max_t, max_it) -> int:
        nb = self._check_buffers(a, b, "Batch.add")
        if len(D) != self.dim:
            raise HydroError(f"Batch.add: {self.dim} spacings expected, got {len(D)}")
        self._after_current_stream()
        slot = lib().hydro_batch_add(self._h, _sizes(nb), (C.c_double * 3)(*[float(x) for x in D] + [0.0] * (3 - len(D))),
                                     float(C_min), float(dt0), float(max_t), int(max_it), _p(a[0]), _p(a[1]), _p(b[0]),
                                     _p(b[1]))
        if slot < 0:
            raise HydroError(f"hydro_batch_add failed ({slot}): {last_error()}")
        self._bufs[slot] = (a, b)
        return slot

    def add_blast(self, a, b, C, radius_frac, p_in, p_out=0.1, rho_in=1.0, rho_out=1.2, max_t=1e30,
                  max_it=100) -> int:
        nb = self._check_buffers(a, b, "Batch.add_blast")
        self._after_current_stream()
        slot = lib().hydro_batch_add_blast(self._h, _sizes(nb), float(C), float(radius_frac), float(rho_in),
                                           float(rho_out), float(p_in), float(p_out), float(max_t), int(max_it),
                                           _p(a[0]), _p(a[1]), _p(b[0]), _p(b[1]))
        if slot < 0:
            raise HydroError(f"hydro_batch_add_blast failed ({slot}): {last_error()}")
        self._bufs[slot] = (a, b)
        return slot

    def remove(self, slot: int) -> None:
        _check(lib().hydro_batch_remove(self._h, int(slot)), "hydro_batch_remove")
        self._bufs.pop(slot, None)

    def advance(self, n_steps: int) -> None:
        _check(lib().hydro_batch_advance(self._h, int(n_steps)), "hydro_batch_advance")

    def sync(self) -> None:
        _check(lib().hydro_batch_sync(self._h), "hydro_batch_sync")

    def status(self, slot: int) -> dict:
        it, t, dt, done, cur = C.c_size_t(), C.c_double(), C.c_double(), C.c_int(), C.c_int()
        rc = lib().hydro_batch_status(self._h, int(slot), C.byref(it), C.byref(t), C.byref(dt), C.byref(done),
                                      C.byref(cur))
        st = {"it": it.value, "t": t.value, "dt": dt.value, "done": bool(done.value), "current": cur.value, "rc": rc}
        if rc != 0:
            st["error"] = last_error()
        return st

    def current(self, slot: int):
        """The (rhoE, mom) pair holding the slot's current state."""
        a, b = self._bufs[slot]
        return a if self.status(slot)["current"] == 0 else b

    def done_flags(self, n: int | None = None):
        import numpy as np

        n = self.capacity if n is None else n
        out = np.zeros(n, np.uint8)
        _check(lib().hydro_batch_done_flags(self._h, out.ctypes.data_as(C.c_void_p), n), "hydro_batch_done_flags")
        return out

    def set_tiling(self, blocks_per_cu: int = 0, tile_rows: int = 0, tile_windows: int = 1) -> None:
        _check(lib().hydro_batch_set_tiling(self._h, int(blocks_per_cu), int(tile_rows), int(tile_windows)),
               "hydro_batch_set_tiling")

    def tiling(self):
        """(blocks per CU, tile rows, tile windows)."""
        bpc, rows, width = C.c_int(), C.c_int(), C.c_int()
        _check(lib().hydro_batch_get_tiling(self._h, C.byref(bpc), C.byref(rows), C.byref(width)),
               "hydro_batch_get_tiling")
        return bpc.value, rows.value, width.value

    def set_graph(self, steps_per_graph: int) -> None:
        _check(lib().hydro_batch_set_graph(self._h, int(steps_per_graph)), "hydro_batch_set_graph")

    def autotune(self, steps_per_trial: int = 4, use_cache: bool = True) -> float:
        ms = C.c_double()
        _check(lib().hydro_batch_autotune(self._h, int(steps_per_trial), 1 if use_cache else 0, C.byref(ms)),
               "hydro_batch_autotune")
        return ms.value

    def close(self):
        if getattr(self, "_h", None):
            lib().hydro_batch_destroy(self._h)
            self._h = None
            self._bufs = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop with continuous batching (include/hydro_hip.h, hydro_server_*)."""

    QUEUED, RUNNING, DONE, RELEASED, FAILED = range(5)

    def __init__(self, hbm_bytes: int, capacity: int = 256, steps_per_tick: int = 50):
        h = lib().hydro_server_create(int(hbm_bytes), int(capacity), int(steps_per_tick))
        if not h:
            raise HydroError(f"hydro_server_create failed: {last_error()}")
        self._h = C.c_void_p(h)
        self._nb = {}

    def now(self) -> float:
        return lib().hydro_server_now(self._h)

    def submit(self, nb, precision=FP64, C=None, radius_frac=0.1, p_in=10.0, p_out=0.1, rho_in=1.0, rho_out=1.2,
               max_t=1e30, max_it=100, keep=False, arrival=-1.0) -> int:
        dim = len(nb)
        C = (0.5 if dim == 2 else 0.4) if C is None else C
        rid = lib().hydro_server_submit(self._h, dim, _sizes(nb), int(precision), float(C), float(radius_frac),
                                        float(rho_in), float(rho_out), float(p_in), float(p_out), float(max_t),
                                        int(max_it), 1 if keep else 0, float(arrival))
        if rid < 0:
            raise HydroError(f"hydro_server_submit failed ({rid}): {last_error()}")
        self._nb[rid] = [int(n) for n in nb]
        return rid

    def tick(self) -> int:
        n = C.c_size_t()
        _check(lib().hydro_server_tick(self._h, C.byref(n)), "hydro_server_tick")
        return n.value

    def pending(self):
        q, r, b = C.c_size_t(), C.c_size_t(), C.c_size_t()
        _check(lib().hydro_server_pending(self._h, C.byref(q), C.byref(r), C.byref(b)), "hydro_server_pending")
        return q.value, r.value, b.value

    def request(self, rid: int) -> dict:
        st, ta, tad, td, it, ts = C.c_int(), C.c_double(), C.c_double(), C.c_double(), C.c_size_t(), C.c_double()
        rc = lib().hydro_server_request(self._h, int(rid), C.byref(st), C.byref(ta), C.byref(tad), C.byref(td),
                                        C.byref(it), C.byref(ts))
        return {"state": st.value, "arrival": ta.value, "admit": tad.value, "done": td.value, "it": it.value,
                "t": ts.value, "rc": rc}

    def result(self, rid: int, rhoE, mom) -> None:
        """Copy a kept result into (rhoE, mom) (device or host tensors / numpy arrays),
        each at least as large as the request's state with ghosts: 2 (rhoE) and dim
        (momentum) float64 values per cell."""
        def ptr(x):
            return C.c_void_p(x.data_ptr()) if hasattr(x, "data_ptr") else x.ctypes.data_as(C.c_void_p)

        nb = self._nb.get(int(rid))
        if nb is None:
            raise HydroError(f"Server.result: unknown request {rid}")
        cells = 1
        for n in nb:
            cells *= n + 2
        for name, dst, per_cell in (("rhoE", rhoE, 2), ("momentum", mom, len(nb))):
            if dst is None:
                continue
            if str(dst.dtype) not in ("torch.float64", "float64"):
                raise HydroError(f"Server.result: float64 {name} destination expected, got {dst.dtype}")
            if _nbytes(dst) < cells * per_cell * 8:
                raise HydroError(f"Server.result: {name} destination holds {_nbytes(dst)} bytes, the result "
                                 f"needs {cells * per_cell * 8}")

        _check(lib().hydro_server_result(self._h, int(rid), ptr(rhoE) if rhoE is not None else None,
                                         ptr(mom) if mom is not None else None), "hydro_server_result")

    def release(self, rid: int) -> None:
        _check(lib().hydro_server_release(self._h, int(rid)), "hydro_server_release")

    def set_limits(self, max_active_cells: float = 0.0, steps_per_tick: int = 50) -> None:
        _check(lib().hydro_server_set_limits(self._h, float(max_active_cells), int(steps_per_tick)),
               "hydro_server_set_limits")

    def set_policy(self, large_seconds: float = 0.0) -> None:
        """Requests whose estimated solve time alone is at least large_seconds run one at
        a time in arrival order as plain resident solves, beside the smaller ones in the
        batched engines (0: one class)."""
        _check(lib().hydro_server_set_policy(self._h, float(large_seconds)), "hydro_server_set_policy")

    def set_backlog(self, large_backlog: int = 0) -> None:
        """While more than large_backlog large requests wait, the oldest waiting ones join
        their batched engines (0: never)."""
        if int(large_backlog) < 0:
            raise HydroError("Server.set_backlog: large_backlog >= 0")
        _check(lib().hydro_server_set_backlog(self._h, int(large_backlog)), "hydro_server_set_backlog")

    def stats(self):
        t, cu = C.c_size_t(), C.c_double()
        _check(lib().hydro_server_stats(self._h, C.byref(t), C.byref(cu)), "hydro_server_stats")
        return t.value, cu.value

    def close(self):
        if getattr(self, "_h", None):
            lib().hydro_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stream_copy(dst, src, stream=None, nontemporal: bool = False) -> None:
    """Plain device copy (achievable-bandwidth reference); nontemporal: with
    non-temporal loads and stores."""
    nbytes = src.numel() * src.element_size()
    if dst.numel() * dst.element_size() < nbytes or dst.device != src.device:
        raise HydroError("stream_copy: dst must be at least as large as src, on the same device")
    fn = lib().hydro_stream_copy_nt if nontemporal else lib().hydro_stream_copy
    _check(fn(_p(dst), _p(src), nbytes, _stream(stream)), "hydro_stream_copy")


def stream_read(src, slots, stream=None) -> None:
    """Read-only stream over src: XOR of every 64-bit word folded into `slots` (256
    uint64 on the device, zeroed by the caller)."""
    nbytes = src.numel() * src.element_size()
    if slots.numel() * slots.element_size() < 256 * 8:
        raise HydroError("stream_read: slots must hold 256 64-bit words")
    _check(lib().hydro_stream_read(_p(src), nbytes, _p(slots), _stream(stream)), "hydro_stream_read")


def stream_fill(dst, pattern: int, stream=None) -> None:
    """Write-only stream: the 64-bit `pattern` to every word of dst."""
    nbytes = dst.numel() * dst.element_size()
    _check(lib().hydro_stream_fill(_p(dst), nbytes, int(pattern) & (2**64 - 1), _stream(stream)), "hydro_stream_fill")


def stream_pattern2d(a, b, stream=None) -> None:
    """The 2D FP64 step's memory schedule without its arithmetic (include/hydro_hip.h,
    hydro_stream_pattern2d): copies the interior cells of the pair `a` = (rhoE, uv) into
    the pair `b` with the step's loads and stores; b's ghost cells are not written."""
    import torch

    _, nb, _ = _state(a[0], a[1], "stream_pattern2d a", dim=2, dtype=torch.float64)
    _state(b[0], b[1], "stream_pattern2d b", dim=2, dtype=torch.float64)
    _like(a[0], "stream_pattern2d b rhoE", b[0])
    _like(a[1], "stream_pattern2d b uv", b[1])
    _check(lib().hydro_stream_pattern2d(_p(a[0]), _p(a[1]), _p(b[0]), _p(b[1]), nb[0], nb[1], _stream(stream)),
           "hydro_stream_pattern2d")


def stream_release(stream=None) -> None:
    """Unbind the plain launchers' Dt reduction state from `stream` (waits for it)."""
    _check(lib().hydro_stream_release(_stream(stream)), "hydro_stream_release")


def stream_states():
    """(streams with a bound reduction state, free pooled states) on the current device."""
    bound, free = C.c_int(), C.c_int()
    _check(lib().hydro_stream_states(C.byref(bound), C.byref(free)), "hydro_stream_states")
    return bound.value, free.value


def tune_set(key: str, value: int) -> None:
    """Launch-geometry knobs ("rows2d", "occ2d", "rows3d"; 0 = per-precision default);
    results do not depend on them."""
    _check(lib().hydro_tune_set(key.encode(), int(value)), f"hydro_tune_set({key})")


def tune_get(key: str) -> int:
    return lib().hydro_tune_get(key.encode())


def library_path() -> str:
    return os.fspath(LIB_PATH)
