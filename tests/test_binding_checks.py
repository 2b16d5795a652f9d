"""The Python binding rejects tensors that do not match what the C ABI will be told,
before any pointer reaches the library: a wrong shape, dtype or device would make a
kernel or a copy run past an allocation.  These checks raise before anything is
launched, so they run without a GPU (on CPU tensors, which the binding would refuse
anyway, after the shape and dtype checks)."""
import numpy as np
import pytest
import torch


def pair2(nx=6, ny=5, dtype=torch.float64):
    return torch.zeros(nx + 2, ny + 2, 2, dtype=dtype), torch.zeros(nx + 2, ny + 2, 2, dtype=dtype)


def pair3(nx=4, ny=3, nz=5, dtype=torch.float64):
    return torch.zeros(nx + 2, ny + 2, nz + 2, 2, dtype=dtype), torch.zeros(nx + 2, ny + 2, nz + 2, 3, dtype=dtype)


def _launch2(hydro, a, b, dt, **kw):
    return hydro.launcher(a[0], a[1], b[0], b[1], dt, 0.01, 1.4, 0.4, 2.5, 1.1, 0.1, 0.1, **kw)


def test_launcher_rejects_float32_next_buffers_with_float64_state(hydro):
    a = pair2()
    b = pair2(dtype=torch.float32)
    with pytest.raises(hydro.HydroError, match="dtype"):
        _launch2(hydro, a, b, torch.zeros(1, dtype=torch.float64))


def test_launcher_rejects_mismatched_shapes(hydro):
    a = pair2(6, 5)
    with pytest.raises(hydro.HydroError, match="shape"):
        _launch2(hydro, a, pair2(6, 4), torch.zeros(1, dtype=torch.float64))
    with pytest.raises(hydro.HydroError, match="momentum of shape"):
        _launch2(hydro, (a[0], torch.zeros(8, 7, 3, dtype=torch.float64)), pair2(), torch.zeros(1, dtype=torch.float64))
    with pytest.raises(hydro.HydroError, match="2D"):
        hydro.launcher(*pair3(), *pair3(), torch.zeros(1, dtype=torch.float64), 0.01, 1.4, 0.4, 2.5, 1.1, 0.1, 0.1)


def test_launcher_rejects_bad_dt_next(hydro):
    a, b = pair2(), pair2()
    with pytest.raises(hydro.HydroError, match="Dt_next dtype"):
        _launch2(hydro, a, b, torch.zeros(1, dtype=torch.float32))
    with pytest.raises(hydro.HydroError, match="one element"):
        _launch2(hydro, a, b, torch.zeros(0, dtype=torch.float64))


def test_cpu_tensors_are_refused_after_the_shape_checks(hydro):
    a, b = pair2(), pair2()
    with pytest.raises(hydro.HydroError, match="device tensor"):
        _launch2(hydro, a, b, torch.zeros(1, dtype=torch.float64))


def test_step_wrappers_need_float64(hydro):
    a, b = pair2(dtype=torch.float32), pair2(dtype=torch.float32)
    with pytest.raises(hydro.HydroError, match="float64"):
        hydro.step2d(a[0], a[1], b[0], b[1], torch.zeros(1, dtype=torch.float32), 0.01, 1.4, 0.4, 2.5, 1.1, 0.1,
                     0.1)
    a3, b3 = pair3(dtype=torch.float32), pair3(dtype=torch.float32)
    with pytest.raises(hydro.HydroError, match="float64"):
        hydro.step3d(a3[0], a3[1], b3[0], b3[1], torch.zeros(1, dtype=torch.float32), 0.01, 1.4, 0.4, 2.5, 1.1, 0.1,
                     0.1, 0.1)


def test_ghost_fill_and_ic_check_the_momentum(hydro):
    rhoE, _ = pair3()
    with pytest.raises(hydro.HydroError, match="momentum of shape"):
        hydro.ghost_fill3d(rhoE, torch.zeros(6, 5, 7, 2, dtype=torch.float64))
    with pytest.raises(hydro.HydroError, match="momentum dtype"):
        hydro.ic(rhoE, torch.zeros(6, 5, 7, 3, dtype=torch.float32), 0.01)
    with pytest.raises(hydro.HydroError, match="entries"):
        hydro.ic(*pair3(), 0.01, NB=[4, 3])
    with pytest.raises(hydro.HydroError, match="rhoE of shape"):
        hydro.total_mass(torch.zeros(8, 7, 3, dtype=torch.float64))


def test_run_advance_checks_precision_and_sizes(hydro):
    run = object.__new__(hydro.Run)  # the checks only; no handle is created
    run.dim, run.nb, run.precision = 2, [6, 5], hydro.FP64
    with pytest.raises(hydro.HydroError, match="float64"):
        run._check_buffers(pair2(dtype=torch.float32), pair2(dtype=torch.float32))
    with pytest.raises(hydro.HydroError, match="interior"):
        run._check_buffers(pair2(7, 5), pair2(7, 5))
    with pytest.raises(hydro.HydroError, match="buffer b"):
        run._check_buffers(pair2(), pair2(dtype=torch.float32))
    run.precision = hydro.FP32
    with pytest.raises(hydro.HydroError, match="float32"):
        run._check_buffers(pair2(), pair2())
    run._check_buffers(pair2(dtype=torch.float32), pair2(dtype=torch.float32))
    run.precision, run.dim, run.nb = hydro.FP32_FLUX, 3, [4, 3, 5]
    run._check_buffers(pair3(), pair3())
    with pytest.raises(hydro.HydroError, match="3D"):
        run._check_buffers(pair2(), pair2())


def test_batch_add_checks_buffers(hydro):
    batch = object.__new__(hydro.Batch)
    batch.dim, batch.precision = 2, hydro.FP32_FLUX
    assert batch._check_buffers(pair2(), pair2(), "Batch.add") == (6, 5)
    with pytest.raises(hydro.HydroError, match="float64"):
        batch._check_buffers(pair2(dtype=torch.float32), pair2(dtype=torch.float32), "Batch.add")
    with pytest.raises(hydro.HydroError, match="shape"):
        batch._check_buffers(pair2(6, 5), pair2(6, 6), "Batch.add")


def test_server_result_checks_the_destination(hydro):
    srv = object.__new__(hydro.Server)
    srv._h = None
    srv._nb = {0: [10, 12]}
    need = 12 * 14 * 2 * 8
    with pytest.raises(hydro.HydroError, match=f"needs {need}"):
        srv.result(0, np.zeros(12 * 14 * 2 - 1), None)
    with pytest.raises(hydro.HydroError, match="float64"):
        srv.result(0, np.zeros(12 * 14 * 2, np.float32), None)
    with pytest.raises(hydro.HydroError, match="contiguous"):
        srv.result(0, np.zeros((12 * 14 * 2, 2))[:, 0], None)
    with pytest.raises(hydro.HydroError, match="unknown request"):
        srv.result(3, np.zeros(10), None)
