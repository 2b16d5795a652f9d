"""hydro_run_noalloc, the resident loop for callers that cannot allocate, against the
allocating entry points it stands for (hydro_run_create / advance / status /
trajectory), bit for bit: two calls in a row, the second on a smaller grid, then a
third in 3D with FP32 state and a reflective X boundary; and the rejection of a max_it
beyond its scratch record.  The entry point has no counterpart in the reference."""
import ctypes as C

import numpy as np
import pytest
import torch

from gpu_helpers import assert_bitwise

pytestmark = pytest.mark.gpu

MAX_IT = 65536  # HYDRO_NOALLOC_MAX_IT


def _state(gpu, nb, dtype, C_cfl):
    shape = tuple(n + 2 for n in nb)
    a = [torch.zeros(shape + (2,), dtype=dtype, device="cuda"),
         torch.zeros(shape + (len(nb),), dtype=dtype, device="cuda")]
    k = gpu.constants(nb, C_cfl, "f64" if dtype == torch.float64 else "f32")
    dt0 = float(gpu.ic(a[0], a[1], k["C_min"], ic_mode=1, radius_frac=0.12))
    b = [torch.zeros_like(a[0]), torch.zeros_like(a[1])]
    torch.cuda.synchronize()
    return a, b, k, dt0


def _noalloc(gpu, nb, D, C_min, dt0, max_t, max_it, precision, bc, a, b):
    L = gpu.lib()
    it, t, dt, cur = C.c_size_t(), C.c_double(), C.c_double(), C.c_int()
    tt = np.zeros(max(max_it, 1))
    td = np.zeros(max(max_it, 1))
    rc = L.hydro_run_noalloc(len(nb), (C.c_size_t * 3)(*nb), (C.c_double * 3)(*[float(x) for x in D]), float(C_min),
                             float(dt0), float(max_t), max_it, precision, (C.c_int * 3)(*bc),
                             C.c_void_p(a[0].data_ptr()), C.c_void_p(a[1].data_ptr()), C.c_void_p(b[0].data_ptr()),
                             C.c_void_p(b[1].data_ptr()), C.byref(it), C.byref(t), C.byref(dt), C.byref(cur),
                             tt.ctypes.data_as(C.c_void_p), td.ctypes.data_as(C.c_void_p),
                             C.c_void_p(torch.cuda.current_stream().cuda_stream))
    return rc, {"it": it.value, "t": t.value, "dt": dt.value, "current": cur.value}, tt[:it.value], td[:it.value]


CASES = [
    # dim, nb, dtype, precision, C, bc, max_t, max_it
    (2, (320, 180), torch.float64, 64, 0.5, (0, 1, 1), 2.0, 120),
    (2, (96, 64), torch.float64, 64, 0.5, (0, 1, 1), 2.0, 60),
    (3, (40, 30, 20), torch.float32, 32, 0.4, (1, 1, 0), 0.05, 90),
]


def test_noalloc_matches_the_resident_run(gpu):
    for dim, nb, dtype, precision, C_cfl, bc, max_t, max_it in CASES:
        print(f"[noalloc] {dim}D {nb} precision {precision} bc {bc} max_it {max_it}", flush=True)
        a, b, k, dt0 = _state(gpu, nb, dtype, C_cfl)
        a2 = [x.clone() for x in a]
        b2 = [torch.zeros_like(x) for x in b]
        run = gpu.Run(list(nb), k["D"], k["C_min"], dt0, max_t, max_it, precision, list(bc))
        run.advance(a, b, max_it)
        st = run.status()
        assert st["rc"] == 0, st
        want_t, want_dt = run.trajectory(st["it"])
        run.close()
        rc, got, got_t, got_dt = _noalloc(gpu, nb, k["D"], k["C_min"], dt0, max_t, max_it, precision, bc, a2, b2)
        assert rc == 0, gpu.last_error()
        assert (got["it"], got["t"], got["dt"], got["current"]) == (st["it"], st["t"], st["dt"], st["current"])
        np.testing.assert_array_equal(got_t, want_t)
        np.testing.assert_array_equal(got_dt, want_dt)
        want = a if st["current"] == 0 else b
        have = a2 if got["current"] == 0 else b2
        torch.cuda.synchronize()
        inner = (slice(1, -1),) * dim
        for x, y, what in ((have[0], want[0], "rhoE"), (have[1], want[1], "momentum")):
            assert_bitwise(x.cpu().numpy()[inner], y.cpu().numpy()[inner], f"{what} {nb}")


def test_noalloc_rejects_max_it_beyond_its_scratch(gpu):
    nb = (16, 16)
    a, b, k, dt0 = _state(gpu, nb, torch.float64, 0.5)
    rc, _, _, _ = _noalloc(gpu, nb, k["D"], k["C_min"], dt0, 1.0, MAX_IT + 1, 64, (0, 1, 1), a, b)
    assert rc == -1
    assert "exceeds" in gpu.last_error()
