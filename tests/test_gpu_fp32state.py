"""GPU parity of the USE_FLOAT build (hydro2d/kernel.hpp:9-23, hydro2d/Makefile:7-9):
every state value and constant in FP32, bit for bit against the oracle's FP32 build.

The constants are computed in FP32 as the reference computes them (gamma - ONE in
float is not float(1.4 - 1.0)): the product side takes them from the package's own
constants(..., "f32"), the oracle side from the oracle's, and
tests/test_oracle_pins.py::test_fp32_constants_are_float_arithmetic pins both to float
arithmetic formed independently in numpy.
"""
import numpy as np
import pytest
import torch

from gpu_helpers import assert_bitwise, interior2, interior3, random_state_2d, random_state_3d, to_dev, to_host

pytestmark = pytest.mark.gpu

F32 = np.float32
SENTINEL = 12345.0


def _k(gpu, nb, C):
    return gpu.constants(nb, C, "f32")


def _launch(gpu, dim, d, n, dtn, k, r):
    fn = gpu.launcher if dim == 2 else gpu.launcher3d
    us = fn(d[0], d[1], n[0], n[1], dtn, k["C_min"], k["gamma"], k["gmo"], k["divgamma"], k["K"], *r)
    assert us > 0


@pytest.mark.parametrize("nb", [(5, 3), (33, 62), (20, 63), (130, 125)])
def test_launcher_f32_random_states_2d(gpu, oracle, nb):
    rhoE, uv = random_state_2d(*nb, seed=7 * nb[0] + nb[1], dtype=F32)
    k = _k(gpu, nb, 0.5)
    r = (F32(0.02), F32(0.02))
    want_r, want_u, want_dt = oracle.launcher2d(rhoE, uv, *r, k["C_min"])
    assert np.isfinite(want_dt)
    d = [to_dev(rhoE), to_dev(uv)]
    n = [torch.full_like(d[0], SENTINEL), torch.full_like(d[1], SENTINEL)]
    dtn = torch.zeros(1, dtype=torch.float32, device="cuda")
    _launch(gpu, 2, d, n, dtn, k, r)
    assert_bitwise(interior2(to_host(n[0])), interior2(want_r), f"rhoE_next {nb}")
    assert_bitwise(interior2(to_host(n[1])), interior2(want_u), f"uv_next {nb}")
    assert_bitwise(to_host(dtn), np.array([want_dt], F32), "Dt_next")


@pytest.mark.parametrize("nb", [(4, 5, 7), (10, 6, 62), (9, 7, 63), (40, 8, 125), (33, 20, 30), (64, 200, 130)])
def test_launcher_f32_random_states_3d(gpu, oracle, nb):
    rhoE, uvw = random_state_3d(*nb, seed=nb[0] * 10000 + nb[1] * 100 + nb[2], dtype=F32)
    k = _k(gpu, nb, 0.4)
    r = (F32(0.015), F32(0.015), F32(0.015))
    want_r, want_u, want_dt = oracle.launcher3d(rhoE, uvw, *r, k["C_min"])
    assert np.isfinite(want_dt)
    d = [to_dev(rhoE), to_dev(uvw)]
    n = [torch.full_like(d[0], SENTINEL), torch.full_like(d[1], SENTINEL)]
    dtn = torch.zeros(1, dtype=torch.float32, device="cuda")
    _launch(gpu, 3, d, n, dtn, k, r)
    assert_bitwise(interior3(to_host(n[0])), interior3(want_r), f"rhoE_next {nb}")
    assert_bitwise(interior3(to_host(n[1])), interior3(want_u), f"uvw_next {nb}")
    assert_bitwise(to_host(dtn), np.array([want_dt], F32), "Dt_next")


@pytest.mark.parametrize("dim,nb,steps", [(2, (100, 37), 25), (3, (20, 20, 20), 12), (3, (31, 17, 70), 12)])
def test_launcher_f32_blast_sequence(gpu, oracle, dim, nb, steps):
    """Host ghost fill and host Dt per step, as in hydro2d/main.cxx:689-815 with
    DATATYPE = float: Dt / DX is formed in float."""
    if dim == 2:
        rhoE, mom, dt = oracle.ic2d(*nb, dtype=F32)
        k = _k(gpu, nb, 0.5)
    else:
        rhoE, mom, dt = oracle.ic3d(*nb, dtype=F32)
        k = _k(gpu, nb, 0.4)
    d = [to_dev(rhoE), to_dev(mom)]
    n = [torch.zeros_like(d[0]), torch.zeros_like(d[1])]
    dtn = torch.zeros(1, dtype=torch.float32, device="cuda")
    nr, nm = np.zeros_like(rhoE), np.zeros_like(mom)
    for it in range(steps):
        if dim == 2:
            oracle.ghost_fill2d(rhoE, mom)
            gpu.ghost_fill2d(d[0], d[1])
        else:
            oracle.ghost_fill3d(rhoE, mom)
            gpu.ghost_fill3d(d[0], d[1])
        r = [F32(dt) / x for x in k["D"]]
        if dim == 2:
            nr, nm, want_dt = oracle.launcher2d(rhoE, mom, *r, k["C_min"], out=(nr, nm))
        else:
            nr, nm, want_dt = oracle.launcher3d(rhoE, mom, *r, k["C_min"], out=(nr, nm))
        _launch(gpu, dim, d, n, dtn, k, r)
        got_dt = to_host(dtn)[0]
        assert got_dt == want_dt, (it, got_dt, want_dt)
        rhoE, nr = nr, rhoE
        mom, nm = nm, mom
        d, n = n, d
        dt = want_dt
    inner = interior2 if dim == 2 else interior3
    assert_bitwise(inner(to_host(d[0])), inner(rhoE), f"rhoE after {steps} steps")
    assert_bitwise(inner(to_host(d[1])), inner(mom), f"momentum after {steps} steps")


@pytest.mark.parametrize("dim,nb,bc,steps", [(2, (60, 60), (0, 1), 150), (2, (40, 70), (0, 0), 100),
                                             (3, (20, 20, 20), (0, 1, 1), 50), (3, (30, 14, 40), (1, 0, 1), 40)])
def test_resident_run_f32(gpu, oracle, dim, nb, bc, steps):
    """The device-resident loop in FP32 state against the oracle's FP32 loop: every
    step's Dt and t, and the final fields."""
    if dim == 2:
        rhoE, mom, dt0 = oracle.ic2d(*nb, dtype=F32)
        k = _k(gpu, nb, 0.5)
    else:
        rhoE, mom, dt0 = oracle.ic3d(*nb, dtype=F32)
        k = _k(gpu, nb, 0.4)
    a = [to_dev(rhoE), to_dev(mom)]
    b = [torch.zeros_like(a[0]), torch.zeros_like(a[1])]
    run = gpu.Run(list(nb), k["D"], k["C_min"], dt0, 2.0, steps, 32, list(bc))
    run.advance(a, b, steps)
    st = run.status()
    assert st["rc"] == 0, st
    tt, td = run.trajectory(st["it"])
    cur = a if st["current"] == 0 else b
    got = (to_host(cur[0]), to_host(cur[1]))
    run.close()
    if dim == 2:
        res = oracle.run2d(rhoE, mom, dt0, 2.0, steps, bc_x=bc[0], bc_y=bc[1])
    else:
        res = oracle.run3d(rhoE, mom, dt0, 2.0, steps, bc_x=bc[0], bc_y=bc[1], bc_z=bc[2])
    assert st["it"] == res["it"] == steps
    np.testing.assert_array_equal(tt, res["traj_t"])
    np.testing.assert_array_equal(td, res["traj_dt"])
    inner = interior2 if dim == 2 else interior3
    assert_bitwise(inner(got[0]), inner(rhoE), "rhoE after run")
    assert_bitwise(inner(got[1]), inner(mom), "momentum after run")


@pytest.mark.parametrize("dim,nb,steps", [(2, (40000, 20000), 3), (3, (512, 512, 512), 5)])
def test_resident_run_f32_full_size_first_steps(gpu, oracle, dim, nb, steps):
    """The FP32-state benchmark sizes themselves (40000 x 20000 with the reference IC,
    512^3 with the parameterised one): the first steps, every interior value and each
    step's t and Dt, equal the oracle's FP32 loop."""
    ic_kw = {} if dim == 2 else {"ic_mode": 1, "radius_frac": 0.1}
    if dim == 2:
        rhoE, mom, dt0 = oracle.ic2d(*nb, dtype=F32)
        k = _k(gpu, nb, 0.5)
    else:
        rhoE, mom, dt0 = oracle.ic3d(*nb, dtype=F32, **ic_kw)
        k = _k(gpu, nb, 0.4)
    a = [to_dev(rhoE), to_dev(mom)]
    b = [torch.zeros_like(a[0]), torch.zeros_like(a[1])]
    run = gpu.Run(list(nb), k["D"], k["C_min"], dt0, 2.0, steps, 32)
    run.advance(a, b, steps)
    st = run.status()
    assert st["rc"] == 0, st
    tt, td = run.trajectory(st["it"])
    cur = a if st["current"] == 0 else b
    got = (to_host(cur[0]), to_host(cur[1]))
    run.close()
    del a, b, cur
    torch.cuda.empty_cache()
    print(f"[fp32 full {nb}] GPU {steps} steps done", flush=True)
    if dim == 2:
        res = oracle.run2d(rhoE, mom, dt0, 2.0, steps, nthreads=16)
    else:
        res = oracle.run3d(rhoE, mom, dt0, 2.0, steps, nthreads=16)
    assert st["it"] == res["it"] == steps
    np.testing.assert_array_equal(tt, res["traj_t"])
    np.testing.assert_array_equal(td, res["traj_dt"])
    inner = interior2 if dim == 2 else interior3
    assert_bitwise(inner(got[0]), inner(rhoE), "rhoE")
    assert_bitwise(inner(got[1]), inner(mom), "momentum")


@pytest.mark.skipif(__import__("os").environ.get("HYDRO_LONG") != "1",
                    reason="minutes of oracle time on 16 host threads; run with HYDRO_LONG=1")
@pytest.mark.parametrize("dim,nb,steps,chunk", [(2, (40000, 20000), 100, 10), (3, (512, 512, 512), 200, 20)])
def test_resident_run_f32_full_size_full_length(gpu, oracle, dim, nb, steps, chunk):
    """The FP32-state benchmark sizes for their configurations' full step counts: every
    interior value and every Dt equal the oracle's FP32 loop, run on the host in chunks
    that carry the state and Dt over (with a progress line per chunk)."""
    import time

    ic_kw = {} if dim == 2 else {"ic_mode": 1, "radius_frac": 0.1}
    if dim == 2:
        rhoE, mom, dt0 = oracle.ic2d(*nb, dtype=F32)
        k = _k(gpu, nb, 0.5)
    else:
        rhoE, mom, dt0 = oracle.ic3d(*nb, dtype=F32, **ic_kw)
        k = _k(gpu, nb, 0.4)
    a = [to_dev(rhoE), to_dev(mom)]
    b = [torch.zeros_like(a[0]), torch.zeros_like(a[1])]
    run = gpu.Run(list(nb), k["D"], k["C_min"], dt0, 2.0, steps, 32)
    run.advance(a, b, steps)
    st = run.status()
    assert st["rc"] == 0 and st["it"] == steps, st
    _, td = run.trajectory(steps)
    cur = a if st["current"] == 0 else b
    got = (to_host(cur[0]), to_host(cur[1]))
    run.close()
    del a, b, cur
    torch.cuda.empty_cache()
    print(f"[fp32 full {nb} x {steps}] GPU done", flush=True)
    dt, tds, t0 = dt0, [], time.perf_counter()
    for c in range(steps // chunk):
        step = oracle.run2d if dim == 2 else oracle.run3d
        res = step(rhoE, mom, dt, 2.0, chunk, nthreads=16)
        assert res["it"] == chunk
        tds.append(np.asarray(res["traj_dt"], np.float64))
        dt = res["dt"]
        print(f"[fp32 full {nb} x {steps}] oracle {(c + 1) * chunk} steps after {time.perf_counter() - t0:.0f} s",
              flush=True)
    np.testing.assert_array_equal(np.concatenate(tds), np.asarray(td, np.float64))
    inner = interior2 if dim == 2 else interior3
    assert_bitwise(inner(got[0]), inner(rhoE), "rhoE")
    assert_bitwise(inner(got[1]), inner(mom), "momentum")
