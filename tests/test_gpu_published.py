"""End-to-end against the published runs, at their real sizes, on the GPU.

For each run in tests/golden/res_logs.json (hydro2d/reference_run/*.res,
hydro3d/reference_run/*.res) the device-resident loop reproduces: the initial mass and
Dt, every per-iteration (t, Dt) pair the log prints (t to 5 and Dt to 4 significant
digits) within the first `exact` iterations -- all 2000 of them for five of the six
runs -- and later ones to 2e-3 relative (in the 170^3 run the published FPGA arithmetic,
which is not strict IEEE, drifts by one unit in the last printed Dt digit after
iteration 1232), the stopping iteration and the final t and mass.  The
two-FPGA runs are one resident domain here; with exact ghost exchange and a global MIN
on Dt they are the same computation (hydro2d/main.cxx:381-424, 781).
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

RUNS = [
    # name, C, exact-match window
    ("3d_170x170x170", 0.4, 1200),
    ("3d_340x170x170x2FPGAS", 0.4, 2000),
    ("3d_2000x170x170", 0.4, 2000),
    ("3d_4000x170x170x2FPGAS", 0.4, 2000),
    ("2d_20000x20000", 0.5, 2000),
    ("2d_40000x20000x2FPGAS", 0.5, 2000),
]


def _close_print(value: float, printed: str, digits: int) -> bool:
    if f"{value:.{digits - 1}e}" == printed:
        return True
    p = float(printed)
    unit = 10.0 ** (math.floor(math.log10(abs(p))) - (digits - 1))
    return abs(value - p) <= 0.5000001 * unit


@pytest.mark.parametrize("name,C,exact", RUNS)
def test_published_run(gpu, oracle, name, C, exact):
    r = oracle.load_res_logs()[name]
    nb = r["nb"]
    print(f"[published] {name} {nb} {len(r['trajectory'])} iterations", flush=True)
    k = gpu.constants(nb, C)
    shape = tuple(n + 2 for n in nb)
    a = [torch.zeros(shape + (2,), dtype=torch.float64, device="cuda"),
         torch.zeros(shape + (len(nb),), dtype=torch.float64, device="cuda")]
    b = [torch.zeros_like(a[0]), torch.zeros_like(a[1])]
    dt0 = gpu.ic(a[0], a[1], k["C_min"])
    m0 = gpu.total_mass(a[0])
    assert f"{m0:.7e}" == r["initial_mass"]
    assert f"{dt0:.2e}" == r["initial_dt"]
    run = gpu.Run(nb, k["D"], k["C_min"], dt0, r["max_t"], r["max_it"])
    run.advance(a, b, r["max_it"])
    st = run.status()
    assert st["rc"] == 0, st
    tt, td = run.trajectory(st["it"])
    assert st["it"] == len(r["trajectory"])
    off = []
    for it, (_, t_p, dt_p) in enumerate(r["trajectory"]):
        ok = _close_print(tt[it], t_p, 5) and _close_print(td[it], dt_p, 4)
        if not ok:
            off.append(it)
            assert it >= exact, (it, tt[it], t_p, td[it], dt_p)
            assert abs(td[it] / float(dt_p) - 1) < 2e-3, (it, td[it], dt_p)
    print(f"[published] {name}: {len(off)} of {st['it']} iterations differ in the last printed digit "
          f"(first at {off[0] if off else None})", flush=True)
    final = r["mass_lines"][-1]
    assert final[0] == st["it"]
    assert f"{st['t']:.4e}" == final[1]
    cur = a if st["current"] == 0 else b
    m1 = gpu.total_mass(cur[0])
    assert f"{m1:.7e}" == final[2]
    assert abs(m1 - m0) / np.prod(nb) < 1e-13
    run.close()
    del a, b, cur
    torch.cuda.empty_cache()
