"""The FP32-flux form (FP64 state; FP32 primitives, fluxes and signal speeds with
hardware reciprocal and square root; FP64 increments, update, Dt and reductions)
against the FP64 scheme, per field, after N steps of the blast."""
import numpy as np
import pytest
import torch

from fields import field_errors, own_scale_velocity_errors

pytestmark = pytest.mark.gpu

TOL = 1e-5  # relative L-inf per field, rho p u v (w)


def _run(gpu, nb, steps, precision, ic_mode=0, radius_frac=0.0):
    dim = len(nb)
    C = 0.5 if dim == 2 else 0.4
    k = gpu.constants(nb, C)
    shape = tuple(n + 2 for n in nb)
    a = [torch.zeros(shape + (2,), dtype=torch.float64, device="cuda"),
         torch.zeros(shape + (dim,), dtype=torch.float64, device="cuda")]
    b = [torch.zeros_like(a[0]), torch.zeros_like(a[1])]
    dt0 = gpu.ic(a[0], a[1], k["C_min"], ic_mode=ic_mode, radius_frac=radius_frac)
    m0 = gpu.total_mass(a[0])
    e0 = gpu.total_energy(a[0])
    run = gpu.Run(nb, k["D"], k["C_min"], dt0, 10.0, steps, precision)
    run.advance(a, b, steps)
    st = run.status()
    assert st["rc"] == 0 and st["it"] == steps, st
    cur = a if st["current"] == 0 else b
    m1 = gpu.total_mass(cur[0])
    e1 = gpu.total_energy(cur[0])
    out = (cur[0].cpu().numpy(), cur[1].cpu().numpy())
    tt, td = run.trajectory(steps)
    run.close()
    # Periodic X and reflective walls (zero wall energy flux) conserve total energy to
    # summation round-off, in both forms.
    assert abs(e1 - e0) / e0 < 1e-13, (e0, e1)
    return out, td, abs(m1 - m0) / m0


@pytest.mark.parametrize("nb,steps", [((400, 300), 300), ((1000, 700), 200)])
def test_fp32_flux_2d_vs_fp64(gpu, nb, steps):
    ref, td_ref, dm_ref = _run(gpu, nb, steps, gpu.FP64)
    got, td_got, dm_got = _run(gpu, nb, steps, gpu.FP32_FLUX)
    errs = field_errors(got, ref)
    own = own_scale_velocity_errors(got, ref)
    print(f"[fp32-flux 2D {nb} x {steps}] " + ", ".join(f"{k}: Linf {v[0]:.2e} L2 {v[1]:.2e}" for k, v in errs.items())
          + f"; velocities against their own max {own}"
          + f"; Dt rel {np.max(np.abs(td_got / td_ref - 1)):.2e}; mass drift {dm_got:.1e} (fp64 {dm_ref:.1e})")
    for k, (linf, _) in errs.items():
        assert linf <= TOL, (k, linf)
    for k, e in own.items():
        assert e is None or e <= TOL, (k, e)
    assert dm_got < 1e-13


@pytest.mark.parametrize("nb,steps", [((64, 64, 64), 100), ((96, 80, 72), 80)])
def test_fp32_flux_3d_vs_fp64(gpu, nb, steps):
    ref, td_ref, dm_ref = _run(gpu, nb, steps, gpu.FP64, ic_mode=1, radius_frac=0.1)
    got, td_got, dm_got = _run(gpu, nb, steps, gpu.FP32_FLUX, ic_mode=1, radius_frac=0.1)
    errs = field_errors(got, ref)
    own = own_scale_velocity_errors(got, ref)
    print(f"[fp32-flux 3D {nb} x {steps}] " + ", ".join(f"{k}: Linf {v[0]:.2e} L2 {v[1]:.2e}" for k, v in errs.items())
          + f"; velocities against their own max {own}"
          + f"; Dt rel {np.max(np.abs(td_got / td_ref - 1)):.2e}; mass drift {dm_got:.1e}")
    for k, (linf, _) in errs.items():
        assert linf <= TOL, (k, linf)
    for k, e in own.items():
        assert e is None or e <= TOL, (k, e)
    assert dm_got < 1e-13


@pytest.mark.parametrize("nb,steps", [((64, 60, 100), 40), ((33, 40, 70), 30), ((128, 330, 130), 20)])
def test_fp32_flux_3d_does_not_depend_on_the_tiling(gpu, nb, steps):
    """The plain 3D FP32-flux step (10-row workgroup tiles, XCD-major order where the
    strips split evenly over the XCDs) and the batched one (6-row tiles, persistent
    workers) give bit-identical states, t and Dt: the per-cell arithmetic does not
    depend on the tiling."""
    C, rf, p_in = 0.4, 0.12, 50.0
    k = gpu.constants(nb, C)
    shape = tuple(n + 2 for n in nb)
    a = [torch.zeros(shape + (2,), dtype=torch.float64, device="cuda"),
         torch.zeros(shape + (3,), dtype=torch.float64, device="cuda")]
    b = [torch.zeros_like(a[0]), torch.zeros_like(a[1])]
    dt0 = gpu.ic(a[0], a[1], k["C_min"], ic_mode=1, radius_frac=rf, p_in=p_in, p_out=0.1)
    run = gpu.Run(nb, k["D"], k["C_min"], dt0, 1e30, steps, gpu.FP32_FLUX)
    run.advance(a, b, steps)
    st = run.status()
    run.close()
    assert st["rc"] == 0 and st["it"] == steps, st
    plain = a if st["current"] == 0 else b

    batch = gpu.Batch(3, gpu.FP32_FLUX, 1)
    c = [torch.zeros_like(a[0]), torch.zeros_like(a[1])]
    d = [torch.zeros_like(a[0]), torch.zeros_like(a[1])]
    torch.cuda.synchronize()
    slot = batch.add_blast(c, d, C, rf, p_in, p_out=0.1, max_it=steps)
    batch.advance(steps)
    bst = batch.status(slot)
    assert bst["rc"] == 0 and bst["it"] == steps, bst
    assert bst["t"] == st["t"] and bst["dt"] == st["dt"], (bst, st)
    cur = batch.current(slot)
    sl = (slice(1, -1),) * 3
    for x, y in zip(plain, cur):
        np.testing.assert_array_equal(x[sl].cpu().numpy().view(np.uint64), y[sl].cpu().numpy().view(np.uint64))
    batch.close()
