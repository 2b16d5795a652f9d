"""Pins of the oracle's FP32 build (the reference's USE_FLOAT, hydro2d/kernel.hpp:9-23)
and of the boundary-mode extension (periodic Y/Z, reflective X), which the reference
does not have (its boundaries are hard-wired: X periodic, hydro2d/main.cxx:349-378; Y/Z
reflective, hydro2d/kernel.cxx:213-251).

FP32: the constants as float arithmetic of hydro2d/main.cxx:482-484, 595-607, formed
here from numpy float32 operations and shown to differ from FP64 constants rounded to
float; the scheme's exact invariants in float (uniform state, mirror symmetry, periodic
translation, wall flux, mass to float round-off); the closed-form flux; Sod against the
exact Riemann solution at float tolerance; and agreement with the pinned FP64 build to
float precision over a blast run.

Boundary modes: translation invariance along a periodic Y (and Z) axis, bit for bit;
axis-swap equivalence (reflective X on a state == reflective Y on its transpose with
rho*u and rho*v exchanged) to one rounding of the update order; and mirror symmetry
with mass conserved when every boundary is reflective.
"""
import math

import numpy as np
import pytest

from riemann_exact import profile

F32 = np.float32
GAMMA = 1.4
P, R = 0, 1  # periodic, reflective


# --------------------------------------------------------------------------- FP32 constants

@pytest.mark.parametrize("nb,C", [((60, 60), 0.5), ((40000, 20000), 0.5), ((170, 170, 170), 0.4),
                                  ((340, 170, 170), 0.4), ((7, 1000, 3), 0.4)])
def test_fp32_constants_are_float_arithmetic(oracle, hydro, nb, C):
    one = np.float32(1.0)
    gamma = np.float32(1.4)
    gmo = np.float32(gamma - one)                         # gamma_minus_one = gamma - ONE
    divgamma = np.float32(one / gmo)                      # ONE / gamma_minus_one
    D = [np.float32(one / np.float32(n + 1)) for n in nb]  # ONE / DATATYPE(NB + 1)
    cmin = np.float32(np.float32(C) * min(D))
    for k in (oracle.constants(nb, C, F32), hydro.constants(nb, C, "f32")):
        for name, want in (("gamma", gamma), ("gmo", gmo), ("divgamma", divgamma), ("C_min", cmin)):
            assert np.float32(k[name]).tobytes() == want.tobytes(), name
        assert [np.float32(d).tobytes() for d in k["D"]] == [d.tobytes() for d in D]
    # The float values are not the double constants rounded to float: 1.4f - 1.0f is
    # 0.39999998f, not float(0.3999999999999999) = 0.4f.
    k64 = oracle.constants(nb, C)
    assert np.float32(k64["gmo"]) != gmo
    assert np.float32(k64["divgamma"]) != divgamma


# --------------------------------------------------------------------------- FP32 invariants

def test_fp32_uniform_state_is_stationary(oracle):
    for u in (F32(0.0), F32(0.7)):
        rhoE, uv = oracle.alloc2d(33, 21, F32)
        rhoE[..., 0] = F32(1.3)
        rhoE[..., 1] = F32(2.0) / F32(F32(GAMMA) - F32(1)) + F32(0.5) * F32(1.3) * u * u
        uv[..., 0] = F32(1.3) * u
        k = oracle.constants((33, 21), 0.5, F32)
        before = (rhoE.copy(), uv.copy())
        oracle.ghost_fill2d(rhoE, uv)
        rn, un, _ = oracle.launcher2d(rhoE, uv, F32(0.3), F32(0.3), k["C_min"])
        assert np.array_equal(rn[1:-1, 1:-1], before[0][1:-1, 1:-1])
        assert np.array_equal(un[1:-1, 1:-1], before[1][1:-1, 1:-1])


def test_fp32_flux_closed_forms(oracle):
    for rho, u, v, p in ((1.0, 0.3, -0.2, 2.0), (0.125, -1.7, 0.4, 0.1)):
        s = np.array([rho, u, v, p], F32)
        F = oracle.flux2(s, s, F32).astype(np.float64)
        rho, u, v, p = (float(x) for x in s)
        E = p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v)
        want = np.array([rho * u, rho * u * u + p, rho * u * v, u * (E + p)])
        np.testing.assert_allclose(F, want, rtol=3e-6, atol=1e-6)
        W = oracle.flux2(np.array([rho, -u, v, p], F32), s, F32)
        assert W[0] == 0 and W[2] == 0 and W[3] == 0  # a wall carries no mass or energy


def test_fp32_blast_mirror_symmetry_and_mass(oracle):
    rhoE, uv, dt0 = oracle.ic2d(90, 70, F32)
    m0 = oracle.mass2d(rhoE)
    oracle.run2d(rhoE, uv, dt0, 2.0, 120)
    I, M = rhoE[1:-1, 1:-1], uv[1:-1, 1:-1]
    assert np.array_equal(I, I[::-1, :]) and np.array_equal(I, I[:, ::-1])
    assert np.array_equal(M[..., 0], -M[::-1, :, 0]) and np.array_equal(M[..., 1], -M[:, ::-1, 1])
    assert abs(oracle.mass2d(rhoE) - m0) / m0 < 1e-6


def test_fp32_periodic_translation(oracle):
    rhoE, uv, dt0 = oracle.ic2d(64, 40, F32)
    rs, us = rhoE.copy(), uv.copy()
    rs[1:-1] = np.roll(rhoE[1:-1], 17, axis=0)
    us[1:-1] = np.roll(uv[1:-1], 17, axis=0)
    oracle.run2d(rhoE, uv, dt0, 2.0, 60)
    oracle.run2d(rs, us, dt0, 2.0, 60)
    assert np.array_equal(np.roll(rhoE[1:-1, 1:-1], 17, axis=0).view(np.uint32), rs[1:-1, 1:-1].view(np.uint32))
    assert np.array_equal(np.roll(uv[1:-1, 1:-1], 17, axis=0).view(np.uint32), us[1:-1, 1:-1].view(np.uint32))


def _rel_linf(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("dim", [2, 3])
def test_fp32_build_tracks_the_fp64_build(oracle, dim):
    """The same blast in FP32 and FP64: the FP32 build stays within float precision of
    the pinned FP64 build (a wrong float literal, constant or sqrt would not)."""
    if dim == 2:
        r64, m64, d64 = oracle.ic2d(96, 72)
        r32, m32, d32 = oracle.ic2d(96, 72, F32)
        a = oracle.run2d(r64, m64, d64, 2.0, 80)
        b = oracle.run2d(r32, m32, d32, 2.0, 80)
    else:
        r64, m64, d64 = oracle.ic3d(40, 36, 32)
        r32, m32, d32 = oracle.ic3d(40, 36, 32, F32)
        a = oracle.run3d(r64, m64, d64, 2.0, 50)
        b = oracle.run3d(r32, m32, d32, 2.0, 50)
    assert abs(float(b["t"]) / a["t"] - 1) < 1e-5
    assert abs(float(b["dt"]) / a["dt"] - 1) < 1e-5
    assert _rel_linf(r32[..., 0], r64[..., 0]) < 2e-5
    assert _rel_linf(r32[..., 1], r64[..., 1]) < 2e-5
    for c in range(dim):
        assert _rel_linf(m32[..., c], m64[..., c]) < 1e-4


def _sod32(oracle, n, t_end=0.1):
    rhoE, uv = oracle.alloc2d(n, 1, F32)
    half = n // 2
    gmo = F32(F32(GAMMA) - F32(1))
    rhoE[1:half + 1, 1] = (F32(1.0), F32(1.0) / gmo)
    rhoE[half + 1:n + 1, 1] = (F32(0.125), F32(0.1) / gmo)
    k = oracle.constants((n, 1), 0.5, F32)
    dx = float(k["D"][0])
    dt0 = F32(k["C_min"] / F32(math.sqrt(GAMMA)))
    used = [float(dt0)]
    res = oracle.run2d(rhoE, uv, dt0, 10.0, 1, nthreads=1)
    while sum(used) < t_end:
        dt = res["dt"]
        res = oracle.run2d(rhoE, uv, dt, 10.0, 1, nthreads=1)
        used.append(float(dt))
    xc = (np.arange(1, n + 1) - 0.5) * dx
    x0 = half * dx
    rho_ex, _, _ = profile(xc, x0, sum(used), (1.0, 0.0, 1.0), (0.125, 0.0, 0.1))
    sel = np.abs(xc - x0) < 0.3
    return np.sum(np.abs(rhoE[1:-1, 1, 0][sel].astype(np.float64) - rho_ex[sel])) * dx


def test_fp32_sod_converges_to_exact_riemann_solution(oracle):
    e200 = _sod32(oracle, 200)
    e800 = _sod32(oracle, 800)
    assert e800 < 0.012
    assert e200 / e800 > 1.6


# --------------------------------------------------------------------------- boundary modes

def _roll_state(a, s, axis):
    """Roll the interior of a ghosted array by s cells along `axis`."""
    b = a.copy()
    idx = tuple(slice(1, -1) for _ in range(a.ndim - 1))
    b[idx] = np.roll(a[idx], s, axis=axis)
    return b


def _random_blast2d(oracle, nx, ny, seed):
    rhoE, uv, dt0 = oracle.ic2d(nx, ny, ic_mode=1, radius_frac=0.15)
    rng = np.random.default_rng(seed)
    uv[1:-1, 1:-1] += rng.normal(0, 0.1, uv[1:-1, 1:-1].shape) * rhoE[1:-1, 1:-1, :1]
    rhoE[1:-1, 1:-1, 1] += 0.5 * (uv[1:-1, 1:-1] ** 2).sum(-1) / rhoE[1:-1, 1:-1, 0]
    return rhoE, uv, dt0


@pytest.mark.parametrize("bc_x", [P, R])
def test_periodic_y_translation_2d(oracle, bc_x):
    rhoE, uv, dt0 = _random_blast2d(oracle, 36, 50, 3)
    rs, us = _roll_state(rhoE, 13, 1), _roll_state(uv, 13, 1)
    oracle.run2d(rhoE, uv, dt0, 2.0, 40, bc_x=bc_x, bc_y=P)
    oracle.run2d(rs, us, dt0, 2.0, 40, bc_x=bc_x, bc_y=P)
    assert np.array_equal(np.roll(rhoE[1:-1, 1:-1], 13, axis=1).view(np.uint64), rs[1:-1, 1:-1].view(np.uint64))
    assert np.array_equal(np.roll(uv[1:-1, 1:-1], 13, axis=1).view(np.uint64), us[1:-1, 1:-1].view(np.uint64))


@pytest.mark.parametrize("axis", [1, 2])
def test_periodic_translation_3d(oracle, axis):
    rhoE, uvw, dt0 = oracle.ic3d(20, 24, 28, ic_mode=1, radius_frac=0.2)
    rng = np.random.default_rng(axis)
    uvw[1:-1, 1:-1, 1:-1] += rng.normal(0, 0.1, uvw[1:-1, 1:-1, 1:-1].shape)
    rhoE[1:-1, 1:-1, 1:-1, 1] += 0.5 * (uvw[1:-1, 1:-1, 1:-1] ** 2).sum(-1) / rhoE[1:-1, 1:-1, 1:-1, 0]
    rs, us = _roll_state(rhoE, 7, axis), _roll_state(uvw, 7, axis)
    bc = dict(bc_x=R, bc_y=P if axis == 1 else R, bc_z=P if axis == 2 else R)
    oracle.run3d(rhoE, uvw, dt0, 2.0, 25, **bc)
    oracle.run3d(rs, us, dt0, 2.0, 25, **bc)
    I = (slice(1, -1),) * 3
    assert np.array_equal(np.roll(rhoE[I], 7, axis=axis).view(np.uint64), rs[I].view(np.uint64))
    assert np.array_equal(np.roll(uvw[I], 7, axis=axis).view(np.uint64), us[I].view(np.uint64))


def test_axis_swap_reflective_x_equals_reflective_y_2d(oracle):
    """Reflective X / periodic Y on a state equals periodic X / reflective Y on its
    transpose with rho*u and rho*v exchanged.  The faces are the same values bit for
    bit (the X and Y face arguments are the same after the exchange); only the update
    adds the X and Y increments in the other order, one rounding per step."""
    rhoE, uv, dt0 = _random_blast2d(oracle, 44, 30, 11)
    rt = np.ascontiguousarray(rhoE.transpose(1, 0, 2))
    ut = np.ascontiguousarray(uv.transpose(1, 0, 2)[..., ::-1])
    a = oracle.run2d(rhoE, uv, dt0, 2.0, 60, bc_x=R, bc_y=P)
    b = oracle.run2d(rt, ut, dt0, 2.0, 60, bc_x=P, bc_y=R)
    assert a["it"] == b["it"] == 60
    assert abs(b["t"] / a["t"] - 1) < 1e-13
    back_r = rt.transpose(1, 0, 2)[1:-1, 1:-1]
    back_u = ut.transpose(1, 0, 2)[1:-1, 1:-1, ::-1]
    assert _rel_linf(back_r[..., 0], rhoE[1:-1, 1:-1, 0]) < 1e-13
    assert _rel_linf(back_r[..., 1], rhoE[1:-1, 1:-1, 1]) < 1e-13
    assert _rel_linf(back_u, uv[1:-1, 1:-1]) < 1e-12


def test_axis_swap_reflective_y_equals_reflective_z_3d(oracle):
    """3D analogue: reflective Y / periodic Z == periodic Y / reflective Z on the Y-Z
    transpose with rho*v and rho*w exchanged (equal to rounding: the Z faces order the
    tangential kinetic terms as (w, v, u) and the update adds Y before Z)."""
    rhoE, uvw, dt0 = oracle.ic3d(18, 22, 26, ic_mode=1, radius_frac=0.2)
    rng = np.random.default_rng(4)
    uvw[1:-1, 1:-1, 1:-1] += rng.normal(0, 0.1, uvw[1:-1, 1:-1, 1:-1].shape)
    rhoE[1:-1, 1:-1, 1:-1, 1] += 0.5 * (uvw[1:-1, 1:-1, 1:-1] ** 2).sum(-1) / rhoE[1:-1, 1:-1, 1:-1, 0]
    rt = np.ascontiguousarray(rhoE.transpose(0, 2, 1, 3))
    ut = np.ascontiguousarray(uvw.transpose(0, 2, 1, 3)[..., [0, 2, 1]])
    a = oracle.run3d(rhoE, uvw, dt0, 2.0, 30, bc_x=P, bc_y=R, bc_z=P)
    b = oracle.run3d(rt, ut, dt0, 2.0, 30, bc_x=P, bc_y=P, bc_z=R)
    assert abs(b["t"] / a["t"] - 1) < 1e-13
    I = (slice(1, -1),) * 3
    back_r = rt.transpose(0, 2, 1, 3)[I]
    back_u = ut.transpose(0, 2, 1, 3)[I][..., [0, 2, 1]]
    assert _rel_linf(back_r[..., 0], rhoE[I][..., 0]) < 1e-13
    assert _rel_linf(back_u, uvw[I]) < 1e-12


def test_all_reflective_blast_is_symmetric_and_conserves_mass(oracle):
    """Reflective on every side: the centred blast keeps its X and Y mirror symmetry bit
    for bit and the walls carry no mass."""
    rhoE, uv, dt0 = oracle.ic2d(80, 60)
    m0 = oracle.mass2d(rhoE)
    oracle.run2d(rhoE, uv, dt0, 2.0, 150, bc_x=R, bc_y=R)
    I, M = rhoE[1:-1, 1:-1], uv[1:-1, 1:-1]
    assert np.array_equal(I, I[::-1, :]) and np.array_equal(I, I[:, ::-1])
    assert np.array_equal(M[..., 0], -M[::-1, :, 0]) and np.array_equal(M[..., 1], -M[:, ::-1, 1])
    assert abs(oracle.mass2d(rhoE) - m0) / m0 < 1e-14


def test_reflective_x_ghosts_mirror_the_adjacent_plane(oracle):
    """The reflective X ghost fill gives the face the wall mirror of hydro2d/kernel.cxx:
    214-229 turned to X: a copy of the adjacent interior plane with rho*u negated."""
    rhoE, uv, _ = _random_blast2d(oracle, 9, 7, 2)
    oracle.ghost_fill2d(rhoE, uv, bc_x=R, bc_y=R)
    assert np.array_equal(rhoE[0], rhoE[1]) and np.array_equal(rhoE[-1], rhoE[-2])
    assert np.array_equal(uv[0, :, 0], -uv[1, :, 0]) and np.array_equal(uv[0, :, 1], uv[1, :, 1])
    assert np.array_equal(uv[-1, :, 0], -uv[-2, :, 0]) and np.array_equal(uv[-1, :, 1], uv[-2, :, 1])
