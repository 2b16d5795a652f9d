"""Input builders and bitwise comparison helpers shared by the GPU parity tests."""
from __future__ import annotations

import numpy as np

GAMMA = 1.4


def random_state_2d(nb_x, nb_y, seed, dtype=np.float64):
    """Random positive-density, positive-pressure states with velocities of both signs
    (both upwind branches of the flux), ghosts included."""
    rng = np.random.default_rng(seed)
    shp = (nb_x + 2, nb_y + 2)
    rho = rng.uniform(0.5, 2.0, shp)
    u = rng.normal(0.0, 1.0, shp)
    v = rng.normal(0.0, 1.0, shp)
    p = rng.uniform(0.1, 10.0, shp)
    E = p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v)
    rhoE = np.stack([rho, E], -1).astype(dtype)
    uv = np.stack([rho * u, rho * v], -1).astype(dtype)
    return np.ascontiguousarray(rhoE), np.ascontiguousarray(uv)


def random_state_3d(nb_x, nb_y, nb_z, seed, dtype=np.float64):
    rng = np.random.default_rng(seed)
    shp = (nb_x + 2, nb_y + 2, nb_z + 2)
    rho = rng.uniform(0.5, 2.0, shp)
    u = rng.normal(0.0, 1.0, shp)
    v = rng.normal(0.0, 1.0, shp)
    w = rng.normal(0.0, 1.0, shp)
    p = rng.uniform(0.1, 10.0, shp)
    E = p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    rhoE = np.stack([rho, E], -1).astype(dtype)
    uvw = np.stack([rho * u, rho * v, rho * w], -1).astype(dtype)
    return np.ascontiguousarray(rhoE), np.ascontiguousarray(uvw)


def bits(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def assert_bitwise(got: np.ndarray, want: np.ndarray, what: str) -> None:
    gb, wb = bits(got), bits(want)
    if not np.array_equal(gb, wb):
        bad = np.argwhere(gb != wb)
        i = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} values differ bitwise; first at {i}: got {got[i]!r} "
                             f"want {want[i]!r}")


def interior2(a):
    return a[1:-1, 1:-1]


def interior3(a):
    return a[1:-1, 1:-1, 1:-1]


def to_dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def to_host(t):
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy()
