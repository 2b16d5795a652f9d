"""Primitive fields of a conserved state and the per-field error measures used to compare
forms of the solver (relative L-inf and relative L2 over interior cells)."""
import numpy as np

GAMMA = 1.4


def primitives(rhoE, mom):
    """rho, p, and the velocities (u, v[, w]) of the interior cells."""
    dim = rhoE.ndim - 1
    sl = (slice(1, -1),) * dim
    rho = rhoE[sl + (0,)].astype(np.float64)
    E = rhoE[sl + (1,)].astype(np.float64)
    vel = [mom[sl + (c,)].astype(np.float64) / rho for c in range(mom.shape[-1])]
    kin = 0.5 * rho * sum(v * v for v in vel)
    p = (GAMMA - 1.0) * (E - kin)
    names = ["rho", "p", "u", "v", "w"][: 2 + len(vel)]
    return dict(zip(names, [rho, p] + vel))


def field_errors(got, want):
    """{field: (rel L-inf, rel L2)}: max|a-b| / max|b| and ||a-b||_2 / ||b||_2."""
    a, b = primitives(*got), primitives(*want)
    out = {}
    for k in b:
        d = a[k] - b[k]
        linf = float(np.max(np.abs(d)) / max(np.max(np.abs(b[k])), 1e-300))
        l2 = float(np.sqrt(np.sum(d * d)) / max(np.sqrt(np.sum(b[k] * b[k])), 1e-300))
        out[k] = (linf, l2)
    return out
