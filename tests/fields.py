"""Primitive fields of a conserved state and the per-field error measures used to compare
forms of the solver (relative L-inf and relative L2 over interior cells).

Scales: rho and p are compared relative to their own maximum magnitude.  Velocity
components are compared relative to the reference's maximum signal speed
max(sqrt(gamma*p/rho) + |u|), the velocity scale of the solve (it sets Dt): a component
that vanishes by symmetry (v in a one-cell-wide channel, u across a one-row periodic
domain) holds only round-off, and an error relative to its own maximum would measure
nothing."""
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


def signal_speed_scale(prims):
    speed = np.sqrt(sum(prims[k] ** 2 for k in ("u", "v", "w") if k in prims))
    return float(np.max(np.sqrt(GAMMA * np.abs(prims["p"]) / prims["rho"]) + speed))


def field_errors(got, want):
    """{field: (rel L-inf, rel L2)} against `want`: max|a-b| / scale and
    ||a-b||_2 / (scale * sqrt(n)), with the scales of the module docstring."""
    a, b = primitives(*got), primitives(*want)
    vscale = max(signal_speed_scale(b), 1e-300)
    out = {}
    for k in b:
        d = a[k] - b[k]
        scale = vscale if k in ("u", "v", "w") else max(float(np.max(np.abs(b[k]))), 1e-300)
        linf = float(np.max(np.abs(d)) / scale)
        l2 = float(np.sqrt(np.mean(d * d)) / scale)
        out[k] = (linf, l2)
    return out


def own_scale_velocity_errors(got, want, floor: float = 1e-2):
    """{component: rel L-inf against the component's own maximum} for the velocity
    components whose maximum in `want` is at least `floor` times the signal-speed scale,
    i.e. that carry a real part of the flow.  A component below that (one that vanishes by
    symmetry, or has nearly died out, as across a 5 x 1 periodic channel after 60 steps)
    is left out (None): against its own tiny maximum the FP32 round-off of the fluxes,
    about 1e-8 of the signal speed, measures nothing about the flow."""
    a, b = primitives(*got), primitives(*want)
    vscale = signal_speed_scale(b)
    out = {}
    for k in ("u", "v", "w"):
        if k not in b:
            continue
        own = float(np.max(np.abs(b[k])))
        out[k] = float(np.max(np.abs(a[k] - b[k])) / own) if own >= floor * vscale else None
    return out
