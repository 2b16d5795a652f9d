"""Seeded synthetic workloads of the batched and served forms.

The PRNG is splitmix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014), so the same seed gives the same blasts on every host.  A
uniform double is the top 53 bits of one output times 2^-53.

Blast parameters follow the reference's blast IC (hydro2d/main.cxx:280-338) with the
parameterised radius and pressure ratio of the larger workloads: rho = 1.0 inside, 1.2
outside, p_out = 0.1, p_in = p_out * ratio, radius = radius_frac * min(nb + 1), zero
initial velocity, X periodic, Y/Z reflective.
"""
from __future__ import annotations

import math

SEED = 0x5EED
_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int = SEED):
        self.state = seed & _MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * 2.0**-53)

    def log_uniform(self, lo: float, hi: float) -> float:
        return math.exp(self.uniform(math.log(lo), math.log(hi)))

    def side(self, lo: int, hi: int) -> int:
        return max(lo, min(hi, int(round(self.log_uniform(lo, hi)))))


def blast_params(rng: SplitMix64, dim: int, side_lo: int, side_hi: int) -> dict:
    nb = tuple(rng.side(side_lo, side_hi) for _ in range(dim))
    return {
        "nb": nb,
        "radius_frac": rng.uniform(0.05, 0.2),
        "p_in": 0.1 * rng.log_uniform(10.0, 1000.0),
        "p_out": 0.1,
        "rho_in": 1.0,
        "rho_out": 1.2,
        "C": 0.5 if dim == 2 else 0.4,
    }


def batched_blasts(n: int = 256, seed: int = SEED, dim: int = 2, side_lo: int = 64, side_hi: int = 2048) -> list:
    """n independent blasts, each side log-uniform in [side_lo, side_hi] (non-square),
    radius uniform in [0.05, 0.2] of the shorter side, pressure ratio log-uniform in
    [10, 1000]."""
    rng = SplitMix64(seed)
    return [blast_params(rng, dim, side_lo, side_hi) for _ in range(n)]


def served_requests(n: int, rate_per_s: float, seed: int = SEED, frac_2d: float = 0.7, side2d=(128, 4096),
                    side3d=(64, 256), steps=(100, 2000), frac_fp32flux: float = 0.5) -> list:
    """A Poisson stream of n mixed solve requests: arrival times (s, exponential gaps at
    rate_per_s), 70% 2D with sides in side2d, 30% 3D with sides in side3d (log-uniform),
    step counts uniform in `steps`, FP32-flux or FP64 at 50/50, blast parameters as in
    batched_blasts."""
    rng = SplitMix64(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += -math.log(1.0 - rng.uniform()) / rate_per_s
        dim = 2 if rng.uniform() < frac_2d else 3
        lo, hi = side2d if dim == 2 else side3d
        p = blast_params(rng, dim, lo, hi)
        p["steps"] = int(rng.uniform(steps[0], steps[1] + 1))
        p["fp32flux"] = rng.uniform() < frac_fp32flux
        p["arrival"] = t
        p["id"] = i
        out.append(p)
    return out


def state_bytes(nb) -> int:
    """Device bytes of one solve: two buffer pairs of FP64 (rhoE: 2, momentum: dim values
    per cell, ghost layer included)."""
    cells = 1
    for n in nb:
        cells *= n + 2
    return 2 * cells * (2 + len(nb)) * 8
