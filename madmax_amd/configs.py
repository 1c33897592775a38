"""Seeded synthetic workloads of the five benchmark configurations.

Randomness comes from a counter-based generator: draw i of stream s under
seed k is SplitMix64(k * 2^32 + s, i), the output finaliser of Steele, Lea &
Flood (OOPSLA 2014) applied to (key + (i+1) * 0x9E3779B97F4A7C15), mapped to
[0, 1) with 53 bits.  Every RVE therefore has a reproducible geometry that
does not depend on how many other RVEs were drawn.

All RVEs are unit cubes with z-aligned cylindrical fibers
(micro/src/micro_struct.c:98-129) and two linear elastic phases
[MATRIX, FIBER] (src/material.c:84-135).
"""
import math

import numpy as np

from .engine import rve_spec

MASK = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


def _mix(z):
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9 & MASK
    z = (z ^ (z >> 27)) * 0x94D049BB133111EB & MASK
    return z ^ (z >> 31)


class Rng:
    """Counter-based uniform generator keyed by (seed, stream)."""

    def __init__(self, seed, stream=0):
        self.key = ((int(seed) & 0xFFFFFFFF) << 32 | (int(stream) & 0xFFFFFFFF)) & MASK
        self.i = 0

    def uniform(self, lo=0.0, hi=1.0):
        self.i += 1
        u = (_mix((self.key + self.i * GOLDEN) & MASK) >> 11) * (1.0 / (1 << 53))
        return lo + (hi - lo) * u

    def normal(self, mu=0.0, sigma=1.0):
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return mu + sigma * math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def choice(self, seq):
        return seq[min(int(self.uniform() * len(seq)), len(seq) - 1)]


E_MATRIX, NU_MATRIX = 1.0e7, 0.25
E_FIBER, NU_FIBER = 1.0e8, 0.3


def c1_spec():
    """10^3 hex8, centred fiber r = 0.25 (160 fiber elements)."""
    return rve_spec(10, [(0.5, 0.5, 0.25)], (E_MATRIX, E_FIBER), (NU_MATRIX, NU_FIBER))


def c1_strain():
    return np.array([1.0e-3, 0, 0, 0, 0, 0])


def c2_spec(seed=2, index=0, n=30):
    """30^3 hex8, centred fiber with r ~ U[0.15, 0.35]."""
    r = Rng(seed, 1000 + index).uniform(0.15, 0.35)
    return rve_spec(n, [(0.5, 0.5, r)], (E_MATRIX, E_FIBER), (NU_MATRIX, NU_FIBER))


def c3_spec(seed=3, n=100, vf=0.4, nfib=4, rmin=0.12, rmax=0.2):
    """100^3 hex8 with 4 non-overlapping z-fibers of radius in [0.12, 0.2]
    whose total cross-section is vf of the unit cell."""
    rng = Rng(seed, 3000)
    for _ in range(10000):
        radii = [rng.uniform(rmin, rmax) for _ in range(nfib)]
        scale = math.sqrt(vf / (math.pi * sum(r * r for r in radii)))
        radii = [r * scale for r in radii]
        if all(rmin <= r <= rmax for r in radii):
            break
    else:
        raise RuntimeError("no radii with the requested volume fraction")
    for _ in range(100000):
        cs = []
        ok = True
        for r in radii:
            placed = False
            for _ in range(1000):
                x, y = rng.uniform(r, 1 - r), rng.uniform(r, 1 - r)
                if all(math.hypot(x - cx, y - cy) >= r + cr for cx, cy, cr in cs):
                    cs.append((x, y, r))
                    placed = True
                    break
            if not placed:
                ok = False
                break
        if ok:
            return rve_spec(n, cs, (E_MATRIX, E_FIBER), (NU_MATRIX, NU_FIBER))
    raise RuntimeError("fiber placement failed")


def c3_strain(seed=3):
    rng = Rng(seed, 3001)
    return np.array([rng.normal(0.0, 1.0e-3) for _ in range(6)])


C4_SIZES = (8, 12, 16, 20, 24, 32)


def c4_spec(seed=4, index=0):
    """One RVE of the uneven batch: n in {8,12,16,20,24,32}, r ~ U[0.1, 0.4],
    E_f/E_m ~ U[5, 20], nu ~ U[0.2, 0.35] per phase."""
    rng = Rng(seed, 4000 + index)
    n = rng.choice(C4_SIZES)
    r = rng.uniform(0.1, 0.4)
    ratio = rng.uniform(5.0, 20.0)
    nu_m = rng.uniform(0.2, 0.35)
    nu_f = rng.uniform(0.2, 0.35)
    return rve_spec(n, [(0.5, 0.5, r)], (E_MATRIX, E_MATRIX * ratio), (nu_m, nu_f))


def c4_specs(seed=4, count=512):
    return [c4_spec(seed, i) for i in range(count)]


C5_SIZES = (12, 20, 30)


def c5_stream(seed=5, nmacro=16, amp=1.0e-3, rate=2000.0, count=None):
    """FE^2 serving stream: one request per gauss point of a nmacro^3 hex8
    macro cube (8 gauss points per element, src/fem.c:324-361 positions),
    strain from a smooth seeded field of amplitude amp, RVE size from
    {12, 20, 30}^3, Poisson arrivals at `rate` requests per second.
    Returns a list of dicts (arrival time, rve spec, strain, key)."""
    rng = Rng(seed, 5000)
    # smooth field: sum of a few low-order Fourier modes per Voigt component
    modes = []
    for v in range(6):
        comp = []
        for _ in range(3):
            k = (rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0))
            comp.append((rng.uniform(-1.0, 1.0), k, rng.uniform(0.0, 2.0 * math.pi)))
        modes.append(comp)
    g = 0.577350269189626
    gp_off = [(-g, -g, -g), (g, -g, -g), (g, g, -g), (-g, g, -g),
              (-g, -g, g), (g, -g, g), (g, g, g), (-g, g, g)]
    out = []
    t = 0.0
    total = nmacro ** 3 * 8
    for idx in range(total if count is None else min(count, total)):
        e, gp = divmod(idx, 8)
        ex, ey, ez = e % nmacro, (e // nmacro) % nmacro, e // (nmacro * nmacro)
        X = [(c + 0.5 + 0.5 * o) / nmacro for c, o in zip((ex, ey, ez), gp_off[gp])]
        strain = np.zeros(6)
        for v in range(6):
            s = 0.0
            for a, k, ph in modes[v]:
                s += a * math.sin(2 * math.pi * (k[0] * X[0] + k[1] * X[1] + k[2] * X[2]) + ph)
            strain[v] = amp * s / 3.0
        t += -math.log(max(rng.uniform(), 1e-300)) / rate
        rr = Rng(seed, 6000 + idx)
        n = rr.choice(C5_SIZES)
        rad = rr.uniform(0.15, 0.35)
        spec = rve_spec(n, [(0.5, 0.5, rad)], (E_MATRIX, E_FIBER), (NU_MATRIX, NU_FIBER))
        out.append(dict(arrival=t, spec=spec, strain=strain, key=idx + 1, n=n))
    return out
