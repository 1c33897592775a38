"""The seeded synthetic workloads of the five benchmark configurations."""
import math

import numpy as np

from madmax_amd import configs, spec_fibers


def test_c1_is_the_plain_method_case():
    s = configs.c1_spec()
    assert (s.nex, s.ney, s.nez) == (10, 10, 10)
    assert spec_fibers(s) == [(0.5, 0.5, 0.25)]
    assert (s.young[0], s.poisson[0], s.young[1], s.poisson[1]) == (1e7, 0.25, 1e8, 0.3)
    assert 3 * 11 ** 3 == 3993


def test_rng_is_counter_based_and_reproducible():
    a = [configs.Rng(7, 3).uniform() for _ in range(3)]
    r = configs.Rng(7, 3)
    b = [r.uniform() for _ in range(3)]
    assert a[0] == b[0] and len(set(b)) == 3
    assert configs.c2_spec(2, 5).fib_r[0] == configs.c2_spec(2, 5).fib_r[0]
    radii = [configs.c2_spec(2, i).fib_r[0] for i in range(200)]
    assert min(radii) >= 0.15 and max(radii) <= 0.35
    assert np.std(radii) > 0.04


def test_c3_fibers_non_overlapping_vf():
    s = configs.c3_spec()
    f = spec_fibers(s)
    assert len(f) == 4 and s.nex == 100
    assert abs(sum(math.pi * r * r for _, _, r in f) - 0.4) < 1e-12
    for i, (x, y, r) in enumerate(f):
        assert 0.12 <= r <= 0.2 and r <= x <= 1 - r and r <= y <= 1 - r
        for (x2, y2, r2) in f[i + 1:]:
            assert math.hypot(x - x2, y - y2) >= r + r2


def test_c4_ranges():
    specs = configs.c4_specs(count=512)
    sizes = {s.nex for s in specs}
    assert sizes == set(configs.C4_SIZES)
    for s in specs:
        assert 0.1 <= s.fib_r[0] <= 0.4
        assert 5.0 <= s.young[1] / s.young[0] <= 20.0
        assert 0.2 <= s.poisson[0] <= 0.35 and 0.2 <= s.poisson[1] <= 0.35


def test_c5_stream():
    st = configs.c5_stream(count=300)
    assert len(st) == 300
    t = [r["arrival"] for r in st]
    assert all(b > a for a, b in zip(t, t[1:]))
    assert {r["n"] for r in st} == set(configs.C5_SIZES)
    amp = max(np.abs(r["strain"]).max() for r in st)
    assert 1e-4 < amp <= 1e-3
    assert len(configs.c5_stream(nmacro=16)) == 32768
