"""Packing preparation (xdem_amd/prepare.py): the random values of the
reference's drivers, restated from minstd_rand0 and libstdc++'s
uniform_real_distribution, against a program that draws them with the C++
standard library in the same expression shape (tests/cpp/rng_probe.cpp,
built by g++ here); and the box scaling of
scaleBoundarySizeToGivenSolidFraction."""
import math
import os
import subprocess

import numpy as np
import pytest

from xdem_amd import prepare as pp

PROBE = os.path.join(os.path.dirname(__file__), "cpp", "rng_probe")


@pytest.mark.parametrize("ar", [1.0 + 1e-9, 1.5, 2.0, 3.7])
def test_draws_match_the_standard_library(ar):
    if not os.path.exists(PROBE):
        subprocess.run(["make", "-C", os.path.dirname(PROBE), "rng_probe"], check=True)
    n = 500
    out = subprocess.run([PROBE, str(n), repr(ar)], capture_output=True, text=True, check=True)
    ref = np.array([[float(v) for v in line.split()] for line in out.stdout.splitlines()])
    got = np.stack(pp.reference_draws(n, ar), axis=1)
    assert ref.shape == got.shape
    assert np.array_equal(ref, got)


def test_minstd_known_values():
    """The 10000th output of minstd_rand0 seeded with 1 is 1043618065
    (C++ standard, [rand.predef])."""
    e = pp.MinstdRand0(1)
    for _ in range(9999):
        e()
    assert e() == 1043618065
    assert pp.MinstdRand0(0).x == 1


def test_generation_scales_to_solid_fraction():
    s = pp.reference_generation(2000, 2.0)
    area = (s.cfg.xmax - s.cfg.xmin) * (s.cfg.ymax - s.cfg.ymin)
    phi = float(np.sum(math.pi * s.r * s.r)) / area
    assert abs(phi - 0.8) < 1e-12
    assert s.cfg.xmax == -s.cfg.xmin and s.cfg.ymax == -s.cfg.ymin
    assert np.all(np.abs(s.x) <= s.cfg.xmax) and np.all(np.abs(s.y) <= s.cfg.ymax)
