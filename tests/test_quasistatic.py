"""The quasi-static compression loop of mainIsoElpsPackTestEntropy.cpp:140-180
(xdem_amd.prepare.QuasiStaticCriterion, apply_strain) on the CPU oracle
against the real reference over 40 steps (tests/golden/quasistatic_*.npz):
the same boundary moves at the same steps, the box to 1e-13, contact sets
identical and positions to 1e-12 at the dumped steps up to step 14 (1e-10 at
step 40 of the dense packing: the reference sums an unordered_map, and the
last-bit differences grow over the steps).  The reference scales
the rattlers' velocities after each step and the engine and the oracle
before the next one; at a dump the oracle's rattler velocities are scaled by
the same factor before the comparison."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.golden_io import QUASI_CASES, QUASI_STEPS, load_quasistatic
from xdem_amd import prepare as pp


def run_loop(world, n, steps, on_step=None):
    """The driver's loop on `world` (Engine or OracleWorld); returns the
    steps at which the boundary moved."""
    crit = pp.QuasiStaticCriterion(n)
    moved = []
    for k in range(1, steps + 1):
        world.step(1)
        ob = world.observables()
        ob = ob[0] if ob.ndim == 2 else ob
        if crit(float(ob[6]), int(ob[1]), float(ob[4])):
            world.apply_strain(crit.STRAIN_RATE, crit.STRAIN_RATE)
            moved.append(k)
        if on_step:
            on_step(k)
    return moved


def rattler_scaled(world, p):
    """Velocities as the reference holds them after the step: rattlers' times 0.95."""
    f = world.forces()
    rat = f["num_contacts"] == 0
    out = dict(p)
    for k in ("vx", "vy", "omega"):
        v = p[k].copy()
        v[rat] = v[rat] * 0.95
        out[k] = v
    return out


@pytest.mark.parametrize("name", QUASI_CASES)
def test_quasi_static_loop_against_reference(name):
    s, box, moved, dumps = load_quasistatic(name)
    w = po.OracleWorld(s)
    w.set_control(**pp.quasi_static_control())

    def check(k):
        b = w.box()
        for q in range(4):
            assert abs(b[q] - box[k - 1][q]) <= 1e-13 * abs(box[k - 1][q]), (name, k, q)
        if k in dumps:
            g = dumps[k]
            c = w.contacts()
            assert np.array_equal(c["i"], g["i"]) and np.array_equal(c["j"], g["j"]), (name, k)
            p = rattler_scaled(w, w.particles())
            bar = 1e-12 if k <= 14 else 1e-10
            for key in ("x", "y", "theta"):
                assert float(np.max(np.abs(p[key] - g[key]))) <= bar, (name, key, k)
            for key in ("vx", "vy", "omega"):
                sc = max(float(np.max(np.abs(g[key]))), 1e-300)
                assert float(np.max(np.abs(p[key] - g[key]))) <= 10 * bar * sc, (name, key, k)

    got = run_loop(w, s.n, QUASI_STEPS, check)
    assert got == [k + 1 for k, m in enumerate(moved) if m]
    if name == "loose":
        assert got == [1, 13, 25, 37]


def test_criterion_waits_after_a_move():
    c = pp.QuasiStaticCriterion(100)
    seq = [c(0.0, 10, 1.0) for _ in range(40)]
    assert [k + 1 for k, m in enumerate(seq) if m] == [1, 13, 25, 37]
    c = pp.QuasiStaticCriterion(100)
    # stressed and well connected: moves only while nearly at rest
    assert c(1e-3, 300, 1.0) is False and c(1e-3, 300, 1e-3) is True
    assert c.crit == 250
