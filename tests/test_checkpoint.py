"""Text checkpoint (xdem_amd/checkpoint.py) against the reference's own
writer and reader: tests/golden/checkpoint_*.npz hold the text of
BaseWorld::print2File_allParticles after 4 steps and the dumps of a world
the reference rebuilt from that text with its stream constructors, after
each of 3 further steps.

Writer: the oracle's state and contacts after the same 4 steps, written by
write_all_particles, give the reference's text (every number equal at the
printed precision, a particle's contacts compared as a set: the reference
prints them in its unordered_map's order).

Reader: the oracle restarted from the reference's text by
read_all_particles follows the reference's restart with identical contact
sets; positions within 1e-11 (rotation 2e-11).  Velocities differ because
the reference integrates with the printed 1/mass and 1/inertia and this
engine with the reciprocals of the printed mass and inertia (up to ~1e-8
relative apart): after k steps by at most 3e-8 times the sum of the
per-step velocity changes."""
import copy
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import checkpoint_check as cc
from tests.golden_io import CHECKPOINT_CASES, CHECKPOINT_K, CHECKPOINT_M, load_checkpoint
from xdem_amd import checkpoint as ck


@pytest.mark.parametrize("name", CHECKPOINT_CASES)
def test_writer_matches_reference_text(name, tmp_path):
    s, text, _ = load_checkpoint(name)
    o = po.OracleWorld(s)
    o.step(CHECKPOINT_K)
    path = str(tmp_path / "ckpt.txt")
    ck.write_all_particles(path, o.export_state(s), o.contacts(), o.contact_branches())
    ours = open(path).read()
    assert ours.split("\n")[0] == text.split("\n")[0]
    cc.compare(cc.parse(ours), cc.parse(text))


def restart_bounds(text, steps):
    """Per restart step: the bound on the velocity difference (3e-8 times
    the accumulated largest per-step change of each velocity component)."""
    ref = cc.parse(text)
    n = len(ref)
    v0 = {k: np.array([ref[p][0][3 + c] for p in range(n)])
          for c, k in ((0, "vx"), (1, "vy"), (3, "omega"))}
    acc = {k: 0.0 for k in v0}
    out = {}
    prev = v0
    for st in sorted(steps):
        cur = {k: steps[st][k] for k in v0}
        for k in v0:
            acc[k] += float(np.max(np.abs(cur[k] - prev[k])))
        out[st] = {k: 3e-8 * acc[k] for k in v0}
        prev = cur
    return out


def check_restart(world, text, steps, label):
    bounds = restart_bounds(text, steps)
    for st in range(1, CHECKPOINT_M + 1):
        world.step(1)
        p, r = world.particles(), steps[st]
        c = world.contacts()
        assert np.array_equal(c["i"], r["i"]) and np.array_equal(c["j"], r["j"]), \
            f"{label}: contact set differs at restart step {st}"
        for k in ("x", "y"):
            assert float(np.max(np.abs(p[k] - r[k]))) <= 1e-11, (label, k, st)
        assert float(np.max(np.abs(p["theta"] - r["theta"]))) <= 2e-11, (label, "theta", st)
        for k in ("vx", "vy", "omega"):
            assert float(np.max(np.abs(p[k] - r[k]))) <= bounds[st][k], (label, k, st)


def restart_state(name, tmp_path):
    """The reference's checkpoint text read back, with the box and contact
    model of the run and the Lees-Edwards shift after CHECKPOINT_K steps."""
    s, text, steps = load_checkpoint(name)
    o = po.OracleWorld(s)
    o.step(CHECKPOINT_K)
    b = o.boundary()
    cfg = copy.deepcopy(s.cfg)
    cfg.total_shift, cfg.total_strain = float(b[0]), float(b[2])
    path = str(tmp_path / "ref.txt")
    with open(path, "w") as f:
        f.write(text)
    return ck.read_all_particles(path, cfg, aspect_ratio=s.aspect), text, steps


@pytest.mark.parametrize("name", CHECKPOINT_CASES)
def test_reader_restart_follows_reference(name, tmp_path):
    rs, text, steps = restart_state(name, tmp_path)
    assert rs.ci.shape[0] == sum(len(v[1]) for v in cc.parse(text).values())
    check_restart(po.OracleWorld(rs), text, steps, name)


def test_reader_accepts_derived_layouts(tmp_path):
    """EllipticParticle / CircularParticle::print2File_particle lines (16 and
    15 fields) read like the base layout; a 16-field line brings its aspect
    ratio."""
    p = str(tmp_path / "d.txt")
    with open(p, "w") as f:
        f.write("2\n")
        f.write("%7d\t" % 0 + "\t".join("%.8e" % v for v in
                (0.5, 0.5, 2.0, 1.0, 2.0, 0.1, 0.2, 0.3, 0.4, 0.39, 0.39, 2.56, 0.01, 100.0))
                + "\t%4d\n" % 1)
        f.write("\t%7d\t" % 1 + "\t".join("%.8e" % (k + 1) for k in range(25)) + "\n")
        f.write("%7d\t" % 1 + "\t".join("%.8e" % v for v in
                (0.6, 0.6, 3.0, 4.0, 0.0, 0.0, 1.0, 0.0, 1.13, 1.13, 0.885, 0.2, 5.0)) + "\t%4d\n" % 0)
    from xdem_amd.state import WorldConfig
    s = ck.read_all_particles(p, WorldConfig(xmin=-10, xmax=10, ymin=-10, ymax=10))
    assert s.n == 2 and s.aspect[0] == 2.0 and s.r[1] == 0.6
    assert (s.x[0], s.y[0], s.vx[0], s.vy[0], s.theta[0], s.omega[0]) == (1.0, 2.0, 0.1, 0.2, 0.3, 0.4)
    assert (s.x[1], s.y[1], s.mass[1], s.inertia[1]) == (3.0, 4.0, 1.13, 0.2)
    assert list(s.ci) == [0] and list(s.cj) == [1]
    # shear overlap = fields 11-12, normal = 1-2, damping = 25
    assert (s.xi_x[0], s.xi_y[0], s.n_x[0], s.n_y[0], s.damping[0]) == (11.0, 12.0, 1.0, 2.0, 25.0)


def test_writer_reader_round_trip(tmp_path):
    """write -> read gives the state back at printed precision, IDs and
    histories unchanged."""
    s, _, _ = load_checkpoint("pb_ellipse")
    o = po.OracleWorld(s)
    o.step(2)
    st = o.export_state(s)
    p = str(tmp_path / "rt.txt")
    ck.write_all_particles(p, st, o.contacts(), o.contact_branches())
    back = ck.read_all_particles(p, s.cfg, aspect_ratio=s.aspect)
    for k in ("x", "y", "theta", "vx", "vy", "omega", "mass", "inertia", "r"):
        a, b = getattr(back, k), getattr(st, k)
        assert np.all(np.abs(a - b) <= 5.01e-9 * np.abs(b) + 1e-300), k
    assert np.array_equal(back.ci, st.ci) and np.array_equal(back.cj, st.cj)
    for k in ("xi_x", "xi_y", "n_x", "n_y", "damping"):
        a, b = getattr(back, k), getattr(st, k)
        assert np.all(np.abs(a - b) <= 5.01e-9 * np.abs(b) + 1e-300), k


@pytest.mark.parametrize("name", ["pb_clumps", "le_poly"])
def test_oracle_binary_state_restart_is_exact(name, tmp_path):
    """The exact binary state (state.write_state of an exported world,
    clump branch vectors as the clumps have turned them included) restarts
    the oracle bit for bit: 9 further steps equal the uninterrupted run."""
    from tests import helpers
    from xdem_amd.state import read_state, write_state
    s = helpers.state(name)
    a = po.OracleWorld(s)
    a.step(12)
    path = str(tmp_path / "s.xdst")
    write_state(path, a.export_state(s))
    b = po.OracleWorld(read_state(path))
    a.step(9)
    b.step(9)
    pa, pb = a.particles(), b.particles()
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k
    ca, cb = a.contacts(), b.contacts()
    for k in ca:
        assert np.array_equal(ca[k], cb[k]), k
