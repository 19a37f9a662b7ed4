"""Text checkpoint from and into the GPU engine, against the reference's
writer and reader (tests/golden/checkpoint_*.npz; the same bars as the
oracle's checks in tests/test_checkpoint.py, plus the engine's own parity
slack: numbers within one unit of the ninth printed digit plus 1e-12
relative, and 1e-10 of each contact field's largest value for values that
cancel to ~0)."""
import numpy as np
import pytest

from tests import checkpoint_check as cc
from tests.golden_io import CHECKPOINT_CASES, CHECKPOINT_K, load_checkpoint
from tests.test_checkpoint import check_restart, restart_state
from xdem_amd import checkpoint as ck

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", CHECKPOINT_CASES)
def test_engine_writes_reference_text(name, tmp_path):
    from xdem_amd.engine import Engine
    s, text, _ = load_checkpoint(name)
    print(f"[gpu] checkpoint writer {name}", flush=True)
    eng = Engine(s)
    eng.step(CHECKPOINT_K)
    path = str(tmp_path / "ckpt.txt")
    ck.engine_checkpoint(eng, path)
    eng.close()
    cc.compare(cc.parse(open(path).read()), cc.parse(text), rel=1e-12, scale_frac=1e-10)


@pytest.mark.parametrize("name", CHECKPOINT_CASES)
def test_engine_restart_follows_reference(name, tmp_path):
    from xdem_amd.engine import Engine
    rs, text, steps = restart_state(name, tmp_path)
    print(f"[gpu] checkpoint restart {name}", flush=True)
    eng = Engine(rs)
    check_restart(eng, text, steps, name)
    eng.close()


HISTORY = ("i", "j", "xi_x", "xi_y", "n_x", "n_y", "damping")


def _same(a, b, label, fields=None):
    """Particles and contact fields (all, or `fields`) equal bit for bit."""
    pa, pb = a.particles(), b.particles()
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), (label, k)
    ca, cb = a.contacts(), b.contacts()
    for k in (fields or ca):
        assert np.array_equal(ca[k], cb[k]), (label, k)


@pytest.mark.parametrize("name", ["le_poly", "pb_clumps", "pb_ellipse"])
def test_binary_sidecar_restart_is_exact(name, tmp_path):
    """Engine -> text checkpoint with its binary sidecar -> Engine: the
    restarted engine follows the uninterrupted one bit for bit (particles
    and every contact field), Lees-Edwards shift and contact history
    included."""
    from xdem_amd.engine import Engine
    from xdem_amd.state import read_state
    from tests import helpers
    s = helpers.state(name)
    print(f"[gpu] binary sidecar restart {name}", flush=True)
    a = Engine(s)
    a.step(12)
    side = str(tmp_path / "ckpt.xdst")
    ck.engine_checkpoint(a, str(tmp_path / "ckpt.txt"), sidecar=side)
    b = Engine(read_state(side))
    _same(a, b, f"{name} at restart", HISTORY)
    a.step(9); b.step(9)
    _same(a, b, f"{name} 9 steps after restart")
    assert np.array_equal(a.boundary(), b.boundary())
    a.close(); b.close()


def test_binary_sidecar_restart_under_strain_control(tmp_path):
    """A box that strain control has moved and rescaled comes back exactly;
    with the same control set again, the restart follows the uninterrupted
    run bit for bit."""
    from xdem_amd.engine import CONTROL_STRAIN, Engine
    from xdem_amd.state import read_state, write_state
    from tests import helpers
    s = helpers.state("pb_ellipse")
    print("[gpu] binary sidecar restart under strain control", flush=True)
    ctl = dict(mode=CONTROL_STRAIN, rate_x=-0.02, rate_y=-0.01)
    a = Engine(s)
    a.set_control(**ctl)
    a.step(10)
    path = str(tmp_path / "st.xdst")
    write_state(path, a.export_states()[0])
    b = Engine(read_state(path))
    b.set_control(**ctl)
    assert np.array_equal(a.box()[:, :4], b.box()[:, :4])
    a.step(8); b.step(8)
    _same(a, b, "strain control")
    assert np.array_equal(a.box()[:, :4], b.box()[:, :4])
    a.close(); b.close()


def test_export_states_splits_packings():
    """export_states of a two-packing engine gives each packing with its own
    contacts, indexed within the packing."""
    from xdem_amd.engine import Engine
    from tests import helpers
    s1, s2 = helpers.state("pb_ellipse"), helpers.state("le_poly")
    print("[gpu] export_states two packings", flush=True)
    e = Engine([s1, s2])
    e.step(6)
    ex = e.export_states()
    c = e.contacts()
    assert [x.n for x in ex] == [s1.n, s2.n]
    assert ex[0].ci.size + ex[1].ci.size == c["i"].size
    for k, x in enumerate(ex):
        solo = Engine(x)
        ref = Engine([s1, s2][k])
        ref.step(6)
        _same(solo, ref, f"packing {k}", HISTORY)
        solo.step(3); ref.step(3)
        _same(solo, ref, f"packing {k}, 3 steps on")
        solo.close(); ref.close()
    e.close()
