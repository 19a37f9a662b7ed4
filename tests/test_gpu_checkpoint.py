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
