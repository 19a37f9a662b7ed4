"""The engine's time-history and shape files against the text the
reference's own writers produce for the same state after the same steps
(tests/golden/history_*.npz).  Same layout (line and field counts, tabs,
printf formats); numbers equal to the printed precision (%le keeps seven
significant digits: values that differ in the last bits may print one unit
apart in the last digit, so numbers are compared to 2e-6 relative, and counts
and IDs exactly); contacts of a particle compared as a set (the reference
prints them in its unordered_map's order)."""
import os

import numpy as np
import pytest

from tests.golden_io import HISTORY_CASES, load_history

pytestmark = pytest.mark.gpu


def _tokens_equal(a, b, where):
    ta, tb = a.split("\t"), b.split("\t")
    assert len(ta) == len(tb), f"{where}: field count {len(ta)} != {len(tb)}"
    for x, y in zip(ta, tb):
        if x == y:
            continue
        assert ("e" in x) == ("e" in y) and ("." in x) == ("." in y), f"{where}: format {x!r} {y!r}"
        fx, fy = float(x), float(y)
        assert abs(fx - fy) <= 2e-6 * max(abs(fy), 1e-30), f"{where}: {x} vs {y}"


def _compare_file(mine, ref, name):
    lm, lr = mine.split("\n"), ref.split("\n")
    assert len(lm) == len(lr), f"{name}: {len(lm)} lines, reference {len(lr)}"
    if name == "output_contact":
        # blocks: count line, then the contacts of the write
        k = 0
        while k < len(lr) and lr[k]:
            assert lm[k] == lr[k], f"{name}: contact count line {k}"
            c = int(lr[k])
            block_m = sorted(lm[k + 1:k + 1 + c], key=lambda s: tuple(int(v) for v in s.split("\t")[:2]))
            block_r = sorted(lr[k + 1:k + 1 + c], key=lambda s: tuple(int(v) for v in s.split("\t")[:2]))
            for q, (a, b) in enumerate(zip(block_m, block_r)):
                assert a.split("\t")[:2] == b.split("\t")[:2], f"{name}: contact ids line {q}"
                _tokens_equal(a, b, f"{name} contact {q}")
            k += 1 + c
        return
    for q, (a, b) in enumerate(zip(lm, lr)):
        _tokens_equal(a, b, f"{name} line {q}")


@pytest.mark.parametrize("name", HISTORY_CASES)
def test_history_files_match_reference_writers(name, tmp_path):
    from xdem_amd.engine import Engine
    s, ref = load_history(name)
    eng = Engine(s)
    eng.history_open(str(tmp_path))
    for _ in range(2):
        eng.step(3)
        eng.history_write()
    eng.history_close()
    mine = {f[:-4]: open(os.path.join(tmp_path, f)).read() for f in os.listdir(tmp_path)}
    assert sorted(mine) == sorted(ref), (sorted(mine), sorted(ref))
    print(f"[gpu] history {name}: {len(mine)} files", flush=True)
    for k in ref:
        _compare_file(mine[k], ref[k], k)
