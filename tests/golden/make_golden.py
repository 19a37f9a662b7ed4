"""Writes the golden fixtures of tests/golden/ by running the real reference.

The reference (Xiaoyu-JIANG/xDEM) is compiled from its own sources by
oracle/Makefile into oracle/_ref/; this script drives it through
oracle/_ref/ref_harness and stores inputs and outputs as compressed npz
files.  Run it in a container where /root/reference is present:

    make -C oracle && python tests/golden/make_golden.py

Fixtures
  world_<case>.npz  initial state (+ optional history) and the reference's
                    particle state, forces, contacts, stress and LE scalars
                    after steps 1, 2, 4 and 8 of takeTimeStep (NUM_THREADS=1)
  pairs.npz         calculateContactGeometry of the reference on random
                    ellipse and circle poses (src/EllipticParticle.cpp:16-285,
                    src/CircularParticle.cpp:67-92)
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402
from tests import helpers  # noqa: E402
from xdem_amd.state import write_state  # noqa: E402

STEPS = 8
DUMP = (1, 2, 4, 8)
CASES = ("le_poly", "pb_ellipse", "pb_circle", "pb_ar1", "pb_clumps", "le_poly_hist")


def state_arrays(s):
    d = {k: getattr(s, k) for k in ("kind", "up_level", "flags", "r", "aspect", "mass",
                                     "inertia", "x", "y", "theta", "vx", "vy", "omega",
                                     "clump_bx", "clump_by", "ci", "cj", "xi_x", "xi_y",
                                     "n_x", "n_y", "damping")}
    d["cfg"] = np.array([s.cfg.world_kind] + list(s.cfg.cfg_array()), np.float64)
    return d


def run_case(name, tmp):
    s = helpers.history_state("le_poly", 3) if name == "le_poly_hist" else helpers.state(name)
    sp = os.path.join(tmp, f"{name}.xst")
    gp = os.path.join(tmp, f"{name}.xdg")
    write_state(sp, s)
    # watchdog: the reference's Newton loop is uncapped and can cycle
    subprocess.run([po.HARNESS, "run", sp, gp, str(STEPS), "1", "1"], check=True,
                   stdout=subprocess.DEVNULL, timeout=120)
    steps = {d["step"]: d for d in po.read_golden(gp)}
    out = {"in_" + k: v for k, v in state_arrays(s).items()}
    for st in DUMP:
        d = steps[st]
        for k, v in d.items():
            if k != "step":
                out[f"s{st}_{k}"] = v
    np.savez_compressed(os.path.join(HERE, f"world_{name}.npz"), **out)
    print(f"{name}: n={s.n} contacts after {STEPS} steps={steps[STEPS]['i'].size}")


def run_pairs(tmp, m=3000, seed=42):
    rng = np.random.default_rng(seed)
    rows = np.zeros((m, 12))
    for k in range(m):
        kind = 1 if k < int(0.85 * m) else 0
        a1, a2 = rng.uniform(0.4, 0.6, 2)
        ar1, ar2 = (rng.uniform(1.0, 2.5, 2) if kind == 1 else (1.0, 1.0))
        if kind == 1 and k % 7 == 0:          # identical shapes, near-parallel
            a2, ar2 = a1, ar1
        z1, z2 = rng.uniform(0, np.pi, 2)
        if kind == 1 and k % 7 == 0:
            z2 = z1 + rng.normal(0, 1e-3)
        ang = rng.uniform(0, 2 * np.pi)
        d = rng.uniform(0.2, 1.05) * (a1 + a2)
        x1, y1 = rng.uniform(-5, 5, 2)
        rows[k] = [kind, a1, ar1, x1, y1, z1, a2, ar2, x1 + d * np.cos(ang),
                   y1 + d * np.sin(ang), z2, 0.0]
    out = po.reference_pairs(rows, tmp)
    np.savez_compressed(os.path.join(HERE, "pairs.npz"), poses=rows, geometry=out)
    print(f"pairs: {m} poses, {int(out[:, 0].sum())} contacts")


def main():
    if not po.have_reference():
        sys.exit("oracle/_ref/ref_harness missing: run `make -C oracle` where the "
                 "reference sources are present")
    with tempfile.TemporaryDirectory() as tmp:
        for c in CASES:
            run_case(c, tmp)
        run_pairs(tmp)


if __name__ == "__main__":
    main()
