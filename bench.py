"""Benchmark: particle-steps/s of the DEM step loop on one MI355X.

Workload (default): 1,000,000 polydisperse ellipses (semi-major axis
U[0.4, 0.6], aspect ratio U[1.0, 2.5], random orientation) at packing
fraction 0.85 under Lees-Edwards simple shear, rate 1e-4; linear
spring-dashpot with kn = 1e4, kt = 0.5 kn, restitution 0.5, friction 0.5;
dt = 0.02 sqrt(m_min / kn); plain FP64 path.  The packing is synthetic and
seeded: jittered-lattice placement with overlap, relaxed on the GPU by the
engine itself (untimed), as the reference's drivers relax random
placements (mainIsoElpsPack_WCCM.cpp:55-91).

One "step" is one takeTimeStep of the whole packing with every input
resident in HBM.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def make_workload(n: int, seed: int, relax_steps: int, precision: int = 0):
    """Seeded 1M-ellipse packing at phi 0.85: jittered-lattice placement with
    overlap, then friction-free relaxation on the GPU engine (untimed), first
    overdamped, then lightly damped, as the reference's drivers relax random
    placements (mainIsoElpsPack_WCCM.cpp:55-91)."""
    from xdem_amd.engine import Engine
    from xdem_amd.state import WORLD_LEES_EDWARDS, WORLD_PERIODIC, make_packing
    s = make_packing(n, phi=0.85, a_range=(0.4, 0.6), ar_range=(1.0, 2.5),
                     world_kind=WORLD_LEES_EDWARDS, shear_rate=1e-4, kn=1e4,
                     kt_ratio=0.5, restitution=0.5, mu=0.5, seed=seed)
    if relax_steps > 0:
        for frac, gamma in ((0.4, 400.0), (0.6, 40.0)):
            r = s.copy()
            r.cfg.world_kind = WORLD_PERIODIC
            r.cfg.shear_rate = 0.0
            r.cfg.mu = 0.0
            r.cfg.global_damping = gamma
            eng = Engine(r)
            eng.step(max(1, int(relax_steps * frac)))
            p = eng.particles()
            eng.close()
            for k in ("x", "y", "theta"):
                setattr(s, k, p[k])
    return s


def cpu_baseline(s, steps: int, threads: int):
    """The CPU oracle (FP64 restatement of the reference, OpenMP) timed on
    the host cores on the same packing."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    from oracle import pyoracle as po
    w = po.OracleWorld(s)
    w.step(1)                      # warm-up (allocations)
    t0 = time.perf_counter()
    w.step(steps)
    dt = time.perf_counter() - t0
    return {"value": s.n * steps / dt, "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{steps} steps of the same {s.n}-particle packing (after 1 warm-up step)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--relax", type=int, default=6000)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine is measured on one GPU; multi-GPU runs "
                          "(y-slab halo exchange) are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from xdem_amd.engine import Engine

    t0 = time.time()
    s = make_workload(args.n, args.seed, args.relax)
    print(f"[bench] workload ready: n={s.n} relax={args.relax} ({time.time() - t0:.1f}s)",
          file=sys.stderr, flush=True)
    eng = Engine(s)
    eng.step(args.warmup)
    torch.cuda.synchronize()
    eng.set_timing(True)
    builds0 = eng.counters()["list_builds"]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    eng.step(args.steps, sync=False)
    eng.sync()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    elapsed = t2 - t1
    kern_ms, launches = eng.timing()
    eng.set_timing(False)
    bytes_step, contacts = eng.traffic()
    cnt = eng.counters()
    ms_per_step = elapsed * 1e3 / args.steps
    value = s.n * args.steps / elapsed
    kern_avg_ms = kern_ms / max(launches, 1)
    # k_contact's algorithmic bytes: each particle's pose, velocity and shape
    # read once (72 B) and its force and torque written once (24 B); each
    # contact's history (xi, n, damping, partner) read and written (88 B)
    kern_bytes = 96.0 * s.n + 88.0 * contacts
    achieved = kern_bytes / (kern_avg_ms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "k_contact_traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("bytes_per_launch")
    out = {
        "metric": "particle-steps/s, 1M ellipses LE shear: plain vs NN-guess GPU; HBM % peak",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "LE shear, 1M polydisperse ellipses (a U[0.4,0.6], AR U[1,2.5]), "
                               "phi 0.85, rate 1e-4, plain FP64 path",
                   "n": s.n, "dt": s.cfg.dt, "contacts": contacts,
                   "coarse_pairs_per_step": cnt["coarse_pairs"] / max(cnt["steps"], 1),
                   "newton_iters_per_pair": cnt["newton_iters"] / max(cnt["coarse_pairs"], 1),
                   "newton_rescues": cnt["rescues"],
                   "pair_list_builds_timed": cnt["list_builds"] - builds0},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_contact", "kernel_avg_ms": kern_avg_ms,
                     "algorithmic_bytes_per_launch": kern_bytes,
                     "step_algorithmic_bytes": bytes_step,
                     "step_hbm_frac": bytes_step * args.steps / elapsed / (HBM_PEAK_GBS * 1e9)},
    }
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
        try:
            out["cpu_baseline"] = cpu_baseline(s, args.cpu_steps, threads)
        except Exception as e:  # reported, never silently replaced
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
