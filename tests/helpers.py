"""Shared builders of small seeded packings for the tests.

Packings are relaxed with the CPU oracle (test infrastructure), the same
protocol the reference's drivers use: overlapping placement, then DEM with
dissipation and no friction (mainIsoElpsPack_WCCM.cpp:55-91).
"""
import functools

import numpy as np

from xdem_amd.state import (WORLD_LEES_EDWARDS, WORLD_PERIODIC, add_clumps,
                            make_packing)
from oracle import pyoracle as po


def relax(s, steps, gdamp=20.0):
    r = s.copy()
    r.cfg.global_damping = gdamp
    r.cfg.mu = 0.0
    r.cfg.restitution = 0.95
    r.cfg.shear_rate = 0.0
    r.cfg.world_kind = WORLD_PERIODIC
    w = po.OracleWorld(r)
    w.step(steps)
    p = w.particles()
    out = s.copy()
    for k, v in p.items():
        setattr(out, k, v)
    return out


@functools.lru_cache(maxsize=None)
def _cached(name):
    if name == "le_poly":
        s = make_packing(400, phi=0.85, a_range=(0.4, 0.6), ar_range=(1.0, 2.5),
                         world_kind=WORLD_LEES_EDWARDS, shear_rate=0.5, seed=1)
        s = relax(s, 300)
        rng = np.random.default_rng(7)
        s.vx = rng.normal(0, 0.05, s.n); s.vy = rng.normal(0, 0.05, s.n)
        s.omega = rng.normal(0, 0.1, s.n)
        s.cfg.total_shift = 3.3
        return s
    if name == "pb_ellipse":
        s = make_packing(500, phi=0.80, a_range=(0.5, 0.5), ar_range=(1.5, 1.5), seed=2)
        s = relax(s, 300)
        rng = np.random.default_rng(8)
        s.vx = rng.normal(0, 0.05, s.n); s.vy = rng.normal(0, 0.05, s.n)
        return s
    if name == "pb_circle":
        s = make_packing(600, phi=0.60, kind="circle", bidisperse=(0.5, 0.7),
                         vel_sigma=0.01, seed=0)
        s = relax(s, 200)
        rng = np.random.default_rng(9)
        s.vx = rng.normal(0, 0.01, s.n); s.vy = rng.normal(0, 0.01, s.n)
        return s
    if name == "pb_ar1":
        s = make_packing(400, phi=0.75, kind="mixed_ar1", a_range=(0.4, 0.6), seed=4)
        return relax(s, 200)
    if name == "pb_clumps":
        s = make_packing(400, phi=0.55, a_range=(0.45, 0.45), ar_range=(1.0, 1.0),
                         kind="mixed_ar1", seed=5)
        s = add_clumps(s, 20, seed=5, sub_r=0.3, spacing=0.35)
        return relax(s, 50, gdamp=5.0)
    raise KeyError(name)


def state(name):
    return _cached(name).copy()


def history_state(name, steps=5):
    """State after `steps` oracle steps, with its live contact history."""
    s = state(name)
    w = po.OracleWorld(s)
    w.step(steps)
    return w.export_state(s)


def compare_step(a, b, scale_f):
    """max abs differences between two read-backs (dicts of arrays)."""
    return {k: float(np.max(np.abs(a[k] - b[k]))) if len(a[k]) else 0.0
            for k in a if k in b}
