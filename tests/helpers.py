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


def pw_f(dx, dy, a1, b1, th1, a2, b2, th2, lam):
    """Perram-Wertheim contact function F(lam) of two ellipses (semi-axes a,
    b, orientation th) whose centres differ by (dx, dy): the ellipses overlap
    iff max over lam in [0, 1] of F is <= 1 (J. Comput. Phys. 58, 409)."""
    c1, s1, c2, s2 = np.cos(th1), np.sin(th1), np.cos(th2), np.sin(th2)
    g1, g2 = a1 * a1 - b1 * b1, a2 * a2 - b2 * b2
    sxx = (1 - lam) * (b1 * b1 + g1 * c1 * c1) + lam * (b2 * b2 + g2 * c2 * c2)
    sxy = (1 - lam) * (g1 * c1 * s1) + lam * (g2 * c2 * s2)
    syy = (1 - lam) * (b1 * b1 + g1 * s1 * s1) + lam * (b2 * b2 + g2 * s2 * s2)
    det = sxx * syy - sxy * sxy
    q = syy * dx * dx - 2 * sxy * dx * dy + sxx * dy * dy
    return lam * (1 - lam) * q / det


def pw_touching_distance(ux, uy, a1, b1, th1, a2, b2, th2):
    """Centre distance along the unit vector (ux, uy) at which the two
    ellipses touch: F scales with the squared distance, so it is
    1 / sqrt(max_lam F(lam; u)) (golden-section search, F is concave in lam)."""
    lo, hi = np.zeros_like(ux), np.ones_like(ux)
    gr = 0.5 * (np.sqrt(5.0) - 1.0)
    for _ in range(90):
        l1, l2 = hi - gr * (hi - lo), lo + gr * (hi - lo)
        up = pw_f(ux, uy, a1, b1, th1, a2, b2, th2, l1) < pw_f(ux, uy, a1, b1, th1, a2, b2, th2, l2)
        lo, hi = np.where(up, l1, lo), np.where(up, hi, l2)
    return 1.0 / np.sqrt(pw_f(ux, uy, a1, b1, th1, a2, b2, th2, 0.5 * (lo + hi)))


def near_touching_poses(n, seed, a_range=(0.4, 0.6), ar_range=(1.0, 2.5)):
    """Random ellipse pairs at centre distance d0 (1 + eps), d0 the touching
    distance, eps = +-10^-U(1, 9): half just apart, half just overlapping.
    Returns dict of arrays (a1, ar1, th1, a2, ar2, th2, dx, dy, apart)."""
    rng = np.random.default_rng(seed)
    a1, a2 = rng.uniform(*a_range, n), rng.uniform(*a_range, n)
    ar1, ar2 = rng.uniform(*ar_range, n), rng.uniform(*ar_range, n)
    th1, th2 = rng.uniform(0, np.pi, n), rng.uniform(0, np.pi, n)
    ang = rng.uniform(0, 2 * np.pi, n)
    ux, uy = np.cos(ang), np.sin(ang)
    d0 = pw_touching_distance(ux, uy, a1, a1 / ar1, th1, a2, a2 / ar2, th2)
    apart = rng.uniform(size=n) < 0.5
    eps = 10.0 ** -rng.uniform(1, 9, n)
    d = d0 * np.where(apart, 1 + eps, 1 - eps)
    return dict(a1=a1, ar1=ar1, th1=th1, a2=a2, ar2=ar2, th2=th2, dx=d * ux, dy=d * uy,
                apart=apart)


def pair_lattice_state(poses, spacing=4.0, world_kind=WORLD_PERIODIC, shear_rate=0.0):
    """A packing of isolated ellipse pairs (one pose per lattice site, far
    enough apart that only the two ellipses of a site can touch)."""
    from xdem_amd.state import KIND_ELLIPSE, State, WorldConfig, particle_properties
    n = len(poses["a1"])
    side = int(np.ceil(np.sqrt(n)))
    k = np.arange(n)
    cx, cy = (k % side + 0.5) * spacing, (k // side + 0.5) * spacing
    x = np.empty(2 * n); y = np.empty(2 * n)
    x[0::2], y[0::2] = cx - 0.5 * poses["dx"], cy - 0.5 * poses["dy"]
    x[1::2], y[1::2] = cx + 0.5 * poses["dx"], cy + 0.5 * poses["dy"]
    r = np.empty(2 * n); ar = np.empty(2 * n); th = np.empty(2 * n)
    r[0::2], r[1::2] = poses["a1"], poses["a2"]
    ar[0::2], ar[1::2] = poses["ar1"], poses["ar2"]
    th[0::2], th[1::2] = poses["th1"], poses["th2"]
    kind = np.full(2 * n, KIND_ELLIPSE, np.int32)
    m, inertia, _ = particle_properties(kind, r, ar)
    L = side * spacing
    cfg = WorldConfig(world_kind=world_kind, xmin=0.0, xmax=L, ymin=0.0, ymax=L,
                      dt=1e-4, shear_rate=shear_rate)
    z = np.zeros(2 * n)
    return State(cfg, kind, np.arange(2 * n, dtype=np.int32), np.zeros(2 * n, np.int32),
                 r, ar, m, inertia, x, y, th, z.copy(), z.copy(), z.copy(), z.copy(), z.copy())


def nan_pose_lattice(n=600, seed=31):
    """pair_lattice_state of n near-touching pairs in the box [0, L) x [-L, 0)
    with five sites holding the pose for which the reference's arithmetic
    gives a NaN contact (test_oracle_math.NAN_POSE).  Those sites sit near
    the pose's own position and are moved by whole units, so both centres
    keep their binades and the centre difference is the pose's to the bit.
    Returns (state, site indices)."""
    from tests.test_oracle_math import NAN_POSE as p
    poses = near_touching_poses(n, seed=seed)
    s = pair_lattice_state(poses)
    side = int(np.ceil(np.sqrt(n)))
    L = s.cfg.ymax
    s.y -= L
    s.cfg.ymin, s.cfg.ymax = -L, 0.0
    sites = []
    for m, row in ((10, 11), (11, 11), (12, 12), (13, 13), (11, 13)):
        k = m + side * row
        cx, cy = (m + 0.5) * 4.0, (row + 0.5) * 4.0 - L
        tx, ty = cx - 48.0, cy + 46.0
        a, b = 2 * k, 2 * k + 1
        s.x[a], s.y[a], s.x[b], s.y[b] = p[2] + tx, p[3] + ty, p[7] + tx, p[8] + ty
        s.r[a], s.aspect[a], s.theta[a] = p[0], p[1], p[4]
        s.r[b], s.aspect[b], s.theta[b] = p[5], p[6], p[9]
        sites.append(k)
    from xdem_amd.state import particle_properties
    s.mass, s.inertia, _ = particle_properties(s.kind, s.r, s.aspect)
    return s, sites
