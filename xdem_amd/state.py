"""Particle/world state containers, the binary state file, and seeded
synthetic packings for the DEM workloads.

The reference builds particles one object at a time
(``EllipticParticle(id, outerRadius, aspectRatio, position, velocity,
rotation, rotationVel, contactModel)``, src/EllipticParticle.h:19-30) and
derives mass and inertia in ``calculateParticleProperties``
(src/EllipticParticle.cpp:7-14, src/CircularParticle.cpp:16-23).  Here a
packing is a set of SoA numpy arrays; :func:`particle_properties` evaluates
the same formulas in the same floating-point order.

State file layout ("XDEMST02", little endian)::

    char[8] magic; int32 n, world_kind; float64 cfg[14]
      cfg = xmin xmax ymin ymax dt kn kt restitution mu global_damping
            shear_rate total_shift total_strain cell_size
    int32 kind[n], up_level[n], flags[n]
    float64 r, aspect, mass, inertia, x, y, theta, vx, vy, omega,
            clump_bx, clump_by   (each [n])
    int32 nc; int32 ci[nc], cj[nc]; float64 xi_x, xi_y, n_x, n_y, damping [nc]
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

KIND_CIRCLE, KIND_ELLIPSE, KIND_CLUMP = 0, 1, 2
SKIP_TIME_INTEGRAL, SKIP_CONTACT_DETECTION, SKIP_MACRO_SUM = 1, 2, 4
WORLD_PERIODIC, WORLD_LEES_EDWARDS = 0, 1

_CFG_KEYS = ("xmin", "xmax", "ymin", "ymax", "dt", "kn", "kt", "restitution",
             "mu", "global_damping", "shear_rate", "total_shift",
             "total_strain", "cell_size")


@dataclasses.dataclass
class WorldConfig:
    """One world's boundary, time step and contact model."""
    world_kind: int = WORLD_PERIODIC
    xmin: float = 0.0
    xmax: float = 0.0
    ymin: float = 0.0
    ymax: float = 0.0
    dt: float = 1e-4
    kn: float = 1e4
    kt: float = 5e3
    restitution: float = 0.5
    mu: float = 0.5
    global_damping: float = 0.0
    shear_rate: float = 0.0
    total_shift: float = 0.0
    total_strain: float = 0.0
    cell_size: float = 0.0
    precision: int = 0
    newton_cap: int = 0          # 0: 200 updates
    strict_newton: int = 0       # 1: error instead of the bisection rescue
    skin: float = 0.0            # pair-list skin; 0: 0.2 * max outer radius
    pair_blocks: int = 0         # workgroups of the pair kernels; 0: 8 per CU

    def cfg_array(self) -> np.ndarray:
        return np.array([getattr(self, k) for k in _CFG_KEYS], dtype=np.float64)


@dataclasses.dataclass
class State:
    """SoA particle state of one world plus optional contact history."""
    cfg: WorldConfig
    kind: np.ndarray
    up_level: np.ndarray
    flags: np.ndarray
    r: np.ndarray
    aspect: np.ndarray
    mass: np.ndarray
    inertia: np.ndarray
    x: np.ndarray
    y: np.ndarray
    theta: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    omega: np.ndarray
    clump_bx: np.ndarray
    clump_by: np.ndarray
    ci: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int32))
    cj: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int32))
    xi_x: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    xi_y: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    n_x: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    n_y: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    damping: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    def copy(self) -> "State":
        d = {f.name: (getattr(self, f.name).copy()
                      if isinstance(getattr(self, f.name), np.ndarray)
                      else dataclasses.replace(getattr(self, f.name)))
             for f in dataclasses.fields(self)}
        return State(**d)

    def with_history(self, ci, cj, xi_x, xi_y, n_x, n_y, damping) -> "State":
        s = self.copy()
        s.ci = np.asarray(ci, np.int32)
        s.cj = np.asarray(cj, np.int32)
        s.xi_x, s.xi_y = np.asarray(xi_x, np.float64), np.asarray(xi_y, np.float64)
        s.n_x, s.n_y = np.asarray(n_x, np.float64), np.asarray(n_y, np.float64)
        s.damping = np.asarray(damping, np.float64)
        return s


def particle_properties(kind, r, aspect, density: float = 1.0):
    """Mass and inertia as the reference computes them.

    Ellipse (src/EllipticParticle.cpp:9-13): volume = pi*r*r/AR,
    I = m*(r*r + r*r/AR/AR)/4.  Circle (src/CircularParticle.cpp:18-22):
    volume = pi*r*r, I = m*r*r/2.  Same evaluation order, float64.
    """
    r = np.asarray(r, np.float64)
    aspect = np.asarray(aspect, np.float64)
    kind = np.asarray(kind)
    vol_e = math.pi * r * r / aspect
    vol_c = math.pi * r * r
    vol = np.where(kind == KIND_CIRCLE, vol_c, vol_e)
    m = vol * density
    i_e = m * (r * r + r * r / aspect / aspect) / 4.0
    i_c = m * r * r / 2.0
    inertia = np.where(kind == KIND_CIRCLE, i_c, i_e)
    return m, inertia, vol


def write_state(path: str, s: State) -> None:
    n = s.n
    with open(path, "wb") as f:
        f.write(b"XDEMST02")
        f.write(np.array([n, s.cfg.world_kind], np.int32).tobytes())
        f.write(s.cfg.cfg_array().tobytes())
        for a in (s.kind, s.up_level, s.flags):
            f.write(np.ascontiguousarray(a, np.int32).tobytes())
        for a in (s.r, s.aspect, s.mass, s.inertia, s.x, s.y, s.theta, s.vx,
                  s.vy, s.omega, s.clump_bx, s.clump_by):
            f.write(np.ascontiguousarray(a, np.float64).tobytes())
        nc = int(s.ci.shape[0])
        f.write(np.array([nc], np.int32).tobytes())
        f.write(np.ascontiguousarray(s.ci, np.int32).tobytes())
        f.write(np.ascontiguousarray(s.cj, np.int32).tobytes())
        for a in (s.xi_x, s.xi_y, s.n_x, s.n_y, s.damping):
            f.write(np.ascontiguousarray(a, np.float64).tobytes())


def read_state(path: str) -> State:
    buf = open(path, "rb").read()
    if buf[:8] != b"XDEMST02":
        raise ValueError(f"{path}: not an XDEMST02 state file")
    off = 8
    n, wk = np.frombuffer(buf, np.int32, 2, off)
    n = int(n)
    off += 8
    cfgv = np.frombuffer(buf, np.float64, 14, off)
    off += 14 * 8
    cfg = WorldConfig(world_kind=int(wk), **{k: float(v) for k, v in zip(_CFG_KEYS, cfgv)})

    def take(dt, cnt):
        nonlocal off
        a = np.frombuffer(buf, dt, cnt, off).copy()
        off += a.nbytes
        return a
    ints = [take(np.int32, n) for _ in range(3)]
    dbl = [take(np.float64, n) for _ in range(12)]
    nc = int(take(np.int32, 1)[0])
    ci, cj = take(np.int32, nc), take(np.int32, nc)
    hist = [take(np.float64, nc) for _ in range(5)]
    return State(cfg, *ints, *dbl, ci, cj, *hist)


# ---------------------------------------------------------------------------
# Synthetic packings (seeded).  The reference's drivers place particles
# uniformly at random WITH overlap and relax them by DEM
# (mainIsoElpsPack_WCCM.cpp:48-91).  Uniform placement occasionally puts two
# centres almost on top of each other, which makes the contact normal
# undefined (0/0) in the reference as well; the generator therefore places
# centres on a randomly jittered lattice (still overlapping at the target
# packing fraction), which the caller then relaxes.

def _lattice_positions(rng, n, xmin, xmax, ymin, ymax, jitter=0.35):
    lx, ly = xmax - xmin, ymax - ymin
    nx = max(1, int(math.ceil(math.sqrt(n * lx / ly))))
    ny = max(1, int(math.ceil(n / nx)))
    sites = rng.permutation(nx * ny)[:n]
    hx, hy = lx / nx, ly / ny
    ix, iy = sites % nx, sites // nx
    x = xmin + (ix + 0.5 + rng.uniform(-jitter, jitter, n)) * hx
    y = ymin + (iy + 0.5 + rng.uniform(-jitter, jitter, n)) * hy
    return x, y


def make_packing(n: int, *, phi: float, kind: str = "ellipse",
                 a_range=(0.5, 0.5), ar_range=(1.5, 1.5), bidisperse=None,
                 world_kind: int = WORLD_PERIODIC, shear_rate: float = 0.0,
                 kn: float = 1e4, kt_ratio: float = 0.5, restitution: float = 0.5,
                 mu: float = 0.5, vel_sigma: float = 0.0, seed: int = 0,
                 dt: Optional[float] = None, dt_factor: float = 0.02,
                 density: float = 1.0) -> State:
    """Seeded packing in a symmetric square box sized for packing fraction phi.

    kind: "circle" (CircularParticle), "ellipse" (EllipticParticle) or
    "mixed_ar1" (ellipses, aspect ratio 1).  bidisperse=(r1, r2) draws radii
    50:50 from the pair.  dt defaults to dt_factor*sqrt(m_min/kn).
    """
    rng = np.random.default_rng(seed)
    if bidisperse is not None:
        r = np.where(rng.permutation(n) < n // 2, bidisperse[0], bidisperse[1]).astype(np.float64)
    else:
        r = rng.uniform(a_range[0], a_range[1], n) if a_range[0] != a_range[1] \
            else np.full(n, float(a_range[0]))
    if kind == "circle":
        k = np.full(n, KIND_CIRCLE, np.int32)
        ar = np.ones(n)
    else:
        k = np.full(n, KIND_ELLIPSE, np.int32)
        ar = np.ones(n) if kind == "mixed_ar1" else (
            rng.uniform(ar_range[0], ar_range[1], n) if ar_range[0] != ar_range[1]
            else np.full(n, float(ar_range[0])))
    m, inertia, vol = particle_properties(k, r, ar, density)
    side = math.sqrt(float(vol.sum()) / phi)
    xmin, xmax, ymin, ymax = -side / 2, side / 2, -side / 2, side / 2
    x, y = _lattice_positions(rng, n, xmin, xmax, ymin, ymax)
    th = rng.uniform(0.0, math.pi, n)
    if vel_sigma > 0:
        vx, vy = rng.normal(0.0, vel_sigma, n), rng.normal(0.0, vel_sigma, n)
    else:
        vx, vy = np.zeros(n), np.zeros(n)
    if dt is None:
        dt = dt_factor * math.sqrt(float(m.min()) / kn)
    cfg = WorldConfig(world_kind=world_kind, xmin=xmin, xmax=xmax, ymin=ymin,
                      ymax=ymax, dt=dt, kn=kn, kt=kt_ratio * kn,
                      restitution=restitution, mu=mu, shear_rate=shear_rate)
    z = np.zeros(n)
    return State(cfg, k, np.arange(n, dtype=np.int32), np.zeros(n, np.int32),
                 r.astype(np.float64), ar.astype(np.float64), m, inertia,
                 x, y, th, vx, vy, z.copy(), z.copy(), z.copy())


def add_clumps(s: State, num_clumps: int, seed: int = 0, sub_r: float = 0.35,
               spacing: float = 0.5) -> State:
    """Replace the first num_clumps*4 particles by 3-ellipse (AR 1) clumps.

    Layout per clump, as in mainClumpParticleTest.cpp:97-140: the clump
    (GluedParticle, SKIP_CONTACT_DETECTION) is followed by its sub-particles,
    which carry up_level = clump ID and SKIP_TIME_INTEGRAL|SKIP_MACRO_SUM.
    The clump's mass/inertia here are analytic (sum of the subs' and
    parallel-axis terms) rather than the reference's grid sampling
    (src/GluedParticle.cpp:7-80); the state carries them explicitly.
    """
    rng = np.random.default_rng(seed)
    s = s.copy()
    for c in range(num_clumps):
        base = 4 * c
        cx, cy = s.x[base], s.y[base]
        ang = rng.uniform(0, 2 * math.pi)
        s.kind[base] = KIND_CLUMP
        s.flags[base] = SKIP_CONTACT_DETECTION
        s.up_level[base] = base
        subs = []
        for k in range(3):
            a = ang + k * 2 * math.pi / 3
            subs.append((cx + spacing * math.cos(a), cy + spacing * math.sin(a)))
        ms = math.pi * sub_r * sub_r
        mc = 3 * ms
        ic = 3 * (ms * sub_r * sub_r / 2.0 + ms * spacing * spacing)
        s.mass[base], s.inertia[base] = mc, ic
        s.r[base], s.aspect[base] = -1.0, 1.0
        s.vx[base] = s.vy[base] = s.omega[base] = 0.0
        for k in range(3):
            p = base + 1 + k
            s.kind[p] = KIND_ELLIPSE
            s.up_level[p] = base
            s.flags[p] = SKIP_TIME_INTEGRAL | SKIP_MACRO_SUM
            s.r[p], s.aspect[p] = sub_r, 1.0
            s.x[p], s.y[p] = subs[k]
            s.clump_bx[p], s.clump_by[p] = s.x[p] - cx, s.y[p] - cy
            s.mass[p], s.inertia[p] = mc, ic
            s.vx[p] = s.vy[p] = s.omega[p] = 0.0
    return s


def make_ensemble(num: int = 256, seed: int = 0, n_range=(2000, 50000), phi: float = 0.8,
                  rate_range=(1e-5, 1e-3), clump_share=(0.05, 0.25)) -> list:
    """Seeded batch of independent packings of uneven size.

    Each packing draws N uniformly from n_range and mixes, particle by
    particle, circles (stored as aspect-ratio-1 ellipses: the reference
    cannot pair a CircularParticle with an EllipticParticle,
    src/EllipticParticle.cpp:30) and ellipses of aspect ratio U[1, 2.5]; half
    of the packings are Lees-Edwards shear cells (rate log-uniform in
    rate_range), the other half periodic cells that also hold 3-disc clumps
    (GluedParticle; the Lees-Edwards world does not honour the clump's
    SKIP_CONTACT_DETECTION, src/LeesEdwardsBoundaryWorld.cpp:141-199).
    """
    rng = np.random.default_rng(seed)
    out = []
    for k in range(num):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        le = bool(rng.random() < 0.5)
        s = make_packing(n, phi=phi, a_range=(0.4, 0.6), ar_range=(1.0, 2.5),
                         world_kind=WORLD_LEES_EDWARDS if le else WORLD_PERIODIC,
                         shear_rate=float(10 ** rng.uniform(np.log10(rate_range[0]),
                                                           np.log10(rate_range[1]))) if le else 0.0,
                         seed=int(rng.integers(1 << 31)))
        circles = rng.random(n) < rng.uniform(0.0, 0.5)
        s.aspect[circles] = 1.0
        s.mass, s.inertia, _ = particle_properties(s.kind, s.r, s.aspect)
        if not le:
            nc = int(n * rng.uniform(*clump_share) / 4)
            s = add_clumps(s, nc, seed=int(rng.integers(1 << 31)), sub_r=0.3, spacing=0.35)
        out.append(s)
    return out


def clump_branches(s: State) -> State:
    """Recompute the clumps' branch vectors (sub-particle centre minus clump
    centre, minimum image) from the current positions, e.g. after a
    relaxation that moved the sub-particles with their clumps."""
    s = s.copy()
    lx, ly = s.cfg.xmax - s.cfg.xmin, s.cfg.ymax - s.cfg.ymin
    idx = np.arange(s.n)
    for c in np.nonzero(s.kind == KIND_CLUMP)[0]:
        subs = np.nonzero((s.up_level == c) & (idx != c))[0]
        bx, by = s.x[subs] - s.x[c], s.y[subs] - s.y[c]
        s.clump_bx[subs] = bx - lx * np.round(bx / lx)
        s.clump_by[subs] = by - ly * np.round(by / ly)
    return s
