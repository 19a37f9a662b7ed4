"""The CPU oracle of the learned root guess, pinned independently.

* BF16 rounding against torch's float32 -> bfloat16 conversion;
* the unrounded gradient against central finite differences of the loss;
* the rounded gradient (the engine's recipe) close to the unrounded one;
* pair features: invariant under a rigid rotation of the pair, under
  d -> -d and under a half turn of either ellipse; the shape ratios exact;
* targets u = log q / log s from the oracle's own quartic roots: q lies
  between 1 and s for poses of the unswapped kind.
"""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as mo
from oracle import pyoracle as po


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    a = np.concatenate([rng.normal(0, 1, 20000), rng.normal(0, 1e-30, 100),
                        rng.normal(0, 1e30, 100), [0.0, -0.0, 1.0, 1.0 + 2 ** -8,
                                                   1.0 + 3 * 2 ** -8, np.inf, -np.inf]])
    a32 = a.astype(np.float32)
    ref = torch.from_numpy(a32).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    assert np.array_equal(mo.bf16(a32), ref)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = mo.init_params(3).astype(np.float64)
    X = rng.uniform(-1, 1, (64, mo.IN))
    T = rng.uniform(0, 1, 64)
    loss, g = mo.loss_and_grad(p, X, T, rounding=False)
    idx = np.concatenate([rng.choice(mo.NPARAM, 200, replace=False), [mo.OFF_B3]])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (mo.loss_and_grad(pp, X, T, False)[0] - mo.loss_and_grad(pm, X, T, False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_rounded_gradient_close_to_exact():
    rng = np.random.default_rng(2)
    p = mo.init_params(4)
    X = rng.uniform(-1, 1, (256, mo.IN))
    T = rng.uniform(0, 1, 256)
    _, g0 = mo.loss_and_grad(p, X, T, rounding=False)
    _, g1 = mo.loss_and_grad(p, X, T, rounding=True)
    assert np.linalg.norm(g1 - g0) <= 0.05 * np.linalg.norm(g0)


def _features(x1, y1, t1, a1, ar1, x2, y2, t2, a2, ar2):
    return mo.root_features(x2 - x1, y2 - y1, np.cos(t1), np.sin(t1), np.cos(t2), np.sin(t2),
                            a1, a1 / ar1, a2, a2 / ar2)


def test_feature_invariances():
    rng = np.random.default_rng(5)
    for _ in range(200):
        a1, a2 = rng.uniform(0.4, 0.6, 2)
        ar1, ar2 = rng.uniform(1.0, 2.5, 2)
        t1, t2, ph = rng.uniform(-np.pi, np.pi, 3)
        dx, dy = rng.normal(0, 0.5, 2)
        f = _features(0, 0, t1, a1, ar1, dx, dy, t2, a2, ar2)
        c, s = np.cos(ph), np.sin(ph)
        fr = _features(0, 0, t1 + ph, a1, ar1, c * dx - s * dy, s * dx + c * dy, t2 + ph, a2, ar2)
        assert np.allclose(f, fr, atol=1e-12)
        assert np.allclose(f, _features(0, 0, t1, a1, ar1, -dx, -dy, t2, a2, ar2), atol=1e-12)
        assert np.allclose(f, _features(0, 0, t1 + np.pi, a1, ar1, dx, dy, t2 - np.pi, a2, ar2),
                           atol=1e-12)
        assert np.allclose(f[5:], [1 / ar1, a2 / a1, a2 / ar2 / a1], rtol=1e-15)
        assert np.all(np.abs(f[:5]) <= 1 + 1e-12)


def test_targets_from_oracle_roots():
    """For poses where the eigenvalues come out in the intended order
    (delta > 0), the plain root q lies in [1, s]: u in [0, 1]."""
    rng = np.random.default_rng(6)
    us, n = [], 0
    while n < 300:
        a1, a2 = rng.uniform(0.4, 0.6, 2)
        ar1, ar2 = rng.uniform(1.0, 2.5, 2)
        t1, t2 = rng.uniform(-np.pi, np.pi, 2)
        ang = rng.uniform(-np.pi, np.pi)
        dist = rng.uniform(0.5, 1.0) * (a1 + a2)
        solved, s, q, it, _ = po.pair_root(a1, ar1, 0.0, 0.0, t1, a2, ar2, dist * np.cos(ang),
                                        dist * np.sin(ang), t2)
        if not solved or not (s > 1.0) or it >= 200:
            continue
        n += 1
        us.append(mo.target(q, s))
    us = np.array(us)
    assert np.all((us >= -1e-9) & (us <= 1 + 1e-9))
    assert us.std() > 0.05           # not a constant target
