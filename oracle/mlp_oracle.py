"""CPU oracle of the learned quartic-root guess -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker of the engine's network kernels; nothing in the
product path imports it.

The network (8-64-64-1, ReLU) and its training loss have no counterpart in
the reference (its Newton iteration starts at sqrt(delta + 1),
src/EllipticParticle.cpp:163).  The oracle restates, in float64 numpy:
  * the pair features (xdem_amd/csrc/mlp.h, root_features);
  * the target u = log q / log s of a pair's plain root q, s = sqrt(delta+1);
  * the forward pass, the mean-squared-error loss and its gradient with the
    engine's rounding points: features, W1, W2 and the layer-1 activations
    rounded to BF16 for the forward products, the layer-2 delta rounded to
    BF16 for the backward product through W2 (straight-through for the
    rounding of the layer-1 activations).  With rounding=False it is the
    exact gradient of the unrounded network, which the tests pin against
    central finite differences.
"""
from __future__ import annotations

import numpy as np

IN, HID = 8, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + HID * IN
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID
NPARAM = OFF_B3 + 1            # 4801


def bf16(a) -> np.ndarray:
    """Round to bfloat16 (nearest, ties to even), returned as float64."""
    a32 = np.ascontiguousarray(a, np.float32)
    b = a32.view(np.uint32).astype(np.uint64)
    nan = (b & 0x7FFFFFFF) > 0x7F800000
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, ((b >> 16) | 0x40) & 0xFFFF, r)
    return (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(HID, IN), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(HID, HID),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def init_params(seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    p[OFF_W1:OFF_B1] = rng.normal(0, np.sqrt(2 / IN), HID * IN)
    p[OFF_W2:OFF_B2] = rng.normal(0, np.sqrt(2 / HID), HID * HID)
    p[OFF_B1:OFF_W2] = rng.normal(0, 0.1, HID)
    p[OFF_B2:OFF_W3] = rng.normal(0, 0.1, HID)
    p[OFF_W3:OFF_B3] = rng.normal(0, np.sqrt(1 / HID), HID)
    p[OFF_B3] = 0.5
    return p


def forward(p, X, rounding=True):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    xb = r(X)
    z1 = xb @ r(W1).T + b1
    h1 = np.maximum(z1, 0.0)
    h1b = r(h1)
    z2 = h1b @ r(W2).T + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ w3 + b3
    return dict(xb=xb, z1=z1, h1=h1, h1b=h1b, z2=z2, h2=h2, y=y)


def loss_and_grad(p, X, T, rounding=True):
    """Mean squared error over the rows and its gradient (float64)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    f = forward(p, X, rounding)
    B = len(T)
    err = f["y"] - np.asarray(T, np.float64)
    loss = float(np.mean(err * err))
    dy = 2.0 / B * err
    g = np.zeros(NPARAM)
    g[OFF_W3:OFF_B3] = f["h2"].T @ dy
    g[OFF_B3] = dy.sum()
    dz2 = (f["h2"] > 0) * dy[:, None] * w3[None, :]
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    g[OFF_W2:OFF_B2] = (dz2.T @ f["h1b"]).ravel()
    dh1 = r(dz2) @ r(W2)
    dz1 = (f["h1"] > 0) * dh1
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    g[OFF_W1:OFF_B1] = (dz1.T @ f["xb"]).ravel()
    return loss, g


def root_features(dx, dy, k1x, k1y, k2x, k2y, a1, b1, a2, b2) -> np.ndarray:
    """Features of an ellipse pair: doubled angles of the centre line d and of
    P2's axis in P1's frame, the doubled angle of d in P2's frame, and the
    ratios b1/a1, a2/a1, b2/a1 (k1, k2: unit major-axis directions)."""
    dl = np.hypot(dx, dy)
    ca, sa = (dx * k1x + dy * k1y) / dl, (dy * k1x - dx * k1y) / dl
    cb, sb = k1x * k2x + k1y * k2y, k1x * k2y - k1y * k2x
    c2a, s2a = ca * ca - sa * sa, 2 * ca * sa
    c2b, s2b = cb * cb - sb * sb, 2 * cb * sb
    return np.stack([c2a, s2a, c2b, s2b, c2a * c2b + s2a * s2b, b1 / a1, a2 / a1, b2 / a1], -1)


def target(q, s):
    """u = log q / log s: 0 at q = 1, 1 at the plain start value s."""
    return np.log(q) / np.log(s)
