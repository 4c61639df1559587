"""CPU restatement of the surrogate network (test infrastructure only).

Restates black-hole_amd/csrc/bh_mlp.hip and host/bh_surrogate.cpp in numpy
float64: Fourier features, 4 x 128 SiLU layers, 5-output head, the loss and
its analytic backward pass.  With `bf16=True` the operands the GPU stores in
bfloat16 (features, activations, weights, output gradients, hidden-layer
gradients) are rounded to bfloat16 (round to nearest even) at the same
points, so GPU results can be compared at tight tolerances; with
`bf16=False` it is a plain FP64 network whose backward pass is pinned by
finite differences (tests/test_mlp_oracle.py).

The surrogate has no counterpart in the reference.
"""
from __future__ import annotations

import numpy as np

N_IN, N_FREQ, N_FF, K_IN, HID, N_LAYERS, N_OUT, OUT_PAD = 6, 16, 192, 224, 128, 5, 5, 16
SIGMA0 = np.float32(1e-3)


def layer_in(l):
    return K_IN if l == 0 else HID


def layer_out(l):
    return OUT_PAD if l == N_LAYERS - 1 else HID


def layout():
    """(W offsets, b offsets, total) in the library's parameter vector."""
    w, b, o = [], [], 0
    for l in range(N_LAYERS):
        w.append(o)
        o += layer_out(l) * layer_in(l)
        b.append(o)
        o += layer_out(l)
    return w, b, o


W_OFF, B_OFF, N_PARAMS = layout()


def unpack(p):
    Ws = [np.asarray(p[W_OFF[l]:W_OFF[l] + layer_out(l) * layer_in(l)], np.float64).reshape(layer_out(l), layer_in(l))
          for l in range(N_LAYERS)]
    bs = [np.asarray(p[B_OFF[l]:B_OFF[l] + layer_out(l)], np.float64) for l in range(N_LAYERS)]
    return Ws, bs


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def freqs():
    """1.25^f * pi/2 evaluated in float32 as the kernels do."""
    out = []
    w = np.float32(1.5707963267948966)
    for _ in range(N_FREQ):
        out.append(w)
        w = np.float32(w * np.float32(1.25))
    return np.array(out, np.float32)


def features(x):
    x = np.asarray(x, np.float32)
    n = x.shape[0]
    F = np.zeros((n, K_IN), np.float64)
    fr = freqs()
    for d in range(N_IN):
        arg = (fr[None, :] * x[:, d:d + 1]).astype(np.float32).astype(np.float64)
        F[:, (d * N_FREQ) * 2:(d * N_FREQ + N_FREQ) * 2:2] = np.sin(arg)
        F[:, (d * N_FREQ) * 2 + 1:(d * N_FREQ + N_FREQ) * 2:2] = np.cos(arg)
    F[:, N_FF:N_FF + N_IN] = x
    return F


def silu(z):
    return z / (1.0 + np.exp(-z))


def forward(p, x, bf=False):
    """Returns head outputs y [n, 16] (bias included) and the cache."""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    Ws, bs = unpack(p)
    Wq = [r(W) for W in Ws]
    X = r(features(x))
    cache = {"X": [X], "Z": []}
    for l in range(N_LAYERS - 1):
        Z = X @ Wq[l].T + bs[l]
        X = r(silu(Z))
        cache["Z"].append(Z)
        cache["X"].append(X)
    y = X @ Wq[-1].T + bs[-1]
    cache["Wq"] = Wq
    return y, cache


CLIP = 20.0            # mrad
REG_W = 1e-3
OUT_PER_RAD = np.float32(1000.0)   # exit outputs in milliradians


def loss_grad(p, x, t, bf=False, tgt=None):
    """Loss sums [regression + log-variance, disk BCE, hit radius] over rows
    and the gradient of the mean loss w.r.t. the parameter vector (library
    layout).  The log-variance head s is fitted by the negative
    log-likelihood 0.1 (s + q e^-s) of an exponential law for the squared
    exit error, q = min(|e|, clip)^2 + 1e-6 with e in mrad (the exit
    outputs' unit, sigma0 = 1 mrad), whose minimum is
    e^s = q, so s estimates log E[q] (exp clamped at |s| <= 20); q is held
    fixed in the gradient; pass `tgt` (q) to fix it explicitly (finite
    differences).  Returns (losses, grad, tgt)."""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    y, c = forward(p, x, bf)
    n = y.shape[0]
    g = np.zeros_like(y)
    reg = (t["escaped"] > 0.5) & (t["band"] < 0.5)
    disk = t["disk"].astype(np.float64)
    y2 = y[:, 2]
    l_bce = np.maximum(y2, 0) - y2 * disk + np.log1p(np.exp(-np.abs(y2)))
    g[:, 2] = (1.0 / (1.0 + np.exp(-y2)) - disk) / n
    e0 = y[:, 0] - (t["d0"].astype(np.float32) * OUT_PER_RAD).astype(np.float64)
    e1 = y[:, 1] - (t["d1"].astype(np.float32) * OUT_PER_RAD).astype(np.float64)
    en2 = e0 ** 2 + e1 ** 2
    c2 = np.minimum(en2, CLIP ** 2)
    inside = en2 <= CLIP ** 2
    if tgt is None:
        tgt = c2 + 1e-6
    s4 = y[:, 4]
    live = np.abs(s4) < 20.0
    qe = tgt * np.exp(-np.clip(s4, -20.0, 20.0))
    en = np.sqrt(en2)
    huber = np.where(inside, en2, 2.0 * CLIP * en - CLIP ** 2)
    l_reg = np.where(reg, REG_W * huber + 0.1 * (s4 + qe), 0.0)
    scale_g = np.where(inside, 1.0, CLIP / np.maximum(en, 1e-300))
    g[:, 0] = np.where(reg, 2.0 * REG_W * e0 * scale_g / n, 0.0)
    g[:, 1] = np.where(reg, 2.0 * REG_W * e1 * scale_g / n, 0.0)
    g[:, 4] = np.where(reg, 0.1 * (1.0 - np.where(live, qe, 0.0)) / n, 0.0)
    hit = reg & (disk > 0.5)
    eh = y[:, 3] - t["hitR"] * np.float32(0.1)
    l_hit = np.where(hit, eh ** 2, 0.0)
    g[:, 3] = np.where(hit, 2.0 * eh / n, 0.0)
    g[:, N_OUT:] = 0.0
    dY = r(g)
    gT = g.copy()
    gT[:, 4] = 0.0               # the log-variance head does not shape the hidden layers
    dYT = r(gT)
    grad = np.zeros(N_PARAMS, np.float64)
    Wq = c["Wq"]
    L = N_LAYERS - 1
    grad[W_OFF[L]:W_OFF[L] + OUT_PAD * HID] = (dY.T @ c["X"][L]).ravel()
    grad[B_OFF[L]:B_OFF[L] + OUT_PAD] = dY.sum(0)
    dA = dYT @ Wq[L]
    for l in range(L - 1, -1, -1):
        Z = c["Z"][l]
        sg = 1.0 / (1.0 + np.exp(-Z))
        dZf = dA * (sg * (1.0 + Z * (1.0 - sg)))
        grad[B_OFF[l]:B_OFF[l] + HID] = dZf.sum(0)
        dZ = r(dZf)
        grad[W_OFF[l]:W_OFF[l] + HID * layer_in(l)] = (dZ.T @ c["X"][l]).ravel()
        if l > 0:
            dA = dZ @ Wq[l]
    losses = np.array([l_reg.sum(), l_bce.sum(), l_hit.sum()])
    return losses, grad, tgt


def mean_loss(p, x, t, tgt, with_var=True):
    """Mean total loss of the FP64 network with the log-variance target
    fixed (for finite differences); with_var=False drops the log-variance
    term, whose gradient does not reach the hidden layers."""
    l, _, _ = loss_grad(p, x, t, bf=False, tgt=tgt)
    total = l.sum()
    if not with_var:
        y, _ = forward(p, x, False)
        reg = (t["escaped"] > 0.5) & (t["band"] < 0.5)
        s4 = y[:, 4]
        total -= np.where(reg, 0.1 * (s4 + tgt * np.exp(-np.clip(s4, -20.0, 20.0))), 0.0).sum()
    return total / x.shape[0]
