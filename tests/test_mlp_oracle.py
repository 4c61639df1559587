"""Pins of the CPU surrogate network (oracle/mlp_oracle.py), CPU only.

* the analytic backward pass equals central finite differences of the FP64
  loss (log-variance target held fixed, as in training) for parameters of
  every layer (weights and biases, all output heads), except that the
  log-variance head's error signal is kept out of the hidden layers: those
  gradients are compared with finite differences of the loss without it;
* the parameter layout matches the library's (bhSurrogateParamCount);
* bfloat16 rounding is round-to-nearest-even on the float32 bit pattern.
"""
import numpy as np
import pytest

import mlp_oracle as mo


def random_case(seed, n=24):
    rng = np.random.default_rng(seed)
    p = np.zeros(mo.N_PARAMS)
    for l in range(mo.N_LAYERS):
        fan = mo.layer_in(l)
        W = rng.uniform(-1, 1, (mo.layer_out(l), fan)) * np.sqrt(6.0 / fan) * (0.3 if l == mo.N_LAYERS - 1 else 1)
        if l == mo.N_LAYERS - 1:
            W[mo.N_OUT:] = 0.0
        p[mo.W_OFF[l]:mo.W_OFF[l] + W.size] = W.ravel()
        p[mo.B_OFF[l]:mo.B_OFF[l] + mo.layer_out(l)] = rng.normal(0, 0.1, mo.layer_out(l))
    x = rng.uniform(-1, 1, (n, 6)).astype(np.float32)
    t = np.zeros(n, dtype=[("d0", "<f4"), ("d1", "<f4"), ("disk", "<f4"), ("hitR", "<f4"), ("escaped", "<f4"),
                           ("band", "<f4"), ("pad", "<f4", (2,))])
    t["d0"] = rng.normal(0, 2e-3, n)
    t["d1"] = rng.normal(0, 2e-3, n)
    t["disk"] = rng.random(n) < 0.4
    t["hitR"] = rng.uniform(2, 20, n)
    t["escaped"] = rng.random(n) < 0.8
    # keep the log-variance head away from its clamp
    p[mo.B_OFF[-1] + 4] = 2.0
    return p, x, t


def test_layout_matches_library(bt):
    assert bt.load().bhSurrogateParamCount() == mo.N_PARAMS


def test_bf16_rounding():
    vals = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -3.14159265, 1e-20, 65504.0], np.float32)
    r = mo.bf16(vals)
    # nearest bfloat16 by brute force over neighbouring bit patterns
    for v, got in zip(vals, r):
        u = np.array([v], np.float32).view(np.uint32)[0]
        lo = np.uint32(u & 0xFFFF0000)
        hi = np.uint32(lo + 0x10000)
        cands = np.array([lo, hi], np.uint32).view(np.float32).astype(np.float64)
        d = np.abs(cands - float(v))
        best = cands[np.argmin(d)] if d[0] != d[1] else cands[int(((lo >> 16) & 1) == 1)]
        assert got == best, (v, got, best)


@pytest.mark.parametrize("seed", [0, 1])
def test_backward_matches_finite_differences(seed):
    p, x, t = random_case(seed)
    _, g, tgt = mo.loss_grad(p, x, t, bf=False)
    rng = np.random.default_rng(100 + seed)
    idx = []
    for l in range(mo.N_LAYERS):
        nW = (mo.N_OUT if l == mo.N_LAYERS - 1 else mo.layer_out(l)) * mo.layer_in(l)
        idx += list(mo.W_OFF[l] + rng.choice(nW, 6, replace=False))
        nb = mo.N_OUT if l == mo.N_LAYERS - 1 else mo.layer_out(l)
        idx += list(mo.B_OFF[l] + rng.choice(nb, 3, replace=False))
    for i in idx:
        h = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        head = i >= mo.W_OFF[mo.N_LAYERS - 1]
        fd = (mo.mean_loss(pp, x, t, tgt, head) - mo.mean_loss(pm, x, t, tgt, head)) / (2 * h)
        scale = max(abs(fd), abs(g[i]), 1e-3 * np.abs(g).max())
        assert abs(fd - g[i]) <= 2e-5 * scale, (i, fd, g[i])
