"""Pins the CPU restatement of the warm-start U-Net (oracle/oracle_unet.cpp).

The network has no reference counterpart; the oracle is pinned by a library
routine (scipy.ndimage.correlate for one 3x3x3 convolution: tap order, zero
padding, channel sums), by construction (a zero head is the identity), and by
finite differences of its FP64 loss (the whole backward pass)."""
import numpy as np
import pytest
from scipy import ndimage


def random_params(orc, C, seed, head_scale=0.5):
    rng = np.random.default_rng(seed)
    n = orc.unet_num_params(C)
    p = np.zeros(n, np.float32)
    # layer by layer, He scale
    io = [(8, C), (C, C), (C, 2 * C), (2 * C, 2 * C), (2 * C, 4 * C), (4 * C, 4 * C), (6 * C, 2 * C),
          (2 * C, 2 * C), (3 * C, C), (C, C)]
    off = 0
    for cin, cout in io:
        nw = cout * 27 * cin
        p[off:off + nw] = rng.normal(0, np.sqrt(2.0 / (27 * cin)), nw)
        off += nw
        p[off:off + cout] = rng.normal(0, 0.05, cout)
        off += cout
    p[off:off + C + 1] = rng.normal(0, head_scale / np.sqrt(C), C + 1)
    return p


def test_conv3_matches_scipy_correlate(orc):
    rng = np.random.default_rng(3)
    x = rng.normal(size=(5, 6, 7, 3)).astype(np.float32)
    W = rng.normal(size=(4, 27, 3)).astype(np.float32)
    b = rng.normal(size=4).astype(np.float32)
    got = orc.conv3(x, W, b)
    want = np.zeros_like(got)
    for co in range(4):
        acc = np.full(x.shape[:3], float(b[co]))
        for ci in range(3):
            k = W[co, :, ci].astype(np.float64).reshape(3, 3, 3)  # tap = (d0+1)*9 + (d1+1)*3 + (d2+1)
            acc += ndimage.correlate(x[..., ci].astype(np.float64), k, mode="constant", cval=0.0)
        want[..., co] = acc
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12)
    relu = orc.conv3(x, W, b, relu=True)
    np.testing.assert_allclose(relu, np.maximum(want, 0), rtol=1e-10, atol=1e-12)


def test_zero_head_is_identity(orc):
    C = 8
    p = random_params(orc, C, 1, head_scale=0.0)
    x = np.random.default_rng(2).uniform(0.0, 0.05, (8, 8, 8)).astype(np.float32)
    for em in (False, True):
        y = orc.unet_forward(p, C, x, float(x.mean()), emulate_bf16=em)
        np.testing.assert_array_equal(y, x)


def test_backward_matches_finite_differences(orc):
    C = 8
    p = random_params(orc, C, 7)
    rng = np.random.default_rng(11)
    x = rng.uniform(0.0, 0.05, (8, 8, 8)).astype(np.float32)
    t = (x + rng.normal(0, 0.005, x.shape)).astype(np.float32)
    s = float(x.mean())
    loss, g = orc.unet_loss_grad(p, C, x, t, s, emulate_bf16=False)
    assert loss > 0
    n = p.size
    # parameters in every layer and the head
    idx = np.unique(np.concatenate([rng.choice(n, 60, replace=False), np.arange(n - C - 1, n)]))
    errs = []
    for i in idx:
        h = np.float32(2e-3) * max(1.0, abs(float(p[i])))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = orc.unet_loss_grad(pp, C, x, t, s, emulate_bf16=False)
        lm, _ = orc.unet_loss_grad(pm, C, x, t, s, emulate_bf16=False)
        fd = (lp - lm) / (float(pp[i]) - float(pm[i]))
        errs.append(abs(fd - g[i]) / (abs(g[i]) + 1e-3 * np.abs(g).max() + 1e-30))
    errs = np.array(errs)
    # ReLU kinks inside +-h can spoil a few differences; the rest must agree
    assert np.median(errs) < 1e-3, errs
    assert np.mean(errs < 1e-2) >= 0.9, errs


def test_bf16_emulation_close_to_fp64(orc):
    C = 8
    p = random_params(orc, C, 5)
    rng = np.random.default_rng(4)
    x = rng.uniform(0.0, 0.05, (8, 8, 8)).astype(np.float32)
    s = float(x.mean())
    y64 = orc.unet_forward(p, C, x, s, emulate_bf16=False)
    y16 = orc.unet_forward(p, C, x, s, emulate_bf16=True)
    d64, d16 = y64 - x, y16 - x
    rel = np.linalg.norm(d16 - d64) / np.linalg.norm(d64)
    assert 0 < rel < 3e-2


def test_oracle_loss_curve_fixture(orc):
    """The committed fixed-seed training curve (tests/golden/make_unet_curve.py)
    is what the oracle reproduces: decreasing, and stable run to run."""
    import json
    import os
    import sys

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    sys.path.insert(0, here)
    import make_unet_curve as mk

    fx = json.load(open(os.path.join(here, "unet_loss_curve.json")))
    p, x, t = mk.tiny_problem(fx["config"])
    curve = mk.adam_curve(p, x, t, fx["config"], mk.oracle_grad(fx["config"]))
    np.testing.assert_allclose(curve, fx["loss"], rtol=1e-9)
    assert all(b < a for a, b in zip(curve, curve[1:]))
