"""Writes tests/golden/unet_loss_curve.json: the loss curve of a fixed-seed
tiny warm-start network (base width 8, 8x8x8 volumes) trained for 20 Adam
steps by the CPU oracle (BF16 emulation, FP32 Adam).  The inputs, targets and
initial parameters come from numpy's default_rng with the seeds stored in the
file.  Run: python tests/golden/make_unet_curve.py"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as orc  # noqa: E402

CFG = dict(C=8, dims=[8, 8, 8], steps=20, lr=3e-4, beta1=0.9, beta2=0.999, eps=1e-8, param_seed=21, data_seed=22)


def tiny_problem(cfg):
    C = cfg["C"]
    rng = np.random.default_rng(cfg["param_seed"])
    io = [(8, C), (C, C), (C, 2 * C), (2 * C, 2 * C), (2 * C, 4 * C), (4 * C, 4 * C), (6 * C, 2 * C),
          (2 * C, 2 * C), (3 * C, C), (C, C)]
    parts = []
    for i, (cin, cout) in enumerate(io):
        w = rng.normal(0, np.sqrt(2.0 / (27 * (1 if i == 0 else cin))), (cout, 27, cin))
        if i == 0:
            w[:, :, 1:] = 0.0  # only the real input channel
        parts += [w.reshape(-1), np.zeros(cout)]
    parts.append(np.zeros(C + 1))  # zero head: the identity at step 0
    p = np.concatenate(parts).astype(np.float32)
    drng = np.random.default_rng(cfg["data_seed"])
    d = cfg["dims"]
    t = drng.uniform(0.0, 0.05, d).astype(np.float32)
    x = (t * 0.8 + drng.normal(0, 0.004, d)).astype(np.float32)  # a blurred-ish, noisy start
    x = np.maximum(x, 0).astype(np.float32)
    return p, x, t


def adam_curve(p, x, t, cfg, grad_fn):
    """FP32 Adam with the product's bias-corrected step (xr_unet_adam_step)."""
    p = p.copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    s = np.float32(x.mean())
    b1, b2 = np.float32(cfg["beta1"]), np.float32(cfg["beta2"])
    curve = []
    for k in range(1, cfg["steps"] + 1):
        loss, g = grad_fn(p, x, t, float(s))
        g = g.astype(np.float32)
        curve.append(loss)
        c1 = 1.0 - cfg["beta1"] ** k
        c2 = 1.0 - cfg["beta2"] ** k
        lr_t = np.float32(cfg["lr"] * np.sqrt(c2) / c1)
        eps_t = np.float32(cfg["eps"] * np.sqrt(c2))
        m = b1 * m + (np.float32(1) - b1) * g
        v = b2 * v + (np.float32(1) - b2) * g * g
        p = (p - lr_t * m / (np.sqrt(v) + eps_t)).astype(np.float32)
    return curve


def oracle_grad(cfg):
    return lambda p, x, t, s: orc.unet_loss_grad(p, cfg["C"], x, t, s, emulate_bf16=True)


if __name__ == "__main__":
    p, x, t = tiny_problem(CFG)
    curve = adam_curve(p, x, t, CFG, oracle_grad(CFG))
    with open(os.path.join(HERE, "unet_loss_curve.json"), "w") as f:
        json.dump({"config": CFG, "loss": curve}, f, indent=1)
    print(curve[0], curve[-1])
