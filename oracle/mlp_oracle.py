"""CPU oracle of the T(rho, e) surrogate: a 4 -> 64 -> 64 -> 1 tanh MLP.

TEST INFRASTRUCTURE ONLY: the checker of the GPU training and inference
kernels (phlegethon_amd/csrc/phlx_mlp.hip).  The surrogate has no
counterpart in the reference; it only supplies the initial guess of the
reference's Newton inversion helm_rhoe_given (source/source.F90:24982-25019).

Network (normalised inputs x, normalised output y = (log10 T - c) / s):
    h1 = tanh(x W1 + b1)          W1 [4, 64]   (FP32 VALU, BF16 weights)
    h2 = tanh(bf16(h1) W2 + b2)   W2 [64, 64]  (BF16 MFMA, FP32 accumulation)
    y  = h2 W3 + b3               W3 [64]      (FP32 VALU, BF16 weights)
The forward rounds the weights and the MFMA input h1 to BF16 (round to
nearest even) exactly where the GPU kernels do; the backward treats the
rounding as identity (straight-through) and uses the rounded values.  With
``quant=False`` the same code is the plain FP64 network, whose analytic
gradients the tests pin against central finite differences.
"""
from __future__ import annotations

import numpy as np

NIN, NH = 4, 64
# flat parameter vector order (shared with the GPU kernels)
SHAPES = [("w1", (NIN, NH)), ("b1", (NH,)), ("w2", (NH, NH)), ("b2", (NH,)),
          ("w3", (NH,)), ("b3", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16(a) -> np.ndarray:
    """Round float32 values to BF16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)
    r = (u + np.uint32(0x7FFF) + ((u >> np.uint32(16)) & np.uint32(1))) & np.uint32(0xFFFF0000)
    return r.view(np.float32)


def unpack(theta: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = theta[o:o + n].reshape(shp)
        o += n
    return out


def pack(p: dict) -> np.ndarray:
    return np.concatenate([np.asarray(p[name]).reshape(-1) for name, _ in SHAPES])


def forward(theta: np.ndarray, x: np.ndarray, quant: bool = True):
    """x [B, 4] -> y [B], cache for backward."""
    dt = np.float32 if quant else np.float64
    p = {k: v.astype(dt) for k, v in unpack(np.asarray(theta)).items()}
    q = bf16 if quant else (lambda a: a)
    x = np.asarray(x, dtype=dt)
    w1, w2, w3 = q(p["w1"]).astype(dt), q(p["w2"]).astype(dt), q(p["w3"]).astype(dt)
    h1 = np.tanh(x @ w1 + p["b1"])
    h1q = q(h1).astype(dt)
    h2 = np.tanh(h1q @ w2 + p["b2"])
    y = h2 @ w3 + p["b3"][0]
    return y, dict(x=x, w1=w1, w2=w2, w3=w3, h1=h1, h1q=h1q, h2=h2)


def loss_and_grad(theta: np.ndarray, x: np.ndarray, t: np.ndarray, quant: bool = True,
                  w: np.ndarray = None):
    """Weighted mean squared error (1/B) sum w (y - t)^2 over the batch (w = 1
    by default) and its gradient w.r.t. theta."""
    y, c = forward(theta, x, quant)
    B = y.shape[0]
    r = y - t
    w = np.ones_like(r) if w is None else np.asarray(w, dtype=r.dtype)
    L = float(np.mean(w * r * r))
    d3 = 2.0 * w * r / B                      # dL/dy
    g = {}
    g["w3"] = c["h2"].T @ d3
    g["b3"] = np.array([d3.sum()])
    d2 = np.outer(d3, c["w3"]) * (1.0 - c["h2"] ** 2)
    g["w2"] = c["h1q"].T @ d2
    g["b2"] = d2.sum(axis=0)
    d1 = (d2 @ c["w2"].T) * (1.0 - c["h1"] ** 2)
    g["w1"] = c["x"].T @ d1
    g["b1"] = d1.sum(axis=0)
    return L, pack(g)


def init_params(seed: int = 1) -> np.ndarray:
    """Xavier-uniform weights, zero biases (the GPU trainer uses the same
    counter-based initialisation, generated on the host)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shp in SHAPES:
        if name.startswith("w"):
            fan_in = shp[0]
            fan_out = shp[1] if len(shp) > 1 else 1
            a = np.sqrt(6.0 / (fan_in + fan_out))
            p[name] = rng.uniform(-a, a, shp)
        else:
            p[name] = np.zeros(shp)
    return pack(p).astype(np.float32)
