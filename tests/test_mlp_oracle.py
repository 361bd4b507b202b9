"""Pins of the surrogate's CPU oracle (oracle/mlp_oracle.py).

- the analytic backward equals central finite differences of the FP64
  network (every parameter block, random batch);
- BF16 rounding is round-to-nearest-even (ties to even, exact for
  representable values, 8 significant bits);
- the quantised forward differs from the FP64 one by BF16-sized amounts.
"""
import numpy as np

import mlp_oracle as M


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e-20],
                 dtype=np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie -> even (mantissa 0)
    assert r[2] == 1.0 + 4 * 2 ** -8        # tie -> even (mantissa 2)
    assert r[3] == 1.0                      # below half
    assert r[4] == -2.5                     # representable
    assert abs(r[5] / 3.0e-20 - 1) < 2 ** -8
    u = np.random.default_rng(0).standard_normal(1000).astype(np.float32)
    assert np.all(np.abs(M.bf16(u) - u) <= np.abs(u) * 2 ** -8)
    assert np.all((M.bf16(u).view(np.uint32) & 0xFFFF) == 0)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    theta = M.init_params(7).astype(np.float64)
    theta += rng.normal(0, 0.05, theta.shape)     # non-zero biases too
    x = rng.uniform(-1, 1, (17, 4))
    t = rng.uniform(-1, 1, 17)
    L, g = M.loss_and_grad(theta, x, t, quant=False)
    idx = np.concatenate([rng.choice(np.arange(o, o + int(np.prod(s))), 6, replace=False)
                          for o, (_, s) in zip(np.cumsum([0] + [int(np.prod(s)) for _, s in
                                                                 M.SHAPES])[:-1], M.SHAPES)
                          if int(np.prod(s)) >= 6] + [np.array([M.NPARAM - 1])])
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (M.loss_and_grad(tp, x, t, quant=False)[0] -
              M.loss_and_grad(tm, x, t, quant=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(fd), (i, fd, g[i])


def test_quantised_forward_close_to_fp64():
    rng = np.random.default_rng(5)
    theta = M.init_params(2)
    x = rng.uniform(-1, 1, (64, 4))
    yq, _ = M.forward(theta, x, quant=True)
    y, _ = M.forward(theta.astype(np.float64), x, quant=False)
    assert np.max(np.abs(yq - y)) < 0.05
    assert np.max(np.abs(yq - y)) > 0.0
