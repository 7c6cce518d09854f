"""Pins the FP64 restatement of the learned-guess network (oracle/mlp.py):
its exact-mode gradient equals central finite differences of its own loss,
for every parameter block; the BF16 mirror mode rounds where documented.
CPU only."""
import numpy as np

from oracle import mlp


def _setup(seed=0, rows=12):
    rng = np.random.default_rng(seed)
    p = np.zeros(mlp.N_PARAMS)
    p[mlp.O_W1:mlp.O_B1] = rng.normal(0, np.sqrt(2 / 72), mlp.N_IN * mlp.N_HID)
    p[mlp.O_B1:mlp.O_W2] = rng.normal(0, 0.1, mlp.N_HID)
    p[mlp.O_W2:mlp.O_B2] = rng.normal(0, np.sqrt(2 / 256), mlp.N_HID * mlp.N_HID)
    p[mlp.O_B2:mlp.O_W3] = rng.normal(0, 0.1, mlp.N_HID)
    p[mlp.O_W3:mlp.O_B3] = rng.normal(0, 0.1, mlp.N_HID * mlp.N_OUT)
    p[mlp.O_B3:] = rng.normal(0, 0.1, mlp.N_OUT)
    x = rng.normal(0, 0.5, (rows, mlp.N_IN))
    t = rng.normal(0, 1, (rows, 5))
    return p, x, t


def test_exact_gradient_matches_finite_differences():
    p, x, t = _setup()
    w = (1.0, 2.0, 0.0, 0.5, 1.0)
    _, g, _ = mlp.forward_backward(p, x, t, w, mode="exact")
    rng = np.random.default_rng(1)
    blocks = [(mlp.O_W1, mlp.O_B1), (mlp.O_B1, mlp.O_W2), (mlp.O_W2, mlp.O_B2), (mlp.O_B2, mlp.O_W3),
              (mlp.O_W3, mlp.O_B3), (mlp.O_B3, mlp.N_PARAMS)]
    h = 1e-6
    for lo, hi in blocks:
        idx = rng.choice(np.arange(lo, hi), size=min(12, hi - lo), replace=False)
        for i in idx:
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (mlp.forward_backward(pp, x, t, w, mode="exact")[0] -
                  mlp.forward_backward(pm, x, t, w, mode="exact")[0]) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1e-3, np.abs(g[lo:hi]).max()), (lo, i, fd, g[i])
    # a random direction through all parameters at once
    v = rng.normal(0, 1, p.size)
    fd = (mlp.forward_backward(p + h * v, x, t, w, mode="exact")[0] -
          mlp.forward_backward(p - h * v, x, t, w, mode="exact")[0]) / (2 * h)
    assert abs(fd - g @ v) < 1e-6 * abs(g @ v)


def test_bf16_rounding_ties_to_even():
    vals = np.array([1.0, 1.00390625, 1.005859375, 1.0078125 + 2 ** -9, -3.3e-3], np.float32)
    r = mlp.bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0                    # halfway, rounds to the even mantissa
    assert r[2] == 1.0078125              # above halfway
    assert r[3] == 1.0078125 or r[3] == 1.015625
    assert abs(r[4] - (-3.3e-3)) < 2 ** -8 * 3.3e-3
    # the mirror gradient stays within BF16 rounding of the exact one
    p, x, t = _setup(2)
    _, ge, _ = mlp.forward_backward(p, x, t, mode="exact")
    _, gm, _ = mlp.forward_backward(p, x, t, mode="mirror")
    assert np.abs(gm - ge).max() < 0.05 * np.abs(ge).max()
