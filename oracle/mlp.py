"""FP64 CPU restatement of the learned-initial-guess network (the BF16
perceptron 80 -> 256 -> 256 -> 16 with GELU of imfcs_mlp.hip).

TEST INFRASTRUCTURE ONLY: the checker of the GPU forward/backward kernels.
The reference has no counterpart (the plugin starts every LM fit from the
fit panel's values, Imaging_FCS_1_52.java:1959-1968), so this is pinned by
mathematics instead: the exact mode's gradient against central finite
differences of its own loss.

``mode="mirror"`` rounds to BF16 (round to nearest even) where the GPU path
stores BF16: the inputs, the weights, both hidden activations and the
back-propagated dY, dZ2, dZ1; everything else is FP64.  ``mode="exact"``
rounds nowhere.
"""
import numpy as np
from scipy.special import erf

N_IN, N_HID, N_OUT, N_USED = 80, 256, 16, 5
O_W1 = 0
O_B1 = O_W1 + N_IN * N_HID
O_W2 = O_B1 + N_HID
O_B2 = O_W2 + N_HID * N_HID
O_W3 = O_B2 + N_HID
O_B3 = O_W3 + N_HID * N_OUT
N_PARAMS = O_B3 + N_OUT


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[O_W1:O_B1].reshape(N_IN, N_HID), p[O_B1:O_W2], p[O_W2:O_B2].reshape(N_HID, N_HID), p[O_B2:O_W3],
            p[O_W3:O_B3].reshape(N_HID, N_OUT), p[O_B3:N_PARAMS])


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward_backward(params, inputs, targets, weights=(1, 1, 1, 1, 1), mode="mirror"):
    """Loss (sum_k w_k (y_k - t_k)^2, mean over rows), its parameter gradient
    (layout of imfcs_mlp_get_params) and the outputs [rows][5]."""
    r = bf16 if mode == "mirror" else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(params)
    W1, W2, W3 = r(W1), r(W2), r(W3)
    X = r(inputs)
    B = X.shape[0]
    w = np.zeros(N_OUT)
    w[:N_USED] = weights
    T = np.zeros((B, N_OUT))
    T[:, :N_USED] = targets
    Z1 = X @ W1 + b1
    H1 = r(gelu(Z1))
    Z2 = H1 @ W2 + b2
    H2 = r(gelu(Z2))
    Y = H2 @ W3 + b3
    E = (Y - T) * (w > 0)
    loss = float((w * E * E).sum() / B)
    dY = r(2.0 * w * E / B)
    g = np.zeros(N_PARAMS)
    g[O_W3:O_B3] = (H2.T @ dY).ravel()
    g[O_B3:N_PARAMS] = dY.sum(0)
    dZ2 = r((dY @ W3.T) * gelu_grad(Z2))
    g[O_W2:O_B2] = (H1.T @ dZ2).ravel()
    g[O_B2:O_W3] = dZ2.sum(0)
    dZ1 = r((dZ2 @ W2.T) * gelu_grad(Z1))
    g[O_W1:O_B1] = (X.T @ dZ1).ravel()
    g[O_B1:O_W2] = dZ1.sum(0)
    return loss, g, Y[:, :N_USED]
