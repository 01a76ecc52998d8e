"""Float64 reference of the flux-surrogate network — TEST INFRASTRUCTURE ONLY.

16 -> 64 -> 64 -> 64 -> 9 with the tanh-form GELU, loss = mean over samples and the 7
learned output slots (0-4, 6, 7) of the squared error; parameters in the flat order
[W1 b1 W2 b2 W3 b3 W4 b4], W_l stored [in][out]. The network has no counterpart in the
reference (it is new work on top of the MHDKernels step); its gradient is pinned by
central finite differences in tests/test_surrogate.py.
"""
import numpy as np

SIZES = [(16, 64), (64, 64), (64, 64), (64, 9)]
LEARNED = np.array([1, 1, 1, 1, 1, 0, 1, 1, 0], dtype=np.float64)
K0, K1 = 0.7978845608028654, 0.044715


def nparams():
    return sum(a * b + b for a, b in SIZES)


def unpack(theta):
    out, o = [], 0
    for a, b in SIZES:
        W = theta[o:o + a * b].reshape(a, b)
        o += a * b
        bb = theta[o:o + b]
        o += b
        out.append((W, bb))
    return out


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(K0 * (z + K1 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(K0 * (z + K1 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * K0 * (1.0 + 3.0 * K1 * z * z)


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def forward(theta, X, round_bf16=False):
    r = bf16 if round_bf16 else (lambda v: v)
    layers = unpack(np.asarray(theta, dtype=np.float64))
    A = np.asarray(X, dtype=np.float64)
    cache = []
    for l, (W, b) in enumerate(layers):
        Z = r(A) @ r(W) + b
        cache.append((A, Z))
        A = gelu(Z) if l < 3 else Z
    return A, cache


def loss_and_grad(theta, X, T, round_bf16=False):
    r = bf16 if round_bf16 else (lambda v: v)
    layers = unpack(np.asarray(theta, dtype=np.float64))
    Y, cache = forward(theta, X, round_bf16)
    n = X.shape[0]
    D = (Y - T) * LEARNED
    loss = float(np.sum(D * D) / (n * 7))
    dZ = 2.0 * D / (n * 7)
    grads = [None] * 4
    for l in range(3, -1, -1):
        A_in, Z = cache[l]
        W, _ = layers[l]
        gW = r(A_in).T @ r(dZ)
        gb = dZ.sum(axis=0)
        grads[l] = (gW, gb)
        if l > 0:
            dA = r(dZ) @ r(W).T
            dZ = dA * gelu_grad(cache[l - 1][1])
    flat = np.concatenate([np.concatenate([g.reshape(-1), b]) for g, b in grads])
    return loss, flat
