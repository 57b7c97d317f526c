"""Oracle of the debris surrogate network (TEST INFRASTRUCTURE: only tests/,
the smoke check and bench.py's CPU leg use it).

The network of include/rsd.h (rsd_mlp_*): 7 -> 64 -> 64 -> 64 -> 64 -> 12,
ReLU hidden layers, linear output; loss = mean over n x 12 of (y - t)^2.
Forward and backward in binary64 with numpy.  With bf16=True the values the
device stores in bfloat16 are rounded at the same points (inputs, weights,
hidden activations, the output error and every dZ), so the device's result
differs only by the order of its FP32 accumulations.  The gradient is pinned by
central finite differences in tests/test_mlp.py."""
import numpy as np

IN, H, OUT = 7, 64, 12
SHAPES = [(H, IN), (H,), (H, H), (H,), (H, H), (H,), (H, H), (H,), (OUT, H), (OUT,)]
NPARAMS = sum(int(np.prod(s)) for s in SHAPES)


def unflatten(p):
    out, k = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(np.asarray(p[k:k + n], np.float64).reshape(s))
        k += n
    return out


def flatten(mats):
    return np.concatenate([np.asarray(m, np.float64).ravel() for m in mats])


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(p, x, bf=False):
    """Outputs (n, 12) before scaling, and the per-layer inputs / activations."""
    W = unflatten(p)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    a = r(x)
    xs = [a]
    for L in range(4):
        z = a @ r(W[2 * L]).T + W[2 * L + 1]
        a = r(np.maximum(z, 0.0))
        xs.append(a)
    y = a @ r(W[8]).T + W[9]
    return y, xs


def loss_grad(p, x, t, bf=False):
    W = unflatten(p)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    n = x.shape[0]
    y, xs = forward(p, x, bf)
    e = y - t
    loss = float(np.mean(e * e))
    dz = r(2.0 * e / (n * OUT))
    grads = [None] * 10
    for L in range(4, -1, -1):
        grads[2 * L] = dz.T @ xs[L]
        grads[2 * L + 1] = dz.sum(0)
        if L > 0:
            dh = dz @ r(W[2 * L])
            dz = r(dh * (xs[L] > 0.0))
    return loss, flatten(grads)
