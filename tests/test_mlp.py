"""The debris surrogate network: the oracle's gradient is pinned by finite
differences; the device forward and the device gradient (BF16 MFMA, FP32
accumulate) match the oracle with the same bfloat16 rounding points; training
on the device lowers the loss of a smooth target."""
import ctypes

import numpy as np
import pytest

import simradar_amd as sa
from oracle import mlp


def _params(seed, scale=1.0):
    g = np.random.default_rng(seed)
    p = []
    for s in mlp.SHAPES:
        fan = s[1] if len(s) == 2 else s[0]
        p.append(g.normal(0, scale / np.sqrt(fan), size=s) if len(s) == 2 else g.normal(0, 0.1, size=s))
    return mlp.flatten(p)


def _data(n, seed):
    g = np.random.default_rng(seed)
    q = g.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    u = g.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    x = np.concatenate([q, u], 1)
    return x


def test_oracle_gradient_matches_finite_differences():
    p = _params(1)
    x = _data(32, 2)
    t = np.random.default_rng(3).normal(size=(32, 12))
    loss, g = mlp.loss_grad(p, x, t)
    rng = np.random.default_rng(4)
    idx = rng.choice(p.size, 60, replace=False)
    h = 1e-7   # small enough that no ReLU changes state
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp.loss_grad(pp, x, t)[0] - mlp.loss_grad(pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_oracle_linear_regime():
    """With positive weights, inputs and biases every ReLU is active and the
    network is the affine map of its matrices."""
    g = np.random.default_rng(5)
    W = [np.abs(g.normal(size=s)) * (0.1 if len(s) == 2 else 0.01) for s in mlp.SHAPES]
    p = mlp.flatten(W)
    x = np.abs(_data(8, 6))
    y, _ = mlp.forward(p, x)
    a = x
    for L in range(4):
        a = a @ W[2 * L].T + W[2 * L + 1]
    lin = a @ W[8].T + W[9]
    np.testing.assert_allclose(y, lin, rtol=1e-12)


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5], np.float32)
    np.testing.assert_array_equal(mlp.bf16(v), [1.0, 1.0, 1.0 + 2 ** -7, -2.5])


# ---------------------------------------------------------------- GPU

def _net(seed=11):
    lib = sa.load()
    m = ctypes.c_void_p()
    assert lib.rsd_mlp_create(0, seed, ctypes.byref(m)) == 0
    return lib, m


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


@pytest.mark.gpu
def test_device_forward_matches_oracle(gpu_ready):
    lib, m = _net()
    p = _params(7).astype(np.float32)
    assert lib.rsd_mlp_set_params(m, sa._ptr(p)) == 0
    x = _f32(_data(1000, 8))
    y = np.zeros((1000, 12), np.float32)
    assert lib.rsd_mlp_forward(m, sa._ptr(x), sa._ptr(y), 1000) == 0
    ref, _ = mlp.forward(p.astype(np.float64), x.astype(np.float64), bf=True)
    err = np.abs(y - ref).max() / np.sqrt(np.mean(ref ** 2))
    assert err < 2e-2, err
    lib.rsd_mlp_destroy(m)


@pytest.mark.gpu
def test_device_gradient_matches_oracle(gpu_ready):
    lib, m = _net()
    p = _params(9).astype(np.float32)
    assert lib.rsd_mlp_set_params(m, sa._ptr(p)) == 0
    n = 1024
    x = _f32(_data(n, 10))
    t = _f32(np.random.default_rng(12).normal(size=(n, 12)))
    g = np.zeros(mlp.NPARAMS, np.float32)
    loss = ctypes.c_double()
    assert lib.rsd_mlp_gradient(m, sa._ptr(x), sa._ptr(t), n, sa._ptr(g), ctypes.byref(loss)) == 0
    lo, gr = mlp.loss_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), bf=True)
    assert abs(loss.value - lo) < 1e-2 * lo
    rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    assert rel < 3e-2, rel
    # every layer's block separately (a transposed or shifted block fails here)
    k = 0
    for s in mlp.SHAPES:
        nn = int(np.prod(s))
        blk = slice(k, k + nn)
        assert np.linalg.norm(g[blk] - gr[blk]) <= 5e-2 * np.linalg.norm(gr[blk]) + 1e-9, s
        k += nn
    lib.rsd_mlp_destroy(m)


@pytest.mark.gpu
def test_device_training_fits_a_smooth_target(gpu_ready):
    lib, m = _net(21)
    n = 65536
    x = _data(n, 22)
    t = np.stack([np.sin(2 * x[:, 0] + x[:, 4]), x[:, 1] * x[:, 5], np.cos(3 * x[:, 2]) * x[:, 6]] * 4, 1)
    t = _f32(t)
    xf = _f32(x)
    assert lib.rsd_mlp_set_data(m, sa._ptr(xf), sa._ptr(t), n) == 0
    g = np.zeros(mlp.NPARAMS, np.float32)
    l0 = ctypes.c_double()
    assert lib.rsd_mlp_gradient(m, sa._ptr(xf[:4096]), sa._ptr(t[:4096]), 4096, sa._ptr(g), ctypes.byref(l0)) == 0
    l1 = ctypes.c_double()
    assert lib.rsd_mlp_train(m, 20, 1024, ctypes.c_float(3e-3), ctypes.byref(l1)) == 0
    assert l1.value < 0.1 * l0.value, (l0.value, l1.value)
    # the trained device parameters reproduce the device loss in the oracle
    p = np.zeros(mlp.NPARAMS, np.float32)
    assert lib.rsd_mlp_get_params(m, sa._ptr(p)) == 0
    lo, _ = mlp.loss_grad(p.astype(np.float64), x[:4096], t[:4096].astype(np.float64), bf=True)
    assert lo < 0.1 * l0.value
    lib.rsd_mlp_destroy(m)
