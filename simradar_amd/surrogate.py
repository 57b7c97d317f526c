"""Debris surrogate network: training from the table path and use in the engine.

    m = fit(radar, debris_type)          # samples from the type's ADM / RCS tables
    radar.set_debris_surrogate(debris_type, m)

fit() draws samples with rsd_debris_samples (uniform orientations and flow
directions, targets read from the tables exactly as the debris kernel reads
them), normalises each output by its mean and standard deviation (kept in the
network as output scale and shift) and trains with the device's BF16 MFMA
kernels (rsd_mlp_train).  The table path stays available (set the surrogate to
None) and judges the network: error() compares both on held-out samples."""
import ctypes

import numpy as np

from . import _ptr, load


class Mlp:
    """A 7 -> 4x64 -> 12 network on one GPU (rsd_mlp_*)."""

    def __init__(self, seed=1, device=0):
        self.lib = load()
        m = ctypes.c_void_p()
        rc = self.lib.rsd_mlp_create(device, seed, ctypes.byref(m))
        if rc:
            raise RuntimeError("rsd_mlp_create failed (%d)" % rc)
        self.m = m

    def _chk(self, rc):
        if rc:
            raise RuntimeError("network: %s" % self.lib.rsd_mlp_last_error(self.m).decode())

    def params(self):
        p = np.zeros(13772, np.float32)
        self._chk(self.lib.rsd_mlp_get_params(self.m, _ptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._chk(self.lib.rsd_mlp_set_params(self.m, _ptr(p)))

    def set_output_scaling(self, scale, shift):
        sc, sh = np.ascontiguousarray(scale, np.float32), np.ascontiguousarray(shift, np.float32)
        self._chk(self.lib.rsd_mlp_set_output_scaling(self.m, _ptr(sc), _ptr(sh)))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], 12), np.float32)
        self._chk(self.lib.rsd_mlp_forward(self.m, _ptr(x), _ptr(y), x.shape[0]))
        return y

    def train(self, x, t, epochs, batch, lr):
        x, t = np.ascontiguousarray(x, np.float32), np.ascontiguousarray(t, np.float32)
        self._chk(self.lib.rsd_mlp_set_data(self.m, _ptr(x), _ptr(t), x.shape[0]))
        loss = ctypes.c_double()
        self._chk(self.lib.rsd_mlp_train(self.m, epochs, batch, ctypes.c_float(lr), ctypes.byref(loss)))
        return loss.value

    def free(self):
        if self.m:
            self.lib.rsd_mlp_destroy(self.m)
            self.m = None


def table_samples(radar, debris_type, n, seed):
    """(x[n, 7], t[n, 12]) from the ADM / RCS tables of a debris type (>= 1)."""
    lib = radar.lib
    adm = rcs = debris_type - 1
    x = np.zeros((n, 7), np.float32)
    t = np.zeros((n, 12), np.float32)
    e = radar.engine()
    if lib.rsd_debris_samples(e, adm, rcs, n, seed, _ptr(x), _ptr(t)):
        raise RuntimeError(lib.rsd_last_error(e).decode())
    return x, t


def state_samples(radar, debris_type):
    """(x[n, 7], t[n, 12]) at the current states of a debris type's bodies:
    the inputs its advance and pulse would give the network, and the table
    path's targets there (rsd_debris_state_samples)."""
    lib = radar.lib
    adm = rcs = debris_type - 1
    n = int(radar.info().counts[debris_type])
    x = np.zeros((max(n, 1), 7), np.float32)
    t = np.zeros((max(n, 1), 12), np.float32)
    got = ctypes.c_size_t()
    inf = radar.info()
    sd = np.ascontiguousarray(radar.sim_desc(), np.float32)
    e = radar.engine()
    if lib.rsd_debris_state_samples(e, debris_type, adm, rcs, int(inf.les_id), _ptr(sd), n, _ptr(x), _ptr(t),
                                    ctypes.byref(got)):
        raise RuntimeError(lib.rsd_last_error(e).decode())
    return x[:got.value], t[:got.value]


def training_set(radar, debris_type, n_samples, seed, state_frac=0.0):
    """n_samples table-path samples: uniform over orientations and flow
    directions, of which a fraction state_frac is replaced by the population's
    current states (repeated to fill their share)."""
    n_state = int(round(n_samples * state_frac))
    x, t = table_samples(radar, debris_type, n_samples - n_state, seed)
    if n_state > 0:
        xs, ts = state_samples(radar, debris_type)
        if len(xs):
            idx = np.random.default_rng(seed).integers(0, len(xs), n_state)
            x = np.concatenate([x, xs[idx]])
            t = np.concatenate([t, ts[idx]])
            perm = np.random.default_rng(seed + 1).permutation(len(x))
            x, t = x[perm], t[perm]
    return x, t


def fit(radar, debris_type, n_samples=1 << 21, epochs=30, batch=4096, lr=2e-3, seed=7, state_frac=0.0):
    """Train a surrogate of one debris type's tables; returns (Mlp, last batch loss).
    state_frac > 0 mixes samples at the population's current states into the
    uniform ones (training_set)."""
    x, t = training_set(radar, debris_type, n_samples, seed, state_frac)
    mean = t.mean(0)
    std = t.std(0)
    std[std < 1e-12] = 1.0
    m = Mlp(seed=seed)
    m.set_output_scaling(std, mean)
    loss = m.train(x, (t - mean) / std, epochs, batch, lr)
    return m, loss


def fit_many(radar, debris_types, n_samples=1 << 21, epochs=30, batch=4096, lr=2e-3, seed=7, state_frac=0.0):
    """fit() for several debris types: the training sets are drawn one after
    another (they use the engine), then the networks train at the same time,
    each on its own stream from its own host thread (a 4096-sample step fills
    a quarter of the device, so the three types' steps overlap).  Each network
    comes out as fit() would make it.  Returns [(Mlp, last batch loss)]."""
    from concurrent.futures import ThreadPoolExecutor
    jobs = []
    for k in debris_types:
        x, t = training_set(radar, k, n_samples, seed, state_frac)
        mean = t.mean(0)
        std = t.std(0)
        std[std < 1e-12] = 1.0
        m = Mlp(seed=seed)
        m.set_output_scaling(std, mean)
        jobs.append((m, x, (t - mean) / std))
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        futs = [ex.submit(m.train, x, t, epochs, batch, lr) for m, x, t in jobs]
        losses = [f.result() for f in futs]
    return [(m, loss) for (m, _, _), loss in zip(jobs, losses)]


def error(m, radar, debris_type, n=65536, seed=99):
    """Relative RMS error of the network against the tables on held-out samples:
    (drag cd, moment cm, scattering amplitudes), each RMS(y - t) / RMS(t)."""
    x, t = table_samples(radar, debris_type, n, seed)
    y = m.forward(x)
    out = []
    for sl in (slice(0, 3), slice(3, 6), slice(6, 12)):
        out.append(float(np.sqrt(np.mean((y[:, sl] - t[:, sl]) ** 2)) / np.sqrt(np.mean(t[:, sl] ** 2))))
    return tuple(out)
