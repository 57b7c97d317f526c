"""Vertically pointing profiler with fixed scatterer positions (concepts V and
F; simple_dbs.c): the host's anchor raster and population, and the oracle's
fp_atts (rs.cl:590-648), pinned without the engine:

* the population is a square raster of zenith offsets (xx, yy) with the
  pattern's angular spacing, one plane of directions repeated at every gate
  range (rs.c:1602-1690, 2143-2157, 4574-4647);
* its random draws follow glibc's rand() stream from srand(seed): age, phase,
  four seeds per scatterer;
* fp_atts under a uniform wind and constant Cn^2: the phase turns by
  -(4 pi / lambda) (p_hat . u) dt per step, the amplitude is 10^cn2, the
  velocity is the wind, the position stays."""
import ctypes

import numpy as np
import pytest

import harness as h
import simradar_amd as sa
from oracle import oracle as orc

FV = sa.CONCEPT_FIXED | sa.CONCEPT_VERTICAL


def _profiler(seed=19760520, gate=30.0, bw=5.0, pattern="D:0,75,50/90,75,50/0,90,50"):
    """simple_dbs.c:31-48, with the synthetic suctvort LES."""
    r = sa.Radar(verb=0, host_only=True)
    r.set_concept(FV)
    r.set_sampling_spacing(gate, 1.0, 1.0)
    r.set_antenna_params(bw, 30.5)
    r.set_tx_params(0.2e-6, 10.0e3)
    r.set_lambda(np.float32(3.0e8 / 915.0e6))
    r.set_prt(np.float32(1.0 / 140.0))
    r.set_random_seed(seed)
    r.set_vel_data_to_config("suctvort")
    r.set_scan_pattern(pattern)
    r.populate()
    return r


def test_profiler_population_is_a_zenith_raster_repeated_at_every_gate():
    r = _profiler()
    p = r.params()
    R, n = int(p.range_count), int(r.info().num_scats)
    assert n % R == 0
    per = n // R
    m = int(round(np.sqrt(per)))
    assert m * m == per and m % 2 == 1            # a square raster centred on the zenith
    pos = r.stream("pos")[:, :3].astype(np.float64)
    rng = np.sqrt((pos ** 2).sum(1)).reshape(R, per)
    want = p.range_start + np.arange(R) * p.range_delta
    np.testing.assert_allclose(rng, np.repeat(want[:, None], per, 1), rtol=1e-6)
    # directions: zenith offset (xx, yy) = (90 - el) (cos az, sin az) on the
    # 1-degree grid (the pattern's spacing), the same in every plane
    d = pos.reshape(R, per, 3) / rng[..., None]
    zen = np.degrees(np.arccos(np.clip(d[..., 2], -1, 1)))
    az = np.arctan2(d[..., 0], d[..., 1])        # x = sin(az), y = cos(az) of the anchor
    xx, yy = zen * np.cos(az), zen * np.sin(az)
    np.testing.assert_allclose(xx, np.round(xx), atol=2e-3)
    np.testing.assert_allclose(yy, np.round(yy), atol=2e-3)
    half = (m - 1) // 2
    grid = {(int(a), int(b)) for a, b in zip(np.round(xx[0]), np.round(yy[0]))}
    assert grid == {(a, b) for a in range(-half, half + 1) for b in range(-half, half + 1)}
    np.testing.assert_allclose(xx, np.repeat(xx[:1], R, 0), atol=2e-3)
    r.free()


def test_fixed_population_draws_follow_the_rand_stream():
    """Per scatterer, in order: age = rand() / RAND_MAX, phase = rand() /
    RAND_MAX * 2 pi, then four raw rand() seeds; rcs starts at (1e-10, 0, 1e-10,
    phase), velocity zero, identity orientation and tumble."""
    seed = 4242
    r = _profiler(seed=seed)
    n = int(r.info().num_scats)
    libc = ctypes.CDLL("libc.so.6")
    libc.rand.restype = ctypes.c_int
    libc.srand(seed)
    k = min(n, 5000)
    draws = np.array([libc.rand() for _ in range(6 * k)], np.int64).reshape(k, 6)
    rmax = np.float32(2147483647)
    aux, rcs, rnd = r.stream("aux"), r.stream("rcs"), r.stream("rnd")
    np.testing.assert_array_equal(aux[:k, 1], draws[:, 0].astype(np.float32) / rmax)
    ph = (draws[:, 1].astype(np.float32) / rmax).astype(np.float64) * 2.0 * np.pi
    np.testing.assert_array_equal(rcs[:k, 3], ph.astype(np.float32))
    np.testing.assert_array_equal(rnd[:k], draws[:, 2:6].astype(np.uint32))
    np.testing.assert_array_equal(rcs[:k, :3], np.tile(np.float32([1e-10, 0.0, 1e-10]), (k, 1)))
    np.testing.assert_array_equal(r.stream("vel")[:k], 0.0)
    np.testing.assert_array_equal(r.stream("ori")[:k], np.tile(np.float32([0, 0, 0, 1]), (k, 1)))
    r.free()


def test_fixed_positions_refuse_debris_and_foreign_populations():
    r = sa.Radar(verb=0, host_only=True)
    h.configure(r, h.Config(n_rain=4096, concept=sa.CONCEPT_FIXED, scan="P:3.0,-1:1:0.01"))
    r.populate()
    assert not (int(r.info().status) & (1 << 4)), "concept F without the profiler's population must not populate"
    r.free()


def _uniform(u, cn2):
    wind = np.tile(np.array([u[0], u[1], u[2], 0.0], np.float32), (4, 4, 4, 1))
    cp = np.tile(np.array([cn2, 0.0, 0.0, 0.0], np.float32), (4, 4, 4, 1))
    les = np.zeros(16, np.float32)
    les[0:3] = 1.0
    les[4:7] = 1.5
    return orc.tex(wind), orc.tex(cp), les


def test_fp_atts_turns_the_phase_by_the_radial_wind():
    g = np.random.default_rng(3)
    u = np.array([3.0, -2.0, 1.5])
    (wt, wk), (ct, ck), les = _uniform(u, -12.5)
    n = 500
    pos = np.stack([g.uniform(-300, 300, n), g.uniform(-300, 300, n), g.uniform(100, 2000, n), np.zeros(n)], -1)
    pos = pos.astype(np.float32)
    ph0 = g.uniform(0, 2 * np.pi, n).astype(np.float32)
    rcs = np.stack([np.full(n, 1e-10), np.zeros(n), np.full(n, 1e-10), ph0], -1).astype(np.float32)
    st = orc.State(pos.copy(), np.zeros((n, 4)), np.tile([0, 0, 0, 1.0], (n, 1)), np.tile([0, 0, 0, 1.0], (n, 1)),
                   rcs, np.ones((n, 4), np.uint32))
    sd = np.zeros(16, np.float32)
    k, dt = np.float32(4 * np.pi / 0.3276), np.float32(1.0 / 140.0)
    sd[4], sd[11] = k, dt
    steps = 20
    for _ in range(steps):
        orc.fp_atts(st, 0, n, wt, ct, les, sd)
    p64 = pos[:, :3].astype(np.float64)
    rad = (p64 / np.linalg.norm(p64, axis=1)[:, None]) @ u
    phi = ph0.astype(np.float64) - steps * float(k) * rad * float(dt)
    wrapped = np.angle(np.exp(1j * phi))
    np.testing.assert_allclose(np.angle(np.exp(1j * (st.rcs[:, 3] - wrapped))), 0.0, atol=2e-4)
    a = 10.0 ** -12.5
    np.testing.assert_allclose(st.rcs[:, 0], a * np.cos(wrapped), rtol=0, atol=a * 3e-4)
    np.testing.assert_allclose(st.rcs[:, 1], a * np.sin(wrapped), rtol=0, atol=a * 3e-4)
    assert np.all(st.rcs[:, 2] == np.float32(-12.5))
    np.testing.assert_allclose(st.vel[:, :3], np.tile(u, (n, 1)), rtol=1e-6)
    assert np.array_equal(st.pos, pos)
