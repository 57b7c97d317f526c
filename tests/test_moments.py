"""Radar moments (Z, V, ZDR, rho_hv; pulse-pair over the IQ history).

The reference has no estimator (only an unused zdr(), rs.c:376-378); the
definition is include/rsd.h's.  The oracle's estimator is pinned by closed
forms: a single scatterer in uniform radial motion (velocity, aliasing,
power, differential reflectivity, rho = 1) and uncorrelated channels
(rho -> 0).  The GPU estimator is compared with the oracle's on the engine's
own IQ, and moments from GPU IQ with moments from oracle IQ (tolerances
0.1 dB Z and 0.05 m/s)."""
import numpy as np
import pytest

import harness as h
import simradar_amd as sa
from oracle import oracle as orc

LAM, PRT = 0.1, 1e-3
VA = LAM / (4 * PRT)     # Nyquist velocity, 25 m/s


def _moving_point(v, a=2.0, b=0.5, r0=2000.0, M=64, R=3, delta=0.3):
    k = np.arange(M)
    ph = -4 * np.pi * (r0 + v * k * PRT) / LAM
    iq = np.zeros((M, R, 4))
    iq[:, :, 0] = (a * np.cos(ph))[:, None]
    iq[:, :, 1] = (a * np.sin(ph))[:, None]
    iq[:, :, 2] = (b * np.cos(ph + delta))[:, None]
    iq[:, :, 3] = (b * np.sin(ph + delta))[:, None]
    return iq


@pytest.mark.parametrize("v", [-24.0, -7.5, 0.0, 3.25, 19.0])
def test_single_scatterer_velocity_power_zdr_rho(v):
    m = orc.moments(_moving_point(v), LAM, PRT, 1.0, 1000.0, 15.0)
    rg = 1000.0 + 15.0 * np.arange(3)
    np.testing.assert_allclose(m[:, 1], v, atol=1e-9)
    np.testing.assert_allclose(m[:, 0], 20 * np.log10(2.0) + 20 * np.log10(rg), atol=1e-9)
    np.testing.assert_allclose(m[:, 2], 20 * np.log10(2.0 / 0.5), atol=1e-9)
    np.testing.assert_allclose(m[:, 3], 1.0, atol=1e-12)


def test_velocity_aliases_into_the_nyquist_interval():
    m = orc.moments(_moving_point(VA + 3.0), LAM, PRT, 1.0, 1000.0, 15.0)
    np.testing.assert_allclose(m[:, 1], -VA + 3.0, atol=1e-9)


def test_gain_shifts_z_only():
    a = orc.moments(_moving_point(5.0), LAM, PRT, 1.0, 1000.0, 15.0)
    b = orc.moments(_moving_point(5.0), LAM, PRT, 10.0, 1000.0, 15.0)
    np.testing.assert_allclose(b[:, 0] - a[:, 0], 20.0, atol=1e-9)
    np.testing.assert_allclose(b[:, 1:], a[:, 1:], atol=1e-12)


def test_uncorrelated_channels_give_low_rho():
    g = np.random.default_rng(4)
    M = 4096
    iq = g.normal(size=(M, 2, 4))
    m = orc.moments(iq, LAM, PRT, 1.0, 1000.0, 15.0)
    assert np.all(m[:, 3] < 4 / np.sqrt(M))
    np.testing.assert_allclose(m[:, 2], 0.0, atol=0.3)


# ---------------------------------------------------------------- GPU

def _record(cfg, pulses, fold=None):
    r = sa.Radar(verb=0)
    h.configure(r, cfg)
    r.populate()
    r.check()
    if fold is not None:
        r.set_fold(fold)
    r.set_moment_pulses(pulses)
    st = h.state_from_radar(r)
    o = h.OracleRun(r, cfg, st)
    sd_adv = r.sim_desc().copy()
    sd_adv[11] = 0.0
    o.advance(sd_adv)
    o.refresh_debris_rcs(r.sim_desc())
    pg, po = [], []
    for k in range(pulses):
        r.make_pulse()
        o.refresh_debris_rcs(r.sim_desc())
        sd = r.sim_desc()
        r.download_pulse_only()
        r.check()
        pg.append(r.pulse_f64())
        po.append(o.pulse(sd))
        r.advance_time()
        o.advance(r.sim_desc())
        r.advance_beam()
    return r, np.array(pg), np.array(po)


@pytest.mark.gpu
def test_gpu_moments_match_oracle_estimator_and_oracle_iq(gpu_ready):
    cfg = h.Config(n_rain=16384, scan="P:3.0,-1:1:0.01")
    M = 24
    r, pg, po = _record(cfg, M)
    p = r.params()
    gain = float(r.lib.RS_radar_gain(r.H))
    dev = r.moments(M).astype(np.float64)
    iq_g = pg.reshape(M, -1, 4)
    iq_o = po.reshape(M, -1, 4)
    mg = orc.moments(iq_g, p.lambda_, p.prt, gain, p.range_start, p.range_delta)
    mo = orc.moments(iq_o, p.lambda_, p.prt, gain, p.range_start, p.range_delta)
    live = np.isfinite(mo[:, 0]) & (mo[:, 0] > np.nanmax(mo[:, 0]) - 60.0)
    assert live.sum() >= 10
    # the device estimator is the oracle estimator on the same IQ
    np.testing.assert_allclose(dev[live, 0], mg[live, 0], atol=1e-4)
    np.testing.assert_allclose(dev[live, 1], mg[live, 1], atol=1e-4)
    np.testing.assert_allclose(dev[live, 2], mg[live, 2], atol=1e-4)
    np.testing.assert_allclose(dev[live, 3], mg[live, 3], atol=1e-5)
    # moments of GPU IQ against moments of oracle IQ
    assert np.max(np.abs(mg[live, 0] - mo[live, 0])) < 0.1
    assert np.max(np.abs(mg[live, 1] - mo[live, 1])) < 0.05
    assert np.max(np.abs(mg[live, 2] - mo[live, 2])) < 0.1
    assert np.max(np.abs(mg[live, 3] - mo[live, 3])) < 1e-3
    r.free()


@pytest.mark.gpu
def test_unread_pulses_fill_the_ring_as_read_ones(gpu_ready):
    """A rain pulse's gate sums are finished by the next rain launch (or by the
    first read): pulses made back to back without a read between them give
    the same IQ ring, moments and last pulse, bit for bit, as pulses read one
    by one."""
    cfg = h.Config(n_rain=16384, debris=[(sa.OBJ_LEAF, 512)], scan="P:3.0,-1:1:0.01")
    M = 12
    runs = []
    for read_each in (True, False):
        r = sa.Radar(verb=0)
        h.configure(r, cfg)
        r.populate()
        r.set_moment_pulses(M)
        for k in range(M):
            r.make_pulse()
            if read_each:
                r.download_pulse_only()
            r.advance_time()
            r.advance_beam()
        r.download_pulse_only()
        r.check()
        runs.append((r.pulse_f64().copy(), r.moments(M).copy()))
        r.free()
    (p_a, m_a), (p_b, m_b) = runs
    assert np.array_equal(h.bits(p_a), h.bits(p_b))
    assert np.array_equal(m_a.view(np.uint32), m_b.view(np.uint32))


@pytest.mark.gpu
def test_gpu_moments_need_enough_pulses(gpu_ready):
    cfg = h.Config(n_rain=4096, scan="P:3.0,-1:1:0.01")
    r = sa.Radar(verb=0)
    h.configure(r, cfg)
    r.populate()
    r.set_moment_pulses(8)
    for _ in range(3):
        r.make_pulse()
        r.advance_time()
    with pytest.raises(RuntimeError):
        r.moments(8)
    r.free()
