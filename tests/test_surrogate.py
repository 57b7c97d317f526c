"""Debris surrogate network in the engine (C5 pattern): the training samples
are the table path's own values (checked against the oracle's debris RCS and
ADM sampling), the trained network tracks the tables on held-out samples far
better than the untrained one, and a debris run with the network reproduces the
table run's reflectivity within 0.5 dB."""
import numpy as np
import pytest

import harness as h
import simradar_amd as sa
from simradar_amd import surrogate as sg
from simradar_amd import workloads as wl
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _qconj(q):
    return np.array([-q[0], -q[1], -q[2], q[3]], np.float32)


def test_table_samples_are_the_table_path(gpu_ready):
    cfg = h.Config(n_rain=4096, debris=[(sa.OBJ_LEAF, 512)], scan="P:3.0,-1:1:0.01")
    r = wl.make(cfg)
    x, t = sg.table_samples(r, 1, 256, 5)
    st = h.state_from_radar(r)
    o = h.OracleRun(r, cfg, st)
    d = o.deb[0]
    (cd, _), (cm, _), (re, _), (im, _) = d["tex"]
    sd = r.sim_desc()
    oc = np.array([-0.5, 0.5, 0.5, 0.5], np.float32)      # o_conj at el = 0, az = 0
    pos = np.array([0.0, 2000.0, 0.0, 0.0], np.float32)
    rcs_err, adm_err = [], []
    ad = d["adm_desc"]
    for k in range(x.shape[0]):
        R = x[k, :4]
        ori = orc.qmul(R, _qconj(oc))
        ss = orc.debris_rcs(pos, ori, re, im, d["rcs_desc"], sd)
        rcs_err.append(np.abs(ss - t[k, 6:10]).max())
        u = x[k, 4:7].astype(np.float32)
        beta = np.float32(np.arccos(np.float64(u[0])))
        alpha = np.float32(np.arctan2(np.float64(u[2]), np.float64(u[1])))
        if alpha < 0:
            alpha, beta = np.float32(np.pi + alpha), -beta
        cu = np.float32(beta * ad[0] + ad[4])
        cv = np.float32(alpha * ad[1] + ad[5])
        c = orc.sample2d(cd, cu, cv)
        m = orc.sample2d(cm, cu, cv)
        adm_err.append(max(np.abs(c[:3] - t[k, 0:3]).max(), np.abs(m[:3] - t[k, 3:6]).max()))
    rms_rcs = np.sqrt(np.mean(t[:, 6:10] ** 2))
    rms_adm = np.sqrt(np.mean(t[:, 0:6] ** 2))
    assert np.mean(np.array(rcs_err) < 1e-3 * rms_rcs) > 0.97, np.sort(rcs_err)[-5:]
    assert np.mean(np.array(adm_err) < 1e-3 * rms_adm) > 0.97, np.sort(adm_err)[-5:]
    r.free()


def test_trained_network_tracks_the_tables(gpu_ready):
    cfg = h.Config(n_rain=4096, debris=[(sa.OBJ_LEAF, 512)], scan="P:3.0,-1:1:0.01")
    r = wl.make(cfg)
    x, t = sg.table_samples(r, 1, 8192, 3)
    untrained = sg.Mlp(seed=7)
    mean, std = t.mean(0), t.std(0)
    untrained.set_output_scaling(std, mean)
    e0 = sg.error(untrained, r, 1)
    m, loss = sg.fit(r, 1, n_samples=1 << 20, epochs=20)
    e1 = sg.error(m, r, 1)
    print("\nheld-out relative RMS error (cd, cm, amplitudes): untrained %s, trained %s" % (e0, e1))
    assert e1[0] < 0.08 and e1[2] < 0.1, e1           # drag and scattering amplitudes
    assert e1[1] < 0.45, e1                           # moments: the table's noise is ~30 % of cm
    assert e1[0] < 0.2 * e0[0] and e1[2] < 0.3 * e0[2], (e0, e1)
    m.free()
    untrained.free()
    r.free()


def _gate_power(pos, rcs, r0, dr, R, comp):
    """Incoherent per-gate power sum |rcs_c|^2 / r^4 of scatterers binned to
    their nearest gate (the expectation the reflectivity estimates)."""
    r = np.linalg.norm(pos[:, :3].astype(np.float64), axis=1)
    g = np.rint((r - r0) / dr).astype(np.int64)
    ok = (g >= 0) & (g < R)
    p = (rcs[:, comp].astype(np.float64) ** 2 + rcs[:, comp + 1].astype(np.float64) ** 2) / r ** 4
    return np.bincount(g[ok], weights=p[ok], minlength=R), np.bincount(g[ok], minlength=R)


def test_surrogate_reflectivity_matches_the_tables(gpu_ready):
    """Same debris geometry, RCS from the tables and from the networks: per-gate
    reflectivity (incoherent power) within 0.5 dB in H and V.  The bodies carry
    uniformly random orientations (a tumbled population; a freshly populated one
    has every body at one orientation, rs.c:4718-4721)."""
    cfg = h.Config(n_rain=256, debris=[(sa.OBJ_LEAF, 65536), (sa.OBJ_METAL_SHEET, 65536)], scan="P:3.0,-1:1:0.01")
    r = wl.make(cfg)
    n0 = int(r.info().counts[0])
    ori = r.stream("ori", copy=False)
    q = np.random.default_rng(3).normal(size=(ori.shape[0] - n0, 4))
    ori[n0:] = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
    r.upload()
    for _ in range(8):
        r.make_pulse()
        r.advance_time()
        r.advance_beam()
    r.make_pulse()
    r.advance_beam()
    r.make_pulse()                      # table refresh at the current poses
    r.download()
    pos = r.stream("pos")[n0:].copy()
    rcs_t = r.stream("rcs")[n0:].copy()
    nets = []
    for k in (1, 2):
        m, _ = sg.fit(r, k, n_samples=1 << 20, epochs=20)
        nets.append(m)
        r.set_debris_surrogate(k, m)
    r.advance_beam()
    r.make_pulse()                      # network refresh, positions unchanged
    r.download()
    assert np.array_equal(r.stream("pos")[n0:], pos)
    rcs_s = r.stream("rcs")[n0:].copy()
    assert np.isfinite(rcs_s).all() and np.isfinite(rcs_t).all()
    p = r.params()
    R = int(p.range_count)
    for comp, name in ((0, "H"), (2, "V")):
        pt, cnt = _gate_power(pos, rcs_t, p.range_start, p.range_delta, R, comp)
        ps, _ = _gate_power(pos, rcs_s, p.range_start, p.range_delta, R, comp)
        live = (cnt >= 100) & (pt > 0)
        dz = np.abs(10 * np.log10(ps[live] / pt[live]))
        print("\n%s: |dZ| over %d gates: mean %.3f dB, max %.3f dB" % (name, live.sum(), dz.mean(), dz.max()))
        assert live.sum() >= 10
        assert dz.mean() < 0.5 and dz.max() < 0.5, dz
    for m in nets:
        m.free()
    r.free()


def test_networks_trained_together_equal_those_trained_alone(gpu_ready):
    """fit_many trains the types' networks at once on their own streams; each
    comes out bit for bit as fit() makes it alone."""
    cfg = h.Config(n_rain=4096, debris=[(sa.OBJ_LEAF, 512), (sa.OBJ_METAL_SHEET, 512), (sa.OBJ_BRICK, 512)],
                   scan="P:3.0,-1:1:0.01")
    r = wl.make(cfg)
    kw = dict(n_samples=16384, epochs=3, batch=4096, seed=11, state_frac=0.25)
    many = sg.fit_many(r, (1, 2, 3), **kw)
    alone, loss = sg.fit(r, 2, **kw)
    assert np.array_equal(many[1][0].params(), alone.params())
    assert many[1][1] == loss
    assert not np.array_equal(many[0][0].params(), many[1][0].params())
    for m, _ in many:
        m.free()
    alone.free()
    r.free()
