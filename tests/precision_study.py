"""Precision study of the C3 geometry (rain in a 3 km x 3 km x 1 km box at 10 km
range, 202 gates of 15 m) against a full-binary64 oracle.

Three engine variants run from one populated state:
  fp32   positions binary32, phase r 4 pi / lambda in binary32 (rs.cl:1096)
  f64ph  positions binary32, phase reduced in binary64 (the engine default)
  hilo   positions as binary32 hi/lo pairs, range and phase in binary64
The oracle advances binary64 positions and velocities with the same drag law
and wind table (orc_el_atts_f64) and forms every pulse in binary64
(ORC_FULL_F64).  Reported per variant: the IQ error (max |a - b| over the RMS
gate magnitude, worst pulse) and the moment errors (Z, V, ZDR, rho_hv over the
recorded pulses, gates within 40 dB of the strongest).  Drops are placed between
100 m and 900 m height and at least 1 m inside the side faces, so that none
leaves the box during the run (the binary64 oracle does not respawn).

    python tests/precision_study.py [n_rain] [pulses]
"""
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import harness as h                       # noqa: E402
import simradar_amd as sa                 # noqa: E402
from simradar_amd import workloads as wl  # noqa: E402
from oracle import oracle as orc          # noqa: E402

VARIANTS = ("fp32", "f64ph", "hilo")


def _start(cfg):
    r = sa.Radar(verb=0)
    wl.configure(r, cfg)
    r.populate()
    r.check()
    pos = r.stream("pos", copy=False)
    sd = r.sim_desc()
    g = np.random.default_rng(77)
    pos[:, 2] = g.uniform(100.0, 900.0, pos.shape[0]).astype(np.float32)
    # and at least 1 m inside the side faces (millions of uniform drops would
    # otherwise put some within the few millimetres they move in a run)
    for a in (0, 1):
        pos[:, a] = np.clip(pos[:, a], sd[8 + a] + 1.0, sd[8 + a] + sd[12 + a] - 1.0)
    r.upload()
    r.check()
    return r


def _rain_rcs(pos64, radius, drop, dd):
    """Drop RCS of rs.cl:268-280 at binary64 positions (truncating table index)."""
    idx = np.clip(np.float32(radius) * dd[0] + dd[1], 0, dd[2]).astype(np.int64)
    t = drop[idx].astype(np.float64)
    rho2 = pos64[:, 0] ** 2 + pos64[:, 1] ** 2
    cb2 = rho2 / (rho2 + pos64[:, 2] ** 2)
    return np.stack([t[:, 0], t[:, 1], t[:, 0] + (t[:, 2] - t[:, 0]) * cb2,
                     t[:, 1] + (t[:, 3] - t[:, 1]) * cb2], 1).astype(np.float32)


def _pieces(n, threads):
    cuts = [n * j // threads for j in range(threads + 1)]
    return [(cuts[j], cuts[j + 1]) for j in range(threads) if cuts[j + 1] > cuts[j]]


def run(n_rain=65536, pulses=20, threads=1):
    """threads > 1 splits the binary64 oracle's population into contiguous
    pieces on host threads (each scatterer's advance is independent; the pulse
    sums the pieces' gate sums in piece order)."""
    cfg = wl.c3(n_rain=n_rain)
    pool = ThreadPoolExecutor(threads) if threads > 1 else None

    def each(fn, n):
        rs = _pieces(n, max(1, threads))
        return list(pool.map(lambda t: fn(*t), rs)) if pool else [fn(*t) for t in rs]
    iq = {}
    for v in VARIANTS:
        r = _start(cfg)
        if v == "fp32":
            r.set_phase_precision(sa.PHASE_F32)
        if v == "hilo":
            r.set_position_precision(1)
        out = []
        for k in range(pulses + 1):
            r.make_pulse()
            r.download_pulse_only()
            r.check()
            out.append(r.pulse_f64())
            r.advance_time()
            r.advance_beam()
        iq[v] = np.array(out[1:])
        r.free()
    # binary64 oracle from the same uploaded state
    r = _start(cfg)
    st = h.state_from_radar(r)
    o = h.OracleRun(r, cfg, st)
    p = r.params()
    R = int(p.range_count)
    pos64 = st.pos[:, :3].astype(np.float64).copy()
    vel64 = st.vel[:, :3].astype(np.float64).copy()
    radius = st.pos[:, 3].copy()
    bounded = bool(cfg.concept & sa.CONCEPT_BOUNDED)
    ref = []
    for k in range(pulses + 1):
        sd = r.sim_desc()
        rcs = _rain_rcs(pos64, radius, o.drop, o.drop_desc)
        if k > 0:
            parts = each(lambda lo, hi: orc.make_pulse(st.pos[lo:hi], rcs[lo:hi], o.angw, o.angw_desc, o.rangew,
                                                       o.rangew_desc, p.range_start, p.range_delta, R, sd, o.kd,
                                                       orc.FULL_F64, pos64=pos64[lo:hi]), pos64.shape[0])
            ref.append(sum(parts[1:], parts[0]))
        r.lib.RS_advance_time(r.H)
        sda = r.sim_desc()
        each(lambda lo, hi: orc.el_atts_f64(pos64[lo:hi], vel64[lo:hi], radius[lo:hi], o.wind_tex, o.les_desc, sda,
                                            bounded), pos64.shape[0])
        r.lib.RS_advance_beam(r.H)
    ref = np.array(ref).reshape(pulses, R, 4)
    gain = float(r.lib.RS_radar_gain(r.H))
    mref = orc.moments(ref, p.lambda_, p.prt, gain, p.range_start, p.range_delta)
    live = np.isfinite(mref[:, 0]) & (mref[:, 0] > np.nanmax(mref[:, 0]) - 40.0)
    rows = {}
    for v in VARIANTS:
        a = iq[v].reshape(pulses, R, 4)
        err = max(h.rel_err(a[k].ravel(), ref[k].ravel()) for k in range(pulses))
        m = orc.moments(a, p.lambda_, p.prt, gain, p.range_start, p.range_delta)
        d = np.abs(m[live] - mref[live])
        rows[v] = dict(iq=err, z=d[:, 0].max(), v=d[:, 1].max(), zdr=d[:, 2].max(), rho=d[:, 3].max())
    r.free()
    if pool:
        pool.shutdown()
    return rows, int(live.sum()), R


def table(rows):
    out = ["| variant | IQ error | Z (dB) | V (m/s) | ZDR (dB) | rho_hv |", "|---|---:|---:|---:|---:|---:|"]
    for v, e in rows.items():
        out.append("| %s | %.2e | %.3g | %.3g | %.3g | %.2e |" % (v, e["iq"], e["z"], e["v"], e["zdr"], e["rho"]))
    return "\n".join(out)


if __name__ == "__main__":
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 65536
    k = int(sys.argv[2]) if len(sys.argv) > 2 else 20
    rows, live, R = run(n, k)
    print("C3 geometry, %d rain, %d pulses, %d gates (%d within 40 dB of the strongest)" % (n, k, R, live))
    print(table(rows))
