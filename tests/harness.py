"""Shared set-up of the test workloads: the engine run (simradar_amd) and the
matching oracle tables and stepping."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import simradar_amd as sa  # noqa: E402
from simradar_amd.workloads import Config, configure, mp_dsd, SEED  # noqa: E402,F401
from oracle import oracle as orc  # noqa: E402


from oracle.run import OracleRun, adm_desc, rcs_desc, state_from_radar  # noqa: E402,F401


def rel_err(a, b):
    """max |a - b| over gates, relative to the RMS gate magnitude of b."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = np.sqrt(np.mean(np.sum(b * b, axis=-1)))
    if scale == 0.0:
        return float(np.max(np.abs(a - b)))
    return float(np.max(np.abs(a - b)) / scale)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def run_pair(cfg, pulses, sort=None, fold=None, launch=None, pre=None, setup=None, oracle_kw=None,
             pulse_mode=None, les_switch=False, on_pulse=None, on_switch=None):
    """Populate an engine run of cfg and step it `pulses` times next to the
    oracle from the same populated state (the driver order of
    simradar.c:1011-1074).  pre(r) runs before RS_populate, setup(r, o) after
    the oracle is made; oracle_kw go to OracleRun; pulse_mode selects the
    oracle's phase precision; les_switch: the oracle follows the host's LES
    volume switches; on_pulse(k, r, o) runs after pulse k is compared;
    on_switch(r, o) after the host loads a new LES volume (before the oracle's
    advance).
    Returns (radar, oracle, engine pulses, oracle pulses)."""
    r = sa.Radar(verb=0)
    configure(r, cfg)
    if pre is not None:
        pre(r)
    r.populate()
    r.check()
    if sort is not None:
        r.set_sort(sort)
    if fold is not None:
        r.set_fold(fold)
    if launch is not None:
        r.set_launch(*launch)
    o = OracleRun(r, cfg, state_from_radar(r), **(oracle_kw or {}))
    sd_adv = r.sim_desc().copy()   # populate's zero-length step (rs.c:4941-4955)
    sd_adv[11] = 0.0
    o.advance(sd_adv)
    o.refresh_debris_rcs(r.sim_desc())
    if setup is not None:
        setup(r, o)
    mode = orc.PHASE_F64 if pulse_mode is None else pulse_mode
    blend = bool(r.info().les_blend)
    frames = [None]

    def follow_blend():
        """The oracle's volumes and weight for the coming advance under time
        interpolation: the active volume vel_idx - 2, the next vel_idx - 1."""
        inf = r.info()
        vn = int(inf.vel_idx)
        assert vn >= 2
        if frames[0] != vn:
            o.set_les_frame(r.les_frame(vn - 2), r.les_cpxx(vn - 2) if o.fixed else None)
            nxt = orc.tex(r.les_frame(vn - 1))
            nc = orc.tex(r.les_cpxx(vn - 1)) if o.fixed else (None, None)
            frames[0] = vn
            frames.append((nxt, nc))
        (wt, wk), (ct, ck) = frames[-1]
        orc.set_blend(wt, ct, float(inf.wind_alpha))
    pg, po = [], []
    try:
        _steps(r, o, pulses, mode, pg, po, on_pulse, les_switch, follow_blend if blend else None, on_switch)
    finally:
        orc.set_blend(None)
    r.download()
    r.check()
    return r, o, np.array(pg), np.array(po)


def _steps(r, o, pulses, mode, pg, po, on_pulse, les_switch, follow_blend, on_switch=None):
    for k in range(pulses):
        if k > 0:
            r.make_pulse()
            o.refresh_debris_rcs(r.sim_desc())
        sd = r.sim_desc()
        r.download_pulse_only()
        r.check()
        pg.append(r.pulse_f64())
        po.append(o.pulse(sd, mode))
        if on_pulse is not None:
            on_pulse(k, r, o)
        vi = int(r.info().vel_idx)
        r.advance_time()
        if follow_blend is not None:
            follow_blend()
        elif les_switch and int(r.info().vel_idx) != vi:
            o.set_les_frame(r.les_frame(vi), r.les_cpxx(vi) if o.fixed else None)
            if on_switch is not None:
                on_switch(r, o)
        o.advance(r.sim_desc())
        r.advance_beam()


def check_state(r, o, streams=("pos", "vel", "ori", "tum")):
    for s in streams:
        a, b = r.stream(s), getattr(o.st, s)
        bad = np.nonzero(np.any(bits(a) != bits(b), axis=1))[0]
        assert bad.size == 0, "%s differs at %d scatterers, first %s: %s vs %s" % (s, bad.size, bad[:3], a[bad[:1]],
                                                                                b[bad[:1]])
    assert np.array_equal(r.stream("rnd"), o.st.rnd), "seeds differ"


def iq_errs(pg, po, first=0):
    return [rel_err(pg[k], po[k]) for k in range(first, len(pg))]
