"""Radar moments of the debris surrogate against the oracle.

The C5 composition (rain plus leaf, plate and brick debris, here at reduced
counts) runs three ways from one populated state:
  oracle     the CPU oracle's table path (binary32 state, binary64 phase and
             per-gate sums): the reference the errors are stated against
  table      the engine's table path (ADM and RCS lookups)
  surrogate  the engine with every debris type's lookups replaced by its trained
             network (drag and moment coefficients in the advance, scattering
             amplitudes in the pulse)
Moments over the run's pulses (pulse-pair Z, V, ZDR, rho_hv, oracle estimator)
are compared gate by gate over the gates within 40 dB of the strongest.
`tumbled` gives every debris body a uniformly random orientation (a population
that has tumbled for a while); otherwise the bodies keep the one orientation
RS_populate gives them (rs.c:4718-4721).  Networks are fitted to `samples`
table-path samples per type, drawn uniformly over orientations and flow
directions, mixed with `state_frac` of samples taken at the population's own
states (rsd_debris_state_samples).

    python tests/surrogate_study.py [n_rain] [n_debris] [pulses]
"""
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor  # noqa: F401  (threads go to OracleRun)

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import harness as h                       # noqa: E402
import simradar_amd as sa                 # noqa: E402
from simradar_amd import surrogate as sg  # noqa: E402
from simradar_amd import workloads as wl  # noqa: E402
from oracle import oracle as orc          # noqa: E402

TYPES = (sa.OBJ_LEAF, sa.OBJ_METAL_SHEET, sa.OBJ_BRICK)


def write_smooth_tables(home, n_alpha=361, n_beta=181):
    """Noise-free versions of the synthetic debris tables (the smooth fits of
    rs_providers.c without the seeded noise) in the reference's file layouts
    (.adm planes over beta in [-180, 180] x alpha in [0, 180], .rcs planes over
    alpha in [-180, 180] x beta in [0, 180]; adm.c:152-225, rcs.c:181-246) under
    home/tables, where the engine's table search finds them."""
    adm_dir, rcs_dir = os.path.join(home, "tables", "adm"), os.path.join(home, "tables", "rcs")
    os.makedirs(adm_dir, exist_ok=True)
    os.makedirs(rcs_dir, exist_ok=True)
    b = np.linspace(-np.pi, np.pi, n_alpha)[None, :]        # x: body-frame beta
    a = np.linspace(0.0, np.pi, n_beta)[:, None]             # y: alpha
    C = 1.0 + 0.3 * np.cos(2.0 * b)
    adm = np.stack([C * np.cos(b) + 0 * a, C * np.sin(b) * np.cos(a), C * np.sin(b) * np.sin(a),
                    0.02 * np.sin(2.0 * b) * np.sin(a), -0.05 * np.sin(2.0 * b) * np.cos(a),
                    0.03 * np.sin(b) * np.sin(2.0 * a)]).astype(np.float32)
    for name in ("plate", "square_plate"):
        with open(os.path.join(adm_dir, name + ".adm"), "wb") as f:
            np.array([n_alpha, n_beta], np.uint16).tofile(f)
            adm.tofile(f)
    al = np.linspace(-np.pi, np.pi, n_alpha)[None, :]
    be = np.linspace(0.0, np.pi, n_beta)[:, None]
    for name, sc in (("leaf", 0.01), ("plate", 0.1), ("brick", 0.05)):
        hh, ph = sc * (0.6 + 0.4 * np.cos(be) ** 2), 0.5 * np.sin(al)
        vv, pv = sc * (0.5 + 0.3 * np.sin(be) ** 2 * np.cos(al) ** 2), -0.3 * np.cos(be)
        hv = 0.1 * sc * np.sin(2.0 * be) * np.sin(al)
        planes = np.stack([hh * np.cos(ph), vv * np.cos(pv), hv, hh * np.sin(ph), vv * np.sin(pv), 0.3 * hv])
        with open(os.path.join(rcs_dir, name + ".rcs"), "wb") as f:
            np.array([n_alpha, n_beta], np.uint16).tofile(f)
            planes.astype(np.float32).tofile(f)


def config(n_rain, n_debris):
    return wl.Config(n_rain=n_rain, debris=[(t, n_debris) for t in TYPES], scan="P:3.0,-1:1:0.01", name="C5s")


def _tumble(r, o=None):
    """Uniformly random orientations for every debris body (same draw each
    call), uploaded to the engine and, with o, given to the oracle as well."""
    n0 = int(r.info().counts[0])
    ori = r.stream("ori", copy=False)
    q = np.random.default_rng(31).normal(size=(ori.shape[0] - n0, 4))
    ori[n0:] = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
    r.upload()
    r.check()
    if o is not None:
        o.st.ori[n0:] = ori[n0:]


def _engine_run(cfg, pulses, tumbled, nets=None, fit=None):
    """The engine alone over the driver order of harness.run_pair; fit(r)
    returns the networks to install after RS_populate (and the tumble)."""
    r = sa.Radar(verb=0)
    wl.configure(r, cfg)
    r.populate()
    r.check()
    if tumbled:
        _tumble(r)
    extra = fit(r) if fit is not None else None
    if extra is not None:
        for k, m in zip((1, 2, 3), extra[0]):
            r.set_debris_surrogate(k, m)
    pg = []
    for k in range(pulses):
        if k > 0:
            r.make_pulse()
        r.download_pulse_only()
        r.check()
        pg.append(r.pulse_f64())
        r.advance_time()
        r.advance_beam()
    r.free()
    return np.array(pg), extra


def fit_networks(r, samples, epochs, state_frac=0.0):
    nets, errs = [], []
    for k in (1, 2, 3):
        m, _ = sg.fit(r, k, n_samples=samples, epochs=epochs, state_frac=state_frac)
        errs.append(sg.error(m, r, k))
        nets.append(m)
    return nets, errs


def run(n_rain=65536, n_debris=8192, pulses=48, tumbled=True, samples=10_000_000, epochs=16, state_frac=0.0,
        threads=1, smooth_home=None):
    """Pulse 0 (made by RS_populate before the tumble) is left out of the
    moments; pulses 1..pulses-1 enter them.  smooth_home: a folder prepared by
    write_smooth_tables, used as HOME (and working directory) so that the debris
    tables are the noise-free ones."""
    if smooth_home is not None:
        old_home, old_cwd = os.environ.get("HOME"), os.getcwd()
        os.environ["HOME"] = smooth_home
        os.chdir(smooth_home)
        try:
            return run(n_rain, n_debris, pulses, tumbled, samples, epochs, state_frac, threads, None)
        finally:
            if old_home is None:
                del os.environ["HOME"]
            else:
                os.environ["HOME"] = old_home
            os.chdir(old_cwd)
    cfg = config(n_rain, n_debris)
    print("study: table path and oracle, %d pulses" % pulses, flush=True)
    t0 = time.perf_counter()
    ra, o, ga, go = h.run_pair(cfg, pulses, setup=(lambda r, oo: _tumble(r, oo)) if tumbled else None,
                               oracle_kw={"threads": threads})
    t_oracle = time.perf_counter() - t0
    p = ra.params()
    gain = float(ra.lib.RS_radar_gain(ra.H))
    R = int(p.range_count)
    ra.free()
    print("study: %.1f s; surrogate path, training on %d samples per type" % (t_oracle, samples), flush=True)
    t0 = time.perf_counter()
    gb, (nets, errs) = _engine_run(cfg, pulses, tumbled, fit=lambda r: fit_networks(r, samples, epochs, state_frac))
    t_train = time.perf_counter() - t0
    for m in nets:
        m.free()
    M = pulses - 1
    mom = {n: orc.moments(x[1:].reshape(M, R, 4), p.lambda_, p.prt, gain, p.range_start, p.range_delta)
           for n, x in (("oracle", go), ("table", ga), ("surrogate", gb))}
    ref = mom["oracle"]
    live = np.isfinite(ref[:, 0]) & (ref[:, 0] > np.nanmax(ref[:, 0]) - 40.0)
    rows = {}
    for n in ("table", "surrogate"):
        d = np.abs(mom[n][live] - ref[live])
        iq = max(h.rel_err((ga if n == "table" else gb)[k], go[k]) for k in range(1, pulses))
        rows[n] = dict(iq=iq, z=d[:, 0].max(), z_mean=d[:, 0].mean(), v=d[:, 1].max(), zdr=d[:, 2].max(),
                       rho=d[:, 3].max())
    return rows, dict(live=int(live.sum()), R=R, heldout=errs, train_s=t_train, oracle_s=t_oracle)


def table(rows):
    out = ["| path | IQ error | max dZ (dB) | mean dZ (dB) | max dV (m/s) | max dZDR (dB) | max d rho_hv |",
           "|---|---:|---:|---:|---:|---:|---:|"]
    for n, e in rows.items():
        out.append("| %s | %.2e | %.3f | %.3f | %.4f | %.3f | %.2e |" % (n, e["iq"], e["z"], e["z_mean"], e["v"],
                                                                       e["zdr"], e["rho"]))
    return "\n".join(out)


if __name__ == "__main__":
    nr = int(sys.argv[1]) if len(sys.argv) > 1 else 65536
    nd = int(sys.argv[2]) if len(sys.argv) > 2 else 8192
    M = int(sys.argv[3]) if len(sys.argv) > 3 else 48
    th = max(1, min(16, len(os.sched_getaffinity(0))))
    import tempfile
    smooth = tempfile.mkdtemp()
    write_smooth_tables(smooth)
    for tables, home in (("synthetic (seeded noise)", None), ("smooth", smooth)):
        for tumbled in (True, False):
            for sf in ((0.0, 0.25) if home is None else (0.25,)):
                rows, info = run(nr, nd, M, tumbled=tumbled, state_frac=sf, threads=th, smooth_home=home)
                print("\n%s tables, %s population, state samples %.0f%%: %d gates, held-out (cd, cm, amp) %s, "
                      "training %.1f s" % (tables, "tumbled" if tumbled else "fresh", 100 * sf, info["live"],
                                           info["heldout"], info["train_s"]))
                print(table(rows), flush=True)
