"""Benchmark of the per-pulse scatterer engine on one MI355X.

Workload (the configuration the headline metric is quoted on): 1,997,824 rain
scatterers + 1,024 leaf debris (1,998,848 in total), Marshall-Palmer 10 mm/h DSD,
suctvort-shaped synthetic LES 176x176x80 with its 10 volumes interpolated in
time (two volumes resident, the wind blended at the pulse's time), concept
D|B|U, 15 m gates (40 gates), 1 deg beam, 0.2 us pulse, PRT 1 ms, lambda 0.1 m,
PPI 3 deg -12..12 in 0.01 deg.  The same workload with the reference's
piecewise-constant volume switch (one volume read per step) is timed beside it
(--no-alt skips it; --les-step makes it the headline).
A step is one pulse: the advance of every scatterer by one PRT and the per-gate
complex H/V sums (RS_advance_time + RS_advance_beam + RS_make_pulse), with the
state resident in HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 50] [--no-cpu]

Other workloads (one JSON line each, not the headline): --workload c3 (8,388,608
rain in a 3 km x 3 km x 1 km box at 10 km range, hi/lo positions, 202 gates),
--workload c4 (64 independent domains of 2^16..2^21 rain, one persistent launch
per round with HIP-graph replay, against the same domains launched one by one),
--workload c5 (2,097,152 rain + 3 x 262,144 debris: the ADM / RCS table path
against the trained surrogate networks, training time reported separately).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PROFILE_STRIDE = 10                 # rain launches per HIP-event timed launch
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_rain_c2_traffic.json")
TRAFFIC_FILE_C3 = os.path.join(ROOT, "profiles", "k_rain_c3_traffic.json")
TRAIN_SAMPLES = 10_000_000          # table-path samples per debris type for the C5 networks
REFERENCE_PULSES_PER_S = 222.3   # README.md:58 (instantaneous, 1,998,848 scatterers, GPU not stated)


def cpu_baseline(r, cfg, pulses, threads, warmup=5, runs=3):
    """The oracle (CPU restatement of the reference kernels: the same kernel
    sequence, dense pass 1 over all gates) timed on a bounded sample of the same
    workload, each population split over `threads` host threads: `warmup`
    untimed pulses, then `runs` runs of `pulses` pulses; the median rate.  With
    the LES volumes interpolated in time the oracle reads both volumes too."""
    from oracle import oracle as orc
    from oracle import run as orun
    st = orun.state_from_radar(r)
    o = orun.OracleRun(r, cfg, st, threads=threads)
    sd = r.sim_desc()
    if cfg.les_interp:
        inf = r.info()
        vn = int(inf.vel_idx)
        o.set_les_frame(r.les_frame(vn - 2))
        nxt = orc.tex(r.les_frame(vn - 1))
        orc.set_blend(nxt[0], None, float(inf.wind_alpha))

    def one():
        o.advance(sd)
        o.refresh_debris_rcs(sd)
        o.pulse(sd)
    for _ in range(warmup):
        one()
    rates = []
    for _ in range(runs):
        t0 = time.perf_counter()
        for _ in range(pulses):
            one()
        rates.append(pulses / (time.perf_counter() - t0))
    if cfg.les_interp:
        orc.set_blend(None)
    return sorted(rates)[len(rates) // 2], rates


def host_cores():
    """Logical CPUs of the host and the CPUs this process may run on."""
    return os.cpu_count(), len(os.sched_getaffinity(0))


def timed_pulses(requested, interval):
    """Pulses in the timed window: the request rounded up to whole range-reorder
    intervals (a reorder runs every `interval` pulses), and at least two of
    them, so the reorder's cost lands in the per-pulse time whatever the
    window's phase."""
    if interval <= 0:
        return max(1, requested)
    n = max(requested, 2 * interval)
    return (n + interval - 1) // interval * interval


def time_c2(r, warmup, steps, e2e_steps):
    """One step = one PRT: RS_advance_time + RS_advance_beam + RS_make_pulse.
    The advance is deferred into the pulse's launch, so every step is one fused
    advance + pulse launch and nothing is left pending between steps.  Returns
    (event ms over the timed steps, wall s, sampled kernel ms, samples,
    end-to-end pulses/s with the per-pulse IQ download or None)."""
    def step():
        r.advance_time()
        r.advance_beam()
        r.make_pulse()

    for _ in range(warmup):
        step()
    r.sync()
    r.check()
    # HIP events around every PROFILE_STRIDE-th rain launch of the timed loop
    # (events on every launch add a few us of queue packets per pulse)
    r.profile(True, stride=PROFILE_STRIDE)
    print("bench: timing %d pulses (whole range-reorder intervals)" % steps, file=sys.stderr, flush=True)
    r.sync()
    t0 = time.perf_counter()
    r.timer_start()
    for _ in range(steps):
        step()
    ev_ms = r.timer_stop()
    r.sync()
    wall = time.perf_counter() - t0
    r.check()
    k_ms, k_n = r.profile_read(3)
    r.profile(False)
    # End to end: the same pulse loop with the blocking per-pulse IQ download
    # that the reference's driver performs when writing IQ (simradar.c:1057-1059).
    e2e = None
    if e2e_steps > 0:
        r.sync()
        t1 = time.perf_counter()
        for _ in range(e2e_steps):
            step()
            r.download_pulse_only()
        e2e = e2e_steps / (time.perf_counter() - t1)
        r.check()
    return ev_ms, wall, k_ms, k_n, e2e


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-pulses", type=int, default=24)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads of the CPU baseline (0: the process's CPUs, at most 16)")
    ap.add_argument("--e2e-steps", type=int, default=200)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--max-blocks", type=int, default=0)
    ap.add_argument("--sort", type=int, default=-1,
                    help="pulses between range reorders of the rain streams (-1: the engine default)")
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c4", "c5"])
    ap.add_argument("--les-step", action="store_true",
                    help="C2 headline with the reference's piecewise-constant LES volume switch instead of the "
                         "time interpolation")
    ap.add_argument("--no-alt", action="store_true", help="C2: skip timing the other LES volume mode")
    ap.add_argument("--c5-serial-train", action="store_true",
                    help="C5: train the three networks one after another (rocprofv3's tracer has crashed on the "
                         "concurrent training threads)")
    ap.add_argument("--c5-samples", type=int, default=TRAIN_SAMPLES,
                    help="training samples per debris type of the C5 networks (fewer: shorter profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build drives a single MI355X per process "
              "(multi-GPU sharding is a design note only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "c3":
        return bench_c3(args)
    if args.workload == "c4":
        return bench_c4(args)
    if args.workload == "c5":
        return bench_c5(args)

    import simradar_amd as sa
    from simradar_amd import workloads as wl

    cfg = wl.c2()
    cfg.les_interp = not args.les_step
    print("bench: populating %s (%d scatterers)" % (cfg.name, cfg.n_total), file=sys.stderr, flush=True)
    r = wl.make(cfg)
    if args.block or args.max_blocks:
        r.set_launch(args.block, args.max_blocks)
    if args.sort >= 0:
        r.set_sort(args.sort)
    R = int(r.params().range_count)

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        ncpu, naff = host_cores()
        print("bench: cpu baseline (5 + 3 x %d pulses of the oracle on %d threads; host has %d CPUs, %d usable)"
              % (args.cpu_pulses, threads, ncpu, naff), file=sys.stderr, flush=True)
        cpu_pps, cpu_runs = cpu_baseline(r, cfg, args.cpu_pulses, threads)
        cpu = {"value": cpu_pps, "unit": "pulses/s", "cores": threads, "kind": "port",
               "host_cpus": ncpu, "usable_cpus": naff, "runs": cpu_runs,
               "sample": "median of 3 runs of %d pulses (after 5 warm-up pulses) of the full %d-scatterer "
                         "workload, oracle C restatement (dense pass 1 over all gates), scatterers split over "
                         "%d host threads" % (args.cpu_pulses, cfg.n_total, threads)}

    steps = timed_pulses(args.steps, r.sort_interval())
    ev_ms, wall, k_ms, k_n, e2e = time_c2(r, args.warmup, steps, args.e2e_steps)
    alt = None
    if not args.no_alt:
        r.free()
        acfg = wl.c2()
        acfg.les_interp = not cfg.les_interp
        print("bench: the %s mode" % ("time-interpolated" if acfg.les_interp else "volume-switch"),
              file=sys.stderr, flush=True)
        r = wl.make(acfg)
        if args.sort >= 0:
            r.set_sort(args.sort)
        a_ms, _, ak_ms, ak_n, _ = time_c2(r, args.warmup, steps, 0)
        alt = {"les_mode": "interpolated" if acfg.les_interp else "volume switch (reference)",
               "pulses_per_s": 1000.0 * steps / a_ms, "kernel_avg_ms": ak_ms / max(1, ak_n),
               "roofline_frac": cfg.bytes_per_pulse() / (ak_ms / max(1, ak_n) * 1e-3) / 1e9 / PEAK_HBM_GBS}
    ms_per_step = ev_ms / steps
    pps = 1000.0 / ms_per_step
    n = cfg.n_total
    kern_ms = k_ms / max(1, k_n)
    # the launch carries the rain and the 1,024 leaves (64 B / 128 B each)
    launch_bytes = cfg.bytes_per_pulse()
    achieved = launch_bytes / (kern_ms * 1e-3) / 1e9
    traffic = None
    traffic_src = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        traffic = t["fetch_bytes"] * t["fetch_factor"] + t["write_bytes"] * t["write_factor"]
        traffic_src = "profiles/k_rain_c2_traffic.json (rocprofv3 FETCH_SIZE / WRITE_SIZE per dispatch, " \
                      "scaled by the k_io copy calibration)"
    line = {
        "metric": "pulses/s & scatterer-updates/s @2M; surrogate vs table at <=0.5 dB Z err",
        "value": pps,
        "unit": "pulses/s",
        "n_gpus": 1,
        "steps": steps,
        "steps_requested": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "wall_ms_per_step": 1000.0 * wall / steps,
        "e2e_pulses_per_s_with_iq_download": e2e,
        "scatterer_updates_per_s": pps * n,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": pps / REFERENCE_PULSES_PER_S,
        "vs_baseline_ref": "README.md:58 222.3 pulses/s (1,998,848 rain scatterers, 30 m gates, GPU not stated)",
        "dtype": "f32 state, f64 phase and gate sums",
        "data": "synthetic (seeded LES / DSD / ADM / RCS tables)",
        "config": {"workload": "C2: 1,997,824 rain + 1,024 leaf debris, %d gates x 15 m, PPI 3 deg -12:12:0.01, "
                               "suctvort LES 176x176x80 (10 volumes), concept DBU" % R,
                   "les_mode": "interpolated in time" if cfg.les_interp else "volume switch every tp (reference)",
                   "scatterers": n, "range_gates": R, "algorithmic_bytes_per_pulse": cfg.bytes_per_pulse()},
        "other_les_mode": alt,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "k_rain<advance,pulse> (rain + leaf debris blocks)", "kernel_avg_ms": kern_ms,
                     "kernel_launches_timed": k_n, "bytes_per_launch": launch_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    r.free()


def _line(metric, value, unit, steps, warmup, ms, dtype, config, extra):
    d = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
         "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
         "data": "synthetic (seeded LES / DSD tables)", "config": config}
    d.update(extra)
    print(json.dumps(d))


def bench_c3(args):
    """8,388,608 rain at 10 km range with hi/lo positions and binary64 phase."""
    import simradar_amd as sa
    from simradar_amd import workloads as wl
    cfg = wl.c3()
    print("bench: populating C3 (%d rain)" % cfg.n_rain, file=sys.stderr, flush=True)
    r = wl.make(cfg)
    r.set_position_precision(1)
    R = int(r.params().range_count)
    steps = timed_pulses(min(args.steps, 2000) if args.steps != 1000 else 2000, r.sort_interval())
    warm = min(args.warmup, 50)

    def step():
        r.advance_time()
        r.advance_beam()
        r.make_pulse()
    for _ in range(warm):
        step()
    r.sync()
    r.profile(True, stride=PROFILE_STRIDE)
    r.timer_start()
    for _ in range(steps):
        step()
    ms = r.timer_stop() / steps
    k_ms, k_n = r.profile_read(3)
    r.profile(False)
    r.check()
    kern = k_ms / max(1, k_n)
    byts = 96 * cfg.n_rain            # pos + vel + lo read and written
    ach = byts / (kern * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE_C3):
        with open(TRAFFIC_FILE_C3) as f:
            t = json.load(f)
        traffic = t["fetch_bytes"] * t["fetch_factor"] + t["write_bytes"] * t["write_factor"]
    _line("pulses/s & scatterer-updates/s (C3, hi/lo positions)", 1000.0 / ms, "pulses/s", steps, warm, ms,
          "f32 state, f32 hi/lo positions, f64 range-to-phase and gate sums",
          {"workload": "C3: 8,388,608 rain, 3 km x 3 km x 1 km box at 10 km range, %d gates x 15 m" % R,
           "scatterers": cfg.n_rain, "range_gates": R},
          {"scatterer_updates_per_s": 1000.0 / ms * cfg.n_rain,
           "roofline": {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": ach / PEAK_HBM_GBS, "traffic": traffic,
                        "traffic_source": "profiles/k_rain_c3_traffic.json" if traffic else None,
                        "kernel": "k_rain<advance,pulse,hilo>",
                        "kernel_avg_ms": kern, "bytes_per_launch": byts}})
    r.free()


def bench_c4(args):
    """64 domains: one persistent launch per round (HIP graph) against the same
    domains launched one by one, both timed over the same rounds (wall clock
    around the rounds plus a final synchronisation)."""
    import simradar_amd as sa
    from simradar_amd import workloads as wl
    cfgs = wl.c4(64)
    total = sum(c.n_rain for c in cfgs)
    print("bench: populating C4 (64 domains, %d rain)" % total, file=sys.stderr, flush=True)
    rs = [wl.make(c) for c in cfgs]
    rounds = 200 if args.steps == 1000 else args.steps

    def plain_round():
        for r in rs:
            r.make_pulse()
        for r in rs:
            r.advance_time()
            r.advance_beam()

    def sync_all():
        for r in rs:
            r.sync()
    for _ in range(10):
        plain_round()
    sync_all()
    t0 = time.perf_counter()
    for _ in range(rounds):
        plain_round()
    sync_all()
    plain_s = (time.perf_counter() - t0) / rounds
    b = sa.Batch(rs)
    for _ in range(40):        # the autotuner's sweep (9 shapes x 3 rounds) and warm-up
        b.round()
    b.sync()
    shape = b.launch_shape()
    t0 = time.perf_counter()
    for _ in range(rounds):
        b.round()
    b.sync()
    batch_s = (time.perf_counter() - t0) / rounds
    b.free()
    for r in rs:
        r.free()
    _line("pulses/s & scatterer-updates/s (C4, 64 batched domains)", 64.0 / batch_s, "pulses/s", rounds, 50,
          batch_s * 1000.0, "f32 state, f64 phase and gate sums",
          {"workload": "C4: 64 domains, rain log-uniform 2^16..2^21 (%d in total), rain rates 1..100 mm/h, "
                       "distinct vortex centres; one persistent launch + HIP graph per round" % total,
           "domains": 64, "scatterers": total, "chunk": shape[0], "blocks_per_cu": shape[1]},
          {"scatterer_updates_per_s": total / batch_s,
           "plain_pulses_per_s": 64.0 / plain_s, "plain_ms_per_round": plain_s * 1000.0,
           "batched_over_plain": plain_s / batch_s})


def bench_c5(args):
    """C5: table lookups against surrogate networks for the three debris types."""
    import simradar_amd as sa
    from simradar_amd import surrogate as sg
    from simradar_amd import workloads as wl
    cfg = wl.c5()
    print("bench: populating C5 (%d scatterers)" % cfg.n_total, file=sys.stderr, flush=True)
    r = wl.make(cfg)
    steps = timed_pulses(500 if args.steps == 1000 else args.steps, r.sort_interval())

    def step():
        r.advance_time()
        r.advance_beam()
        r.make_pulse()

    def timed():
        for _ in range(20):
            step()
        r.sync()
        r.timer_start()
        for _ in range(steps):
            step()
        return r.timer_stop() / steps
    ms_table = timed()
    t0 = time.perf_counter()
    kw = dict(n_samples=args.c5_samples, epochs=16, state_frac=0.25)
    if args.c5_serial_train:
        nets = [sg.fit(r, k, **kw)[0] for k in (1, 2, 3)]
    else:
        nets = [m for m, _ in sg.fit_many(r, (1, 2, 3), **kw)]
    train_s = time.perf_counter() - t0
    errs = [sg.error(m, r, k) for k, m in zip((1, 2, 3), nets)]
    for k, m in zip((1, 2, 3), nets):
        r.set_debris_surrogate(k, m)
    ms_surr = timed()
    _line("pulses/s & scatterer-updates/s (C5, debris surrogate)", 1000.0 / ms_surr, "pulses/s", steps, 20,
          ms_surr, "f32 physics, bf16 network (fp32 accumulate), f64 phase and gate sums",
          {"workload": "C5: 2,097,152 rain + leaf, plate, brick at 262,144 each; surrogate networks 7-4x64-12",
           "scatterers": cfg.n_total},
          {"scatterer_updates_per_s": 1000.0 / ms_surr * cfg.n_total,
           "table_pulses_per_s": 1000.0 / ms_table, "surrogate_over_table": ms_table / ms_surr,
           "training_s": train_s, "training_samples_per_type": args.c5_samples,
           "training_set": "uniform orientations and flow directions, a quarter at the population's states",
           "heldout_rel_err_cd_cm_amp": errs})
    for m in nets:
        m.free()
    r.free()


if __name__ == "__main__":
    main()
