"""Workload configurations of the engine (the run set-ups the benchmark and the
tests use), applied through the RS_* API the way simradar.c does
(simradar.c:812-955).

C1  131,072 rain, 15 m gates, Marshall-Palmer 10 mm/h, suctvort LES, 100 pulses
C2  1,997,824 rain + 1,024 leaf debris (1,998,848), 1,000 pulses
C3  8,388,608 rain, 3 km x 3 km x 1 km box at 10 km range (precision study)
C5  2,097,152 rain + 262,144 each of leaf, metal sheet (plate) and brick
"""
import numpy as np

from . import (CONCEPT_BOUNDED, CONCEPT_DRAGGED, CONCEPT_UNIFORM_DSD, OBJ_BRICK, OBJ_LEAF, OBJ_METAL_SHEET,
               PHASE_F64, Radar)

SEED = 19760520
DBU = CONCEPT_DRAGGED | CONCEPT_BOUNDED | CONCEPT_UNIFORM_DSD


def mp_dsd(rain_rate_mm_h=10.0, d_mm=None, n0=8000.0):
    """Marshall-Palmer N(D) = N0 exp(-Lambda D), Lambda = 4.1 R^-0.21 mm^-1, on
    diameter bins D; returns (N per bin as float32, D in metres as float32)."""
    if d_mm is None:
        d_mm = np.arange(0.5, 5.0001, 0.5)
    lam = 4.1 * rain_rate_mm_h ** -0.21
    nd = n0 * np.exp(-lam * np.asarray(d_mm, np.float64))
    return nd.astype(np.float32), (np.asarray(d_mm, np.float64) * 1e-3).astype(np.float32)


class Config:
    def __init__(self, n_rain=131072, debris=(), concept=DBU, gate=15.0, scan="P:3.0,-12:12:0.01", seed=SEED,
                 prt=1e-3, phase=PHASE_F64, rain_rate=10.0, box=None, name="custom", range_cap=None,
                 les="suctvort", les_interp=False):
        self.n_rain = n_rain
        self.debris = list(debris)
        self.concept = concept
        self.gate = gate
        self.scan = scan
        self.seed = seed
        self.prt = prt
        self.phase = phase
        self.rain_rate = rain_rate
        self.box = box          # None: RS_suggest_scan_domain; else 8 floats (origin r,a,e,w, size r,a,e,w)
        self.range_cap = range_cap  # None: the reference's 10 km cap on the padded extent
        self.les = les          # LES provider config ("suctvort" or "suctvort:cx,cy,seed")
        self.les_interp = les_interp  # LES volumes interpolated in time (else switched every tp)
        self.name = name

    @property
    def n_total(self):
        return self.n_rain + sum(c for _, c in self.debris)

    def bytes_per_pulse(self, hilo=False):
        """Algorithmic HBM bytes of one pulse: read + write of pos and vel per rain
        scatterer (64 B), of pos, vel, ori and tum per debris scatterer (128 B),
        +32 B per scatterer when split hi/lo positions are carried."""
        nd = sum(c for _, c in self.debris)
        b = 64 * self.n_rain + 128 * nd
        if hilo:
            b += 32 * self.n_total
        return b


def c1(scan="P:3.0,-12:12:0.01"):
    return Config(n_rain=131072, scan=scan, name="C1")


def c2(scan="P:3.0,-12:12:0.01"):
    return Config(n_rain=1997824, debris=[(OBJ_LEAF, 1024)], scan=scan, name="C2")


C3_BOX = (8500.0, -8.6, 0.0, 0.0, 3000.0, 17.2, 5.7, 0.0)   # 3 km x 3 km x 1 km at 10 km


def c3(n_rain=8388608, scan="P:3.0,-1:1:0.01"):
    """Precision study: rain in a 3 km x 3 km x 1 km box at 10 km range (202
    gates of 15 m); needs the 10 km extent cap lifted."""
    return Config(n_rain=n_rain, scan=scan, box=C3_BOX, range_cap=20000.0, name="C3")


def c4(n_domains=64, seed=4242, scale=1.0):
    """Batching: independent domains with rain counts log-uniform in
    [2^16, 2^21] (times `scale`, for reduced test sizes), rain rates log-uniform
    in [1, 100] mm/h and vortex centres moved up to 200 m from the grid centre,
    each with its own turbulence seed (seeded, reproducible)."""
    g = np.random.default_rng(seed)
    out = []
    for d in range(n_domains):
        n = int(round(2.0 ** g.uniform(16.0, 21.0) * scale))
        n = max(256, n)
        rate = float(10.0 ** g.uniform(0.0, 2.0))
        cx, cy = g.uniform(-200.0, 200.0, 2)
        out.append(Config(n_rain=n, rain_rate=rate, seed=SEED + d, name="C4-%02d" % d,
                          les="suctvort:%.1f,%.1f,%d" % (cx, cy, 1000 + d)))
    return out


def c5(scan="P:3.0,-12:12:0.01"):
    return Config(n_rain=2097152, debris=[(OBJ_LEAF, 262144), (OBJ_METAL_SHEET, 262144), (OBJ_BRICK, 262144)],
                  scan=scan, name="C5")


def configure(r, cfg):
    """Apply a configuration to a Radar before RS_populate."""
    r.set_concept(cfg.concept)
    r.set_prt(cfg.prt)
    r.set_lambda(0.1)
    r.set_antenna_params(1.0, 44.5)
    r.set_tx_params(0.2e-6, 50.0e3)
    r.set_sampling_spacing(cfg.gate, 1.0, 1.0)
    nd, d = mp_dsd(cfg.rain_rate)
    r.set_dsd(nd, d, 1)
    r.set_random_seed(cfg.seed)
    r.set_scan_pattern(cfg.scan)
    r.set_vel_data_to_config(cfg.les)
    if cfg.les_interp:
        r.set_les_time_interpolation(1)
    for kind, count in cfg.debris:
        r.add_debris(kind, count)
    if cfg.range_cap:
        r.lib.RS_set_range_cap(r.H, float(cfg.range_cap))
    box = r.suggest_scan_domain() if cfg.box is None else np.asarray(cfg.box, np.float32)
    r.set_scan_box(box)
    r.set_population(cfg.n_rain)
    r.set_phase_precision(cfg.phase)
    return r


def make(cfg, verb=0):
    """Configured and populated Radar for cfg."""
    r = Radar(verb=verb)
    configure(r, cfg)
    r.populate()
    r.check()
    return r
