"""Oracle stepping of an engine run (TEST INFRASTRUCTURE): builds the oracle's
own tables (drop RCS, antenna pattern, range weight, LES descriptor) for the
configuration of a Radar and advances / pulses the reference-layout streams.
Used by tests/ and by bench.py's cpu_baseline leg only."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import oracle as orc

CONCEPT_DRAGGED = 1 << 1
CONCEPT_TRANSPARENT = 1 << 2
CONCEPT_UNIFORM_DSD = 1 << 4
CONCEPT_FIXED = 1 << 5


def adm_desc(nb, na, phys):
    """rs.c:3778-3783 + 3746-3758 restated: x = beta over [-pi, pi] (nb), y = alpha
    over [0, pi] (na), +0.5 offsets; slots 12-15 = 1/(In Ln) xyz and Ta."""
    d = np.zeros(16, np.float32)
    xs = np.float32(np.float32(nb - 1) / (2.0 * np.pi))
    d[0] = xs
    d[1] = np.float32(np.float32(na - 1) / np.pi)
    d[4] = np.float32(np.pi * np.float64(xs) + 0.5)
    d[5] = 0.5
    d[8] = nb - 1
    d[9] = na - 1
    d[12:16] = phys
    return d


def rcs_desc(na, nb):
    """rs.c:3987-3992 + 3960-3968 restated: x = alpha over [-pi, pi] (na), y = beta
    over [0, pi] (nb), no +0.5."""
    d = np.zeros(16, np.float32)
    xs = np.float32(np.float32(na - 1) / (2.0 * np.pi))
    d[0] = xs
    d[1] = np.float32(np.float32(nb - 1) / np.pi)
    d[4] = np.float32(np.pi * np.float64(xs))
    d[5] = 0.0
    d[8] = na - 1
    d[9] = nb - 1
    return d


class OracleRun:
    """CPU oracle stepping the same run from a given state (driver order of
    simradar.c:1011-1074: make_pulse, download, advance_time, advance_beam).

    threads > 1 splits every population into that many contiguous ranges run
    on concurrent host threads (the C calls release the interpreter lock): the
    advance is unchanged bit for bit, the pulse sums the per-range gate sums in
    range order (binary64 reassociation only).  Tests use one thread."""

    def __init__(self, r, cfg, state, threads=1, wind=None, rangew=None, angw=None):
        """wind = (float32 [nz, ny, nx, 4] texels, les_desc16) replaces the LES
        frame 0 / stretched-grid descriptor; rangew / angw = (table, desc4)
        replace the triangle range weight and the standard antenna pattern."""
        self.threads = max(1, int(threads))
        self._pool = ThreadPoolExecutor(self.threads) if self.threads > 1 else None
        self.r = r
        self.cfg = cfg
        self.st = state
        info = r.info()
        self.counts = [int(c) for c in info.counts]
        self.origins = [int(o) for o in info.origins]
        p = r.params()
        self.params = p
        sd = r.sim_desc()
        self.fixed = bool(cfg.concept & CONCEPT_FIXED) and not cfg.concept & CONCEPT_DRAGGED
        self.cpxx_tex = None
        if wind is None:
            self.set_les_frame(r.les_frame(0), r.les_cpxx(0) if self.fixed else None)
        else:
            self.wind_tex, self._wind_keep = orc.tex(wind[0])
            self.les_desc = np.asarray(wind[1], np.float32)
        dsd_r, dsd_pdf, _ = r.dsd()
        uniform = bool(cfg.concept & CONCEPT_UNIFORM_DSD)
        self.drop, self.drop_desc = orc.drop_rcs_table(sd[4], sd[6], bool(cfg.concept & CONCEPT_TRANSPARENT),
                                                       uniform, dsd_r, dsd_pdf)
        if angw is None:
            self.angw, self.angw_desc = orc.angular_weight_standard(p.antenna_bw_rad)
        else:
            self.angw, self.angw_desc = np.asarray(angw[0], np.float32), np.asarray(angw[1], np.float32)
        if rangew is None:
            self.rangew = np.array([0.0, 1.0, 0.0], np.float32)
            self.rangew_desc = orc.table1d_desc(-p.dr, p.dr, 3)
        else:
            self.rangew, self.rangew_desc = np.asarray(rangew[0], np.float32), np.asarray(rangew[1], np.float32)
        self.kd = 4.0 * np.pi / float(np.float32(p.lambda_))
        self.deb = []
        for t in range(len(cfg.debris)):
            cd, cm, re, im = (r.debris_table(t, w) for w in range(4))
            ad = r.descriptor(16 + t)
            rd = r.descriptor(32 + t)
            tx = [orc.tex(a) for a in (cd, cm, re, im)]
            self.deb.append(dict(tex=tx, adm_desc=ad, rcs_desc=rd))

    def set_les_frame(self, frame, cpxx=None):
        """Wind of later advances: one suctvort-geometry LES frame (stretched
        grid, les.c:164-170), as the host uploads at a volume switch; cpxx its
        (cn2, p, _, _) volume for the fixed-position background."""
        if cpxx is not None:
            self.cpxx_tex, self._cpxx_keep = orc.tex(cpxx)
        self.wind_tex, self._wind_keep = orc.tex(frame)
        self.les_desc = orc.les_desc(True, frame.shape[2], frame.shape[1], frame.shape[0], 2.0, 2.0, 2.0,
                                     1.0212, 1.0212, 1.05, 50.0)

    def _ranges(self, b, e):
        n = self.threads
        cuts = [b + (e - b) * j // n for j in range(n + 1)]
        return [(cuts[j], cuts[j + 1]) for j in range(n) if cuts[j + 1] > cuts[j]]

    def _run(self, fn, b, e):
        rs = self._ranges(b, e)
        if self._pool is None or len(rs) < 2:
            return [fn(lo, hi) for lo, hi in rs]
        return list(self._pool.map(lambda t: fn(*t), rs))

    def advance(self, sim_desc):
        n0 = self.counts[0]
        if self.fixed:      # rs.c:5343-5347: fp_atts under F without D
            self._run(lambda lo, hi: orc.fp_atts(self.st, lo, hi, self.wind_tex, self.cpxx_tex, self.les_desc,
                                                 sim_desc), 0, n0)
        else:
            atts = orc.el_atts if self.cfg.concept & CONCEPT_DRAGGED else orc.bg_atts
            self._run(lambda lo, hi: atts(self.st, lo, hi, self.wind_tex, self.les_desc, self.drop, self.drop_desc,
                                          sim_desc), 0, n0)
        for t, d in enumerate(self.deb):
            k = t + 1
            b, e = self.origins[k], self.origins[k] + self.counts[k]
            (cd, _), (cm, _), (re, _), (im, _) = d["tex"]
            self._run(lambda lo, hi: orc.db_atts(self.st, lo, hi, self.wind_tex, self.les_desc, cd, cm,
                                                 d["adm_desc"], re, im, d["rcs_desc"], sim_desc), b, e)

    def refresh_debris_rcs(self, sim_desc):
        for t, d in enumerate(self.deb):
            k = t + 1
            b, e = self.origins[k], self.origins[k] + self.counts[k]
            _, _, (re, _), (im, _) = d["tex"]
            self._run(lambda lo, hi: orc.db_rcs(self.st, lo, hi, re, im, d["rcs_desc"], sim_desc), b, e)

    def pulse(self, sim_desc, mode=orc.PHASE_F64):
        p = self.params
        pos, rcs = self.st.pos, self.st.rcs
        parts = self._run(lambda lo, hi: orc.make_pulse(pos[lo:hi], rcs[lo:hi], self.angw, self.angw_desc,
                                                        self.rangew, self.rangew_desc, p.range_start,
                                                        p.range_delta, int(p.range_count), sim_desc, self.kd,
                                                        mode), 0, pos.shape[0])
        out = parts[0]
        for q in parts[1:]:
            out = out + q
        return out


def state_from_radar(r):
    return orc.State(r.stream("pos"), r.stream("vel"), r.stream("ori"), r.stream("tum"), r.stream("rcs"),
                     r.stream("rnd"), r.stream("aux"))
