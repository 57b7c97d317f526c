"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg import
this module.  Arrays are numpy float32 (n, 4) streams / uint32 (n, 4) seeds in
the reference's layout; descriptors are float32 arrays of 16 or 4 slots.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

PHASE_F32, PHASE_F64, FULL_F64 = 0, 1, 2
MATH = {"log": 1, "sin": 2, "cos": 3, "atan2": 4, "acos": 5, "pow": 6, "exp": 7}

_lib = None


class Tex(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int)]


class F4(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float), ("w", ctypes.c_float)]

    def np(self):
        return np.array([self.x, self.y, self.z, self.w], np.float32)


def _f4(v):
    v = np.asarray(v, np.float32).ravel()
    return F4(*[float(x) for x in v[:4]])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("oracle not built: %s (make -C oracle)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        P, F, D, I, S = ctypes.c_void_p, ctypes.c_float, ctypes.c_double, ctypes.c_int, ctypes.c_size_t
        T = ctypes.POINTER(Tex)
        sig = {
            "orc_rand4": (None, [P, P]),
            "orc_el_atts": (None, [P, P, P, P, S, S, T, P, P, P, P]),
            "orc_bg_atts": (None, [P, P, P, P, S, S, T, P, P, P, P]),
            "orc_db_atts": (None, [P, P, P, P, P, P, S, S, T, P, T, T, P, T, T, P, P]),
            "orc_fp_atts": (None, [P, P, P, S, S, T, T, P, P]),
            "orc_set_blend": (None, [T, T, F]),
            "orc_set_dff": (None, [P, ctypes.c_int, P]),
            "orc_dff_place_probe": (ctypes.c_int, [F4, P, P]),
            "orc_dff_icdf_from_pdf": (ctypes.c_int, [P, ctypes.c_int, P, ctypes.c_int]),
            "orc_db_rcs": (None, [P, P, P, S, S, T, T, P, P]),
            "orc_scat_sig_aux": (None, [P, P, P, P, S, P, P, P]),
            "orc_gate_ranges": (None, [F, F, I, P]),
            "orc_make_pulse": (None, [P, P, P, P, S, P, P, P, P, F, F, I, P, D, I]),
            "orc_scatterer_signal": (None, [P, P, P, P, P, D, I, P, P, P]),
            "orc_read_table": (F, [P, F, F]),
            "orc_populate": (None, [P, P, P, P, P, P, P, P, I, P, ctypes.c_uint, I, P, P, I]),
            "orc_drop_rcs_table": (None, [P, P, F, F, I, I, I, P, P]),
            "orc_angular_weight_standard": (None, [P, P, F]),
            "orc_table1d_desc": (None, [P, F, F, I]),
            "orc_radar_gain": (F, [F, F]),
            "orc_el_atts_f64": (None, [P, P, P, S, T, P, P, I]),
            "orc_les_desc": (None, [P, I, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, F, F, F, F, F, F, F]),
            "orc_math": (None, [I, P, P, P, S]),
            "orc_moments": (None, [P, I, I, D, D, D, D, D, P]),
            "orc_wind_index": (F4, [F4, P, P]),
            "orc_wind_index_table": (None, [P, P, P, P, P]),
            "orc_sample3d": (F4, [T, F, F, F]),
            "orc_sample2d": (F4, [T, F, F]),
            "orc_ellipsoid_rcs": (F4, [F4, P, P]),
            "orc_debris_rcs": (F4, [F4, F4, T, T, P, P]),
            "orc_qmul": (F4, [F4, F4]),
            "orc_quat2abg": (None, [F4, P, P, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            if args:
                fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def f32(a, shape=None):
    a = np.ascontiguousarray(a, np.float32)
    return a if shape is None else a.reshape(shape)


def tex(data):
    """Texture from a (nz, ny, nx, 4) or (ny, nx, 4) float32 array (kept alive by caller)."""
    d = np.ascontiguousarray(data, np.float32)
    if d.ndim == 4:
        return Tex(d.ctypes.data_as(ctypes.c_void_p), d.shape[2], d.shape[1], d.shape[0]), d
    return Tex(d.ctypes.data_as(ctypes.c_void_p), d.shape[1], d.shape[0], 1), d


def math(fn, a, b=None):
    a = f32(a).ravel()
    b = a if b is None else f32(b).ravel()
    out = np.empty_like(a)
    lib().orc_math(MATH[fn], _p(a), _p(b), _p(out), a.size)
    return out


def moments(iq, lam, prt, gain, range_start, range_delta):
    """Pulse-pair moments of iq[M, R, 4] (Ih, Qh, Iv, Qv): (R, 4) = (Z dB, V m/s,
    ZDR dB, rho_hv), binary64."""
    q = np.ascontiguousarray(iq, np.float64)
    M, R = q.shape[0], q.shape[1]
    out = np.zeros((R, 4), np.float64)
    lib().orc_moments(_p(q), M, R, lam, prt, gain, range_start, range_delta, _p(out))
    return out


def wind_index(pos, les_desc, sim_desc):
    return lib().orc_wind_index(_f4(pos), _p(f32(les_desc)), _p(f32(sim_desc))).np()


def wind_index_table(les_desc, sim_desc):
    """(rows[768, 4], ih[3], jm[3]) of the stretched-grid index."""
    rows = np.zeros((768, 4), np.float32)
    ih = np.zeros(3, np.float32)
    jm = np.zeros(3, np.float32)
    lib().orc_wind_index_table(_p(f32(les_desc)), _p(f32(sim_desc)), _p(rows), _p(ih), _p(jm))
    return rows, ih, jm


def sample3d(t, u, v, w):
    return lib().orc_sample3d(ctypes.byref(t), u, v, w).np()


def sample2d(t, u, v):
    return lib().orc_sample2d(ctypes.byref(t), u, v).np()


def ellipsoid_rcs(pos, table, desc):
    return lib().orc_ellipsoid_rcs(_f4(pos), _p(f32(table)), _p(f32(desc))).np()


def debris_rcs(pos, ori, re_tex, im_tex, rcs_desc, sim_desc):
    return lib().orc_debris_rcs(_f4(pos), _f4(ori), ctypes.byref(re_tex), ctypes.byref(im_tex),
                                _p(f32(rcs_desc)), _p(f32(sim_desc))).np()


def qmul(a, b):
    return lib().orc_qmul(_f4(a), _f4(b)).np()


def quat2abg(q):
    a, b, g = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    lib().orc_quat2abg(_f4(q), ctypes.byref(a), ctypes.byref(b), ctypes.byref(g))
    return a.value, b.value, g.value


def rand4(seeds):
    s = np.ascontiguousarray(seeds, np.uint32).reshape(4).copy()
    r = np.zeros(4, np.float32)
    lib().orc_rand4(_p(s), _p(r))
    return s, r


class State:
    """Reference-layout streams of one population."""

    def __init__(self, pos, vel, ori, tum, rcs, rnd, aux=None):
        self.pos = np.ascontiguousarray(pos, np.float32).copy()
        self.vel = np.ascontiguousarray(vel, np.float32).copy()
        self.ori = np.ascontiguousarray(ori, np.float32).copy()
        self.tum = np.ascontiguousarray(tum, np.float32).copy()
        self.rcs = np.ascontiguousarray(rcs, np.float32).copy()
        self.rnd = np.ascontiguousarray(rnd, np.uint32).copy()
        self.aux = None if aux is None else np.ascontiguousarray(aux, np.float32).copy()

    @property
    def n(self):
        return self.pos.shape[0]


def el_atts(st, begin, end, wind_tex, les_desc, drop, drop_desc, sim_desc):
    lib().orc_el_atts(_p(st.pos), _p(st.vel), _p(st.rcs), _p(st.rnd), begin, end, ctypes.byref(wind_tex),
                      _p(f32(les_desc)), _p(f32(drop)), _p(f32(drop_desc)), _p(f32(sim_desc)))


def bg_atts(st, begin, end, wind_tex, les_desc, drop, drop_desc, sim_desc):
    lib().orc_bg_atts(_p(st.pos), _p(st.vel), _p(st.rcs), _p(st.rnd), begin, end, ctypes.byref(wind_tex),
                      _p(f32(les_desc)), _p(f32(drop)), _p(f32(drop_desc)), _p(f32(sim_desc)))


_blend_keep = None


def set_blend(wind_next=None, cpxx_next=None, alpha=0.0):
    """Blend later advances' wind (and Cn^2) towards the next LES volume with
    weight alpha (textures from tex()); wind_next None switches it off."""
    global _blend_keep
    _blend_keep = (wind_next, cpxx_next)
    lib().orc_set_blend(None if wind_next is None else ctypes.byref(wind_next),
                        None if cpxx_next is None else ctypes.byref(cpxx_next), float(alpha))


def fp_atts(st, begin, end, wind_tex, cpxx_tex, les_desc, sim_desc):
    lib().orc_fp_atts(_p(st.pos), _p(st.vel), _p(st.rcs), begin, end, ctypes.byref(wind_tex), ctypes.byref(cpxx_tex),
                      _p(f32(les_desc)), _p(f32(sim_desc)))


def db_atts(st, begin, end, wind_tex, les_desc, cd_tex, cm_tex, adm_desc, re_tex, im_tex, rcs_desc, sim_desc):
    lib().orc_db_atts(_p(st.pos), _p(st.ori), _p(st.vel), _p(st.tum), _p(st.rcs), _p(st.rnd), begin, end,
                      ctypes.byref(wind_tex), _p(f32(les_desc)), ctypes.byref(cd_tex), ctypes.byref(cm_tex),
                      _p(f32(adm_desc)), ctypes.byref(re_tex), ctypes.byref(im_tex), _p(f32(rcs_desc)),
                      _p(f32(sim_desc)))


def set_dff(icdf, desc16):
    """Debris flux field of later db_atts calls (rs.cl:795-836); None clears it."""
    if icdf is None:
        lib().orc_set_dff(None, 0, None)
    else:
        icdf = f32(icdf)
        lib().orc_set_dff(_p(icdf), icdf.size, _p(f32(desc16)))


def dff_place(r4, sim_desc):
    """Respawn position (x, y, z, w) of one draw r4 under the field of set_dff."""
    out = np.zeros(4, np.float32)
    rc = lib().orc_dff_place_probe(_f4(r4), _p(f32(sim_desc)), _p(out))
    return None if rc else out


def dff_icdf_from_pdf(pdf, n=2048):
    """Inverse-CDF table from a cell PDF (rs.c:3247-3320); None on the error returns."""
    pdf = f32(pdf)
    tab = np.zeros(n, np.float32)
    return tab if lib().orc_dff_icdf_from_pdf(_p(pdf), pdf.size, _p(tab), n) == 0 else None


def db_rcs(st, begin, end, re_tex, im_tex, rcs_desc, sim_desc):
    lib().orc_db_rcs(_p(st.pos), _p(st.ori), _p(st.rcs), begin, end, ctypes.byref(re_tex), ctypes.byref(im_tex),
                     _p(f32(rcs_desc)), _p(f32(sim_desc)))


def scat_sig_aux(pos, rcs, aux, angw, angw_desc, sim_desc):
    pos, rcs = f32(pos), f32(rcs)
    aux = f32(aux).copy()
    sig = np.zeros_like(pos)
    lib().orc_scat_sig_aux(_p(sig), _p(aux), _p(pos), _p(rcs), pos.shape[0], _p(f32(angw)), _p(f32(angw_desc)),
                           _p(f32(sim_desc)))
    return sig, aux


def gate_ranges(start, delta, R):
    out = np.zeros(R, np.float32)
    lib().orc_gate_ranges(start, delta, R, _p(out))
    return out


def make_pulse(pos, rcs, angw, angw_desc, rangew, rangew_desc, range_start, range_delta, R, sim_desc,
               wave_number_d, mode=PHASE_F64, pos64=None):
    pos, rcs = f32(pos), f32(rcs)
    out = np.zeros((R, 4), np.float64)
    p64 = None if pos64 is None else np.ascontiguousarray(pos64, np.float64)
    lib().orc_make_pulse(_p(out), _p(pos), _p(p64), _p(rcs), pos.shape[0], _p(f32(angw)), _p(f32(angw_desc)),
                         _p(f32(rangew)), _p(f32(rangew_desc)), range_start, range_delta, R, _p(f32(sim_desc)),
                         wave_number_d, mode)
    return out


def populate(counts, sim_desc, seed, dsd_r=None, dsd_cdf=None, uniform=False):
    n = int(sum(counts))
    arrs = [np.zeros((n, 4), np.float32) for _ in range(6)]
    rnd = np.zeros((n, 4), np.uint32)
    c = np.zeros(8, np.uint64)
    c[:len(counts)] = counts
    nd = 0 if dsd_r is None else len(dsd_r)
    lib().orc_populate(*[_p(a) for a in arrs], _p(rnd), c.ctypes.data_as(ctypes.c_void_p), len(counts),
                       _p(f32(sim_desc)), seed, nd, _p(None if dsd_r is None else f32(dsd_r)),
                       _p(None if dsd_cdf is None else f32(dsd_cdf)), int(uniform))
    pos, vel, ori, tum, aux, rcs = arrs
    return dict(pos=pos, vel=vel, ori=ori, tum=tum, aux=aux, rcs=rcs, rnd=rnd)


def drop_rcs_table(wave_number, drop_scale, transparent=False, uniform=False, dsd_r=None, dsd_pdf=None):
    t = np.zeros((96, 4), np.float32)
    d = np.zeros(4, np.float32)
    nd = 0 if dsd_r is None else len(dsd_r)
    lib().orc_drop_rcs_table(_p(t), _p(d), wave_number, drop_scale, int(transparent), int(uniform), nd,
                             _p(None if dsd_r is None else f32(dsd_r)), _p(None if dsd_pdf is None else f32(dsd_pdf)))
    return t, d


def angular_weight_standard(bw_rad):
    w = np.zeros(32, np.float32)
    d = np.zeros(4, np.float32)
    lib().orc_angular_weight_standard(_p(w), _p(d), bw_rad)
    return w, d


def table1d_desc(start, delta, n):
    d = np.zeros(4, np.float32)
    lib().orc_table1d_desc(_p(d), start, delta, n)
    return d


def les_desc(stretched, nx, ny, nz, ax, ay, az, rx, ry, rz, tr):
    d = np.zeros(16, np.float32)
    lib().orc_les_desc(_p(d), int(stretched), nx, ny, nz, ax, ay, az, rx, ry, rz, tr)
    return d


def radar_gain(gain_dbi, watt):
    return float(lib().orc_radar_gain(gain_dbi, watt))


def read_table(table, index_last, index):
    t = f32(table)
    return float(lib().orc_read_table(_p(t), index_last, index))


def el_atts_f64(p3, v3, radius, wind_tex, les_desc, sim_desc, bounded):
    lib().orc_el_atts_f64(_p(p3), _p(v3), _p(f32(radius)), p3.shape[0], ctypes.byref(wind_tex), _p(f32(les_desc)),
                          _p(f32(sim_desc)), int(bounded))
