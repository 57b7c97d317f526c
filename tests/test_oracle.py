"""Pins of the CPU oracle: every function is checked against something the
mathematics or the reference fixes independently of the oracle's own code
(binary64 / mpmath references, closed forms, the reference's own known-answer
configuration and MATLAB round trip, glibc's rand stream)."""
import ctypes
import math

import numpy as np
import pytest
from scipy import integrate, optimize, special

from oracle import oracle as orc


def ulps(got, ref):
    ref32 = np.asarray(ref, np.float64).astype(np.float32)
    sp = np.spacing(np.abs(ref32)).astype(np.float64)
    return np.abs(np.asarray(got, np.float64) - np.asarray(ref, np.float64)) / sp


# ---------------------------------------------------------------- elementary functions

@pytest.mark.parametrize("fn,gen,ref,bound", [
    ("log", lambda g: np.exp(g.uniform(-22, 8, 200000)), np.log, 1.0),
    ("sin", lambda g: g.uniform(-40, 40, 200000), np.sin, 2.0),
    ("cos", lambda g: g.uniform(-40, 40, 200000), np.cos, 2.0),
    ("acos", lambda g: g.uniform(-1, 1, 200000), np.arccos, 3.0),
])
def test_math_ulp_bounds(fn, gen, ref, bound):
    g = np.random.default_rng(7)
    x = gen(g).astype(np.float32)
    got = orc.math(fn, x)
    e = ulps(got, ref(x.astype(np.float64)))
    assert e.max() <= bound, "%s max error %.2f ulp" % (fn, e.max())


def test_math_atan2_quadrants_and_specials():
    g = np.random.default_rng(8)
    y = g.normal(size=200000).astype(np.float32)
    x = g.normal(size=200000).astype(np.float32)
    e = ulps(orc.math("atan2", y, x), np.arctan2(y.astype(np.float64), x.astype(np.float64)))
    assert e.max() <= 3.0
    ys = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, 0.0, 2.0], np.float32)
    xs = np.array([0.0, 0.0, -0.0, -0.0, 0.0, 0.0, -3.0, 2.0], np.float32)
    got = orc.math("atan2", ys, xs)
    want = np.arctan2(ys.astype(np.float64), xs.astype(np.float64)).astype(np.float32)
    assert np.array_equal(np.signbit(got), np.signbit(want))
    np.testing.assert_allclose(got, want, rtol=1e-7)


def test_sin_large_arguments_absolute_error():
    # the binary32 phase of the reference reaches 4 pi r / lambda ~ 1.3e6 rad at 10 km
    g = np.random.default_rng(9)
    x = g.uniform(-1.3e6, 1.3e6, 100000).astype(np.float32)
    err = np.abs(orc.math("sin", x) - np.sin(x.astype(np.float64)))
    assert err.max() < 2e-7


# ---------------------------------------------------------------- LCG (rs.cl:35-44)

def test_lcg_closed_form():
    g = np.random.default_rng(3)
    s0 = g.integers(1, 2 ** 31 - 1, size=4, dtype=np.uint64)
    s = s0.astype(np.uint32)
    for n in range(1, 400):
        s, r = orc.rand4(s)
        want = [(int(v) * pow(16807, n, 2 ** 31)) % (2 ** 31) for v in s0]
        assert [int(v) for v in s] == want
        np.testing.assert_array_equal(r, np.array(want, np.float32) * np.float32(2.0 ** -31))


def test_lcg_is_not_park_miller():
    s, _ = orc.rand4(np.array([2 ** 30, 1, 1, 1], np.uint32))
    assert int(s[0]) == (2 ** 30 * 16807) % (2 ** 31) != (2 ** 30 * 16807) % (2 ** 31 - 1)


# ---------------------------------------------------------------- sampler (rs.cl:8)

def test_trilinear_reproduces_linear_fields():
    nx, ny, nz = 7, 5, 4
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    data = np.stack([1.0 + 0.5 * i - 0.25 * j + 2.0 * k, -3.0 + i * 0.125, 0.75 * k + j, 0.0 * i], -1).astype(np.float32)
    t, keep = orc.tex(data)
    g = np.random.default_rng(1)
    for _ in range(300):
        u, v, w = g.uniform(0.5, nx - 0.5), g.uniform(0.5, ny - 0.5), g.uniform(0.5, nz - 0.5)
        got = orc.sample3d(t, u, v, w)
        a, b, c = u - 0.5, v - 0.5, w - 0.5     # texel i sits at coordinate i + 0.5
        want = [1.0 + 0.5 * a - 0.25 * b + 2.0 * c, -3.0 + 0.125 * a, 0.75 * c + b, 0.0]
        np.testing.assert_allclose(got, want, rtol=2e-6, atol=2e-6)


def test_trilinear_clamps_to_edge():
    data = np.arange(3 * 2 * 2 * 4, dtype=np.float32).reshape(2, 2, 3, 4)
    t, keep = orc.tex(data)
    np.testing.assert_array_equal(orc.sample3d(t, -5.0, -7.0, -1.0), data[0, 0, 0])
    np.testing.assert_array_equal(orc.sample3d(t, 50.0, 70.0, 10.0), data[1, 1, 2])
    np.testing.assert_array_equal(orc.sample3d(t, 1.5, 0.5, 0.5), data[0, 0, 1])


def test_stretched_wind_index_maps_grid_nodes_to_texels():
    # |x| = a (r^m - 1) / (r - 1) maps to coordinate (n - 1)/2 + m (rs.c:3058-3061);
    # the box holds every node (the index tables cover the box)
    d = orc.les_desc(True, 176, 176, 80, 2.0, 2.0, 2.0, 1.0212, 1.0212, 1.05, 50.0)
    sd = np.zeros(16, np.float32)
    sd[8], sd[9], sd[12], sd[13] = -500.0, 600.0, 1000.0, 1100.0
    sd[10], sd[14] = 0.0, 2600.0
    cx, cy = -500.0 + 500.0, 600.0 + 550.0
    for m in (0, 1, 5, 40, 85):
        x = 2.0 * (1.0212 ** m - 1.0) / 0.0212
        z = 2.0 * (1.05 ** m - 1.0) / 0.05
        c = orc.wind_index([cx + x, cy - x, z, 0.0], d, sd)
        np.testing.assert_allclose(c[:3], [87.5 + m, 87.5 - m, m], atol=2e-3)


def test_uniform_wind_index_is_fma():
    d = np.zeros(16, np.float32)
    d[0:3] = [0.01, 0.02, 0.03]
    d[4:7] = [1.5, 2.5, 0.0]
    c = orc.wind_index([100.0, -50.0, 20.0, 0.0], d, np.zeros(16, np.float32))
    np.testing.assert_allclose(c[:3], [2.5, 1.5, 0.6], rtol=1e-6)


# ------------------------------------------------------- pulse known answers (rs.cl:1050-1441)

def _kat_inputs(n):
    k = np.arange(n)
    pos = np.stack([(k % 9) - 4.0, k / 90000.0 + 1.0, np.full(n, 0.1), np.ones(n)], -1).astype(np.float32)
    rcs = np.tile(np.array([1.0, 0.5, 1.0, 0.5], np.float32), (n, 1))
    return pos, rcs


def test_make_pulse_known_answer_test():
    """test_make_pulse.c: pop inputs (rs.cl:1444-1453), 9 gates at 1 + 0.25 k m,
    range weights {0,1,0} over +-25 m, angular {0,.7,1,.7,0} over +-20 deg, beam y;
    the CPU loop of test_make_pulse.c:553-583 restated in binary64 numpy, held to
    the reference's 1e-3."""
    n = 1079296
    pos, rcs = _kat_inputs(n)
    rw = np.array([0.0, 1.0, 0.0], np.float32)
    xs = np.float32(1.0) / np.float32(25.0)
    rwd = np.array([xs, np.float32(25.0) * xs, 2.0, 0.0], np.float32)
    aw = np.array([0.0, 0.7, 1.0, 0.7, 0.0], np.float32)
    d10 = np.float32(10.0 * 0.01745329251994)
    awd = np.array([np.float32(1.0) / d10, np.float32(20.0 * 0.01745329251994) * (np.float32(1.0) / d10), 4.0, 0.0], np.float32)
    sd = np.zeros(16, np.float32)
    sd[1] = 1.0
    sd[4] = np.float32(4.0 * np.pi / 0.1)
    got = orc.make_pulse(pos, rcs, aw, awd, rw, rwd, 1.0, 0.25, 9, sd, 4.0 * np.pi / 0.1, orc.PHASE_F32)
    # restated CPU loop (binary64)
    p = pos[:, :3].astype(np.float64)
    r = np.sqrt((p ** 2).sum(1))
    ang = np.arccos(p[:, 1] / r)

    def read_table(table, last, idx):
        fl = np.floor(idx)
        i0 = np.clip(fl, 0, last).astype(int)
        i1 = np.clip(fl + 1, 0, last).astype(int)
        v = table[i0] + (idx - fl) * (table[i1] - table[i0])
        v = np.where(idx <= 0, table[0], v)
        return np.where(fl >= last, table[int(last)], v)

    wa = read_table(aw.astype(np.float64), 4, ang * float(awd[0]) + float(awd[1]))
    ph = r * float(sd[4])
    att = r ** -2
    sig = np.stack([np.cos(ph), -np.sin(ph), np.cos(ph), -np.sin(ph)], -1)
    a = rcs.astype(np.float64)
    s = np.stack([a[:, 0] * sig[:, 0] - a[:, 1] * sig[:, 1], a[:, 1] * sig[:, 0] + a[:, 0] * sig[:, 1],
                  a[:, 2] * sig[:, 2] - a[:, 3] * sig[:, 3], a[:, 3] * sig[:, 2] + a[:, 2] * sig[:, 3]], -1) * att[:, None]
    want = np.zeros((9, 4))
    for ir in range(9):
        rg = ir * 0.25 + 1.0
        wr = read_table(rw.astype(np.float64), 2, (r - rg) * float(xs) + 1.0)
        want[ir] = ((wr * wa)[:, None] * s).sum(0)
    scale = np.sqrt(np.mean((want ** 2).sum(1)))
    assert np.max(np.abs(got - want)) / scale < 1e-3


def test_single_scatterer_closed_form():
    """One scatterer on the beam axis at a gate centre: pulse = rcs e^{-j 4 pi r / lambda} / r^2."""
    r0 = 2000.0
    pos = np.array([[0.0, r0, 0.0, 1e-3]], np.float32)
    rcs = np.array([[2.0, 0.5, 1.5, -0.25]], np.float32)
    aw, awd = orc.angular_weight_standard(np.float32(np.pi / 180))
    rw = np.array([0.0, 1.0, 0.0], np.float32)
    rwd = orc.table1d_desc(-30.0, 30.0, 3)
    sd = np.zeros(16, np.float32)
    sd[1] = 1.0
    sd[4] = np.float32(4 * np.pi / 0.1)
    kd = 4.0 * np.pi / 0.1
    out = orc.make_pulse(pos, rcs, aw, awd, rw, rwd, r0 - 45.0, 15.0, 7, sd, kd, orc.PHASE_F64)
    e = np.exp(-1j * kd * r0) / r0 ** 2
    h = (2.0 + 0.5j) * e
    v = (1.5 - 0.25j) * e
    np.testing.assert_allclose(out[3], [h.real, h.imag, v.real, v.imag], rtol=2e-6, atol=1e-13)
    # triangle: neighbours at +-15 m of a 30 m half-width weight 1/2, +-30 m weight 0
    np.testing.assert_allclose(out[2], 0.5 * out[3], rtol=1e-5, atol=1e-14)
    np.testing.assert_allclose(out[4], 0.5 * out[3], rtol=1e-5, atol=1e-14)
    assert np.all(out[[0, 1, 5, 6]] == 0.0)


def test_pulse_superposition():
    g = np.random.default_rng(4)
    n = 4000
    pos = np.stack([g.uniform(-200, 200, n), g.uniform(1900, 2300, n), g.uniform(10, 400, n), np.full(n, 1e-3)], -1).astype(np.float32)
    rcs = g.normal(size=(n, 4)).astype(np.float32)
    aw, awd = orc.angular_weight_standard(np.float32(np.pi / 180))
    rw = np.array([0.0, 1.0, 0.0], np.float32)
    rwd = orc.table1d_desc(-30.0, 30.0, 3)
    sd = np.zeros(16, np.float32)
    sd[1] = np.float32(np.cos(0.05))
    sd[2] = np.float32(np.sin(0.05))
    sd[4] = np.float32(4 * np.pi / 0.1)
    args = (aw, awd, rw, rwd, 1850.0, 15.0, 40, sd, 4 * np.pi / 0.1, orc.PHASE_F64)
    full = orc.make_pulse(pos, rcs, *args)
    a = orc.make_pulse(pos[:1500], rcs[:1500], *args)
    b = orc.make_pulse(pos[1500:], rcs[1500:], *args)
    np.testing.assert_allclose(full, a + b, rtol=1e-12, atol=1e-25)


def test_gate_ranges_repeated_addition():
    rk = orc.gate_ranges(1845.0, 15.0, 40)
    r = np.float32(1845.0)
    for k in range(40):
        assert rk[k] == r
        r = np.float32(r + np.float32(15.0))


# ---------------------------------------------------------------- drop RCS (rs.c:5922-6029)

def test_drop_rcs_is_rayleigh_spheroid():
    """alpha = V eps0 (eps_r - 1) / (1 + L (eps_r - 1)) with the depolarisation
    factor L_z of an oblate spheroid from its defining integral."""
    wave = np.float32(4 * np.pi / 0.1)
    t, d = orc.drop_rcs_table(wave, 1.0)
    eps0 = float(np.float32(8.85418782e-12))
    k0 = float(np.float32(wave * np.float32(0.5)))
    sc = k0 * k0 / (4.0 * np.pi * eps0) * float(np.float32(1e-9))
    erm1 = 78.669 + 18.2257j     # eps_r - 1 of water (rs.c:5928)
    for i in (0, 10, 45, 95):
        D = 0.5 + 0.1 * i
        rba = 1.0048 + 0.0057e-1 * D - 2.628e-2 * D ** 2 + 3.682e-3 * D ** 3 - 1.677e-4 * D ** 4
        c, a = rba, 1.0                      # axis ratio c/a of the oblate drop
        Lz = a * a * c / 2.0 * integrate.quad(lambda s: 1.0 / ((s + c * c) ** 1.5 * (s + a * a)), 0, np.inf,
                                               epsabs=1e-14, epsrel=1e-12)[0]
        Lx = (1 - Lz) / 2
        V = np.pi * D ** 3 / 6.0
        ax = V * eps0 * erm1 / (1 + Lx * erm1) * sc
        az = V * eps0 * erm1 / (1 + Lz * erm1) * sc
        np.testing.assert_allclose(t[i], [ax.real, ax.imag, az.real, az.imag], rtol=2e-6)
    # index = radius * 20000 - 5: the 0.5 mm drop (r = 0.25 mm) is row 0
    np.testing.assert_allclose(d[:3], [np.float32(1) / np.float32(0.05e-3), -5.0, 95.0], rtol=1e-6)


def test_angular_weight_is_bessel_pattern():
    bw = np.float32(np.pi / 180)
    w, d = orc.angular_weight_standard(bw)
    assert w[0] == 1.0 and w[31] == 0.0
    b = np.float32(1.27 * np.pi / bw)
    a = np.arange(1, 31) * np.float32(np.pi / 360)
    c = b * np.sin(a.astype(np.float32))
    np.testing.assert_allclose(w[1:31], 8 * special.jv(2, c.astype(np.float64)) / c.astype(np.float64) ** 2, rtol=1e-5, atol=1e-7)


# ---------------------------------------------------------------- quaternions / Euler

def _qmul64(l, r):
    return np.array([l[3] * r[0] - l[2] * r[1] + l[1] * r[2] + l[0] * r[3],
                     l[3] * r[1] + l[2] * r[0] + l[1] * r[3] - l[0] * r[2],
                     l[3] * r[2] + l[2] * r[3] - l[1] * r[0] + l[0] * r[1],
                     l[3] * r[3] - l[2] * r[2] - l[1] * r[1] - l[0] * r[0]])


def _r323(a, b, g):
    rz = lambda x: np.array([0, 0, np.sin(0.5 * x), np.cos(0.5 * x)])
    ry = lambda x: np.array([0, np.sin(0.5 * x), 0, np.cos(0.5 * x)])
    return _qmul64(_qmul64(rz(a), ry(b)), rz(g))


def test_quat2abg_round_trip():
    """matlab/rcs_abg_quat_test.m: r323quat(alpha, beta, gamma) -> quat2abg recovers
    the angles (canonical case pi/2, 0.2, 0 plus random ones off the poles)."""
    cases = [(0.5 * np.pi, 0.2, 0.0)]
    g = np.random.default_rng(12)
    cases += [(g.uniform(-3.1, 3.1), g.uniform(0.05, np.pi - 0.05), g.uniform(-3.1, 3.1)) for _ in range(200)]
    for a, b, c in cases:
        got = orc.quat2abg(_r323(a, b, c).astype(np.float32))
        np.testing.assert_allclose(got, (a, b, c), atol=5e-4)


def test_quat_mult_matches_hamilton_product():
    g = np.random.default_rng(2)
    for _ in range(100):
        a, b = g.normal(size=4), g.normal(size=4)
        np.testing.assert_allclose(orc.qmul(a, b), _qmul64(a.astype(np.float32).astype(np.float64),
                                                          b.astype(np.float32).astype(np.float64)), rtol=1e-5, atol=1e-6)


# ---------------------------------------------------------------- populate (rs.c:4652-4799)

def test_populate_follows_glibc_rand_stream():
    libc = ctypes.CDLL(None)
    libc.srand(1)
    seq = [libc.rand() for _ in range(8)]
    assert seq[:3] == [1804289383, 846930886, 1681692777]        # glibc TYPE_3, seed 1
    sd = np.zeros(16, np.float32)
    sd[8:11] = [-400.0, 1600.0, 0.0]
    sd[12:15] = [800.0, 900.0, 850.0]
    st = orc.populate([3], sd, 1)
    f = np.float32
    assert st["pos"][0, 0] == f(f(seq[0]) / f(2 ** 31)) * f(800.0) + f(-400.0)
    assert st["pos"][0, 2] == f(f(seq[2]) / f(2 ** 31)) * f(10.0)
    np.testing.assert_array_equal(st["rnd"][0], seq[4:8])
    np.testing.assert_array_equal(st["ori"][0], [0.5, -0.5, -0.5, 0.5])


def test_populate_bounds_and_dsd_frequencies():
    sd = np.zeros(16, np.float32)
    sd[8:11] = [-400.0, 1600.0, 5.0]
    sd[12:15] = [800.0, 900.0, 850.0]
    r = np.array([0.5e-3, 1e-3, 2e-3], np.float32)
    cdf = np.array([0.0, 0.5, 0.8], np.float32)
    st = orc.populate([200000], sd, 7, r, cdf)
    p = st["pos"]
    assert np.all(p[:, 0] >= -400) and np.all(p[:, 0] < 400) and np.all(p[:, 2] >= 5) and np.all(p[:, 2] < 15.0001)
    frac = np.array([(p[:, 3] == v).mean() for v in r])
    np.testing.assert_allclose(frac, [0.5, 0.3, 0.2], atol=5 * np.sqrt(0.25 / 200000))
    assert np.all(st["rnd"] < 2 ** 31)


# ---------------------------------------------------------------- drag model (rs.cl:653-739)

def _uniform_wind(u, v, w):
    data = np.array([[[[u, v, w, 0.0]]]], np.float32)
    t, keep = orc.tex(data)
    d = np.zeros(16, np.float32)
    d[0:3] = 1e-3
    return t, keep, d


def test_el_atts_relaxes_to_terminal_velocity():
    """Under still air a drop settles at v_t with 0.5 rho Cd(Re) (A/m) v_t^2 = g,
    Cd = 24/Re + 6/(1 + sqrt Re) + 0.4 (the drag law of rs.cl:706-709)."""
    t, keep, les = _uniform_wind(0.0, 0.0, 0.0)
    radius = 1e-3
    drop = np.zeros((96, 4), np.float32)
    dd = np.array([20000.0, -5.0, 95.0, 1.0], np.float32)
    sd = np.zeros(16, np.float32)
    sd[8:11] = [-1e4, -1e4, -1e5]
    sd[12:15] = [2e4, 2e4, 2e5]
    sd[11] = 0.01
    sd[5] = np.array([2], np.uint32).view(np.float32)[0]
    st = orc.State([[0.0, 0.0, 0.0, radius]], [[0.0, 0.0, 0.0, 0.0]], [[0, 0, 0, 1]], [[0, 0, 0, 1]],
                   [[0, 0, 0, 0]], np.array([[1, 2, 3, 4]], np.uint32))
    for _ in range(3000):
        orc.el_atts(st, 0, 1, t, les, drop, dd, sd)

    def resid(v):
        re = 6.7308e4 * 2 * radius * v
        cd = 24 / re + 6 / (1 + np.sqrt(re)) + 0.4
        return 0.5 * 1.225 * cd * (0.003006012 / radius) * v * v - 9.8

    vt = optimize.brentq(resid, 0.1, 30.0)
    np.testing.assert_allclose(-st.vel[0, 2], vt, rtol=2e-3)


def test_el_atts_respawn_draw():
    """Leaving the box draws one LCG step and places the drop at
    fma(r, size, origin) with zero velocity (rs.cl:685-692); rcs is kept."""
    t, keep, les = _uniform_wind(0.0, 0.0, 0.0)
    drop = np.zeros((96, 4), np.float32)
    dd = np.array([20000.0, -5.0, 95.0, 1.0], np.float32)
    sd = np.zeros(16, np.float32)
    sd[8:11] = [-100.0, 900.0, 0.0]
    sd[12:15] = [200.0, 300.0, 400.0]
    sd[11] = 1e-3
    sd[5] = np.array([2], np.uint32).view(np.float32)[0]
    seed = np.array([[11, 222, 3333, 44444]], np.uint32)
    st = orc.State([[0.0, 1000.0, 0.0005, 1e-3]], [[0.0, 0.0, -1.0, 0.0]], [[0, 0, 0, 1]], [[0, 0, 0, 1]],
                   [[7.0, 8.0, 9.0, 10.0]], seed)
    orc.el_atts(st, 0, 1, t, les, drop, dd, sd)
    s2, r = orc.rand4(seed[0])
    np.testing.assert_array_equal(st.rnd[0], s2)
    want = [np.float32(np.float64(r[i]) * sd[12 + i] + sd[8 + i]) for i in range(3)]
    np.testing.assert_array_equal(st.pos[0, :3], np.array(want, np.float32))
    np.testing.assert_array_equal(st.vel[0], [0, 0, 0, 0])
    np.testing.assert_array_equal(st.rcs[0], [7.0, 8.0, 9.0, 10.0])


# ---------------------------------------------------------------- background and debris (rs.cl:533-580, 744-919)

def _box_sd(dt, concept=0):
    sd = np.zeros(16, np.float32)
    sd[8:11] = [-1e4, -1e4, -1e4]
    sd[12:15] = [2e4, 2e4, 2e4]
    sd[11] = dt
    sd[5] = np.array([concept], np.uint32).view(np.float32)[0]
    return sd


def test_bg_atts_is_advect_then_sample():
    """bg_atts: pos += vel * dt without fma (rs.cl:551), then vel = wind at the
    new position (rs.cl:573); rcs from the drop table."""
    t, keep, les = _uniform_wind(3.0, -2.0, 0.5)
    drop = np.zeros((96, 4), np.float32)
    dd = np.array([20000.0, -5.0, 95.0, 1.0], np.float32)
    sd = _box_sd(0.01)
    v0 = np.array([[1.25, 0.5, -4.0, 0.0]], np.float32)
    p0 = np.array([[10.0, 20.0, 30.0, 1e-3]], np.float32)
    st = orc.State(p0, v0, [[0, 0, 0, 1]], [[0, 0, 0, 1]], [[0, 0, 0, 0]], np.array([[1, 2, 3, 4]], np.uint32))
    orc.bg_atts(st, 0, 1, t, les, drop, dd, sd)
    dt = np.float32(0.01)
    want = p0[0, :3] + np.float32(v0[0, :3] * dt)          # binary32 product, then binary32 sum
    assert np.array_equal(st.pos[0, :3], want.astype(np.float32))
    np.testing.assert_array_equal(st.vel[0, :3], [3.0, -2.0, 0.5])


def _const_tex2(val, nx=9, ny=7):
    return orc.tex(np.tile(np.array(val, np.float32), (ny, nx, 1)))


def _debris_step(ori, cd, cm, u=(6.0, 0.0, 0.0), dt=0.01, ta=0.05, inln=(2.0, 3.0, 4.0), concept=0):
    wt, wk, les = _uniform_wind(*u)
    cdt, cdk = _const_tex2(cd)
    cmt, cmk = _const_tex2(cm)
    ret, rek = _const_tex2([0.0, 0.0, 0.0, 0.0])
    imt, imk = _const_tex2([0.0, 0.0, 0.0, 0.0])
    ad = np.zeros(16, np.float32)
    ad[0:2] = 1.0
    ad[4:6] = 1.0
    ad[8:10] = [8, 6]
    ad[12:15] = inln
    ad[15] = ta
    rd = np.zeros(16, np.float32)
    rd[0:2] = 1.0
    rd[4:6] = 1.0
    sd = _box_sd(dt, concept)
    st = orc.State([[5.0, 100.0, 50.0, 0.0]], [[0, 0, 0, 0]], [ori], [[0, 0, 0, 1]], [[0, 0, 0, 0]],
                   np.array([[11, 12, 13, 14]], np.uint32))
    orc.db_atts(st, 0, 1, wt, les, cdt, cmt, ad, ret, imt, rd, sd)
    return st


def test_db_atts_force_is_the_drag_coefficient_rotated_into_the_world():
    """dudt = Ta |u_r|^2 rot(cd, ori) + g (rs.cl:195-261) and pos moves with the
    old velocity: a body at rest in wind u = (6, 0, 0) with a constant table."""
    ta, dt, c = 0.05, 0.01, 0.3
    st = _debris_step([0, 0, 0, 1], [c, 0, 0, 0], [0, 0, 0, 0], ta=ta, dt=dt)
    np.testing.assert_allclose(st.vel[0, :3], [ta * 36 * c * dt, 0.0, -9.8 * dt], rtol=1e-6)
    np.testing.assert_array_equal(st.pos[0, :3], [5.0, 100.0, 50.0])
    # the same body yawed +90 deg about z: body x is world y (q v q*, rs.cl:81-85)
    s = np.float32(np.sqrt(0.5))
    st = _debris_step([0, 0, s, s], [c, 0, 0, 0], [0, 0, 0, 0], ta=ta, dt=dt)
    np.testing.assert_allclose(st.vel[0, :3], [0.0, ta * 36 * c * dt, -9.8 * dt], rtol=1e-5, atol=1e-9)


def test_db_atts_torque_tumbles_by_radians_of_the_moment():
    """dw = radians(Ta |u_r|^2 inv_inln cm) (rs.cl:243) and the tumble quaternion
    from sincos(dw dt) (rs.cl:889-897): a pure x moment gives (sin a, 0, 0, cos a)."""
    ta, dt, m, inl = 0.05, 0.01, 0.7, 2.0
    st = _debris_step([0, 0, 0, 1], [0, 0, 0, 0], [m, 0, 0, 0], ta=ta, dt=dt, inln=(inl, 0.0, 0.0))
    a = np.radians(ta * 36 * inl * m) * dt
    np.testing.assert_allclose(st.tum[0], [np.sin(a), 0.0, 0.0, np.cos(a)], rtol=1e-5, atol=1e-7)


def test_db_atts_respawn_resets_the_body():
    """Leaving the box: xy from the LCG over the box, z = z0 + 2 m, unit random
    orientation, zero velocity, identity tumble, rcs 0 (rs.cl:790-873)."""
    wt, wk, les = _uniform_wind(0.0, 0.0, 0.0)
    z = _const_tex2([0.0, 0.0, 0.0, 0.0])
    ad = np.zeros(16, np.float32)
    ad[8:10] = [8, 6]
    sd = np.zeros(16, np.float32)
    sd[8:11] = [0.0, 0.0, 0.0]
    sd[12:15] = [100.0, 200.0, 50.0]
    sd[11] = 0.01
    st = orc.State([[50.0, 100.0, 60.0, 0.0]], [[1, 2, 3, 0]], [[0, 0, 0, 1]], [[0.1, 0, 0, 0.99]], [[1, 1, 1, 1]],
                   np.array([[11, 12, 13, 14]], np.uint32))
    orc.db_atts(st, 0, 1, wt, les, z[0], z[0], ad, z[0], z[0], ad, sd)
    seeds, r = orc.rand4(np.array([11, 12, 13, 14], np.uint32))
    np.testing.assert_array_equal(st.pos[0, :2], [np.float32(r[0] * 100.0), np.float32(r[1] * 200.0)])
    assert st.pos[0, 2] == 2.0
    np.testing.assert_array_equal(st.vel[0], 0.0)
    np.testing.assert_array_equal(st.tum[0], [0, 0, 0, 1])
    np.testing.assert_array_equal(st.rcs[0], 0.0)
    assert abs(np.linalg.norm(st.ori[0].astype(np.float64)) - 1.0) < 1e-5
    assert not np.array_equal(st.rnd[0], [11, 12, 13, 14])


def test_debris_rcs_of_an_isotropic_scatterer_ignores_orientation():
    """hh = vv = c, hv = 0 is invariant under the gamma rotation of the
    scattering matrix (rs.cl:340-347): every pose returns (c, c) in H and V."""
    g = np.random.default_rng(3)
    re_t, rk = _const_tex2([0.7, 0.7, 0.0, 0.0])
    im_t, ik = _const_tex2([-0.2, -0.2, 0.0, 0.0])
    rd = np.zeros(16, np.float32)
    rd[0:2] = 1.0
    rd[4:6] = 1.0
    sd = _box_sd(0.01)
    for _ in range(50):
        q = g.normal(size=4).astype(np.float32)
        q /= np.linalg.norm(q)
        p = g.uniform(-1e3, 1e3, 4).astype(np.float32)
        p[2] = abs(p[2])
        out = orc.debris_rcs(p, q, re_t, im_t, rd, sd)
        np.testing.assert_allclose(out, [0.7, -0.2, 0.7, -0.2], rtol=2e-6, atol=1e-7)
