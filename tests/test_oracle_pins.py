"""Pins of the oracle's table lookups and respawn placement, each against a
route to the same quantity that does not share the oracle's code:

* debris RCS (rs.cl:282-359): the pose is built from elementary rotations
  (r323quat of matlab/common.m:29 and the line-of-sight frame as three axis
  rotations), the table is a ramp whose bilinear sample is known in closed form,
  and the scattering-matrix rotation is a 2 x 2 matrix product;
* ADM lookup inside db_atts (rs.cl:195-261): the body-frame flow direction is
  formed with rotation matrices and the drag coefficient is read back from the
  velocity change;
* table descriptors of the host runtime (rs.c:3746-3783 ADM with +0.5,
  rs.c:3960-3992 RCS without): where the end angles land in texel coordinates;
* sample2d (rs.cl:8 sampler): linear fields are reproduced exactly, edges clamp;
* pow / exp of the flux-field placement against binary64, and the placement
  itself (rs.cl:795-836) against the closed forms of its three grid kinds.
"""
import numpy as np
import pytest

import harness as h
import simradar_amd as sa
from oracle import oracle as orc
from oracle.run import rcs_desc


# ----------------------------------------------------------------- quaternions in binary64

def _qm(l, r):
    """Hamilton product, quaternions as (x, y, z, w) (matlab/common.m:1-5)."""
    return np.array([l[3] * r[0] - l[2] * r[1] + l[1] * r[2] + l[0] * r[3],
                     l[3] * r[1] + l[2] * r[0] + l[1] * r[3] - l[0] * r[2],
                     l[3] * r[2] + l[2] * r[3] - l[1] * r[0] + l[0] * r[1],
                     l[3] * r[3] - l[2] * r[2] - l[1] * r[1] - l[0] * r[0]])


def _rx(a):
    return np.array([np.sin(0.5 * a), 0.0, 0.0, np.cos(0.5 * a)])


def _ry(a):
    return np.array([0.0, np.sin(0.5 * a), 0.0, np.cos(0.5 * a)])


def _rz(a):
    return np.array([0.0, 0.0, np.sin(0.5 * a), np.cos(0.5 * a)])


def _conj(q):
    return np.array([-q[0], -q[1], -q[2], q[3]])


def _r323(a, b, c):
    """matlab/common.m:29"""
    return _qm(_qm(_rz(a), _ry(b)), _rz(c))


def _los_conj(el, az):
    """The conjugated line-of-sight frame of a scatterer at (el, az): the three
    axis rotations y(90 deg), x(-el - 90 deg), z(az).  Its conjugate turns body x
    onto the direction from the scatterer back to the radar."""
    return _qm(_qm(_ry(0.5 * np.pi), _rx(-el - 0.5 * np.pi)), _rz(az))


def _qmat(q):
    """Rotation matrix of v -> q v q*."""
    x, y, z, w = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_line_of_sight_frame_points_body_x_at_the_radar():
    """The frame whose conjugate the kernel multiplies in: conj(o_conj) maps
    body x onto -p/|p| for every scatterer position p."""
    g = np.random.default_rng(4)
    for _ in range(50):
        p = g.normal(size=3) * [800.0, 800.0, 200.0] + [0.0, 2000.0, 400.0]
        el, az = np.arctan2(p[2], np.hypot(p[0], p[1])), np.arctan2(p[0], p[1])
        bx = _qmat(_conj(_los_conj(el, az))) @ [1.0, 0.0, 0.0]
        np.testing.assert_allclose(bx, -p / np.linalg.norm(p), atol=1e-12)


# ----------------------------------------------------------------- debris RCS pose mapping

NA, NB = 73, 37          # 5 deg tables: alpha over [-pi, pi], beta over [0, pi]
# ramp coefficients per channel: value(i, j) = c0 + ci i + cj j at texel (i, j)
RE_RAMP = np.array([[0.20, 0.010, 0.003], [-0.10, 0.004, 0.012], [0.05, -0.002, 0.006], [0.0, 0.0, 0.0]])
IM_RAMP = np.array([[-0.03, 0.001, -0.004], [0.02, -0.003, 0.002], [0.01, 0.0025, -0.001], [0.0, 0.0, 0.0]])


def _ramp_table(ramp, nx, ny):
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    return np.stack([ramp[c, 0] + ramp[c, 1] * i + ramp[c, 2] * j for c in range(4)], axis=-1).astype(np.float32)


def _ramp_at(ramp, ti, tj):
    """Bilinear sample of a ramp at texel coordinates (ti, tj) inside the table."""
    return ramp[:, 0] + ramp[:, 1] * ti + ramp[:, 2] * tj


def _expected_rcs(alpha, beta, gamma, cross):
    d = rcs_desc(NA, NB).astype(np.float64)
    # sampler: texel centres at i + 0.5, clamped to the edge texels (beta = 0
    # lands at -0.5 without the half-texel offset and reads row 0)
    ti = min(max(alpha * d[0] + d[4] - 0.5, 0.0), NA - 1.0)
    tj = min(max(beta * d[1] + d[5] - 0.5, 0.0), NB - 1.0)
    re, im = _ramp_at(RE_RAMP, ti, tj), _ramp_at(IM_RAMP, ti, tj)
    sc = np.array([[re[0] + 1j * im[0], re[2] + 1j * im[2]], [re[2] + 1j * im[2], re[1] + 1j * im[1]]])
    cg, sg = np.cos(gamma), np.sin(gamma)
    T = np.array([[cg, -sg], [sg, cg]])
    S = T @ sc @ T.T
    h, v = S[0, 0], S[1, 1]
    if cross:
        h, v = h + S[1, 0], v + S[0, 1]
    return np.array([h.real, h.imag, v.real, v.imag])


def _tables():
    re, im = _ramp_table(RE_RAMP, NA, NB), _ramp_table(IM_RAMP, NA, NB)
    (rt, rk), (it, ik) = orc.tex(re), orc.tex(im)
    return rt, it, (rk, ik)


def _sd(concept):
    sd = np.zeros(16, np.float32)
    sd[5] = np.array([concept], np.uint32).view(np.float32)[0]
    return sd


def _pose(alpha, beta, gamma, el, az):
    """ori such that the kernel's relative quaternion, after its (x, z, -y, w)
    axis shuffle, is r323quat(alpha, beta, gamma): the shuffle is undone with
    the permutation matrix it applies to the vector part."""
    q = _r323(alpha, beta, gamma)
    P = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, -1.0, 0.0]])    # (x, y, z) -> (x, z, -y)
    R = np.concatenate([P.T @ q[:3], q[3:]])
    return _qm(R, _conj(_los_conj(el, az)))


@pytest.mark.parametrize("cross", [False, True])
def test_debris_rcs_reads_the_table_at_the_pose_angles(cross):
    """Bodies posed by r323quat(alpha, beta, gamma) relative to the line of
    sight read the RCS table at (alpha, beta) and rotate the scattering matrix
    by gamma (S = T Sc T^T); with cross-pol, vh / hv add into H / V."""
    rt, it, keep = _tables()
    g = np.random.default_rng(31)
    sd = _sd(1 if cross else 0)
    rd = rcs_desc(NA, NB)
    for _ in range(300):
        a, b, c = g.uniform(-3.0, 3.0), g.uniform(0.06, np.pi - 0.06), g.uniform(-3.1, 3.1)
        p = np.array([g.uniform(-900, 900), g.uniform(1500, 3000), g.uniform(5, 900)])
        el, az = np.arctan2(p[2], np.hypot(p[0], p[1])), np.arctan2(p[0], p[1])
        ori = _pose(a, b, c, el, az)
        got = orc.debris_rcs(np.append(p, 0.0).astype(np.float32), ori.astype(np.float32), rt, it, rd, sd)
        np.testing.assert_allclose(got, _expected_rcs(a, b, c, cross), atol=3e-5,
                                   err_msg="pose (%.3f, %.3f, %.3f) at el %.3f az %.3f" % (a, b, c, el, az))


def test_debris_rcs_swapped_axes_would_be_caught():
    """The ramp pins the axes: reading the table at (beta, alpha) instead of
    (alpha, beta) moves the expected values far outside the tolerance."""
    a, b, c = 1.1, 0.4, 0.3
    good, swapped = _expected_rcs(a, b, c, False), _expected_rcs(b, a, c, False)
    assert np.max(np.abs(good - swapped)) > 1e-2


def test_debris_rcs_q_and_minus_q_agree():
    """matlab/rcs_abg_quat_test.m: a quaternion and its negative are the same
    rotation, so a non-isotropic table returns the same H and V for both."""
    rt, it, keep = _tables()
    g = np.random.default_rng(5)
    sd = _sd(1)
    rd = rcs_desc(NA, NB)
    for _ in range(100):
        q = g.normal(size=4)
        q /= np.linalg.norm(q)
        p = np.array([g.uniform(-500, 500), g.uniform(1500, 2500), g.uniform(5, 500), 0.0], np.float32)
        a = orc.debris_rcs(p, q.astype(np.float32), rt, it, rd, sd)
        b = orc.debris_rcs(p, (-q).astype(np.float32), rt, it, rd, sd)
        np.testing.assert_allclose(a, b, atol=1e-6)


def test_debris_rcs_gimbal_lumps_alpha_into_gamma():
    """beta within 1 deg of 0: the table is read at (0, 0) and the whole turn
    about the line of sight, alpha + gamma, rotates the scattering matrix
    (rs.cl:321-325)."""
    rt, it, keep = _tables()
    g = np.random.default_rng(6)
    sd = _sd(0)
    rd = rcs_desc(NA, NB)
    for _ in range(50):
        a, c = g.uniform(-1.5, 1.5), g.uniform(-1.5, 1.5)
        p = np.array([g.uniform(-500, 500), g.uniform(1500, 2500), g.uniform(5, 500)])
        el, az = np.arctan2(p[2], np.hypot(p[0], p[1])), np.arctan2(p[0], p[1])
        ori = _pose(a, 0.0, c, el, az)
        got = orc.debris_rcs(np.append(p, 0.0).astype(np.float32), ori.astype(np.float32), rt, it, rd, sd)
        np.testing.assert_allclose(got, _expected_rcs(0.0, 0.0, a + c, False), atol=3e-5)


# ----------------------------------------------------------------- ADM lookup inside db_atts

def _adm_desc_ref(nb, na):
    from oracle.run import adm_desc
    return adm_desc(nb, na, [2.0, 3.0, 4.0, 0.05])


CD_RAMP = np.array([[0.10, 0.004, 0.001], [-0.05, 0.002, 0.006], [0.30, -0.001, 0.003], [0.0, 0.0, 0.0]])


def test_db_atts_reads_the_drag_table_at_the_body_frame_flow_angles():
    """A body at rest in a uniform wind: beta = acos(u.x), alpha = atan2(u.z,
    u.y) of the body-frame flow direction, folded into alpha in [0, pi) (beta
    negated); the table is sampled at (beta, alpha) with the +0.5 texel offset,
    so grid angles hit texel centres.  The drag coefficient read back from the
    velocity change is the ramp at those angles."""
    nb, na = 73, 37
    ad = _adm_desc_ref(nb, na)
    cd_tab = _ramp_table(CD_RAMP, nb, na)
    (cdt, cdk) = orc.tex(cd_tab)
    (cmt, cmk) = orc.tex(np.zeros((na, nb, 4), np.float32))
    (ret, rek) = orc.tex(np.zeros((4, 4, 4), np.float32))
    wind = np.tile(np.array([6.0, -2.0, 1.5, 0.0], np.float32), (4, 4, 4, 1))
    wt, wk = orc.tex(wind)
    les = np.zeros(16, np.float32)
    les[0:3] = 1.0
    les[4:7] = 1.5
    les[7] = np.array([0], np.uint32).view(np.float32)[0]
    sd = np.zeros(16, np.float32)
    sd[8:11] = [-1000.0, -1000.0, 0.0]
    sd[12:15] = [2000.0, 2000.0, 1000.0]
    dt, ta = np.float32(0.01), 0.05
    sd[11] = dt
    u = wind[0, 0, 0, :3].astype(np.float64)
    g = np.random.default_rng(9)
    for _ in range(200):
        q = g.normal(size=4)
        q /= np.linalg.norm(q)
        q32 = q.astype(np.float32)
        st = orc.State([[5.0, 10.0, 50.0, 0.0]], [[0, 0, 0, 0]], [q32], [[0, 0, 0, 1]], [[0, 0, 0, 0]],
                       np.array([[11, 12, 13, 14]], np.uint32))
        orc.db_atts(st, 0, 1, wt, les, cdt, cmt, ad, ret, ret, ad, sd)
        M = _qmat(q32.astype(np.float64))
        uh = M.T @ (u / np.linalg.norm(u))
        beta, alpha = np.arccos(uh[0]), np.arctan2(uh[2], uh[1])
        if alpha < 0.0:
            alpha, beta = alpha + np.pi, -beta
        ti = (beta + np.pi) * (nb - 1) / (2 * np.pi)
        tj = alpha * (na - 1) / np.pi
        want = _ramp_at(CD_RAMP, ti, tj)[:3]
        dv = st.vel[0, :3].astype(np.float64) / float(dt) - [0.0, 0.0, -9.8]
        got = M.T @ dv / (ta * np.dot(u, u))
        np.testing.assert_allclose(got, want, atol=2e-4)


# ----------------------------------------------------------------- host descriptors

def test_host_adm_and_rcs_descriptors_place_the_end_angles():
    """RS_set_adm_data (rs.c:3746-3783): beta -pi and +pi land on the centres of
    the first and last texel (0.5 and nb - 0.5), alpha 0 and pi likewise.
    RS_set_rcs_data (rs.c:3960-3992) has no half-texel offset: alpha -pi lands on
    0 and +pi on na - 1, beta 0 on 0 and pi on nb - 1."""
    r = sa.Radar(verb=0, host_only=True)
    h.configure(r, h.Config(n_rain=4096, debris=[(sa.OBJ_LEAF, 256)]))
    r.populate()
    cd = r.debris_table(0, 0)
    re = r.debris_table(0, 2)
    nb, na = cd.shape[1], cd.shape[0]
    ad = r.descriptor(16).astype(np.float64)
    np.testing.assert_allclose([-np.pi * ad[0] + ad[4], np.pi * ad[0] + ad[4]], [0.5, nb - 0.5], atol=2e-4)
    np.testing.assert_allclose([ad[5], np.pi * ad[1] + ad[5]], [0.5, na - 0.5], atol=2e-4)
    na2, nb2 = re.shape[1], re.shape[0]
    rd = r.descriptor(32).astype(np.float64)
    np.testing.assert_allclose([-np.pi * rd[0] + rd[4], np.pi * rd[0] + rd[4]], [0.0, na2 - 1], atol=2e-4)
    np.testing.assert_allclose([rd[5], np.pi * rd[1] + rd[5]], [0.0, nb2 - 1], atol=2e-4)
    r.free()


# ----------------------------------------------------------------- sample2d

def test_bilinear_reproduces_linear_fields_and_clamps():
    """Interior samples of a linear field are exact up to binary32 rounding;
    outside [0.5, n - 0.5] the edge texels are returned (clamp to edge)."""
    nx, ny = 19, 11
    ramp = np.array([[1.0, 0.25, -0.5], [-2.0, 0.125, 0.75], [0.5, -0.375, 0.0625], [3.0, 0.0, 1.0]])
    t, k = orc.tex(_ramp_table(ramp, nx, ny))
    g = np.random.default_rng(2)
    for _ in range(500):
        u, v = g.uniform(0.5, nx - 0.5), g.uniform(0.5, ny - 0.5)
        u32, v32 = np.float32(u), np.float32(v)
        got = orc.sample2d(t, u32, v32)
        np.testing.assert_allclose(got, _ramp_at(ramp, float(u32) - 0.5, float(v32) - 0.5), atol=2e-5)
    for u, v, ti, tj in [(-3.0, 4.7, 0.0, 4.2), (nx + 2.0, 4.7, nx - 1, 4.2), (7.3, -1.0, 6.8, 0.0),
                         (7.3, ny + 5.0, 6.8, ny - 1), (0.2, 0.1, 0.0, 0.0)]:
        np.testing.assert_allclose(orc.sample2d(t, np.float32(u), np.float32(v)), _ramp_at(ramp, ti, tj),
                                   atol=2e-5)


# ----------------------------------------------------------------- pow / exp and flux-field placement

def test_pow_and_exp_against_binary64():
    """exp(y log a) of the stretched flux-field placement: within a few ulp of
    binary64 pow for the grid ratios and cell offsets it is used with."""
    g = np.random.default_rng(3)
    a = g.uniform(0.9, 1.1, 200000).astype(np.float32)
    y = g.uniform(-200, 200, 200000).astype(np.float32)
    got = orc.math("pow", a, y).astype(np.float64)
    want = np.power(a.astype(np.float64), y.astype(np.float64))
    ulp = np.spacing(want.astype(np.float32)).astype(np.float64)
    err = np.abs(got - want) / ulp
    # t = y log a carries |t| 2^-24 from its own rounding and |t| 2^-23 from
    # log's ulp; e^t turns an absolute error of t into a relative one, and a
    # result ulp is at least 2^-24 of the result: at most 3 |t| + 2 ulp with
    # the exp's own rounding
    bound = 2.0 + 3.0 * np.abs(y.astype(np.float64) * np.log(a.astype(np.float64)))
    assert np.all(err <= bound), float(np.max(err - bound))
    x = g.uniform(-80, 80, 200000).astype(np.float32)
    e = np.abs(orc.math("exp", x).astype(np.float64) - np.exp(x.astype(np.float64)))
    assert np.max(e / np.spacing(np.exp(x).astype(np.float32))) <= 2.0


def _dff_desc(xs, xo, xm, kind, count):
    """Descriptor layout of rs.c:3392-3404."""
    d = np.zeros(16, np.float32)
    d[0] = d[1] = xs
    d[4] = d[5] = xo
    d[2] = d[6] = d[10] = -1.0
    d[8] = d[9] = xm
    d[13] = np.float32(xm) + np.float32(1.0)
    d[14] = kind
    d[15] = np.float32(count) - np.float32(1.0)
    return d


def _box():
    sd = np.zeros(16, np.float32)
    sd[8:11] = [-450.0, 1800.0, 0.0]
    sd[12:15] = [900.0, 700.0, 400.0]
    return sd


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_flux_field_placement_closed_forms(kind):
    """A uniform 5 x 5 PDF (icdf linear in r.x): the draw picks cell
    c = 25 r.x, column c mod 5, row floor(c / 5) + r.y, and the cell goes
    (0) onto a uniform grid of spacing s and origin at the box corner,
    (1) onto the stretched grid: offset rel from the centre cell maps to
        sign(rel) s (1 - q^|rel|) from the box centre, the node positions
        a (q^|i| - 1) / (q - 1) of a grid of first spacing a = -s (q - 1),
    (2) onto the whole box (checker board: col / 5 of the width)."""
    ncell = 5
    pdf = np.full(ncell * ncell, 1.0 / (ncell * ncell), np.float32)
    icdf = orc.dff_icdf_from_pdf(pdf)
    s, q = {0: (12.0, 0.0), 1: (-40.0, 1.05), 2: (1.0 / ncell, 0.0)}[kind]
    desc = _dff_desc(s, q, ncell - 1, float(kind), icdf.size)
    sd = _box()
    orc.set_dff(icdf, desc)
    try:
        g = np.random.default_rng(kind)
        for _ in range(300):
            r = g.uniform(0.0, 1.0, 4).astype(np.float32)
            got = orc.dff_place(r, sd).astype(np.float64)
            c = float(np.interp(float(r[0]) * (icdf.size - 1), np.arange(icdf.size), icdf.astype(np.float64)))
            row = np.floor(c / ncell)
            col, row = c - row * ncell, row + float(r[1])
            if kind == 0:
                want = [sd[8] + s * col, sd[9] + s * row]
            elif kind == 1:
                rel = np.array([col, row]) - 0.5 * (ncell - 1) - 0.5
                want = np.sign(rel) * s * (1.0 - q ** np.abs(rel)) + [sd[8] + 0.5 * sd[12], sd[9] + 0.5 * sd[13]]
                a = -s * (q - 1.0)
                k = 3
                node = a * (q ** k - 1.0) / (q - 1.0)          # the stretched grid's k-th node
                assert abs(s * (1.0 - q ** k) - node) < 1e-9
            else:
                want = [sd[8] + col * sd[12] / ncell, sd[9] + row * sd[13] / ncell]
            np.testing.assert_allclose(got[:2], want, rtol=2e-5, atol=2e-3)
            assert got[2] == 2.0
    finally:
        orc.set_dff(None, None)


# ----------------------------------------------------------------- LES time interpolation

def test_time_interpolated_wind_is_linear_in_the_weight():
    """Two uniform volumes u0, u1 and weight alpha: bg_atts sets the velocity
    to u0 + alpha (u1 - u0) (the sampler's lerp, within one rounding)."""
    u0, u1 = np.array([3.0, -2.0, 0.5]), np.array([-1.0, 4.0, 2.5])
    t0 = np.tile(np.append(u0, 0.0).astype(np.float32), (3, 3, 3, 1))
    t1 = np.tile(np.append(u1, 0.0).astype(np.float32), (3, 3, 3, 1))
    (w0, k0), (w1, k1) = orc.tex(t0), orc.tex(t1)
    les = np.zeros(16, np.float32)
    les[0:3] = 1.0
    sd = np.zeros(16, np.float32)
    sd[8:11] = [-100.0, -100.0, 0.0]
    sd[12:15] = [200.0, 200.0, 100.0]
    sd[11] = 0.01
    drop = np.zeros((96, 4), np.float32)
    dd = np.array([20000.0, -5.0, 95.0, 1.0], np.float32)
    try:
        for alpha in (0.0, 0.25, 0.5, 0.9, 1.0):
            orc.set_blend(w1, None, alpha)
            st = orc.State([[1.0, 2.0, 3.0, 1e-3]], [[0, 0, 0, 0]], [[0, 0, 0, 1]], [[0, 0, 0, 1]], [[0, 0, 0, 0]],
                           np.ones((1, 4), np.uint32))
            orc.bg_atts(st, 0, 1, w0, les, drop, dd, sd)
            np.testing.assert_allclose(st.vel[0, :3], u0 + alpha * (u1 - u0), rtol=1e-6, atol=1e-6)
        orc.set_blend(None)
        st = orc.State([[1.0, 2.0, 3.0, 1e-3]], [[0, 0, 0, 0]], [[0, 0, 0, 1]], [[0, 0, 0, 1]], [[0, 0, 0, 0]],
                       np.ones((1, 4), np.uint32))
        orc.bg_atts(st, 0, 1, w0, les, drop, dd, sd)
        np.testing.assert_array_equal(st.vel[0, :3], u0.astype(np.float32))
    finally:
        orc.set_blend(None)


def test_quat2abg_unit_component_rounded_past_one():
    """A radar-relative quaternion within a degree of the identity whose w
    rounds past 1 (a product of unit quaternions): rs.cl:323's acos(w) would
    be NaN; the oracle (and engine) take w as 1, so gamma = 0, and any w inside
    [-1, 1] gives the unclamped result."""
    w = np.nextafter(np.float32(1.0), np.float32(2.0))
    a, b, g = orc.quat2abg([1e-5, 0.0, 2e-5, float(w)])
    assert (a, b) == (0.0, 0.0) and g == 0.0
    a, b, g = orc.quat2abg([0.0, 0.0, 0.01, 0.99995])
    want = 2.0 * np.arccos(np.float32(0.99995))
    assert a == 0.0 and b == 0.0 and abs(g - want) < 2e-6
