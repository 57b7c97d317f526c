"""GPU parity of the engine's less-travelled branches against the oracle on
identical seeded inputs: the background advance without drag (bg_atts), the
DSD drawn from its CDF with unbounded drag, cross-polar debris returns,
uniform-grid winds (cube27 / cube125), non-triangle range-weight tables in both
gate folds, the binary32 phase, the signal / aux download streams, the pulse
test's known-answer inputs (test_make_pulse.c) through the C-ABI, the
transparent-rain superposition of test_wx_db_compare.sh, the LES volume switch,
a batched domain and a debris pose whose radar-relative quaternion rounds past
the identity.  State streams bit-exact, IQ within 1e-4 of the RMS gate
magnitude (tests/harness.py rel_err)."""
import ctypes

import numpy as np
import pytest

import harness as h
import simradar_amd as sa
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

IQ_TOL = 1e-4
CROSS_POL = 1 << 0
DRAGGED = sa.CONCEPT_DRAGGED
BOUNDED = sa.CONCEPT_BOUNDED
UNIFORM_DSD = sa.CONCEPT_UNIFORM_DSD
TRANSPARENT = 1 << 2


def _assert_iq(pg, po, first=0):
    errs = h.iq_errs(pg, po, first)
    assert max(errs) < IQ_TOL, errs


def test_background_advance_without_drag(gpu_ready):
    """Concept B | U without D: rain follows bg_atts (pos += vel dt without
    fma, vel = wind at the new position; rs.cl:533-580)."""
    cfg = h.Config(n_rain=8192, concept=BOUNDED | UNIFORM_DSD, scan="P:3.0,-1:1:0.01")
    r, o, pg, po = h.run_pair(cfg, 10)
    h.check_state(r, o, ("pos", "vel"))
    _assert_iq(pg, po)
    r.free()


def test_cdf_dsd_draw_and_unbounded_drag(gpu_ready):
    """Concept D alone: drop sizes drawn from the DSD's CDF at populate
    (rs.c:4775-4799) and no velocity bound (rs.cl:716 skipped)."""
    cfg = h.Config(n_rain=8192, concept=DRAGGED, scan="P:3.0,-1:1:0.01")
    r, o, pg, po = h.run_pair(cfg, 10)
    h.check_state(r, o, ("pos", "vel"))
    radii = np.unique(np.abs(r.stream("pos")[:, 3]))
    assert radii.size > 3, radii          # several DSD bins drawn
    _assert_iq(pg, po)
    r.free()


def test_cross_polar_debris(gpu_ready):
    """Concept bit 0: vh and hv add into the H and V returns of every debris
    body (rs.cl:350-356), for three debris types riding in the rain launch."""
    cfg = h.Config(n_rain=4096, debris=[(sa.OBJ_LEAF, 700), (sa.OBJ_METAL_SHEET, 500), (sa.OBJ_BRICK, 300)],
                   concept=DRAGGED | BOUNDED | UNIFORM_DSD | CROSS_POL, scan="P:3.0,-1:1:0.01")
    r, o, pg, po = h.run_pair(cfg, 6)
    h.check_state(r, o)
    n0 = o.counts[0]
    np.testing.assert_allclose(r.stream("rcs")[n0:], o.st.rcs[n0:], rtol=1e-5, atol=1e-9)
    _assert_iq(pg, po)
    r.free()


def _cube(n, v, origin_mode, sd):
    """cube27 (n = 3, v = 10 m/s) and cube125 (n = 5, v = 0.5 m/s) tables of
    rs.c:3159-3234 restated: a linear field of the texel indices, and the
    uniform-grid descriptor (scale n / size, origin 1 or the box's corner,
    maximum n - 1)."""
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    off = v if n == 3 else 2.0 * v
    tex = np.stack([i * v - off, j * v - off, k * v - off, np.zeros_like(i, float)], -1).astype(np.float32)
    d = np.zeros(16, np.float32)
    size, org = sd[12:15], sd[8:11]
    d[0:3] = np.float32(n) / size
    if origin_mode == "one":
        d[4:7] = [1.0, 1.0, 0.0]
    else:
        d[4:7] = -org * d[0:3]
    d[8:11] = n - 1
    d[3] = 1000.0
    return tex, d


@pytest.mark.parametrize("n", [3, 5])
def test_uniform_grid_wind(gpu_ready, n):
    """RS_set_vel_data_to_cube27 / cube125: the uniform-grid index fma(pos,
    scale, origin) (rs.cl:173) and trilinear sampling of a linear field."""
    def pre(r):
        r.lib.RS_set_vel_data_to_cube27(r.H) if n == 3 else r.lib.RS_set_vel_data_to_cube125(r.H)

    cfg = h.Config(n_rain=8192, scan="P:3.0,-1:1:0.01")
    probe = sa.Radar(verb=0, host_only=True)
    h.configure(probe, cfg)
    probe.populate()
    sd = probe.sim_desc()
    probe.free()
    tex, desc = _cube(n, 10.0 if n == 3 else 0.5, "one" if n == 3 else "corner", sd)
    r, o, pg, po = h.run_pair(cfg, 8, pre=pre, oracle_kw={"wind": (tex, desc)})
    h.check_state(r, o, ("pos", "vel"))
    _assert_iq(pg, po)
    r.free()


@pytest.mark.parametrize("table", ["cosine", "bell", "dense"])
@pytest.mark.parametrize("sort", [0, 32])
def test_range_weight_tables(gpu_ready, table, sort):
    """Tables whose ends are zero take the wave-shared window (its LDS weight
    pairs): a 9-entry raised cosine over +-1.5 gates and a 5-entry bell (0, .7,
    1, .7, 0) over +-1 gate; a table whose ends do not vanish weighs every gate
    (the dense fold)."""
    if table == "cosine":
        w = (0.5 - 0.5 * np.cos(2 * np.pi * np.arange(9) / 8)).astype(np.float32)
        w[0] = w[-1] = 0.0
        half = 22.5
    elif table == "bell":
        w = np.array([0.0, 0.7, 1.0, 0.7, 0.0], np.float32)
        half = 15.0
    else:
        w = np.array([0.2, 0.6, 1.0, 0.6, 0.2], np.float32)
        half = 15.0

    def pre(r):
        r.set_range_weight(w, -half, 2 * half / (len(w) - 1))

    cfg = h.Config(n_rain=12000, debris=[(sa.OBJ_LEAF, 256)], scan="P:3.0,-1:1:0.01")
    desc = orc.table1d_desc(-half, np.float32(2 * half / (len(w) - 1)), len(w))
    r, o, pg, po = h.run_pair(cfg, 5, pre=pre, sort=sort, oracle_kw={"rangew": (w, desc)})
    h.check_state(r, o)
    _assert_iq(pg, po)
    r.free()


def _f32_fma(a, b, c):
    """fl32(a b + c): the binary32 product is exact in binary64 and, for the
    magnitudes used here, so is the sum."""
    return np.float32(np.float64(a) * np.float64(b) + np.float64(c))


@pytest.mark.parametrize("fold", [sa.FOLD_WINDOW, sa.FOLD_WAVE32])
def test_range_weight_on_table_nodes(gpu_ready, fold):
    """Scatterers exactly on the nodes of a 5-entry bell (0, .7, 1, .7, 0) over
    +-15 m and a few ulps either side.  The table read (rs.cl:1192-1206) takes
    its upper entry at floor(fma(x, s0, s1 + 1)); at x = -7.5 m, x0 =
    fma(x, s0, s1) rounds to just below 1 while that index rounds up to 3, so
    the weight is mix(W[0], W[2], 0.99999994), not W[1].  Both gate folds give
    the oracle's pulse."""
    lib = sa.load()
    R, start, delta = 16, np.float32(1000.0), np.float32(15.0)
    rw = np.array([0.0, 0.7, 1.0, 0.7, 0.0], np.float32)
    rwd = orc.table1d_desc(-15.0, np.float32(7.5), 5)
    gr = orc.gate_ranges(start, delta, R)
    rr = []
    for g in range(R):
        for off in (-15.0, -7.5, 0.0, 7.5, 15.0):
            r0 = np.float32(gr[g] + np.float32(off))
            for u in range(-3, 4):
                rr.append(np.nextafter(r0, np.float32(np.inf if u > 0 else -np.inf)) if u else r0)
                for _ in range(abs(u) - 1):
                    rr[-1] = np.nextafter(rr[-1], np.float32(np.inf if u > 0 else -np.inf))
    rr = np.array(rr, np.float32)
    n = rr.size
    # the case above occurs: the upper index lies two entries above the lower
    x = rr[:, None] - gr[None, :]
    x0 = np.clip(_f32_fma(x, rwd[0], rwd[1]), 0.0, rwd[2])
    x1 = np.clip(_f32_fma(x, rwd[0], rwd[1] + np.float32(1.0)), 0.0, rwd[2])
    assert np.any(np.floor(x1) > np.floor(x0) + 1)
    pos = np.stack([np.zeros(n), rr, np.zeros(n), np.full(n, 1e-3)], -1).astype(np.float32)
    rcs = np.tile(np.array([1.0, 0.5, 1.0, 0.5], np.float32), (n, 1))
    aw = np.array([0.0, 0.7, 1.0, 0.7, 0.0], np.float32)
    awd = orc.table1d_desc(np.float32(-0.35), np.float32(0.175), 5)
    drop = np.tile(np.array([1.0, 0.5, 1.0, 0.5], np.float32), (96, 1))
    ddesc = np.array([20000.0, -5.0, 95.0, 1.0], np.float32)
    sd = np.zeros(16, np.float32)
    sd[1] = 1.0
    sd[4] = np.float32(4.0 * np.pi / 0.1)
    sd[8:11] = [-2000.0, -2000.0, -2000.0]
    sd[12:15] = [4000.0, 4000.0, 4000.0]
    want = orc.make_pulse(pos, rcs, aw, awd, rw, rwd, float(start), float(delta), R, sd, 4.0 * np.pi / 0.1,
                          orc.PHASE_F32)
    e = ctypes.c_void_p()
    assert lib.rsd_open(0, ctypes.byref(e)) == 0
    try:
        def ok(rc):
            assert rc == 0, lib.rsd_last_error(e).decode()
        origins = (ctypes.c_size_t * 8)(*([0] * 8))
        counts = (ctypes.c_size_t * 8)(*([n] + [0] * 7))
        ok(lib.rsd_alloc_state(e, n, origins, counts))
        z = np.zeros((n, 4), np.float32)
        seeds = np.ones((n, 4), np.uint32)
        ok(lib.rsd_upload_state(e, pos.ctypes.data, z.ctypes.data, z.ctypes.data, z.ctypes.data, z.ctypes.data,
                                rcs.ctypes.data, seeds.ctypes.data))
        ok(lib.rsd_set_drop_rcs(e, drop.ctypes.data, 96, ddesc.ctypes.data))
        ok(lib.rsd_set_range_weight(e, rw.ctypes.data, 5, rwd.ctypes.data))
        ok(lib.rsd_set_angular_weight(e, aw.ctypes.data, 5, awd.ctypes.data))
        ok(lib.rsd_set_sort(e, 0))
        ok(lib.rsd_set_fold(e, fold))
        p = sa.RSDPulse()
        p.sim_desc[:] = sd.tolist()
        p.two_over_lambda = 2.0 / 0.1
        p.range_start, p.range_delta, p.range_count = float(start), float(delta), R
        p.phase_mode = sa.PHASE_F32
        ok(lib.rsd_make_pulse(e, ctypes.byref(p)))
        out = np.zeros(4 * R, np.float64)
        ok(lib.rsd_read_pulse(e, out.ctypes.data, R))
    finally:
        lib.rsd_close(e)
    assert h.rel_err(out.reshape(R, 4), want) < IQ_TOL, h.rel_err(out.reshape(R, 4), want)


def test_binary32_phase(gpu_ready):
    """RSD_PHASE_F32: phase r (4 pi / lambda) in binary32 as rs.cl:1096; the
    oracle's binary32 phase gives the same IQ."""
    cfg = h.Config(n_rain=8192, debris=[(sa.OBJ_LEAF, 256)], scan="P:3.0,-1:1:0.01", phase=sa.PHASE_F32)
    r, o, pg, po = h.run_pair(cfg, 6, pulse_mode=orc.PHASE_F32)
    h.check_state(r, o)
    _assert_iq(pg, po)
    r.free()


@pytest.mark.parametrize("phase", ["f32", "f64"])
def test_signal_and_aux_downloads(gpu_ready, phase):
    """RS_download's sig and aux streams are scat_sig_aux's of the last pulse
    (rs.cl:1050-1110): range, age, bin and angular weight in aux, rcs e^{-j
    phase} / r^2 in sig.  After the next advance they still belong to that
    pulse while the positions have moved."""
    f32 = phase == "f32"
    cfg = h.Config(n_rain=6000, debris=[(sa.OBJ_LEAF, 300)], scan="P:3.0,-1:1:0.01",
                   phase=sa.PHASE_F32 if f32 else sa.PHASE_F64)
    seen = {}

    def on_pulse(k, r, o):
        if k != 3:
            return
        r.download()
        r.check()
        sd = r.sim_desc()
        seen["sig"], seen["aux"], seen["pos"] = r.stream("sig"), r.stream("aux"), r.stream("pos")
        sig, aux = orc.scat_sig_aux(o.st.pos, o.st.rcs, o.st.aux, o.angw, o.angw_desc, sd)
        np.testing.assert_array_equal(seen["pos"], o.st.pos)
        np.testing.assert_allclose(seen["aux"][:, 0], aux[:, 0], rtol=1e-7)
        np.testing.assert_array_equal(seen["aux"][:, 1:3], aux[:, 1:3])
        np.testing.assert_allclose(seen["aux"][:, 3], aux[:, 3], atol=2e-5)
        if not f32:
            # binary64 phase of the binary32 range: rcs e^{-j 4 pi r / lambda} / r^2
            # with the binary32 wavelength of the radar parameters
            rr = aux[:, 0].astype(np.float64)
            e = np.exp(-1j * (4 * np.pi / float(np.float32(0.1))) * rr) / rr ** 2
            x = o.st.rcs.astype(np.float64)
            sig = np.stack([((x[:, 0] + 1j * x[:, 1]) * e).real, ((x[:, 0] + 1j * x[:, 1]) * e).imag,
                            ((x[:, 2] + 1j * x[:, 3]) * e).real, ((x[:, 2] + 1j * x[:, 3]) * e).imag], -1)
        scale = np.max(np.abs(sig))
        assert np.max(np.abs(seen["sig"] - sig)) < 2e-5 * scale

    r, o, pg, po = h.run_pair(cfg, 4, on_pulse=on_pulse,
                              pulse_mode=orc.PHASE_F32 if f32 else orc.PHASE_F64)
    # run_pair ended with an advance after pulse 3 and a download
    assert not np.array_equal(r.stream("pos"), seen["pos"])
    np.testing.assert_array_equal(r.stream("sig"), seen["sig"])
    np.testing.assert_array_equal(r.stream("aux"), seen["aux"])
    r.free()


def test_make_pulse_known_answer_through_the_engine(gpu_ready):
    """test_make_pulse.c (rs.cl:1444-1453 inputs): 1,079,296 scatterers on a
    9 x 120,000 lattice 1..13 m from the radar, 9 gates at 1 + 0.25 k m, a
    +-25 m triangle (every scatterer weighs into all 9 gates, more than the
    8-gate wave window holds), angular table {0, .7, 1, .7, 0} over +-20 deg,
    binary32 phase.  The engine's pulse equals the oracle's within 1e-4 and
    the reference's own CPU loop within its 1e-3 (test_make_pulse.c:645)."""
    lib = sa.load()
    n = 1079296
    k = np.arange(n)
    pos = np.stack([(k % 9) - 4.0, k / 90000.0 + 1.0, np.full(n, 0.1), np.full(n, 1e-3)], -1).astype(np.float32)
    rcs = np.tile(np.array([1.0, 0.5, 1.0, 0.5], np.float32), (n, 1))
    rw = np.array([0.0, 1.0, 0.0], np.float32)
    xs = np.float32(1.0) / np.float32(25.0)
    rwd = np.array([xs, np.float32(25.0) * xs, 2.0, 0.0], np.float32)
    aw = np.array([0.0, 0.7, 1.0, 0.7, 0.0], np.float32)
    d10 = np.float32(10.0 * 0.01745329251994)
    awd = np.array([np.float32(1.0) / d10, np.float32(20.0 * 0.01745329251994) * (np.float32(1.0) / d10), 4.0, 0.0],
                   np.float32)
    # drop table whose every row is (1, 0.5 | 1, 0.5): the rain rcs is the KAT's
    drop = np.tile(np.array([1.0, 0.5, 1.0, 0.5], np.float32), (96, 1))
    ddesc = np.array([20000.0, -5.0, 95.0, 1.0], np.float32)
    sd = np.zeros(16, np.float32)
    sd[1] = 1.0
    sd[4] = np.float32(4.0 * np.pi / 0.1)
    sd[8:11] = [-100.0, -100.0, -100.0]
    sd[12:15] = [200.0, 200.0, 200.0]
    want_o = orc.make_pulse(pos, rcs, aw, awd, rw, rwd, 1.0, 0.25, 9, sd, 4.0 * np.pi / 0.1, orc.PHASE_F32)
    for sort in (0, 1):
        e = ctypes.c_void_p()
        assert lib.rsd_open(0, ctypes.byref(e)) == 0
        try:
            origins = (ctypes.c_size_t * 8)(*([0] * 8))
            counts = (ctypes.c_size_t * 8)(*([n] + [0] * 7))

            def ok(rc):
                assert rc == 0, lib.rsd_last_error(e).decode()
            ok(lib.rsd_alloc_state(e, n, origins, counts))
            z = np.zeros((n, 4), np.float32)
            seeds = np.ones((n, 4), np.uint32)
            # an uploaded rcs stream holds until the next advance recomputes it
            ok(lib.rsd_upload_state(e, pos.ctypes.data, z.ctypes.data, z.ctypes.data, z.ctypes.data, z.ctypes.data,
                                    rcs.ctypes.data, seeds.ctypes.data))
            ok(lib.rsd_set_drop_rcs(e, drop.ctypes.data, 96, ddesc.ctypes.data))
            ok(lib.rsd_set_range_weight(e, rw.ctypes.data, 3, rwd.ctypes.data))
            ok(lib.rsd_set_angular_weight(e, aw.ctypes.data, 5, awd.ctypes.data))
            ok(lib.rsd_set_sort(e, sort))
            p = sa.RSDPulse()
            p.sim_desc[:] = sd.tolist()
            p.two_over_lambda = 2.0 / 0.1
            p.range_start, p.range_delta, p.range_count = 1.0, 0.25, 9
            p.phase_mode = sa.PHASE_F32
            ok(lib.rsd_make_pulse(e, ctypes.byref(p)))
            out = np.zeros(36, np.float64)
            ok(lib.rsd_read_pulse(e, out.ctypes.data, 9))
        finally:
            lib.rsd_close(e)
        got = out.reshape(9, 4)
        assert h.rel_err(got, want_o) < IQ_TOL, (sort, h.rel_err(got, want_o))
        # the reference's CPU loop (test_make_pulse.c:553-583) in binary64
        pp = pos[:, :3].astype(np.float64)
        rr = np.sqrt((pp ** 2).sum(1))
        ang = np.arccos(pp[:, 1] / rr)
        idx = ang * float(awd[0]) + float(awd[1])
        fl = np.floor(idx)
        i0, i1 = np.clip(fl, 0, 4).astype(int), np.clip(fl + 1, 0, 4).astype(int)
        wa = aw[i0] + (idx - fl) * (aw[i1].astype(np.float64) - aw[i0])
        ph = rr * float(sd[4])
        s = np.exp(-1j * ph) / rr ** 2
        want = np.zeros((9, 4))
        for ir in range(9):
            x = (rr - (ir * 0.25 + 1.0)) * float(xs) + 1.0
            wr = np.interp(x, [0.0, 1.0, 2.0], [0.0, 1.0, 0.0])
            t = wr * wa * s
            want[ir] = [(t * (1 + 0.5j)).real.sum(), (t * (1 + 0.5j)).imag.sum(),
                        (t * (1 + 0.5j)).real.sum(), (t * (1 + 0.5j)).imag.sum()]
        assert h.rel_err(got, want) < 1e-3


def test_transparent_rain_superposition(gpu_ready):
    """test_wx_db_compare.sh: IQ(rain + debris) = IQ(rain alone) + IQ(rain +
    debris with concept T), where concept T zeroes the drop table (rs.c:5948-
    5949) so only the debris return.  Populations from one seed share the rain
    state, so the three runs step the same scatterers."""
    base = DRAGGED | BOUNDED | UNIFORM_DSD
    runs = {}
    for name, debris, concept in (("both", [(sa.OBJ_LEAF, 1024)], base), ("rain", [], base),
                                  ("debris", [(sa.OBJ_LEAF, 1024)], base | TRANSPARENT)):
        cfg = h.Config(n_rain=8192, debris=debris, concept=concept, scan="P:3.0,-1:1:0.01")
        r, o, pg, po = h.run_pair(cfg, 5)
        h.check_state(r, o)
        _assert_iq(pg, po)
        runs[name] = (pg, r.stream("pos")[:8192].copy())
        r.free()
    assert np.array_equal(runs["both"][1], runs["rain"][1])
    assert np.array_equal(runs["both"][1], runs["debris"][1])
    both, rain, deb = runs["both"][0], runs["rain"][0], runs["debris"][0]
    assert np.max(np.abs(deb)) > 0.0 and np.max(np.abs(rain)) > 0.0
    for k in range(len(both)):
        assert h.rel_err(rain[k] + deb[k], both[k]) < IQ_TOL


def test_les_volume_switch_against_the_oracle(gpu_ready):
    """Every tp = 2 s of simulated time the host loads the next LES volume
    into the other slot (rs.c:5189-5203); with a PRT of 10 ms that is every 200
    pulses.  Across two switches the state stays bit-exact with the oracle
    stepping the same volumes."""
    cfg = h.Config(n_rain=4096, debris=[(sa.OBJ_LEAF, 256)], scan="P:3.0,-1:1:0.01", prt=0.01)
    switches = []

    def on_pulse(k, r, o):
        switches.append(int(r.info().vel_idx))

    r, o, pg, po = h.run_pair(cfg, 430, les_switch=True, on_pulse=on_pulse)
    assert len(set(switches)) >= 3, sorted(set(switches))
    h.check_state(r, o)
    _assert_iq(pg, po)
    r.free()


def test_batched_domain_against_the_oracle(gpu_ready):
    """Two rain domains of different sizes in one persistent launch per round
    (RS_batch_round): each steps bit-exact with its own oracle and its pulses
    agree within 1e-4."""
    from simradar_amd import workloads as wl
    cfgs = wl.c4(n_domains=2, seed=17, scale=1.0 / 16)
    rs = [wl.make(c) for c in cfgs]
    os_ = []
    for r, c in zip(rs, cfgs):
        o = h.OracleRun(r, c, h.state_from_radar(r))
        sd = r.sim_desc().copy()
        sd[11] = 0.0
        o.advance(sd)
        os_.append(o)
    b = sa.Batch(rs)
    b.set_launch(1024, 4)
    for _ in range(6):
        sds = [r.sim_desc() for r in rs]
        b.round()
        for r, o, sd in zip(rs, os_, sds):
            r.download_pulse_only()
            r.check()
            assert h.rel_err(r.pulse_f64(), o.pulse(sd)) < IQ_TOL
            o.advance(sd)
    for r, o in zip(rs, os_):
        r.download()
        r.check()
        h.check_state(r, o, ("pos", "vel"))
    b.free()
    for r in rs:
        r.free()


def test_fixed_position_profiler(gpu_ready):
    """Concepts F | V (simple_dbs.c): a vertically pointing profiler over
    fixed scatterers on a zenith raster, DBS beams at 75 and 90 deg.  Each
    advance is fp_atts (rs.cl:590-648): velocity = wind, phase turned by the
    radial wind, rcs = 10^cn2 e^{j phi} from the Cn^2 volume.  pos / vel / rcs
    bit-exact; H and V IQ each within 1e-4 (the V half of a fixed scatterer's
    rcs carries (cn2, phi), as the reference's)."""
    cfg = h.Config(n_rain=0, concept=sa.CONCEPT_FIXED | sa.CONCEPT_VERTICAL,
                   scan="D:0,75,3/90,75,3/0,90,3")
    r, o, pg, po = h.run_pair(cfg, 8)
    assert o.fixed and int(r.info().num_scats) > 100000
    h.check_state(r, o, ("pos", "vel", "rcs"))
    assert np.max(np.abs(pg[..., :2])) > 0.0
    for half in (slice(0, 2), slice(2, 4)):
        errs = [h.rel_err(pg[k][:, half], po[k][:, half]) for k in range(len(pg))]
        assert max(errs) < IQ_TOL, errs
    r.free()


@pytest.mark.parametrize("debris", [False, True])
def test_les_time_interpolation_against_the_oracle(gpu_ready, debris):
    """RS_set_les_time_interpolation: volumes k and k + 1 resident in the two
    slots and the wind lerp(sample_k, sample_k+1, (t - k tp) / tp) for rain and
    debris; across two volume boundaries (PRT 10 ms, tp 2 s) the state stays
    bit-exact with the oracle blending the same volumes."""
    cfg = h.Config(n_rain=4096, debris=[(sa.OBJ_LEAF, 256)] if debris else [], scan="P:3.0,-1:1:0.01", prt=0.01,
                   les_interp=True)
    alphas = []

    def on_pulse(k, r, o):
        alphas.append(float(r.info().wind_alpha))

    r, o, pg, po = h.run_pair(cfg, 430, on_pulse=on_pulse)
    assert r.info().les_blend == 1
    a = np.array(alphas[1:])
    assert a.min() < 0.01 and a.max() > 0.99 and np.sum(np.diff(a) < 0) == 2, "two volume boundaries"
    h.check_state(r, o)
    _assert_iq(pg, po)
    r.free()


def _radar_rel_w(p, o):
    """w of R = ori (x) o_conj(el, az) (rs.cl:287-315) in the oracle's binary32
    arithmetic."""
    f = np.float32
    one = lambda fn, *x: orc.math(fn, *[np.array([v], f) for v in x])[0]
    el = one("atan2", p[2], np.sqrt(f(f(p[0] * p[0]) + f(p[1] * p[1]))))
    az = one("atan2", p[0], p[1])
    a1, a2 = f(f(0.5) * f(el + f(np.pi / 2))), f(f(0.5) * az)
    se, ce, sa_, ca = one("sin", a1), one("cos", a1), one("sin", a2), one("cos", a2)
    s2 = f(0.70710677)
    oc = [f(f(f(-se * ca) + f(ce * sa_)) * s2), f(f(f(ce * ca) + f(se * sa_)) * s2),
          f(f(f(se * ca) + f(ce * sa_)) * s2), f(f(f(ce * ca) + f(-se * sa_)) * s2)]
    return orc.qmul(o, oc)[3]


def test_debris_at_the_identity_pose_stays_finite(gpu_ready):
    """A leaf straight ahead of the radar (az ~ 0, el ~ 0) near the orientation
    RS_populate gives every body: its radar-relative quaternion is the identity
    up to rounding and w rounds past 1, where the reference's acos (rs.cl:323)
    gives NaN (and, through the dense pass 1, NaN in every gate).  Engine and
    oracle take w as 1: the refreshed rcs and the pulse are finite and agree."""
    pos = np.float32([-0.4720836, 1925.8678, 1.2334673, 0.0])
    ori = np.float32([0.50010043, -0.49989828, -0.5000516, 0.49994978])
    assert _radar_rel_w(pos, ori) > 1.0          # the pose exercises the clamp
    cfg = h.Config(n_rain=4096, debris=[(sa.OBJ_LEAF, 64)], scan="P:3.0,-1:1:0.01")
    r = sa.Radar(verb=0)
    h.configure(r, cfg)
    r.populate()
    r.check()
    n0 = int(r.info().counts[0])
    for s, v in (("pos", pos), ("ori", ori)):
        r.stream(s, copy=False)[n0] = v
    r.upload()
    o = h.OracleRun(r, cfg, h.state_from_radar(r))
    r.advance_beam()
    r.make_pulse()                  # the beam moved: debris rcs refreshed at the placed pose
    sd = r.sim_desc()
    r.download_pulse_only()
    r.check()
    g = r.pulse_f64()
    o.refresh_debris_rcs(sd)
    ref = o.pulse(sd, orc.PHASE_F64)
    r.download()
    rcs = r.stream("rcs")
    assert np.all(np.isfinite(rcs)) and np.all(np.isfinite(o.st.rcs))
    np.testing.assert_array_equal(rcs[n0].view(np.uint32), o.st.rcs[n0].view(np.uint32))
    assert np.all(np.isfinite(g)) and np.all(np.isfinite(ref))
    assert h.rel_err(g, ref) < IQ_TOL
    r.free()
