"""GPU parity: the HIP engine against the CPU oracle on identical seeded inputs.

Integer / RNG work and the position / velocity / orientation / seed streams are
compared bit for bit; per-gate IQ within 1e-4 of the RMS gate magnitude (FP32
signal terms, FP64 phase and accumulation on both sides); rcs streams within
2e-6 relative (the engine evaluates cos^2 of the elevation in closed form)."""
import numpy as np
import pytest

import harness as h
import simradar_amd as sa
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

IQ_TOL = 1e-4


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("fn,gen", [
    (1, lambda g: np.exp(g.uniform(-21, 0, 100000))),
    (2, lambda g: g.uniform(-20, 20, 100000)),
    (3, lambda g: g.uniform(-20, 20, 100000)),
    (5, lambda g: g.uniform(-1, 1, 100000)),
    (6, lambda g: np.exp(g.uniform(-30, 30, 100000))),
])
def test_device_math_matches_oracle(gpu_ready, fn, gen):
    g = np.random.default_rng(fn + 11)
    a = gen(g).astype(np.float32)
    dev = sa.math_probe(fn, a)
    if fn == 6:
        ref = np.sqrt(a)
    else:
        ref = orc.math({1: "log", 2: "sin", 3: "cos", 5: "acos"}[fn], a)
    bad = int(np.sum(_bits(dev) != _bits(ref)))
    assert bad == 0, "device function %d differs from the oracle at %d of %d inputs" % (fn, bad, a.size)


def test_device_atan2_div_match(gpu_ready):
    g = np.random.default_rng(5)
    y = g.normal(size=100000).astype(np.float32)
    x = g.normal(size=100000).astype(np.float32)
    assert np.array_equal(_bits(sa.math_probe(4, y, x)), _bits(orc.math("atan2", y, x)))
    assert np.array_equal(_bits(sa.math_probe(7, y, x)), _bits(y / x))


def _run(cfg, pulses, sort=None, reupload_at=None, fold=None, launch=None, setup=None, pre=None):
    r = sa.Radar(verb=0)
    h.configure(r, cfg)
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
    st = h.state_from_radar(r)     # host streams still hold the populated state
    o = h.OracleRun(r, cfg, st)
    sd_adv = r.sim_desc().copy()   # populate's zero-length step (rs.c:4941-4955)
    sd_adv[11] = 0.0
    o.advance(sd_adv)
    o.refresh_debris_rcs(r.sim_desc())
    if setup is not None:
        setup(r, o)
    pg, po = [], []
    for k in range(pulses):
        if k == reupload_at:
            # reference-order round trip of every stream through the host
            r.download()
            r.check()
            r.upload()
            r.check()
        if k > 0:
            r.make_pulse()
            o.refresh_debris_rcs(r.sim_desc())
        sd = r.sim_desc()
        r.download_pulse_only()
        r.check()
        pg.append(r.pulse_f64())
        po.append(o.pulse(sd))
        r.advance_time()
        o.advance(r.sim_desc())
        r.advance_beam()
    r.download()
    r.check()
    return r, o, np.array(pg), np.array(po)


def _check_state(r, o, streams=("pos", "vel", "ori", "tum")):
    for s in streams:
        a, b = r.stream(s), getattr(o.st, s)
        bad = np.nonzero(np.any(_bits(a) != _bits(b), axis=1))[0]
        assert bad.size == 0, "%s differs at %d scatterers, first %s: %s vs %s" % (s, bad.size, bad[:3], a[bad[:1]], b[bad[:1]])
    assert np.array_equal(r.stream("rnd"), o.st.rnd), "seeds differ"


def test_rain_step_parity_small(gpu_ready):
    cfg = h.Config(n_rain=8192, scan="P:3.0,-1:1:0.01")
    r, o, pg, po = _run(cfg, 12)
    _check_state(r, o, ("pos", "vel"))
    np.testing.assert_allclose(r.stream("rcs"), o.st.rcs, rtol=2e-6, atol=0)
    errs = [h.rel_err(pg[k], po[k]) for k in range(len(pg))]
    assert max(errs) < IQ_TOL, errs
    r.free()


def test_rain_c1_population(gpu_ready):
    """C1 population (131,072 rain), 20 pulses of the default sweep."""
    cfg = h.Config(n_rain=131072, scan="P:3.0,-12:12:0.01")
    r, o, pg, po = _run(cfg, 20)
    _check_state(r, o, ("pos", "vel"))
    errs = [h.rel_err(pg[k], po[k]) for k in range(len(pg))]
    assert max(errs) < IQ_TOL, errs
    r.free()


def test_debris_step_parity(gpu_ready):
    """Rain + leaf debris: 6-DoF debris state bit-exact, IQ within tolerance."""
    cfg = h.Config(n_rain=8192, debris=[(sa.OBJ_LEAF, 512)], scan="P:3.0,-1:1:0.01")
    r, o, pg, po = _run(cfg, 8)
    _check_state(r, o)
    n0 = o.counts[0]
    np.testing.assert_allclose(r.stream("rcs")[n0:], o.st.rcs[n0:], rtol=1e-5, atol=1e-9)
    errs = [h.rel_err(pg[k], po[k]) for k in range(len(pg))]
    assert max(errs) < IQ_TOL, errs
    r.free()


def test_three_debris_types_ride(gpu_ready):
    """Rain + leaf, metal sheet and brick: all three table-driven types ride in
    the rain launch; 6-DoF state bit-exact, IQ within tolerance."""
    cfg = h.Config(n_rain=8192, debris=[(sa.OBJ_LEAF, 700), (sa.OBJ_METAL_SHEET, 500), (sa.OBJ_BRICK, 300)],
                   scan="P:3.0,-1:1:0.01")
    r, o, pg, po = _run(cfg, 6)
    _check_state(r, o)
    errs = [h.rel_err(pg[k], po[k]) for k in range(len(pg))]
    assert max(errs) < IQ_TOL, errs
    r.free()


def _lift_debris(r, o):
    """Move every other debris body above the box (on the GPU and in the oracle)
    so that the next step respawns it."""
    r.download()
    r.check()
    n0 = o.counts[0]
    pos = r.stream("pos", copy=False)
    pos[n0::2, 2] = 1.0e5
    o.st.pos[n0::2, 2] = 1.0e5
    r.upload()
    r.check()
    return n0


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


@pytest.mark.parametrize("mode", ["board", "stretched"])
def test_debris_flux_field_respawn(gpu_ready, mode):
    """Concept C: respawned debris are placed through the flux-field inverse CDF
    (rs.cl:795-836).  "board" goes through the host runtime
    (RS_set_debris_flux_field_to_checker_board(9), rs.c:3462-3504); "stretched"
    sets a stretched-grid field through the C-ABI (rsd_set_dff).  State bit-exact
    with the oracle; after the first step every lifted body has respawned, and
    on the board almost all land in cells that carry flux."""
    pdf = np.zeros(81, np.float32)
    pdf[::2] = 100.0
    s = np.float32(0.0)
    for v in pdf:
        s = np.float32(s + v)
    pdf = (pdf / s).astype(np.float32)
    icdf = orc.dff_icdf_from_pdf(pdf)
    ninth = np.float32(1.0) / np.float32(9.0)
    lifted = {}

    def pre(r):
        # concepts enter the simulation descriptor when the population is made
        r.set_concept(lifted["concept"])

    def setup(r, o):
        assert int(np.asarray(r.sim_desc()[5], np.float32).view(np.uint32)) & 128
        if mode == "board":
            r.set_debris_flux_field_to_checker_board(9)
            desc = _dff_desc(ninth, -1.0, 8.0, 2.0, icdf.size)
        else:
            desc = _dff_desc(-400.0, 1.05, 8.0, 1.0, icdf.size)
            rc = sa.load().rsd_set_dff(r.engine(), icdf.ctypes.data, icdf.size, desc.ctypes.data)
            assert rc == 0, sa.load().rsd_last_error(r.engine())
        r.check()
        orc.set_dff(icdf, desc)
        lifted["n0"] = _lift_debris(r, o)

    cfg = h.Config(n_rain=4096, debris=[(sa.OBJ_LEAF, 2048)], scan="P:3.0,-1:1:0.01")
    r0 = sa.Radar(verb=0)      # the configuration's own concepts, plus C
    h.configure(r0, cfg)
    r0.populate()
    r0.check()
    lifted["concept"] = int(np.asarray(r0.sim_desc()[5], np.float32).view(np.uint32)) | 128
    r0.free()
    try:
        r, o, pg, po = _run(cfg, 1, setup=setup, pre=pre)
        _check_state(r, o)
        sd = r.sim_desc()
        pos = r.stream("pos")[lifted["n0"]::2]
        assert np.all(pos[:, 2] == np.float32(2.0)), "lifted bodies not respawned at MIN_HEIGHT"
        if mode == "board":
            ix = np.floor((pos[:, 0] - sd[8]) / (sd[12] / 9.0)).astype(int)
            iy = np.floor((pos[:, 1] - sd[9]) / (sd[13] / 9.0)).astype(int)
            assert np.mean((iy * 9 + ix) % 2 == 0) > 0.95
        r.free()
        r, o, pg, po = _run(cfg, 5, setup=setup, pre=pre)
        _check_state(r, o)
        # pulse 0 on the engine was made by the population, before the lift
        errs = [h.rel_err(pg[k], po[k]) for k in range(1, len(pg))]
        assert max(errs) < IQ_TOL, errs
        r.free()
    finally:
        orc.set_dff(None, None)


def test_concept_c_flux_field_follows_the_les_volumes(gpu_ready):
    """Concept C with the LES velocity (rs.c:3106-3111, 3565-3626): every LES
    volume the host loads rebuilds the debris flux field from the volume's
    second level and uploads it to that volume's slot, and respawned debris are
    placed through the active slot's field (dff_icdf[les_id], rs.c:5358).  Over
    a volume switch (every 200 pulses at a PRT of 10 ms), with debris lifted
    above the box on alternate pulses around it so that they respawn under
    both fields, the state stays bit-exact with the oracle taking each field
    from the host."""
    cfg = h.Config(n_rain=4096, debris=[(sa.OBJ_LEAF, 1024)], scan="P:3.0,-1:1:0.01", prt=0.01,
                   concept=h.Config().concept | sa.CONCEPT_DFF)
    fields = []

    def take_field(r):
        f = r.debris_flux_field(int(r.info().les_id))
        assert f is not None, "concept C left no flux field"
        orc.set_dff(*f)
        fields.append(f[0].copy())

    def setup(r, o):
        take_field(r)

    def on_pulse(k, r, o):
        if 190 <= k <= 210 and k % 2 == 0:
            _lift_debris(r, o)

    def on_switch(r, o):
        take_field(r)

    try:
        r, o, pg, po = h.run_pair(cfg, 215, setup=setup, on_pulse=on_pulse, les_switch=True, on_switch=on_switch)
        _check_state(r, o)
        assert len(fields) == 2, "expected one volume switch"
        assert not np.array_equal(fields[0], fields[1]), "the new volume's field equals the old one"
        errs = [h.rel_err(pg[k], po[k]) for k in range(len(pg))]
        assert max(errs) < IQ_TOL, errs
        r.free()
    finally:
        orc.set_dff(None, None)


def test_rain_sorted_every_pulse(gpu_ready):
    """Spatial reorder before every pulse: streams still come back bit-exact in
    the reference order and the gate sums are unchanged."""
    cfg = h.Config(n_rain=8192, debris=[(sa.OBJ_LEAF, 512)], scan="P:3.0,-1:1:0.01")
    r, o, pg, po = _run(cfg, 8, sort=1)
    _check_state(r, o)
    errs = [h.rel_err(pg[k], po[k]) for k in range(len(pg))]
    assert max(errs) < IQ_TOL, errs
    r.free()


def test_sorted_state_upload_roundtrip(gpu_ready):
    """Download + upload while the rain streams are reordered restores the
    reference order first; the run continues bit-exact."""
    cfg = h.Config(n_rain=8192, debris=[(sa.OBJ_LEAF, 256)], scan="P:3.0,-1:1:0.01")
    r, o, pg, po = _run(cfg, 10, sort=3, reupload_at=5)
    _check_state(r, o)
    errs = [h.rel_err(pg[k], po[k]) for k in range(len(pg))]
    assert max(errs) < IQ_TOL, errs
    r.free()


@pytest.mark.parametrize("reupload_at", [None, 8])
def test_range_sort_interval(gpu_ready, reupload_at):
    """Range reorders every 6 pulses over 20 pulses, with and without a
    download / upload round trip between two reorders: state bit-exact, IQ
    within tolerance."""
    cfg = h.Config(n_rain=20000, debris=[(sa.OBJ_LEAF, 300)], scan="P:3.0,-1:1:0.01")
    r, o, pg, po = _run(cfg, 20, sort=6, reupload_at=reupload_at)
    _check_state(r, o)
    errs = [h.rel_err(pg[k], po[k]) for k in range(len(pg))]
    assert max(errs) < IQ_TOL, errs
    r.free()


@pytest.mark.parametrize("fold", [0, 1, 2, 3])
@pytest.mark.parametrize("sort", [0, 1])
def test_fold_strategies(gpu_ready, fold, sort):
    """Every gate-sum strategy, on range-sorted and on reference-order streams."""
    cfg = h.Config(n_rain=16384, debris=[(sa.OBJ_LEAF, 256)], scan="P:3.0,-1:1:0.01")
    r, o, pg, po = _run(cfg, 4, sort=sort, fold=fold)
    _check_state(r, o)
    errs = [h.rel_err(pg[k], po[k]) for k in range(len(pg))]
    assert max(errs) < IQ_TOL, errs
    r.free()


@pytest.mark.parametrize("launch", [(256, 0), (256, 37), (128, 1016), (64, 3)])
def test_launch_shapes(gpu_ready, launch):
    """Block size and grid (0: two resident rounds; grids that are not a
    multiple of 8; fewer blocks than chunks of a wave) change only which block
    sums which scatterers: state bit-exact, IQ within tolerance."""
    cfg = h.Config(n_rain=20000, debris=[(sa.OBJ_LEAF, 300)], scan="P:3.0,-1:1:0.01")
    r, o, pg, po = _run(cfg, 5, sort=2, launch=launch)
    _check_state(r, o)
    errs = [h.rel_err(pg[k], po[k]) for k in range(len(pg))]
    assert max(errs) < IQ_TOL, errs
    r.free()


def test_threaded_oracle_is_the_oracle(gpu_ready):
    """The CPU baseline's threaded oracle: the advance is the single-thread
    oracle's bit for bit, the pulse differs only by binary64 reassociation."""
    cfg = h.Config(n_rain=6000, debris=[(sa.OBJ_LEAF, 300)], scan="P:3.0,-1:1:0.01")
    r = sa.Radar(verb=0)
    h.configure(r, cfg)
    r.populate()
    o1 = h.OracleRun(r, cfg, h.state_from_radar(r))
    o4 = h.OracleRun(r, cfg, h.state_from_radar(r), threads=4)
    for _ in range(3):
        sd = r.sim_desc()
        o1.advance(sd)
        o4.advance(sd)
        o1.refresh_debris_rcs(sd)
        o4.refresh_debris_rcs(sd)
        a, b = o1.pulse(sd), o4.pulse(sd)
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a))
        r.advance_beam()
    for s in ("pos", "vel", "ori", "tum", "rcs"):
        assert np.array_equal(_bits(getattr(o1.st, s)), _bits(getattr(o4.st, s))), s
    assert np.array_equal(o1.st.rnd, o4.st.rnd)
    r.free()


def test_unsorted_engine_parity(gpu_ready):
    """Sorting switched off: the stored order is the reference order throughout."""
    cfg = h.Config(n_rain=8192, scan="P:3.0,-1:1:0.01")
    r, o, pg, po = _run(cfg, 6, sort=0)
    _check_state(r, o, ("pos", "vel"))
    errs = [h.rel_err(pg[k], po[k]) for k in range(len(pg))]
    assert max(errs) < IQ_TOL, errs
    r.free()


def test_device_fast_acos_accuracy(gpu_ready):
    """The pulse path's acos (no division) stays within 3e-7 rad of binary64."""
    g = np.random.default_rng(21)
    c = np.concatenate([g.uniform(-1, 1, 100000), 1 - np.exp(g.uniform(-20, -1, 50000))]).astype(np.float32)
    err = np.abs(sa.math_probe(10, c).astype(np.float64) - np.arccos(c.astype(np.float64)))
    assert err.max() < 3e-7, err.max()


def test_wave_transpose_sum_layout(gpu_ready):
    """The gate window's wave reduction: lanes 2j and 2j + 1 of every wave end
    with the sum over the wave's 64 lanes of value j = x + j y (integers, so the
    sums are exact whatever the order)."""
    g = np.random.default_rng(5)
    n = 256
    x = g.integers(-1000, 1000, n).astype(np.float32)
    y = g.integers(-50, 50, n).astype(np.float32)
    out = sa.math_probe(11, x, y)
    lane = np.arange(n) % 64
    wave = np.arange(n) // 64
    j = lane >> 1
    xs = x.astype(np.float64).reshape(-1, 64).sum(1)
    ys = y.astype(np.float64).reshape(-1, 64).sum(1)
    want = xs[wave] + j * ys[wave]
    assert np.array_equal(out.astype(np.float64), want), (out[:16], want[:16])
    with pytest.raises(RuntimeError):
        sa.math_probe(11, x[:100], y[:100])


def test_les_frame_switch_is_stream_ordered(gpu_ready):
    """The LES frame switch (every 2 s of simulated time, here every 200 pulses
    of 10 ms) copies the next frame on the engine's stream while the queue is
    full: across two switches the state matches, bit for bit, a run that waits
    for the device before every step."""
    cfg = h.Config(n_rain=8192, debris=[(sa.OBJ_LEAF, 512)], scan="P:3.0,-1:1:0.01", prt=0.01)
    runs = []
    for drain in (False, True):
        r = sa.Radar(verb=0)
        h.configure(r, cfg)
        r.populate()
        r.check()
        for _ in range(450):
            r.make_pulse()
            if drain:
                r.sync()
            r.advance_time()
            r.advance_beam()
        r.make_pulse()
        r.download()
        r.check()
        runs.append({s: r.stream(s).copy() for s in ("pos", "vel", "ori", "tum")})
        runs[-1]["iq"] = r.pulse_f64()
        r.free()
    for s in ("pos", "vel", "ori", "tum"):
        assert np.array_equal(_bits(runs[0][s]), _bits(runs[1][s])), s
    assert np.array_equal(runs[0]["iq"], runs[1]["iq"])


def test_file_tables_parity(gpu_ready, tmp_path, monkeypatch):
    """Rain + leaves on tables read from reference-format files (LES records,
    .adm, .rcs) instead of the synthetic ones: state bit-exact, IQ within
    tolerance against the oracle on the same tables."""
    from test_tables import _write_adm, _write_les, _write_rcs
    g = np.random.default_rng(8)
    home = tmp_path / "home"
    (home / "tables" / "adm").mkdir(parents=True)
    (home / "tables" / "rcs").mkdir(parents=True)
    for n in ("plate", "square_plate"):
        _write_adm(str(home / "tables" / "adm" / (n + ".adm")), 73, 37, g)
    for n in ("leaf", "plate", "brick", "woodboard"):
        _write_rcs(str(home / "tables" / "rcs" / (n + ".rcs")), 73, 37, g)
    # the suctvort grid (the scan box follows the LES extent), one volume
    _write_les(str(home / "Documents" / "tables" / "les" / "suctvort"), 176, 176, 80, files=1, nvol=1, g=g)
    monkeypatch.setenv("HOME", str(home))
    monkeypatch.chdir(tmp_path)
    cfg = h.Config(n_rain=8192, debris=[(sa.OBJ_LEAF, 512)], scan="P:3.0,-1:1:0.01")
    r, o, pg, po = _run(cfg, 6)
    assert r.les_frame(0).shape == (80, 176, 176, 4)
    assert r.debris_table(0, 0).shape == (37, 73, 4)
    _check_state(r, o)
    errs = [h.rel_err(pg[k], po[k]) for k in range(len(pg))]
    assert max(errs) < IQ_TOL, errs
    r.free()
