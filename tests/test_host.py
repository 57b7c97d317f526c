"""Host-side checks that need no GPU: the C-ABI library exports every symbol its
headers declare, a reference-style C caller compiles and links against it and
fails the way RS_init does without a device, and the plain-C RS_* runtime's
geometry, population and tables equal the oracle's restatements of rs.c."""
import os
import re
import subprocess

import numpy as np
import pytest

import harness as h
import simradar_amd as sa
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    """Function names declared at file scope in a header (comments removed)."""
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set()
    for m in re.finditer(r"^[A-Za-z_][\w \t\*]*?\b((?:RS|rsd)_\w+)\s*\([^;{]*\)\s*;", text, flags=re.M):
        names.add(m.group(1))
    return names


@pytest.mark.parametrize("header", ["rs.h", "rsd.h"])
def test_library_exports_every_declared_symbol(header):
    names = _declared(header)
    assert len(names) > (50 if header == "rs.h" else 25), names
    lib = sa.load()
    missing = sorted(n for n in names if not hasattr(lib, n))
    assert not missing, missing


def test_python_bindings_cover_the_headers():
    declared = _declared("rs.h") | _declared("rsd.h")
    bound = set(sa.EXPORTED_RS) | set(sa.EXPORTED_RSD)
    assert set(sa.EXPORTED_RSD) <= declared
    assert {n for n in declared if n.startswith("rsd_")} <= bound


def test_reference_style_caller_compiles_and_reports_no_device(tmp_path):
    exe = str(tmp_path / "rs_caller")
    subprocess.run(["gcc", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c", "rs_caller.c"), "-L", os.path.join(ROOT, "simradar_amd"),
                    "-lsimradar_amd", "-Wl,-rpath," + os.path.join(ROOT, "simradar_amd"), "-lm", "-o", exe],
                   check=True)
    if sa.gpu_count() > 0:
        pytest.skip("a GPU is present: tests/test_c_caller_gpu.py runs the caller on it")
    p = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert p.returncode == 3, (p.returncode, p.stdout, p.stderr)
    assert "No GPU devices found" in p.stderr


def _host_radar(cfg):
    r = sa.Radar(verb=0, host_only=True)
    h.configure(r, cfg)
    return r


def test_c1_geometry_matches_the_suggested_domain():
    """RS_suggest_scan_domain + 15 m gates give R = 40 over [1850.25, 2420.25] m
    (rs.c:1592-1597, 5822-5904 with the suctvort LES geometry)."""
    r = _host_radar(h.Config(n_rain=131072))
    p = r.params()
    assert p.range_count == 40
    assert abs(p.range_start - 1850.25) < 0.02 and abs(p.range_end - 2420.25) < 0.02
    assert p.range_delta == 15.0
    sd = r.sim_desc()
    assert abs(sd[4] - 4 * np.pi / 0.1) < 1e-3          # 4 pi / lambda (rs.c:1462)
    assert sd[11] == np.float32(1e-3)                    # PRT = dt
    assert np.all(sd[12:15] > 0)
    r.free()


def test_host_populate_equals_oracle_populate():
    """RS_populate's host loop (glibc rand, DSD bin draw, seeds) reproduces the
    oracle's restatement of rs.c:4430-4958 bit for bit."""
    cfg = h.Config(n_rain=8192, debris=[(sa.OBJ_LEAF, 512)])
    r = _host_radar(cfg)
    r.populate()
    sd = r.sim_desc()
    inf = r.info()
    counts = [int(c) for c in inf.counts][:int(inf.num_types)]
    dr, dpdf, dcdf = r.dsd()
    st = orc.populate(counts, sd, cfg.seed, dr, dcdf, bool(cfg.concept & sa.CONCEPT_UNIFORM_DSD))
    for s in ("pos", "vel", "ori", "tum", "rnd"):
        a = r.stream(s)
        assert np.array_equal(a.view(np.uint32), np.ascontiguousarray(st[s]).view(np.uint32)), s
    r.free()


def test_host_tables_equal_oracle_restatements():
    """Drop RCS table (rs.c:5922-6029), angular weight (rs.c:2826-2851), range
    weight descriptor (rs.c:2620-2623) and radar gain (rs.c:5086)."""
    cfg = h.Config(n_rain=4096)
    r = _host_radar(cfg)
    r.populate()
    sd = r.sim_desc()
    p = r.params()
    dr, dpdf, _ = r.dsd()
    t, d = orc.drop_rcs_table(sd[4], sd[6], False, bool(cfg.concept & sa.CONCEPT_UNIFORM_DSD), dr, dpdf)
    np.testing.assert_array_equal(r.descriptor(4), d)
    w, wd = orc.angular_weight_standard(p.antenna_bw_rad)
    np.testing.assert_array_equal(r.descriptor(3), wd)
    np.testing.assert_array_equal(r.descriptor(2)[:3], orc.table1d_desc(-p.dr, p.dr, 3)[:3])
    g = float(r.lib.RS_radar_gain(r.H))
    assert g == pytest.approx(orc.radar_gain(p.antenna_gain_dbi, p.tx_power_watt), rel=1e-6)
    r.free()


def test_les_descriptor_is_the_stretched_grid_restatement():
    """les_desc slots of RS_set_vel_data_to_LES_table (rs.c:2944-2975, 3058-3061)."""
    r = _host_radar(h.Config(n_rain=4096))
    r.populate()
    frame = r.les_frame(0)
    ref = orc.les_desc(True, frame.shape[2], frame.shape[1], frame.shape[0], 2.0, 2.0, 2.0, 1.0212, 1.0212, 1.05, 50.0)
    np.testing.assert_array_equal(r.descriptor(1), ref)
    r.free()
