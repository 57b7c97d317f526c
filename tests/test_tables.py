"""Table ingest: the providers read the reference's file formats (.adm
adm.c:152-225, .rcs rcs.c:181-246, LES grid and data records les.c:318-384,
474-517) from the folders the reference searches, and fall back to the
synthetic tables when none is found.  The files here are written by the test
in those layouts (no data files ship with the reference)."""
import os

import numpy as np
import pytest

import simradar_amd as sa

V0 = 100.0


def _write_adm(path, nb, na, g):
    planes = g.normal(size=(6, na, nb)).astype(np.float32)
    with open(path, "wb") as f:
        np.array([nb, na], np.uint16).tofile(f)
        planes.tofile(f)
    return planes


def _write_rcs(path, na, nb, g):
    planes = g.normal(size=(6, nb, na)).astype(np.float32)
    with open(path, "wb") as f:
        np.array([na, nb], np.uint16).tofile(f)
        planes.tofile(f)
    return planes


def _write_les(folder, nx, ny, nz, files, nvol, g):
    os.makedirs(folder)
    nn = nx * ny * nz
    pad = np.zeros(2, np.float32)
    with open(os.path.join(folder, "fort.10_2"), "wb") as f:
        np.array([1, nx, ny, nz], np.uint32).tofile(f)
        for _ in range(3):
            pad.tofile(f)
            g.uniform(size=nn).astype(np.float32).tofile(f)
    vols = []
    for k in range(files):
        with open(os.path.join(folder, "LES_mean_1_6_fnum%d.dat" % (k + 1)), "wb") as f:
            np.array([7], np.uint32).tofile(f)
            for v in range(nvol):
                np.array([0.5 * (k * nvol + v)], np.float32).tofile(f)
                pad.tofile(f)
                planes = g.normal(scale=0.1, size=(5, nz, ny, nx)).astype(np.float32)
                for p in planes:
                    p.tofile(f)
                    pad.tofile(f)
                vols.append(planes)
    return vols


@pytest.fixture
def tables_home(tmp_path, monkeypatch):
    home = tmp_path / "home"
    (home / "tables" / "adm").mkdir(parents=True)
    (home / "tables" / "rcs").mkdir(parents=True)
    monkeypatch.setenv("HOME", str(home))
    monkeypatch.chdir(tmp_path)
    return home / "tables"


def test_debris_tables_are_read_from_adm_and_rcs_files(tables_home):
    g = np.random.default_rng(3)
    adm = {n: _write_adm(str(tables_home / "adm" / (n + ".adm")), 37, 19, g) for n in ("plate", "square_plate")}
    rcs = {n: _write_rcs(str(tables_home / "rcs" / (n + ".rcs")), 41, 23, g) for n in ("leaf", "plate", "brick", "woodboard")}
    r = sa.Radar(host_only=True)
    r.add_debris(sa.OBJ_LEAF, 16)          # ADM "plate" + RCS "leaf" (obj.c:120-124)
    r.add_debris(sa.OBJ_METAL_SHEET, 16)   # ADM "square_plate" + RCS "plate"
    for t, (a, c) in enumerate((("plate", "leaf"), ("square_plate", "plate"))):
        cd, cm = r.debris_table(t, 0), r.debris_table(t, 1)
        re, im = r.debris_table(t, 2), r.debris_table(t, 3)
        assert cd.shape == (19, 37, 4) and re.shape == (23, 41, 4)
        for c3 in range(3):
            np.testing.assert_array_equal(cd[..., c3], adm[a][c3])
            np.testing.assert_array_equal(cm[..., c3], adm[a][3 + c3])
            np.testing.assert_array_equal(re[..., c3], rcs[c][c3])
            np.testing.assert_array_equal(im[..., c3], rcs[c][3 + c3])
    r.free()


def test_debris_tables_fall_back_to_synthetic_without_files(tmp_path, monkeypatch):
    monkeypatch.setenv("HOME", str(tmp_path))
    monkeypatch.chdir(tmp_path)
    r = sa.Radar(host_only=True)
    r.add_debris(sa.OBJ_LEAF, 16)
    assert r.debris_table(0, 0).shape == (181, 361, 4)
    r.free()


def test_les_frames_are_read_from_the_reference_records(tables_home):
    g = np.random.default_rng(4)
    nx, ny, nz = 12, 10, 6
    # the LES search list has $HOME/Documents/tables but not $HOME/tables (les.c:84-91)
    les_dir = tables_home.parent / "Documents" / "tables" / "les" / "suctvort"
    vols = _write_les(str(les_dir), nx, ny, nz, files=2, nvol=10, g=g)
    r = sa.Radar(host_only=True)
    r.set_vel_data_to_config("suctvort")
    for k in (0, 1, 13, 19, 0):
        fr = r.les_frame(k)
        assert fr.shape == (nz, ny, nx, 4)
        u, v, w, p, t = vols[k]
        np.testing.assert_array_equal(fr[..., 0], u * np.float32(V0))
        np.testing.assert_array_equal(fr[..., 1], v * np.float32(V0))
        np.testing.assert_array_equal(fr[..., 2], w * np.float32(V0))
        np.testing.assert_array_equal(fr[..., 3], t * np.float32(V0 * V0))
    r.free()
