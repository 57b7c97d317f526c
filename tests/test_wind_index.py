"""Stretched-grid wind index through per-axis cubic tables (CPU).

Every advance evaluates the index c = s8 + copysign(s0 log1p(s4 |r|), r) of
rs.cl:130-151 from 256 cubic pieces per axis (include/rsd.h
rsd_wind_index_table).  Pinned here: the engine's host builder equals the
oracle's restatement bit for bit; each piece reproduces binary64 P at its four
Chebyshev nodes; over positions inside the box the tables stay within 3e-5
texel of binary64 s8 + copysign(s0 log1p(s4 |r|), r), which is the size of the
binary32 form's own rounding there (about one ulp of indices up to 176)."""
import numpy as np
import pytest

import harness as h
import simradar_amd as sa
from simradar_amd import workloads as wl
from oracle import oracle as orc

LES = orc.les_desc(True, 176, 176, 80, 2.0, 2.0, 2.0, 1.0212, 1.0212, 1.05, 50.0)
COS1, COS3 = np.cos(np.pi / 8), np.cos(3 * np.pi / 8)
NODES = np.array([0.5 + 0.5 * COS1, 0.5 + 0.5 * COS3, 0.5 - 0.5 * COS3, 0.5 - 0.5 * COS1])


def _box(cfg):
    r = sa.Radar(verb=0, host_only=True)
    wl.configure(r, cfg)
    sd = r.sim_desc().copy()
    r.free()
    return sd


CONFIGS = {"c2": lambda: wl.c2(), "c3": lambda: wl.c3(n_rain=65536), "c1": lambda: h.Config(n_rain=131072)}


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_engine_tables_equal_the_oracle_restatement(name):
    sd = _box(CONFIGS[name]())
    rows, ih, jm = sa.wind_index_table(LES, sd)
    orows, oih, ojm = orc.wind_index_table(LES, sd)
    assert np.array_equal(rows.view(np.uint32), orows.view(np.uint32))
    assert np.array_equal(ih.view(np.uint32), oih.view(np.uint32)) and np.array_equal(jm, ojm)
    assert np.all(jm == 255) and np.all(np.isfinite(rows))
    # the pieces cover the box on every axis: 256 / ih >= the extent from the origin
    cx, cy = np.float32(sd[8] + np.float32(0.5) * sd[12]), np.float32(sd[9] + np.float32(0.5) * sd[13])
    ext = [max(abs(sd[8] - cx), abs(sd[8] + sd[12] - cx)), max(abs(sd[9] - cy), abs(sd[9] + sd[13] - cy)),
           max(abs(sd[10]), abs(sd[10] + sd[14]))]
    assert np.all(256.0 / ih.astype(np.float64) >= np.array(ext, np.float64))


def test_pieces_reproduce_the_function_at_their_nodes():
    """Each piece's cubic in f equals s0 log1p(s4 (j + f) / ih) at the four
    Chebyshev nodes up to the binary32 rounding of its coefficients."""
    sd = _box(wl.c2())
    rows, ih, _ = orc.wind_index_table(LES, sd)
    for a in range(3):
        s0, s4 = float(LES[a]), float(LES[4 + a])
        for j in (0, 1, 17, 128, 255):
            c = rows[256 * a + j].astype(np.float64)
            for f in NODES:
                want = s0 * np.log1p(s4 * (j + f) / float(ih[a]))
                got = ((c[3] * f + c[2]) * f + c[1]) * f + c[0]
                assert abs(got - want) <= 4 * np.spacing(np.float32(max(abs(want), 1e-3))), (a, j, f, got, want)


@pytest.mark.parametrize("name", ["c2", "c3"])
def test_index_within_tolerance_of_binary64(name):
    sd = _box(CONFIGS[name]())
    g = np.random.default_rng(3)
    n = 4000
    P = np.stack([sd[8] + g.uniform(0, 1, n) * sd[12], sd[9] + g.uniform(0, 1, n) * sd[13],
                  sd[10] + g.uniform(0, 1, n) * sd[14]], 1).astype(np.float32)
    # corners and the origin lines as well
    P = np.concatenate([P, np.array([[sd[8], sd[9], sd[10]], [sd[8] + sd[12], sd[9] + sd[13], sd[10] + sd[14]],
                                     [sd[8] + 0.5 * sd[12], sd[9] + 0.5 * sd[13], sd[10]]], np.float32)])
    cx, cy = np.float32(sd[8] + np.float32(0.5) * sd[12]), np.float32(sd[9] + np.float32(0.5) * sd[13])
    rel = np.stack([P[:, 0] - cx, P[:, 1] - cy, P[:, 2]], 1).astype(np.float32).astype(np.float64)
    exact = np.stack([np.copysign(float(LES[a]), rel[:, a]) * np.log1p(float(LES[4 + a]) * np.abs(rel[:, a]))
                      + float(LES[8 + a]) for a in range(3)], 1)
    got = np.array([orc.wind_index([p[0], p[1], p[2], 0.0], LES, sd)[:3] for p in P], np.float64)
    err = np.abs(got - exact).max(0)
    print("\n%s: max |c - c64| per axis (texels): %s" % (name, err))
    assert np.all(err < 3e-5), err


def test_out_of_box_positions_stay_finite_and_in_the_table():
    """A position beyond the box (a dead lane's placeholder) extrapolates the
    last piece: finite, never an index outside the rows."""
    sd = _box(wl.c2())
    for p in ([0.0, 0.0, 0.0], [1e6, -1e6, 1e5], [sd[8] - 5000.0, sd[9] + 9000.0, -3.0]):
        c = orc.wind_index([p[0], p[1], p[2], 0.0], LES, sd)
        assert np.all(np.isfinite(c[:3]))
