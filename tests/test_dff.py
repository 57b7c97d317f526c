"""Debris flux field (concept C): the inverse-CDF table built from a cell PDF
(rs.c:3247-3320) by the host runtime and by the oracle, pinned by closed forms
(uniform and two-cell PDFs invert linearly, empty cells are never entered), by
the reference's own 3 x 3 example (rs.c:3431-3459: icdf = [4, 5]) and by the
reference's error return; product and oracle tables agree bit for bit.  The
respawn placement that reads the table (rs.cl:795-836) is checked on the GPU
against the oracle (test_gpu_parity.py)."""
import ctypes

import numpy as np
import pytest

import simradar_amd as sa
from oracle import oracle as orc

N = 2048


def _product_icdf(pdf, n=N):
    lib = ctypes.CDLL(sa.LIB_PATH)
    fn = lib.rs_dff_icdf_from_pdf
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    pdf = np.ascontiguousarray(pdf, np.float32)
    tab = np.zeros(n, np.float32)
    rc = fn(pdf.ctypes.data, pdf.size, tab.ctypes.data, n)
    return tab if rc == 0 else None


def _normalised(w):
    """PDF as the reference normalises it: binary32 sum in cell order, then divide."""
    w = np.asarray(w, np.float32)
    s = np.float32(0.0)
    for v in w:
        s = np.float32(s + v)
    return (w / s).astype(np.float32)


def _board(c):
    w = np.zeros(c * c, np.float32)
    w[::2] = 100.0
    return _normalised(w)


def test_uniform_pdf_inverts_linearly():
    for cells in (1, 4, 81):
        tab = orc.dff_icdf_from_pdf(_normalised(np.ones(cells)))
        x = np.arange(N) / (N - 1)
        np.testing.assert_allclose(tab, cells * x, atol=2e-5 * cells)
        assert tab[0] == 0.0 and tab[-1] == cells


def test_two_cell_pdf_closed_form():
    tab = orc.dff_icdf_from_pdf(np.array([0.25, 0.75], np.float32))
    x = np.arange(N) / (N - 1)
    want = np.where(x <= 0.25, x / 0.25, 1.0 + (x - 0.25) / 0.75)
    np.testing.assert_allclose(tab, want, atol=1e-5)


def test_empty_cells_are_not_entered():
    tab = orc.dff_icdf_from_pdf(_board(9))
    assert np.all(np.diff(tab) >= 0)
    cell = np.floor(tab).astype(int)
    inside = tab != np.floor(tab)
    assert np.all(cell[inside] % 2 == 0), "table value inside an empty cell"
    assert tab[0] == 0.0 and tab[-1] == 81.0


def test_leading_empty_cells_start_at_first_flux():
    tab = orc.dff_icdf_from_pdf(np.array([0, 0, 0, 0, 1, 0, 0, 0, 0], np.float32))
    # all flux in cell 4: the reference's own hand-made table for this field is [4, 5]
    x = np.arange(N) / (N - 1)
    np.testing.assert_allclose(tab, 4.0 + x, atol=1e-6)


def test_bad_pdf_is_refused():
    assert orc.dff_icdf_from_pdf(np.array([0.6, 0.6, 0.0, 0.0], np.float32)) is None
    assert _product_icdf(np.array([0.6, 0.6, 0.0, 0.0], np.float32)) is None
    # a sum past 1 in one of the last two cells is clamped, not refused
    assert _product_icdf(np.array([0.5, 0.5000001], np.float32)) is not None


@pytest.mark.parametrize("kind", ["uniform", "board", "random", "sparse", "les"])
def test_product_table_equals_oracle(kind):
    rng = np.random.default_rng(7)
    if kind == "uniform":
        pdf = _normalised(np.ones(64))
    elif kind == "board":
        pdf = _board(9)
    elif kind == "random":
        pdf = _normalised(rng.random(400))
    elif kind == "sparse":
        w = rng.random(300)
        w[rng.random(300) < 0.7] = 0.0
        pdf = _normalised(w)
    else:
        u = rng.normal(size=(3, 50 * 40)).astype(np.float32)
        pdf = _normalised((u * u).sum(axis=0))
    a, b = _product_icdf(pdf), orc.dff_icdf_from_pdf(pdf)
    assert a is not None and b is not None
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_concept_c_builds_the_field_from_the_second_les_level():
    """With concept C, loading an LES volume (RS_set_vel_data_to_config here)
    builds the flux field from the volume (rs.c:3106-3111): |u|^2 of its second
    level (cell nx ny + k) times the stretched cell area a r^|i - m| per axis,
    normalised (rs.c:3565-3626).  The first field fills both LES slots.  The
    table inverts that PDF's CDF (binary64 here) within the table's resolution,
    and not the first level's."""
    import harness as h
    r = sa.Radar(verb=0, host_only=True)
    h.configure(r, h.Config(n_rain=4096, scan="P:3.0,-1:1:0.01", concept=h.Config().concept | sa.CONCEPT_DFF))
    f0, f1 = r.debris_flux_field(0), r.debris_flux_field(1)
    assert f0 is not None and f1 is not None
    assert np.array_equal(f0[0], f1[0]) and np.array_equal(f0[1], f1[1])
    icdf, desc = f0
    assert desc[14] == 1.0          # stretched grid: scale a / (1 - r) in slot 0, ratio r in slot 4
    frame = r.les_frame(0).astype(np.float64)      # (nz, ny, nx, 4)
    nz, ny, nx, _ = frame.shape
    assert icdf.size == N and desc[13] == nx and desc[8] == nx - 1 and desc[9] == ny - 1
    rx, ry = float(desc[4]), float(desc[5])
    ax, ay = float(desc[0]) * (1.0 - rx), float(desc[1]) * (1.0 - ry)
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny))
    area = (ax * rx ** np.abs(ix - 0.5 * (nx - 1))) * (ay * ry ** np.abs(iy - 0.5 * (ny - 1)))

    def pdf(level):
        u = frame[level, :, :, :3]
        w = (u * u).sum(-1) * area
        return (w / w.sum()).ravel()

    def quantile_error(p):
        # entry j of the table inverts the CDF (linear inside each cell) at ~ j / N
        c = np.concatenate([[0.0], np.cumsum(p)])
        cell = np.minimum(np.floor(icdf).astype(np.int64), p.size - 1)
        at = c[cell] + (icdf - cell) * p[cell]
        return np.max(np.abs(at - np.arange(N) / float(N)))

    e1, e0 = quantile_error(pdf(1)), quantile_error(pdf(0))
    assert e1 <= 1.5 / N, e1        # one entry: the table's sampling of the CDF
    assert e0 > 3.0 / N, e0         # the first level's field is another table
    r.free()
