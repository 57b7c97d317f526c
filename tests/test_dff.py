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
