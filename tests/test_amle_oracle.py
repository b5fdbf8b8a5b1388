"""The AMLE oracle (oracle/amle_oracle.c) against the reference amle's
golden vectors and against what the mathematics fixes, without a GPU."""
import numpy as np
import pytest

from helpers import (amle_trace_lines, bits_equal, load_amle_golden, oracle,
                     oracle_amle, ptr)

CASES = load_amle_golden()


def test_golden_file_covers_every_neighbourhood_and_edge_case():
    nns = {c["nn"] for c in CASES}
    assert nns == {4, 6, 8, 16, 24, 32}
    assert any(c["init"] is not None for c in CASES)
    assert any(c["mask"] is not None for c in CASES)
    assert any(np.isnan(c["x"]).all() for c in CASES)
    assert any(not np.isnan(c["x"]).any() for c in CASES)
    assert any(np.isinf(c["x"]).any() for c in CASES)
    assert any(c["niter"] == 100 for c in CASES)


@pytest.mark.parametrize("k", range(len(CASES)))
def test_oracle_matches_reference_golden(k):
    c = CASES[k]
    y = oracle_amle(c["x"], c["nn"], c["niter"], c["init"])
    assert bits_equal(y, c["out"]), (c["nn"], c["niter"], c["x"].shape)
    # the exact bits too, NaN signs included (same float operations on x86)
    assert np.array_equal(y.view(np.uint32), c["out"].view(np.uint32))


@pytest.mark.parametrize("k", range(len(CASES)))
def test_oracle_trace_matches_reference_stderr(k):
    c = CASES[k]
    _, d = oracle_amle(c["x"], c["nn"], c["niter"], c["init"], diag=True)
    ref = c["trace"].replace("-nan", "nan")
    assert amle_trace_lines(d) == ref


def test_strip_converges_to_linear_interpolation():
    # one row: only the east / west neighbours are in the image, so the
    # update is their average and the fixed point is the straight line
    x = np.full((1, 9), np.nan, np.float32)
    x[0, 0], x[0, 8] = 2.0, 10.0
    y = oracle_amle(x, 4, 2000)
    assert np.allclose(y[0], np.linspace(2, 10, 9), atol=1e-5)


def test_column_uses_south_and_north():
    x = np.full((7, 1), np.nan, np.float32)
    x[0, 0], x[6, 0] = -3.0, 3.0
    y = oracle_amle(x, 4, 2000)
    assert np.allclose(y[:, 0], np.linspace(-3, 3, 7), atol=1e-5)


@pytest.mark.parametrize("nn", [4, 8, 16, 32])
def test_constant_data_fills_exactly(nn):
    x = np.full((13, 17), 4.5, np.float32)
    x[3:10, 2:15] = np.nan
    assert np.array_equal(oracle_amle(x, nn, 50), np.full_like(x, 4.5))


@pytest.mark.parametrize("nn", [4, 6, 8, 16, 24, 32])
def test_max_principle_and_known_pixels(nn):
    rng = np.random.default_rng(nn)
    x = rng.standard_normal((21, 19)).astype(np.float32)
    x[rng.random(x.shape) < 0.6] = np.nan
    y = oracle_amle(x, nn, 40)
    known = ~np.isnan(x)
    assert np.array_equal(y[known], x[known])
    assert np.nanmin(x) <= y.min() and y.max() <= np.nanmax(x)


def test_start_is_the_maximum_and_init_replaces_it():
    x = np.array([[1.0, np.nan, 7.0]], np.float32)
    assert oracle_amle(x, 4, 0)[0, 1] == 7.0
    init = np.array([[0.0, -2.0, 0.0]], np.float32)
    assert oracle_amle(x, 4, 0, init)[0, 1] == -2.0


def test_weights_balance_the_slopes():
    # one hole in a 3x3 window, 8 neighbours: the minimum on a diagonal
    # (squared length 2), the maximum on an axis (1).  The update balances
    # (vmax - u) / wmax = (u - vmin) / wmin, i.e. u = (wmax vmin + wmin vmax)
    # / (wmax + wmin); swapping the weights would give 10/3 instead of 20/3.
    x = np.full((3, 3), 5.0, np.float32)
    x[1, 1] = np.nan
    x[0, 0] = 0.0          # (-1, -1): squared length 2
    x[1, 2] = 10.0         # (+1, 0): squared length 1
    u = float(oracle_amle(x, 8, 1)[1, 1])
    assert abs((10.0 - u) / 1.0 - (u - 0.0) / 2.0) < 1e-5
    assert abs(u - 20.0 / 3.0) < 1e-6


def test_first_minimum_in_table_order_sets_the_weight():
    # equal minima at (+1,0) (weight 1) and (+1,+1) (weight 2): the first in
    # the table's order, (+1,0), is the minimum, so u = (2*0 + 1*9)/(2+1)
    # with the maximum 9 at (-1,-1) (weight 2)
    x = np.full((3, 3), 5.0, np.float32)
    x[1, 1] = np.nan
    x[1, 2] = 0.0
    x[2, 2] = 0.0
    x[0, 0] = 9.0
    u = float(oracle_amle(x, 8, 1)[1, 1])
    assert u == np.float32((np.float32(2) * 0 + np.float32(1) * 9) / np.float32(3))


def test_raster_order_is_in_place():
    # two holes in a row: the second sees the first's new value
    x = np.array([[0.0, np.nan, np.nan, 8.0]], np.float32)
    y = oracle_amle(x, 4, 1)
    first = (0.0 + 8.0) / 2          # neighbours 0 and the start value 8
    assert y[0, 1] == first
    assert y[0, 2] == (first + 8.0) / 2


def test_no_neighbour_gives_nan_and_all_nan_gives_minus_inf():
    one = np.full((1, 1), np.nan, np.float32)
    assert np.isnan(oracle_amle(one, 4, 1)[0, 0])
    allnan = np.full((3, 4), np.nan, np.float32)
    assert (oracle_amle(allnan, 4, 5) == -np.inf).all()


def test_sweeps_entry_matches_driver():
    rng = np.random.default_rng(5)
    x = rng.standard_normal((15, 22)).astype(np.float32)
    x[rng.random(x.shape) < 0.5] = np.nan
    y = oracle_amle(x, 8, 6)
    u = np.where(np.isnan(x), np.nanmax(x), x).astype(np.float32)
    hole = np.isnan(x).astype(np.uint8)
    oracle().orc_amle_sweeps(ptr(u), ptr(hole), 22, 15, 8, 6)
    assert np.array_equal(u, y)
