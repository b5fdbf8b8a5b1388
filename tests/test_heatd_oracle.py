"""The heatd oracle (oracle/heatd_oracle.c) against the reference heatd's
golden vectors and against what the mathematics fixes, without a GPU."""
import numpy as np
import pytest

from helpers import load_heatd_golden, oracle_heatd

CASES = load_heatd_golden()


def test_golden_file_covers_fields_specials_and_channels():
    assert any(c["A"] is not None for c in CASES)
    assert any(c["A"] is None for c in CASES)
    assert any(c["x"].shape[0] == 3 for c in CASES)
    assert any(not np.isfinite(c["x"]).all() for c in CASES)
    assert {c["n"] for c in CASES} >= {1, 2, 5}


@pytest.mark.parametrize("k", range(len(CASES)))
def test_oracle_matches_reference_golden(k):
    c = CASES[k]
    y, x_after = oracle_heatd(c["x"], c["n"], np.float32(c["t"]), c["A"])
    assert np.array_equal(y.view(np.uint32), c["out"].view(np.uint32))
    # the reference leaves its input equal to the output
    assert np.array_equal(x_after.view(np.uint32), y.view(np.uint32))


def test_constant_image_is_a_fixed_point():
    x = np.full((2, 9, 11), 3.25, np.float32)
    rng = np.random.default_rng(0)
    A = rng.uniform(0.1, 2, (9, 11, 3)).astype(np.float32)
    assert np.array_equal(oracle_heatd(x, 7, 0.25)[0], x)
    assert np.array_equal(oracle_heatd(x, 7, 0.25, A)[0], x)


def test_linear_ramp_is_fixed_inside():
    # the 5-point Laplacian of a ramp is exactly zero away from the borders
    j, i = np.mgrid[0:12, 0:15].astype(np.float32)
    x = 3 * i - 2 * j
    y = oracle_heatd(x, 1, 0.25)[0]
    assert np.array_equal(y[1:-1, 1:-1], x[1:-1, 1:-1])
    assert not np.array_equal(y, x)       # the clamped borders move


def test_mass_is_conserved_with_clamped_borders():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((40, 33)).astype(np.float32)
    y = oracle_heatd(x, 10, 0.2)[0]
    assert abs(y.astype(np.float64).sum() - x.astype(np.float64).sum()) < 1e-3
    assert y.std() < x.std()              # and it smooths


def test_identity_field_equals_isotropic_scheme():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((3, 17, 13)).astype(np.float32)
    A = np.zeros((17, 13, 3), np.float32)
    A[..., 0] = 1
    A[..., 2] = 1
    a = oracle_heatd(x, 4, 0.25, A)[0]
    b = oracle_heatd(x, 4, 0.25)[0]
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_mixed_term_of_the_field():
    # u = i*j: the corner weights +-b/2 give sum b/2 * 4 = 2b, divided by 4:
    # L(u) = b/2 inside; u_xx = u_yy = 0 so a and c drop out
    j, i = np.mgrid[0:9, 0:10].astype(np.float32)
    u = i * j
    A = np.zeros((9, 10, 3), np.float32)
    A[..., 0], A[..., 1], A[..., 2] = 0.7, 0.5, 1.3
    y = oracle_heatd(u, 1, 1.0, A)[0]
    assert np.allclose((y - u)[1:-1, 1:-1], 0.25, atol=1e-5)


def test_quadratic_gives_the_second_derivatives():
    # u = i^2: a*(u(i-1)+u(i+1)-2u) = 2a, divided by 4: L = a/2
    j, i = np.mgrid[0:8, 0:9].astype(np.float32)
    A = np.zeros((8, 9, 3), np.float32)
    A[..., 0], A[..., 2] = 0.6, 1.4
    y = oracle_heatd(i * i, 1, 1.0, A)[0]
    assert np.allclose((y - i * i)[1:-1, 1:-1], 0.3, atol=1e-5)
    y = oracle_heatd(j * j, 1, 1.0, A)[0]
    assert np.allclose((y - j * j)[1:-1, 1:-1], 0.7, atol=1e-5)


def test_infinite_corner_makes_nan_through_zero_weight():
    # the isotropic scheme multiplies the corners by 0: an infinite corner
    # poisons the result, as in the reference
    x = np.zeros((3, 3), np.float32)
    x[0, 0] = np.inf
    y = oracle_heatd(x, 1, 0.25)[0]
    assert np.isnan(y[1, 1])
    assert y[2, 2] == 0.0


def test_steps_compose():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 14, 9)).astype(np.float32)
    y3 = oracle_heatd(x, 3, 0.3)[0]
    y = x
    for _ in range(3):
        y = oracle_heatd(y, 1, 0.3)[0]
    assert np.array_equal(y3, y)
