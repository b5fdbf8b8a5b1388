"""plambda's 3x3 Laplacian operators ("x,l", "x,lr"; plambda.c:2589-2689
with the boundary rules of getpixel.c), the producer of simpois's guidance
field: the oracle restatement against the golden vectors written by the
reference binary and against closed forms."""
import numpy as np
import pytest

from helpers import bits_equal, load_plambda_golden, oracle_plambda_lap

GOLDEN = load_plambda_golden()


@pytest.mark.parametrize("k", range(len(GOLDEN)))
def test_oracle_matches_reference_plambda(k):
    x, want, variant, mode = GOLDEN[k]
    got = oracle_plambda_lap(x, variant, mode)
    assert bits_equal(got, want), (x.shape, variant, mode)


@pytest.mark.parametrize("variant", [0, 1])
def test_laplacian_of_a_quadratic_is_four(variant):
    # both stencils are exact on quadratics: Lap(x^2 + y^2) = 4 inside
    h, w = 9, 12
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    out = oracle_plambda_lap(x * x + y * y, variant, 1)
    assert np.all(out[1:-1, 1:-1] == 4.0)


@pytest.mark.parametrize("mode,want", [(0, -4.0 * 3.5), (1, 0.0), (2, 0.0),
                                       (3, 0.0), (4, np.nan)])
def test_boundary_rules_on_one_pixel(mode, want):
    # every neighbour of a 1x1 image is outside: zero, the pixel itself
    # (nearest, reflection, periodic) or NaN
    got = oracle_plambda_lap(np.full((1, 1), 3.5, np.float32), 0, mode)[0, 0]
    assert (np.isnan(got) and np.isnan(want)) or got == want


def test_reflection_rule_is_the_reference_one():
    # getpixel.c:74-83: -1 -> 1 but w -> w-1 (the right edge repeats)
    x = np.array([[1, 10, 100]], np.float32)
    # row: centre weight -4, E and W neighbours weight 1; rows above and
    # below reflect onto row 0 (h = 1): N = S = the pixel itself
    got = oracle_plambda_lap(x, 0, 2)[0]
    assert got[0] == 10 + 10 + 1 + 1 - 4 * 1      # W of i=0 reflects to i=1
    assert got[2] == 100 + 10 + 100 - 4 * 100 + 100   # E of i=2 repeats i=2
