"""The command lines end to end on TIFF and PNG files: the reference
tutorial's float TIFF with holes and colour PNG, PNG masks, TIFF / PNG /
NPY outputs and TIFF on stdin, against the NPY path and the oracles."""
import os
import subprocess

import numpy as np
import pytest

from codec_ref import png_decode, png_encode, tiff_decode
from helpers import (AMLE_BIN, HEATD_BIN, ROOT, SIMPOIS_BIN, bits_equal, oracle_amle,
                     oracle_heatd, oracle_solve)

pytestmark = pytest.mark.gpu

TUT = os.path.join(ROOT, "tests", "golden", "tutorial")
X_TIF = os.path.join(TUT, "X.tif")


def run(args, **kw):
    print("run", " ".join(os.path.basename(a) for a in args[:1]) + " " + " ".join(args[1:]),
          flush=True)
    p = subprocess.run(args, capture_output=True, timeout=120, **kw)
    assert p.returncode == 0, p.stderr
    return p


def npy(path):
    return np.load(path).astype(np.float32)


def q8(v):
    d = 0.5 + v.astype(np.float64)
    return np.where(d > 0, np.where(d < 255, np.floor(d), 255), 0)


def test_simpois_reads_tiff_and_writes_every_format(tmp_path):
    for out in ("o.npy", "o.tif", "o.png"):
        run([SIMPOIS_BIN, "-i", X_TIF, "-o", str(tmp_path / out)])
    ref = npy(tmp_path / "o.npy")
    assert ref.shape == (500, 500, 1) and np.isfinite(ref).all()
    assert bits_equal(tiff_decode((tmp_path / "o.tif").read_bytes()), ref)
    assert np.array_equal(png_decode((tmp_path / "o.png").read_bytes()), q8(ref))
    # the same fill as the oracle on the decoded image (Gauss-Seidel pyramid)
    x = tiff_decode(open(X_TIF, "rb").read())[..., 0]
    want, _ = oracle_solve(x)
    assert bits_equal(ref[..., 0], want.reshape(x.shape))


def test_simpois_tiff_on_stdin(tmp_path):
    with open(X_TIF, "rb") as f:
        p = run([SIMPOIS_BIN, "-o", str(tmp_path / "a.npy")], stdin=f)
    run([SIMPOIS_BIN, "-i", X_TIF, "-o", str(tmp_path / "b.npy")])
    assert bits_equal(npy(tmp_path / "a.npy"), npy(tmp_path / "b.npy"))
    assert p.stderr == b""


def test_png_mask_makes_holes(tmp_path):
    rng = np.random.default_rng(3)
    x = (rng.standard_normal((40, 50, 1)) * 10).astype(np.float32)
    np.save(tmp_path / "x.npy", x)
    m = np.where(rng.random((40, 50, 1)) < 0.4, 255, 0)
    (tmp_path / "m.png").write_bytes(png_encode(m, 0, 8))
    run([SIMPOIS_BIN, "-i", str(tmp_path / "x.npy"), "-m", str(tmp_path / "m.png"),
         "-o", str(tmp_path / "o.npy")])
    g = x[..., 0].copy()
    g[m[..., 0] > 0] = np.nan
    want, _ = oracle_solve(g)
    assert bits_equal(npy(tmp_path / "o.npy")[..., 0], want.reshape(g.shape))


def test_amle_on_the_tutorial_tiff(tmp_path):
    env = dict(os.environ, AMLE_NN="4", AMLE_NITER="30", SIMPOIS_HIP_QUIET="1")
    run([AMLE_BIN, X_TIF, "", str(tmp_path / "o.tif")], env=env)
    x = tiff_decode(open(X_TIF, "rb").read())[..., 0]
    got = tiff_decode((tmp_path / "o.tif").read_bytes())[..., 0]
    assert bits_equal(got, oracle_amle(x, 4, 30))


def test_heatd_on_a_colour_png(tmp_path):
    lc = os.path.join(TUT, "lenac.png")
    run([HEATD_BIN, "-n", "3", lc, str(tmp_path / "o.tif")])
    x = png_decode(open(lc, "rb").read()).astype(np.float32)
    want, _ = oracle_heatd(np.ascontiguousarray(x.transpose(2, 0, 1)), 3, 0.25)
    got = tiff_decode((tmp_path / "o.tif").read_bytes())
    assert bits_equal(got.transpose(2, 0, 1), want)


def test_simpois_with_transformed_names(tmp_path):
    """IIO_TRANS applies to every image simpois reads, and to a
    "TRANS[...]:" name twice (inside, before the name's own list, and after
    it, as iio's recursive read does); "one:" / "zero:" names work as masks.
    The fill equals the oracle's on the transformed arrays."""
    rng = np.random.default_rng(11)
    x = (rng.standard_normal((48, 48)) * 10).astype(np.float32)
    m = (rng.random((48, 48)) < 0.3).astype(np.float32)
    np.save(tmp_path / "x.npy", x)
    np.save(tmp_path / "m.npy", m)
    env = dict(os.environ, IIO_TRANS="flip=transpose")
    run([SIMPOIS_BIN, "-i", str(tmp_path / "x.npy"),
         "-m", "TRANS[flip=leftright]:" + str(tmp_path / "m.npy"),
         "-o", str(tmp_path / "o.npy")], env=env)
    g = x.T.copy()
    g[m.T[:, ::-1].T > 0] = np.nan          # transposed, flipped, transposed
    want, _ = oracle_solve(g)
    assert bits_equal(npy(tmp_path / "o.npy")[..., 0], want.reshape(g.shape))
    # a mask with no hole keeps the input; one with every pixel a hole gives 0
    run([SIMPOIS_BIN, "-i", str(tmp_path / "x.npy"), "-m", "zero:48x48",
         "-o", str(tmp_path / "z.npy")])
    assert bits_equal(npy(tmp_path / "z.npy")[..., 0], x)
    run([SIMPOIS_BIN, "-i", str(tmp_path / "x.npy"), "-m", "one:48x48",
         "-o", str(tmp_path / "a.npy")])
    assert not npy(tmp_path / "a.npy").any()
