"""iio's image names on the reading side, without a GPU: ``zero:``, ``one:``
and ``constant:`` images, ``TRANS[...]:`` names and the ``IIO_TRANS``
environment variable (flips, axis permutations, crops and pads in the
samples' own type, a pipe through a command), through the library's
``iio_read_image_float_split`` / ``iio_read_image_float``, bit for bit
against what the reference's readers return for the same names
(``tests/golden/iio_names.npz``, written by ``make_golden_iio.py``)."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

import imscript_amd
from helpers import ROOT, SIMPOIS_BIN

GOLD = np.load(os.path.join(ROOT, "tests", "golden", "iio_names.npz"))
CASES = json.loads(str(GOLD["cases"]))
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "simpois")

lib = imscript_amd.lib
_libc = []


def _free(p):
    if not _libc:
        _libc.append(ctypes.CDLL(None))
        _libc[0].free.argtypes = [ctypes.c_void_p]
    _libc[0].free(ctypes.cast(p, ctypes.c_void_p))


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("iio")
    paths = {}
    for k in GOLD.files:
        if k.startswith("in_"):
            paths[k[3:]] = str(d / (k[3:] + ".npy"))
            np.save(paths[k[3:]], GOLD[k])
    return paths


def read(reader, name):
    w, h, pd = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(1)
    if reader == "split":
        p = lib.iio_read_image_float_split(name.encode(), ctypes.addressof(w),
                                           ctypes.addressof(h), ctypes.addressof(pd))
    else:
        p = lib.iio_read_image_float(name.encode(), ctypes.addressof(w),
                                     ctypes.addressof(h))
    a = np.ctypeslib.as_array(p, (pd.value, h.value, w.value)).copy()
    _free(p)
    return a


@pytest.mark.parametrize("i", range(len(CASES)),
                         ids=[c["reader"] + ":" + c["name"] + ("|" + c["env"] if c["env"] else "")
                              for c in CASES])
def test_name_reads_as_the_reference_reads_it(i, files, monkeypatch):
    c = CASES[i]
    monkeypatch.delenv("IIO_TRANS", raising=False)
    if c["env"]:
        monkeypatch.setenv("IIO_TRANS", c["env"])
    got = read(c["reader"], c["name"].format(**files))
    want = GOLD["out_%d" % i]
    assert got.shape == want.shape
    # the same bits, NaN payloads of padded samples included
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_names_cover_every_transformation():
    names = " ".join(c["name"] + " " + c["env"] for c in CASES)
    for op in ("flip=", "pipe=", "x=", "y=", "w=", "h=", "zero:", "one:", "constant:"):
        assert op in names
    assert any(c["env"] and c["name"].startswith("TRANS[") for c in CASES)


@pytest.mark.parametrize("name", ["zero:7", "zero:x5", "one:", "constant:3:4",
                                  "constant:abc:2x2"])
def test_bad_names_exit_like_iio(name, tmp_path):
    """a name iio refuses ends the program with status -1 and an IIO_ERROR
    line before any solve (the reference's too, where it is built)"""
    for exe in (SIMPOIS_BIN, REF_BIN):
        if exe == REF_BIN and not os.path.exists(REF_BIN):
            continue
        p = subprocess.run([exe, "-i", name, "-o", str(tmp_path / "o.npy")],
                           capture_output=True, timeout=60)
        assert p.returncode == 255, (exe, p.returncode, p.stderr)
        assert b"IIO_ERROR" in p.stderr


def test_empty_crop_is_an_error(tmp_path):
    """a crop that leaves no pixel: iio's assert aborts; the drop-in reports
    an IIO_ERROR and exits with status -1"""
    p = subprocess.run([SIMPOIS_BIN, "-i", "TRANS[x=9]:zero:4x4", "-o",
                        str(tmp_path / "o.npy")], capture_output=True, timeout=60)
    assert p.returncode == 255 and b"IIO_ERROR" in p.stderr
