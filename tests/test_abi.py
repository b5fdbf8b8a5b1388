"""The C ABI and the host side of the command line, without a GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import imscript_amd
from helpers import ROOT, SIMPOIS_BIN

HEADERS = [os.path.join(ROOT, "include", f) for f in sorted(os.listdir(
    os.path.join(ROOT, "include"))) if f.endswith(".h")]


def header_functions():
    names = []
    for hdr in HEADERS:
        text = open(hdr).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while")))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(imscript_amd.lib_path), "build the library first"
    h = ctypes.CDLL(imscript_amd.lib_path)
    declared = header_functions()
    assert len(declared) >= 10
    for name in declared:
        assert hasattr(h, name), name
    assert declared == imscript_amd.exported_symbols()


def test_params_default_follow_reference_and_env():
    p = imscript_amd.default_params()
    assert (p.tstep, p.niter, p.scale, p.cgit) == (np.float32(0.25), 10, 99, 0)


def run_cli(args, **kw):
    env = dict(os.environ)
    env.update(kw.pop("env", {}))
    return subprocess.run([SIMPOIS_BIN, *args], capture_output=True,
                          timeout=60, env=env, **kw)


@pytest.mark.parametrize("flag,start", [("-h", b"usage:"), ("--help", b"Simpois is"),
                                        ("--version", b"simpois 1.0"),
                                        ("-?", b"a simple Poisson")])
def test_cli_help_flags(flag, start):
    p = run_cli([flag])
    assert p.returncode == 0 and p.stdout.startswith(start)


def test_cli_leftover_argument_exits_1():
    p = run_cli(["-n", "3", "stray"])
    assert p.returncode == 1 and b"Usage:" in p.stderr


def test_cli_unreadable_input_exits_255(tmp_path):
    p = run_cli(["-i", str(tmp_path / "missing.npy"), "-o", str(tmp_path / "o.npy")])
    assert p.returncode == 255 and b"IIO_ERROR" in p.stderr
    bad = tmp_path / "bad.npy"
    bad.write_bytes(b"garbage!")
    p = run_cli(["-i", str(bad), "-o", str(tmp_path / "o.npy")])
    assert p.returncode == 255


def test_cli_size_mismatch_returns_fprintf_count(tmp_path):
    # simpois.c:405 returns fprintf's character count (26)
    np.save(tmp_path / "g.npy", np.zeros((4, 5), np.float32))
    np.save(tmp_path / "f.npy", np.zeros((5, 5), np.float32))
    p = run_cli(["-i", str(tmp_path / "g.npy"), "-f", str(tmp_path / "f.npy"),
                 "-o", str(tmp_path / "o.npy")])
    assert p.returncode == len("input sizes mismatch (i,f)")
    np.save(tmp_path / "m.npy", np.zeros((3, 5), np.float32))
    p = run_cli(["-i", str(tmp_path / "g.npy"), "-m", str(tmp_path / "m.npy"),
                 "-o", str(tmp_path / "o.npy")])
    assert p.returncode == len("input sizes mismatch (i,m)")


AMLE_BIN = os.path.join(ROOT, "imscript_amd", "lib", "amle")


@pytest.mark.parametrize("flag,start", [("-h", b"usage:"), ("--help", b"Interpolate by AMLE"),
                                        ("--version", b"amle 1.0"),
                                        ("-?", b"Absolutely Minimizing")])
def test_amle_cli_help_flags(flag, start):
    p = subprocess.run([AMLE_BIN, flag], capture_output=True, timeout=60)
    assert p.returncode == 0 and p.stdout.startswith(start)


def test_amle_cli_too_many_arguments_exits_1():
    p = subprocess.run([AMLE_BIN, "a", "b", "c", "d"], capture_output=True, timeout=60)
    assert p.returncode == 1 and b"[in [mask [out]]]" in p.stderr


def test_amle_cli_unreadable_input_exits_255(tmp_path):
    p = subprocess.run([AMLE_BIN, str(tmp_path / "missing.npy")], capture_output=True,
                       timeout=60)
    assert p.returncode == 255 and b"IIO_ERROR" in p.stderr


HEATD_BIN = os.path.join(ROOT, "imscript_amd", "lib", "heatd")


def test_heatd_cli_bad_positional_count_prints_usage():
    p = subprocess.run([HEATD_BIN, "a", "b", "c"], capture_output=True, timeout=60)
    assert p.returncode != 0 and b"[in [out]]" in p.stderr


def test_heatd_cli_bad_field_dimensions(tmp_path):
    np.save(tmp_path / "x.npy", np.zeros((4, 5, 1), np.float32))
    np.save(tmp_path / "A.npy", np.zeros((4, 5, 2), np.float32))
    p = subprocess.run([HEATD_BIN, "-A", str(tmp_path / "A.npy"), str(tmp_path / "x.npy"),
                        str(tmp_path / "y.npy")], capture_output=True, timeout=60)
    assert p.returncode != 0 and b"ERROR: bad A dimensions(5 4 2) ((5 4))" in p.stderr
