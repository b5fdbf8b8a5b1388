"""The Gauss-Seidel kernels' own phase functions (imscript_amd/csrc/sp_gs.h)
run on the CPU under AddressSanitizer, against the oracle's raster sweeps.

This checks, without a GPU, the tile geometry, the LDS and image indexing,
the wavefront schedule (every tile of a wavefront loads before any tile
stores) and the write-back rule of the device kernels.
"""
import os
import shutil
import subprocess

import pytest

from helpers import ROOT

BUILD = os.path.join(ROOT, "tests", "cpu", "_build")


@pytest.fixture(scope="module")
def emulator():
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    # one build per process (pytest-xdist workers build side by side)
    build = os.path.join(BUILD, f"p{os.getpid()}")
    os.makedirs(build, exist_ok=True)
    san = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
    objs = []
    for src in ("simpois_oracle.c", "amle_oracle.c"):
        obj = os.path.join(build, src.replace(".c", "_asan.o"))
        subprocess.run(["gcc", "-O1", "-g", "-std=gnu99", "-ffp-contract=off", *san,
                        "-c", os.path.join(ROOT, "oracle", src), "-o", obj],
                       check=True)
        objs.append(obj)
    exe = os.path.join(build, "gs_emulate")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-ffp-contract=off", *san,
                    "-o", exe, os.path.join(ROOT, "tests", "cpu", "gs_emulate.cpp"),
                    *objs, "-lm"], check=True)
    yield exe
    shutil.rmtree(build, ignore_errors=True)


CASES = [
    # W H T tstep seed holefrac withf mode
    ("40 30 5 0.25 1 0.4 0 tiles"),
    ("150 140 40 0.25 2 0.3 1 tiles"),
    ("150 140 40 0.4 3 0.5 0 tiles"),
    ("200 3 33 0.25 1 0.5 1 tiles"),
    ("33 1 50 0.25 3 0.5 1 tiles"),
    ("1 1 7 0.25 1 1 1 tiles"),
    ("70 45 17 -0.08 4 0.3 0 tiles"),
    ("100 90 20 -0.08 7 0.4 1 tiles"),
    ("1 9 7 -0.08 1 0.5 1 tiles"),
    ("50 37 6 -1000.01 5 0.3 0 tiles"),
    ("97 83 9 -1000.02 8 0.4 1 tiles"),
    ("40 30 5 0.25 1 0.4 0 small"),
    ("70 45 17 -0.08 4 0.3 1 small"),
    ("50 37 6 -1000.01 5 0.3 1 small"),
    ("2 2 100 -0.08 3 0.5 1 small"),
    # seeds from 100: a few infinite and 3e37 known values
    ("300 260 40 0.25 105 0.4 1 tiles"),
    ("260 250 70 0.25 107 0.5 0 tiles"),
    ("60 50 30 0.25 108 0.4 1 small"),
]


@pytest.mark.parametrize("args", CASES)
def test_kernel_phases_match_raster_gauss_seidel(emulator, args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    p = subprocess.run([emulator, *args.split()], capture_output=True,
                       text=True, timeout=600, env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.startswith("OK")


# every compiled tile shape (sp_gs.h SP_GS_SHAPES), sizes spanning several
# tiles in each skewed direction and a partial last pass in time
SHAPE_CASES = (
    [f"150 140 40 0.25 {11 + v} 0.45 1 tiles {v}" for v in range(12)]
    + [f"200 3 37 0.25 {21 + v} 0.5 0 tiles {v}" for v in range(12)]
    + [f"100 90 20 -0.08 {31 + v} 0.4 1 tiles {v}" for v in range(4)]
    + [f"97 83 9 -1000.02 {41 + v} 0.4 1 tiles {v}" for v in range(3)]
)


@pytest.mark.parametrize("args", SHAPE_CASES)
def test_every_tile_shape_matches_raster_gauss_seidel(emulator, args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    p = subprocess.run([emulator, *args.split()], capture_output=True,
                       text=True, timeout=600, env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.startswith("OK")


def test_unknown_tile_shape_is_refused(emulator):
    p = subprocess.run([emulator, *"40 30 5 0.25 1 0.4 0 tiles 99".split()],
                       capture_output=True, text=True, timeout=60,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"))
    assert p.returncode == 2 and "NO SHAPE" in p.stdout


# whole-level kernel at the limits of its slot table and LDS budget
SMALL_LIMIT_CASES = [
    "1 8000 9 0.25 51 0.5 1 small",      # W = 1: 8000 slots, 8 per thread
    "8000 1 9 0.25 52 0.5 1 small",
    "89 89 30 0.25 53 0.4 1 small",      # framed 91x91 cells = 66 KB of u + c
    "63 61 25 -0.08 54 0.4 1 small",
    "40 40 7 -1000.02 55 0.4 1 small",
]


# the dependence-ordered launch (k_gs_flow): random tile orders in which
# each tile waits only for its existing predecessors at skewed distance
# {0, 1}^3, up to 64 tiles between load and store
FLOW_CASES = (
    [f"150 140 40 0.25 {61 + v} 0.45 1 flow {v}" for v in (0, 3, 6, 10, 11)]
    + ["200 3 37 0.25 21 0.5 0 flow 0", "300 260 40 0.25 105 0.4 1 flow 0",
       "100 90 20 -0.08 31 0.4 1 flow 0", "100 90 20 -0.08 32 0.4 1 flow 3",
       "97 83 9 -1000.02 41 0.4 1 flow 2", "1 9 7 -0.08 1 0.5 1 flow 0"]
)


@pytest.mark.parametrize("args", FLOW_CASES)
def test_dependence_ordered_tiles_match_raster_gauss_seidel(emulator, args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    p = subprocess.run([emulator, *args.split()], capture_output=True,
                       text=True, timeout=600, env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.startswith("OK")


# whole R = 1 level by unknown pixel (k_gs_cells): sizes up to its
# 16384-pixel limit, all-hole planes (8 pixels per thread and parity),
# single rows and columns, a 1x1 level, infinite and 3e37 known values
CELLS_CASES = [
    "40 30 5 0.25 1 0.4 0 cells",
    "1 1 7 0.25 1 1 1 cells",
    "2 3 100 0.25 4 0.5 1 cells",
    "128 128 60 0.25 3 0.3 1 cells",
    "128 128 30 0.25 8 1.0 1 cells",
    "300 54 40 0.4 5 0.5 1 cells",
    "16384 1 9 0.25 6 0.5 1 cells",
    "1 16384 9 0.25 7 0.5 1 cells",
    "61 67 45 0.25 108 0.4 1 cells",
    "64 64 200 0.25 9 0.3 0 cells",
]


@pytest.mark.parametrize("args", CELLS_CASES)
def test_whole_level_kernel_by_unknown_pixel(emulator, args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    p = subprocess.run([emulator, *args.split()], capture_output=True,
                       text=True, timeout=600, env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.startswith("OK")


# a fixed point (a sweep that changes no pixel's bits) ends the level's
# steps early; the values still match all T raster sweeps bit for bit.
# W H T ..., and the step the level is left at: a level without unknowns
# after its first sweep's W + H - 1 steps, a Laplace fill of 20x20 after
# 163 of its 6037 steps
CELLS_FIXED_CASES = [
    ("40 30 50 0.25 1 0.0 1 cells", 69, 167),
    ("20 20 3000 0.25 2 0.3 0 cells", 163, 6037),
]


@pytest.mark.parametrize("args,left,total", CELLS_FIXED_CASES)
def test_whole_level_kernel_stops_at_a_fixed_point(emulator, args, left, total):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    p = subprocess.run([emulator, *args.split()], capture_output=True,
                       text=True, timeout=600, env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.startswith("OK")
    assert f"CELLS LEFT AT STEP {left} OF {total}" in p.stderr


def test_whole_level_kernel_by_unknown_pixel_limits(emulator):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    p = subprocess.run([emulator, *"129 128 5 0.25 1 0.4 1 cells".split()],
                       capture_output=True, text=True, timeout=60, env=env)
    assert p.returncode == 1 and "TOO MANY CELLS" in p.stdout
    p = subprocess.run([emulator, *"40 30 5 -0.08 1 0.4 1 cells".split()],
                       capture_output=True, text=True, timeout=60, env=env)
    assert p.returncode == 2 and "CELLS NEED R = 1" in p.stdout


@pytest.mark.parametrize("args", SMALL_LIMIT_CASES)
def test_whole_level_kernel_limits(emulator, args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    p = subprocess.run([emulator, *args.split()], capture_output=True,
                       text=True, timeout=600, env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.startswith("OK")


AMLE_CASES = [
    # W H T nn seed holefrac [variant]
    "40 30 5 4 1 0.5",
    "150 140 40 4 2 0.7",
    "200 3 33 4 3 0.5",
    "3 170 21 4 4 0.6",
    "1 1 7 4 1 1",
    "1 9 5 4 2 0.6",
    "61 47 9 1 5 0.5",
    "61 47 9 2 6 0.5",
    "61 47 9 3 7 0.5",
    "70 45 17 6 4 0.6",
    "70 45 17 8 8 0.6",
    "2 2 9 8 3 0.5",
    "66 59 11 16 5 0.6",
    "63 71 10 24 6 0.6",
    "80 77 12 32 9 0.7",
    "5 4 6 32 2 0.8",
    "90 70 23 4 101 0.5",
    "47 33 9 4 103 0.3",
    "61 58 12 8 105 0.5",
    "55 51 9 32 107 0.5",
]


@pytest.mark.parametrize("args", AMLE_CASES)
def test_amle_tiles_match_raster_sweeps(emulator, args):
    """AMLE tile steps (sp_amle.h) on every neighbourhood size, against
    amle.c's raster-order sweeps restated in oracle/amle_oracle.c."""
    r = subprocess.run([emulator, "amle", *args.split()], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr


# the dependence-ordered launch of the AMLE sweeps: random tile orders
# (each tile waits only for its existing predecessors) on every radius
AMLE_FLOW_CASES = ["150 140 20 4 3 0.5 0", "150 140 20 4 103 0.5 6",
                   "120 100 15 8 4 0.5 0", "100 90 12 16 5 0.5 0",
                   "90 80 9 32 6 0.5 0", "97 83 9 2 7 0.5 0", "1 1 5 4 8 1 0"]


@pytest.mark.parametrize("args", AMLE_FLOW_CASES)
def test_dependence_ordered_amle_tiles_match_raster_sweeps(emulator, args):
    r = subprocess.run([emulator, "amle", *args.split(), "flow"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr


@pytest.mark.parametrize("var", range(15))
def test_every_amle_tile_shape_matches_raster_sweeps(emulator, var):
    r = subprocess.run([emulator, "amle", "150", "140", "40", "4", str(var), "0.6",
                        str(var)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr
