"""The grouped tile order of the trailing updates (csrc/chol_tiles.h),
compiled for the host with g++: a bijection between [0, T) and the lower
tile set for square and rectangular (narrow and bulk) updates and group
widths 1 .. 8; a planted mistake in the order is caught."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "gpumode_cholesky_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("tiles") / "tile_order_check")
    subprocess.run(["g++", "-O1", "-std=c++17", f"-I{CSRC}", os.path.join(HERE, "tile_order_check.cpp"), "-o", exe],
                   check=True)
    return exe


def test_grouped_order_is_a_bijection(checker):
    shapes = [(1, 1, 4), (3, 3, 4), (5, 2, 4), (128, 128, 4), (128, 4, 4), (127, 9, 4), (64, 64, 1), (64, 63, 8),
              (33, 7, 3), (16, 16, 8), (250, 200, 4), (2, 1, 4)]
    args = [str(x) for s in shapes for x in s]
    r = subprocess.run([checker, *args], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout


def test_grouped_order_planted_mistake_is_caught(checker, tmp_path):
    # the same check must reject an order that skips the triangle part
    src = open(os.path.join(CSRC, "chol_tiles.h")).read()
    bad = src.replace("if (r < tri) {", "if (false) {")
    assert bad != src
    inc = tmp_path / "inc"
    inc.mkdir()
    (inc / "chol_tiles.h").write_text(bad)
    exe = str(tmp_path / "bad_check")
    subprocess.run(["g++", "-O1", "-std=c++17", f"-I{inc}", os.path.join(HERE, "tile_order_check.cpp"), "-o", exe],
                   check=True)
    r = subprocess.run([exe, "20", "12", "4"], capture_output=True, text=True)
    assert r.returncode != 0
