"""The autotuner end to end on small shapes: every candidate route of a
shape is timed, the fastest becomes its table row, and a table written by
the tuner routes the library as stated (chol_route)."""
import json

import pytest
import torch

from gpumode_cholesky_amd import _lib, autotune, harness

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def test_time_candidates_covers_every_route():
    a = harness.config_input(8, 192, 77, device=DEV)
    cands = autotune.candidates(192)
    assert cands == [-1, -2, 128]
    try:
        times = autotune.time_candidates(a, "mixed", cands, reps=2)
    finally:
        autotune.load_table()
    assert sorted(times) == sorted(cands)
    assert all(t > 0 for t in times.values())


def test_tune_writes_rows_that_route(tmp_path):
    shapes = [(4, 256), (2, 640)]
    try:
        rows = autotune.tune(shapes, "plain", reps=2, log=lambda m: None)
    finally:
        autotune.load_table()
    assert [(r[0], r[2], r[4]) for r in rows] == [(256, 4, 0), (640, 2, 0)]
    for r in rows:
        assert r[5] in autotune.candidates(r[0])
    path = tmp_path / "t.json"
    path.write_text(json.dumps({"rows": rows}))
    try:
        autotune.load_table(str(path))
        lib = _lib.load()
        import ctypes
        want = {-1: 2, -2: 4}  # CHOL_ROUTE_LDS, CHOL_ROUTE_FUSED
        for n, _, batch, _, mode, nbo in rows:
            out = ctypes.c_int64(0)
            route = lib.chol_route(batch, n, mode, ctypes.byref(out))
            if nbo > 0:
                assert route == 3 and out.value == nbo, (n, batch, route, out.value)  # CHOL_ROUTE_BLOCKED
            else:
                assert route == want[nbo], (n, batch, route)
    finally:
        autotune.load_table()
