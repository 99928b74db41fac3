"""The C ABI without a GPU: the library loads, exports every symbol the
header declares, and rejects bad arguments before any launch."""
import ctypes
import os
import re

import pytest
import torch

from gpumode_cholesky_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "chol_amd.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(chol_\w+)\s*\(", text)))


def test_header_symbols_exported():
    lib = _lib.load()
    names = _declared()
    assert "chol_potrf_f32" in names
    for name in names:
        assert hasattr(lib, name), f"{name} declared but not exported"
    # and the Python binding knows every declared symbol
    assert sorted(_lib.SIGNATURES) == names


def test_abi_version():
    assert _lib.load().chol_abi_version() == 2


def test_tuning_table_rejects_bad_rows():
    import numpy as np
    lib = _lib.load()
    bad_mode = np.array([[300, 300, 1, 8, 5, 256]], dtype=np.int64)
    assert lib.chol_set_tuning(bad_mode.ctypes.data, 1) == _lib.E_MODE
    bad_nbo = np.array([[300, 300, 1, 8, 0, 200]], dtype=np.int64)
    assert lib.chol_set_tuning(bad_nbo.ctypes.data, 1) == _lib.E_ARG
    lds_big = np.array([[300, 300, 1, 8, -1, -1]], dtype=np.int64)
    assert lib.chol_set_tuning(lds_big.ctypes.data, 1) == _lib.E_ARG
    from gpumode_cholesky_amd import autotune
    autotune.load_table()


def test_noalloc_rejects_over_limit_before_any_launch():
    lib = _lib.load()
    rc = lib.chol_potrf_f32_noalloc(FAKE, ctypes.c_void_p(0x2000), 1025, 300, 90000, 300, 0, None, None)
    assert rc == _lib.E_WORKSPACE
    assert "scratch" in _lib.last_error()
    assert lib.chol_potrf_f32_noalloc(FAKE, FAKE, 0, 300, 0, 300, 0, None, None) == 0


FAKE = ctypes.c_void_p(0x1000)  # never dereferenced: every call below fails validation


def _potrf(**kw):
    args = dict(A=FAKE, L=ctypes.c_void_p(0x2000), batch=1, n=8, bs=64, ld=8, mode=0,
                info=ctypes.c_void_p(0), ws=ctypes.c_void_p(0), wsb=0, stream=ctypes.c_void_p(0))
    args.update(kw)
    return _lib.load().chol_potrf_f32(args["A"], args["L"], args["batch"], args["n"], args["bs"],
                                      args["ld"], args["mode"], args["info"], args["ws"], args["wsb"],
                                      args["stream"])


def test_rejects_negative_sizes():
    assert _potrf(batch=-1) == _lib.E_ARG
    assert "batch" in _lib.last_error()
    assert _potrf(n=-3) == _lib.E_ARG


def test_empty_problem_is_a_no_op():
    assert _potrf(batch=0) == 0
    assert _potrf(n=0) == 0


def test_rejects_bad_mode_and_strides():
    assert _potrf(mode=7) == _lib.E_MODE
    assert _potrf(ld=4) == _lib.E_ARG
    assert _potrf(batch=2, bs=10) == _lib.E_ARG
    assert _potrf(A=ctypes.c_void_p(0)) == _lib.E_ARG


def test_rejects_short_workspace():
    lib = _lib.load()
    need = lib.chol_workspace_bytes(4, 300, 0)
    assert need >= 4 * 128 * 128 * 4  # one 128x128 diagonal-block inverse per matrix
    assert _potrf(batch=4, n=300, ld=300, bs=300 * 300) == _lib.E_WORKSPACE
    assert _potrf(batch=4, n=300, ld=300, bs=300 * 300, ws=ctypes.c_void_p(0x3000),
                  wsb=need - 1) == _lib.E_WORKSPACE
    assert "workspace" in _lib.last_error()
    # whole-matrix paths (registers, LDS) need none
    assert lib.chol_workspace_bytes(4096, 32, 0) == 0
    assert lib.chol_workspace_bytes(1024, 64, 0) == 0


def test_tuning_table_validation():
    import numpy as np
    lib = _lib.load()
    # rows {n_lo, n_hi, batch_lo, batch_hi, mode, nbo}
    ok = np.array([[65, 256, 1, 10**9, -1, -1], [257, 4096, 1, 8, 0, 256], [257, 4096, 1, 8, 1, 512]],
                  dtype=np.int64)
    assert lib.chol_set_tuning(ok.ctypes.data, 3) == 0
    nbo = ctypes.c_int64(0)
    assert lib.chol_route(4, 300, 0, ctypes.byref(nbo)) == 3 and nbo.value == 256
    assert lib.chol_route(4, 300, 1, ctypes.byref(nbo)) == 3 and nbo.value == 512
    assert lib.chol_route(4, 200, 1, ctypes.byref(nbo)) == 2
    bad = np.array([[300, 512, 1, 10, -1, -1]], dtype=np.int64)  # LDS route beyond n = 256
    assert lib.chol_set_tuning(bad.ctypes.data, 1) == _lib.E_ARG
    bad = np.array([[300, 512, 1, 10, 0, 100]], dtype=np.int64)  # not a multiple of 128
    assert lib.chol_set_tuning(bad.ctypes.data, 1) == _lib.E_ARG
    bad = np.array([[64, 512, 1, 10, 0, -2]], dtype=np.int64)  # fused route down to n = 64
    assert lib.chol_set_tuning(bad.ctypes.data, 1) == _lib.E_ARG
    fused = np.array([[65, 700, 100, 1000, -1, -2]], dtype=np.int64)
    assert lib.chol_set_tuning(fused.ctypes.data, 1) == 0
    assert lib.chol_route(640, 512, 0, ctypes.byref(nbo)) == 4 and nbo.value == 0
    assert lib.chol_route(640, 100, 1, ctypes.byref(nbo)) == 4
    assert lib.chol_route(99, 512, 1, ctypes.byref(nbo)) == 3
    assert lib.chol_route(640, 32, 1, ctypes.byref(nbo)) == 1
    assert lib.chol_set_tuning(None, 0) == 0
    # built-in defaults: the fused kernel where a batch of mid orders fills
    # the GPU one to a few times over, the blocked path outside
    assert lib.chol_route(256, 300, 0, None) == 4 and lib.chol_route(4096, 300, 1, None) == 4
    assert lib.chol_route(512, 512, 1, None) == 4 and lib.chol_route(640, 512, 0, None) == 3
    assert lib.chol_route(255, 300, 0, None) == 3 and lib.chol_route(512, 128, 0, None) == 3
    assert lib.chol_route(1024, 200, 0, None) == 3 and lib.chol_route(8192, 200, 0, None) == 2
    from gpumode_cholesky_amd import autotune
    autotune.load_table()


def test_custom_kernel_refuses_cpu_tensors():
    from gpumode_cholesky_amd import custom_kernel
    a = torch.eye(4).unsqueeze(0)
    with pytest.raises(ValueError, match="HIP device"):
        custom_kernel(a)


def test_custom_kernel_refuses_bad_dtype_and_shape():
    from gpumode_cholesky_amd import cholesky
    with pytest.raises(TypeError):
        cholesky(torch.eye(4, dtype=torch.float64))
    with pytest.raises(ValueError):
        cholesky(torch.zeros(2, 3, 4), mode="plain")
    with pytest.raises(ValueError):
        cholesky(torch.eye(3), mode="fp16")
