"""chol_potrf_f32_noalloc against chol_potrf_f32, through the C ABI.

The allocation-free entry point runs the same kernels with its workspace in
the library's device array and every launch on the caller's stream, so its
factor must equal chol_potrf_f32's bit for bit.  Each case makes two calls
in a row on the shared array, the second with a smaller input, and compares
both; the size limit is rejected before anything is launched.
"""
import pytest
import torch

from gpumode_cholesky_amd import _lib, harness

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)
LIMIT = 1024  # matrices of order > 64 per call (64 MiB / 64 KiB)


def _stream():
    return torch.cuda.current_stream(DEV).cuda_stream


def _reference(a, mode):
    lib = _lib.load()
    b, n, _ = a.shape
    L = torch.empty_like(a)
    info = torch.full((b,), -1, dtype=torch.int32, device=DEV)
    wsn = lib.chol_workspace_bytes(b, n, mode)
    ws = torch.empty((max(wsn, 1),), dtype=torch.uint8, device=DEV)
    rc = lib.chol_potrf_f32(a.data_ptr(), L.data_ptr(), b, n, n * n, n, mode, info.data_ptr(), ws.data_ptr(),
                            wsn, _stream())
    assert rc == 0, _lib.last_error()
    return L, info


def _noalloc(a, mode):
    lib = _lib.load()
    b, n, _ = a.shape
    L = torch.full_like(a, float("nan"))
    info = torch.full((b,), -1, dtype=torch.int32, device=DEV)
    rc = lib.chol_potrf_f32_noalloc(a.data_ptr(), L.data_ptr(), b, n, n * n, n, mode, info.data_ptr(), _stream())
    assert rc == 0, _lib.last_error()
    return L, info


CASES = [((640, 512), (16, 512)), ((8, 2048), (2, 1100)), ((LIMIT, 300), (3, 300)), ((1, 4096), (1, 1000)),
         ((4096, 32), (1000, 17)), ((256, 128), (3, 100)), ((60, 1024), (4, 1024))]


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("first,second", CASES, ids=lambda c: f"{c[0]}x{c[1]}")
def test_noalloc_matches_allocating_entry(first, second, mode):
    for i, (b, n) in enumerate((first, second)):
        print(f"[noalloc] call {i + 1}: {b}x{n} mode={mode}", flush=True)
        a = harness.config_input(b, n, 3300 + 7 * i + n, device=DEV)
        ref, ref_info = _reference(a, mode)
        out, info = _noalloc(a, mode)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), f"call {i + 1} ({b}x{n}, mode {mode}) differs from chol_potrf_f32"
        assert torch.equal(info, ref_info) and (info == 0).all().item()
        del a, ref, out


def test_noalloc_non_spd_info():
    a = harness.generate_input(5, 300, 2, 77, device=DEV)
    a[2, 150, 150] = -1.0
    out, info = _noalloc(a, 0)
    torch.cuda.synchronize()
    assert info.cpu().tolist() == [0, 0, 151, 0, 0]
    assert torch.isnan(out[2, 150, 150]).item()


def test_noalloc_rejects_over_limit():
    lib = _lib.load()
    a = torch.zeros((1, 300, 300), device=DEV)
    rc = lib.chol_potrf_f32_noalloc(a.data_ptr(), a.data_ptr() + 4, LIMIT + 1, 300, 0, 300, 0, None, _stream())
    assert rc == _lib.E_WORKSPACE
    assert "scratch" in _lib.last_error() and "1024" in _lib.last_error()
    torch.cuda.synchronize()
    # orders <= 64 take no scratch, so their batch has no such limit
    b = harness.config_input(8192, 64, 5, device=DEV)
    out, info = _noalloc(b, 0)
    ref, _ = _reference(b, 0)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
