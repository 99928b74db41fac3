"""Out-of-range stores: every buffer a call writes (L, info, the workspace)
and the input A sit inside larger allocations whose margins hold a
sentinel pattern; after the call (our kernels only, nothing else on the
GPU) the margins must be unchanged and A untouched.  Covers every route
(register n <= 64, LDS, blocked with partial last blocks: 300 = 2*128+44,
320 = 2*128+64, 1100 = 8*128+76, fused one-workgroup-per-matrix), both
forms, unaligned leading dimensions, the ragged entry point (its orders
65..512 on the fused kernel, the larger ones in lock step), the triangular
inverse and the backward."""
import numpy as np
import pytest
import torch

from gpumode_cholesky_amd import _lib, autotune, harness

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)
PAD = 1 << 20            # elements of margin on each side (4 MiB of floats)
SENT_F = -1234.5678      # float sentinel
SENT_B = 0xA5            # byte sentinel of the workspace margins
SENT_I = -77             # int32 sentinel of info


class Guarded:
    """A device view of `count` elements inside a buffer with PAD-element
    margins on both sides filled with the sentinel."""

    def __init__(self, count, dtype, sentinel, fill=None):
        self.count = count
        self.sentinel = sentinel
        self.buf = torch.full((PAD + max(count, 1) + PAD,), sentinel, dtype=dtype, device=DEV)
        self.view = self.buf[PAD:PAD + count]
        if fill is not None:
            self.view.copy_(fill.reshape(-1))

    def ptr(self):
        return self.view.data_ptr()

    def margins_intact(self):
        lo = self.buf[:PAD]
        hi = self.buf[PAD + self.count:]
        return bool((lo == self.sentinel).all().item()) and bool((hi == self.sentinel).all().item())


def _stream():
    return torch.cuda.current_stream(DEV).cuda_stream


def _potrf_guarded(a, ld, mode):
    """chol_potrf_f32 on guarded A (rows padded to ld with NaN), L, info, ws."""
    batch, n, _ = a.shape
    lib = _lib.load()
    Ap = torch.full((batch, n, ld), float("nan"), device=DEV)
    Ap[:, :, :n] = a
    A = Guarded(batch * n * ld, torch.float32, SENT_F, Ap)
    L = Guarded(batch * n * ld, torch.float32, SENT_F)
    info = Guarded(batch, torch.int32, SENT_I)
    wsn = lib.chol_workspace_bytes(batch, n, mode)
    ws = Guarded(wsn, torch.uint8, SENT_B)
    a_before = A.buf.clone()
    print(f"[canary] potrf {batch}x{n} ld={ld} mode={mode} ws={wsn}", flush=True)
    rc = lib.chol_potrf_f32(A.ptr(), L.ptr(), batch, n, n * ld, ld, mode, info.ptr(), ws.ptr(), wsn, _stream())
    torch.cuda.synchronize()
    assert rc == 0, _lib.last_error()
    assert torch.equal(torch.nan_to_num(A.buf, nan=7.0), torch.nan_to_num(a_before, nan=7.0)), "A was written"
    for name, g in (("L", L), ("info", info), ("workspace", ws)):
        assert g.margins_intact(), f"{batch}x{n} ld={ld} mode={mode}: store outside {name}"
    Lv = L.view.reshape(batch, n, ld)
    assert (info.view == 0).all().item()
    assert (Lv[:, :, n:] == SENT_F).all().item(), "store into the row padding of L"
    return Lv[:, :, :n]


CASES = [(3, n) for n in (1, 5, 16, 31, 32, 33, 63, 64, 65, 100, 127, 128, 129, 200, 255, 256, 257, 300, 320,
                          383, 512, 640, 1100)] + [(1, 300), (8, 320), (1, 1100), (40, 300), (300, 129)]


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("batch,n", CASES)
def test_no_store_outside_buffers(batch, n, mode):
    a = harness.generate_input(batch, n, 2, 8800 + n, device=DEV)
    for ld in (n, n + 3):
        L = _potrf_guarded(a, ld, mode)
        r = harness.fp64_residual(a, L.contiguous())
        assert r < 1e-5, (batch, n, ld, r)


@pytest.mark.parametrize("n", [65, 100, 128, 200, 256])
def test_no_store_outside_buffers_lds_route(n):
    try:
        autotune.set_table([[n, n, 1, 1 << 20, -1]])
        a = harness.generate_input(3, n, 2, 8900 + n, device=DEV)
        _potrf_guarded(a, n, 0)
        _potrf_guarded(a, n + 1, 0)
    finally:
        autotune.load_table()


@pytest.mark.parametrize("n", [65, 100, 128, 129, 200, 256, 300, 320, 383, 512, 700])
def test_no_store_outside_buffers_fused_route(n):
    try:
        autotune.set_table([[n, n, 1, 1 << 20, -2]])
        a = harness.generate_input(3, n, 2, 9200 + n, device=DEV)
        for mode in (0, 1):
            for ld in (n, n + 3):
                L = _potrf_guarded(a, ld, mode)
                r = harness.fp64_residual(a, L.contiguous())
                assert r < 1e-5, (n, ld, mode, r)
    finally:
        autotune.load_table()


@pytest.mark.parametrize("n", [300, 320, 1100])
@pytest.mark.parametrize("nbo", [128, 384])
def test_no_store_outside_buffers_outer_widths(n, nbo):
    try:
        autotune.set_table([[n, n, 1, 1 << 20, nbo]])
        a = harness.generate_input(3, n, 2, 9000 + n, device=DEV)
        for mode in (0, 1):
            _potrf_guarded(a, n, mode)
    finally:
        autotune.load_table()


def test_no_store_outside_buffers_ragged():
    lib = _lib.load()
    ns = [300, 5, 320, 64, 1100, 33, 129, 256, 65, 17]
    mats = [harness.generate_input(1, n, 2, 9100 + i, device=DEV)[0] for i, n in enumerate(ns)]
    # every output inside its own guarded buffer
    outs = [Guarded(n * n, torch.float32, SENT_F) for n in ns]
    n_host = np.array(ns, dtype=np.int64)
    a_host = np.array([m.data_ptr() for m in mats], dtype=np.uint64)
    l_host = np.array([o.ptr() for o in outs], dtype=np.uint64)
    wsn = lib.chol_ragged_workspace_bytes(n_host.ctypes.data, len(ns))
    ws = Guarded(wsn, torch.uint8, SENT_B)
    info = Guarded(len(ns), torch.int32, SENT_I)
    for mode in (0, 1):
        print(f"[canary] ragged {ns} mode={mode}", flush=True)
        rc = lib.chol_potrf_ragged_f32(a_host.ctypes.data, l_host.ctypes.data, n_host.ctypes.data, None, len(ns),
                                       mode, info.ptr(), ws.ptr(), wsn, _stream())
        torch.cuda.synchronize()
        assert rc == 0, _lib.last_error()
        assert ws.margins_intact() and info.margins_intact()
        for n, o, a in zip(ns, outs, mats):
            assert o.margins_intact(), f"ragged n={n}: store outside its factor"
            assert harness.fp64_residual(a[None], o.view.reshape(1, n, n)) < 1e-5


@pytest.mark.parametrize("batch,n", [(3, 300), (3, 320), (2, 1100), (5, 65)])
def test_no_store_outside_buffers_inverse_and_backward(batch, n):
    lib = _lib.load()
    a = harness.spectral_input(batch, n, 1e2, 9200 + n, device=DEV)
    Lf = torch.linalg.cholesky(a.cpu()).to(DEV)  # factor from the host (LAPACK)
    gL = torch.tril(torch.randn(batch, n, n, device=DEV))
    X = Guarded(batch * n * n, torch.float32, SENT_F)
    wsn = lib.chol_tri_inverse_workspace_bytes(batch, n)
    ws = Guarded(wsn, torch.uint8, SENT_B)
    print(f"[canary] tri_inverse {batch}x{n}", flush=True)
    rc = lib.chol_tri_inverse_f32(Lf.data_ptr(), X.ptr(), batch, n, ws.ptr(), wsn, _stream())
    torch.cuda.synchronize()
    assert rc == 0, _lib.last_error()
    assert X.margins_intact() and ws.margins_intact()
    G = Guarded(batch * n * n, torch.float32, SENT_F)
    wsn = lib.chol_backward_workspace_bytes(batch, n)
    ws2 = Guarded(wsn, torch.uint8, SENT_B)
    for have in (0, 1):
        print(f"[canary] backward {batch}x{n} have_linv={have}", flush=True)
        rc = lib.chol_backward_f32(Lf.data_ptr(), gL.data_ptr(), G.ptr(), batch, n,
                                   X.ptr() if have else None, have, ws2.ptr(), wsn, _stream())
        torch.cuda.synchronize()
        assert rc == 0, _lib.last_error()
        assert G.margins_intact() and ws2.margins_intact() and X.margins_intact()


def test_no_store_outside_buffers_ragged_padded_rows():
    # L rows padded to a multiple of 4 (ld_host), as cholesky_ragged passes
    # them: the padding columns [n, ld) of every row keep the sentinel.
    lib = _lib.load()
    ns = [301, 7, 1101, 64, 130, 66, 255, 1024, 97]
    lds = [((n + 3) // 4 * 4 if n > 64 else n) for n in ns]
    mats = [harness.generate_input(1, n, 2, 9300 + i, device=DEV)[0] for i, n in enumerate(ns)]
    outs = [Guarded(n * ld, torch.float32, SENT_F) for n, ld in zip(ns, lds)]
    n_host = np.array(ns, dtype=np.int64)
    ld_host = np.array(lds, dtype=np.int64)
    a_host = np.array([m.data_ptr() for m in mats], dtype=np.uint64)
    l_host = np.array([o.ptr() for o in outs], dtype=np.uint64)
    wsn = lib.chol_ragged_workspace_bytes(n_host.ctypes.data, len(ns))
    ws = Guarded(wsn, torch.uint8, SENT_B)
    for mode in (0, 1):
        print(f"[canary] ragged padded {ns} mode={mode}", flush=True)
        rc = lib.chol_potrf_ragged_f32(a_host.ctypes.data, l_host.ctypes.data, n_host.ctypes.data,
                                       ld_host.ctypes.data, len(ns), mode, None, ws.ptr(), wsn, _stream())
        torch.cuda.synchronize()
        assert rc == 0, _lib.last_error()
        assert ws.margins_intact()
        for n, ld, o, a in zip(ns, lds, outs, mats):
            assert o.margins_intact(), f"ragged n={n}: store outside its factor"
            Lv = o.view.reshape(n, ld)
            assert (Lv[:, n:] == SENT_F).all().item(), f"ragged n={n}: store into the row padding"
            assert harness.fp64_residual(a[None], Lv[:, :n].contiguous()[None]) < 1e-5
