"""Callers on several host threads and streams at once.

The blocked route shares one library-owned side stream per device between
its callers (created on the first blocked call, with two events); each
call's fork / join steps are enqueued under a per-device lock.  Two host
threads, each on its own torch stream, factor different shapes in a loop;
every result must equal the single-threaded factor bit for bit.
"""
import threading

import pytest
import torch

from gpumode_cholesky_amd import cholesky, harness

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_two_threads_two_streams(mode):
    jobs = [harness.config_input(2, 2048, 51, device=DEV), harness.config_input(640, 512, 52, device=DEV),
            harness.config_input(1, 4096, 53, device=DEV), harness.config_input(4096, 32, 54, device=DEV)]
    refs = [cholesky(a, mode=mode) for a in jobs]
    torch.cuda.synchronize()
    results = {}
    errors = []

    def worker(tid, order):
        try:
            s = torch.cuda.Stream(device=DEV)
            outs = []
            with torch.cuda.stream(s):
                for _ in range(3):
                    for j in order:
                        outs.append((j, cholesky(jobs[j], mode=mode)))
            s.synchronize()
            results[tid] = outs
        except Exception as exc:  # surfaced below
            errors.append(exc)

    threads = [threading.Thread(target=worker, args=(0, [0, 1, 2, 3])),
               threading.Thread(target=worker, args=(1, [2, 3, 1, 0]))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for tid, outs in results.items():
        for j, out in outs:
            assert torch.equal(out, refs[j]), f"thread {tid}, job {j} differs from the single-threaded factor"


@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_ragged_beside_dense_threads(mode):
    # A ragged call (lock step on the caller's stream, zero fill, fused class
    # and n <= 64 queues on the side stream) on one thread while another
    # thread's blocked calls use the same side stream: every result equals
    # its single-threaded factor bit for bit.
    from gpumode_cholesky_amd.ragged import cholesky_ragged

    sizes = [1100, 700, 513, 512, 300, 129, 65, 64, 33, 17, 5, 900, 257]
    mats = [harness.generate_input(1, n, 2, 7100 + i, device=DEV)[0] for i, n in enumerate(sizes)]
    dense = [harness.config_input(2, 2048, 61, device=DEV), harness.config_input(640, 512, 62, device=DEV)]
    ref_r = cholesky_ragged(mats, mode=mode)
    ref_d = [cholesky(a, mode=mode) for a in dense]
    torch.cuda.synchronize()
    results, errors = {}, []

    def ragged_worker():
        try:
            s = torch.cuda.Stream(device=DEV)
            with torch.cuda.stream(s):
                outs = [cholesky_ragged(mats, mode=mode) for _ in range(3)]
            s.synchronize()
            results["ragged"] = outs
        except Exception as exc:
            errors.append(exc)

    def dense_worker():
        try:
            s = torch.cuda.Stream(device=DEV)
            with torch.cuda.stream(s):
                outs = [(j, cholesky(dense[j], mode=mode)) for _ in range(3) for j in (0, 1)]
            s.synchronize()
            results["dense"] = outs
        except Exception as exc:
            errors.append(exc)

    threads = [threading.Thread(target=ragged_worker), threading.Thread(target=dense_worker)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for outs in results["ragged"]:
        for i, (L, R) in enumerate(zip(outs, ref_r)):
            assert torch.equal(L, R), f"ragged matrix {i} (n={sizes[i]}) differs from the single-threaded call"
    for j, out in results["dense"]:
        assert torch.equal(out, ref_d[j]), f"dense job {j} differs from the single-threaded factor"
