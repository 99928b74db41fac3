"""hipGraph replay of the factorisation: every replay must equal the eager
result bit for bit, with eager work and synchronisations between replays.
The cases include blocked shapes whose eager calls use the library's side
stream (640 x 512 zero fill; 2 x 2048 and 1 x 4096 bulk updates), which
capture on the capturing stream alone, and an eager call from another
thread while a capture is in progress."""
import threading

import pytest
import torch

from gpumode_cholesky_amd import cholesky, harness
from gpumode_cholesky_amd.graphs import GraphedCholesky

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


@pytest.mark.parametrize("branches", [False, True])
@pytest.mark.parametrize("batch,n,mode", [(4096, 32, "plain"), (1024, 64, "plain"), (64, 256, "plain"),
                                          (4, 1024, "plain"), (4, 1024, "mixed"), (640, 512, "plain"),
                                          (2, 2048, "mixed"), (1, 4096, "mixed"), (1, 4096, "plain"),
                                          (4, 256, "mixed"), (512, 300, "mixed"), (512, 300, "plain")])
def test_graph_replays_match_eager(batch, n, mode, branches):
    # (512, 300): the fused route (its default for this shape).  Batches of
    # 256 and more take dense inputs: the spectral generator's batched QR
    # is slow there, and replay equality does not depend on the family.
    print(f"[graph] {batch}x{n} {mode} branches={branches}", flush=True)

    def gen(seed):
        if batch >= 256:
            return harness.dense_input(batch, n, 1.0, seed, device=DEV)
        return harness.config_input(batch, n, seed, device=DEV)

    inputs = [gen(900 + i) for i in range(2)]
    eager = [cholesky(a, mode=mode) for a in inputs]
    g = GraphedCholesky(inputs, mode=mode, branches=branches)
    for rep in range(3):
        outs = g.replay()
        torch.cuda.synchronize()
        for i, out in enumerate(outs):
            assert torch.equal(out, eager[i]), f"replay {rep}, input {i} differs"
        # eager work between replays
        other = cholesky(gen(5), mode=mode)
        torch.cuda.synchronize()
        del other
    # refresh an input in place: the replay factors the new contents
    fresh = gen(777)
    inputs[0].copy_(fresh)
    outs = g.replay()
    torch.cuda.synchronize()
    assert torch.equal(outs[0], cholesky(fresh, mode=mode))
    assert torch.equal(outs[1], eager[1])


def test_eager_call_on_other_thread_during_capture():
    # A blocked call from another host thread (its own stream, using the
    # side stream) while this thread captures a blocked call: both results
    # must equal their eager factors.
    a = harness.config_input(2, 2048, 41, device=DEV)
    b = harness.config_input(1, 4096, 42, device=DEV)
    ref_a = cholesky(a, mode="mixed")
    ref_b = cholesky(b, mode="mixed")
    torch.cuda.synchronize()
    got = {}

    def worker():
        s = torch.cuda.Stream(device=DEV)
        with torch.cuda.stream(s):
            outs = [cholesky(b, mode="mixed") for _ in range(4)]
        s.synchronize()
        got["b"] = outs

    t = threading.Thread(target=worker)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, capture_error_mode="thread_local"):
        t.start()
        out_a = cholesky(a, mode="mixed")
    t.join()
    g.replay()
    torch.cuda.synchronize()
    assert torch.equal(out_a, ref_a)
    for o in got["b"]:
        assert torch.equal(o, ref_b)
