"""hipGraph replay of the factorisation: every replay must equal the eager
result bit for bit, with eager work and synchronisations between replays."""
import pytest
import torch

from gpumode_cholesky_amd import cholesky, harness
from gpumode_cholesky_amd.graphs import GraphedCholesky

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


@pytest.mark.parametrize("batch,n,mode", [(4096, 32, "plain"), (1024, 64, "plain"), (64, 256, "plain"),
                                          (4, 1024, "plain"), (4, 1024, "mixed")])
def test_graph_replays_match_eager(batch, n, mode):
    inputs = [harness.config_input(batch, n, 900 + i, device=DEV) for i in range(2)]
    eager = [cholesky(a, mode=mode) for a in inputs]
    g = GraphedCholesky(inputs, mode=mode)
    for rep in range(3):
        for i in range(len(g)):
            out = g.replay(i)
            torch.cuda.synchronize()
            assert torch.equal(out, eager[i]), f"replay {rep} of graph {i} differs"
        # eager work between replays
        other = cholesky(harness.config_input(batch, n, 5, device=DEV), mode=mode)
        torch.cuda.synchronize()
        del other
    # refresh an input in place: the replay factors the new contents
    fresh = harness.config_input(batch, n, 777, device=DEV)
    inputs[0].copy_(fresh)
    out = g.replay(0)
    torch.cuda.synchronize()
    assert torch.equal(out, cholesky(fresh, mode=mode))
