"""hipGraph capture and replay of the factorisation for fixed shapes.

A ``GraphedCholesky`` captures ``cholesky`` once per (static) input buffer
and replays the captured launch sequence afterwards, removing the host-side
cost of the Python/ctypes call (the kernels themselves are unchanged).  The
caller refreshes the input buffers in place and calls ``replay``.  The
library allocates nothing on the hot path and records no memset, so its
launches capture as plain kernel nodes; the output tensors belong to the
graph's memory pool and are overwritten by every replay.

The reference captures its split32 launch train with torch.cuda.CUDAGraph
and a shared pool (submission.py:2606-2661, :2716-2728).
"""
from __future__ import annotations

import torch

from .cholesky import cholesky


class GraphedCholesky:
    """One captured graph per input buffer of ``inputs`` (same device)."""

    def __init__(self, inputs, mode: str = "plain"):
        self.inputs = list(inputs)
        self.mode = mode
        self.graphs = []
        self.outputs = []
        dev = self.inputs[0].device
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):  # warm-up outside capture
            for a in self.inputs:
                cholesky(a, mode=mode)
        torch.cuda.current_stream(dev).wait_stream(side)
        torch.cuda.synchronize(dev)
        pool = torch.cuda.graph_pool_handle()
        for a in self.inputs:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, pool=pool):
                out = cholesky(a, mode=mode)
            self.graphs.append(g)
            self.outputs.append(out)

    def replay(self, i: int) -> torch.Tensor:
        """Replay graph ``i`` on the current stream; returns its output buffer."""
        self.graphs[i].replay()
        return self.outputs[i]

    def __len__(self):
        return len(self.graphs)
