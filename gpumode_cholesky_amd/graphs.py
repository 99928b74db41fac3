"""hipGraph capture and replay of the factorisation for fixed shapes.

``GraphedCholesky`` captures ``cholesky`` over a list of static input
buffers into ONE graph (one kernel sequence per input, back to back) and
replays it: a single graph launch then stands for len(inputs)
factorisations, so the host cost of launching is paid once per sequence
instead of once per call.  The caller refreshes the input buffers in place
and calls ``replay``; the outputs belong to the graph's memory pool and are
overwritten by every replay.  The library allocates nothing and records no
memset, so its launches capture as plain kernel nodes; under capture the
blocked route keeps all of its work on the capturing stream.

The reference captures its split32 launch train with torch.cuda.CUDAGraph
and a shared pool, one graph per rotated input (submission.py:2606-2661,
:2716-2728).
"""
from __future__ import annotations

import torch

from .cholesky import cholesky


class GraphedCholesky:
    """One captured graph over every input buffer of ``inputs`` (one device)."""

    def __init__(self, inputs, mode: str = "plain"):
        self.inputs = list(inputs)
        self.mode = mode
        dev = self.inputs[0].device
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):  # warm-up outside capture
            for a in self.inputs:
                cholesky(a, mode=mode)
        torch.cuda.current_stream(dev).wait_stream(side)
        torch.cuda.synchronize(dev)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.outputs = [cholesky(a, mode=mode) for a in self.inputs]

    def replay(self):
        """Replay the sequence on the current stream; returns the output
        buffers (one per input, in order)."""
        self.graph.replay()
        return self.outputs

    def __len__(self):
        return len(self.inputs)
