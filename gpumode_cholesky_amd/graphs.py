"""hipGraph capture and replay of the factorisation for fixed shapes.

``GraphedCholesky`` captures ``cholesky`` over a list of static input
buffers into ONE graph and replays it: a single graph launch stands for
len(inputs) factorisations, so the host cost of launching is paid once per
sequence instead of once per call.  The caller refreshes the input buffers
in place and calls ``replay``; the outputs belong to the graph's memory
pool and are overwritten by every replay.  The library allocates nothing
and records no memset, so its launches capture as plain kernel nodes;
under capture the blocked route keeps all of its work on the capturing
stream.

``branches=True`` captures each input's factorisation on a stream of its
own (forked from and joined back into the capture), so the graph holds
independent branches that the GPU may run side by side: a win when one
call leaves most CUs idle (a few matrices of a few hundred), a loss when
each call already fills the GPU.

The reference captures its split32 launch train with torch.cuda.CUDAGraph
and a shared pool, one graph per rotated input (submission.py:2606-2661,
:2716-2728).
"""
from __future__ import annotations

import torch

from .cholesky import cholesky


class GraphedCholesky:
    """One captured graph over every input buffer of ``inputs`` (one device)."""

    def __init__(self, inputs, mode: str = "plain", branches: bool = False):
        self.inputs = list(inputs)
        self.mode = mode
        self.branches = branches
        dev = self.inputs[0].device
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):  # warm-up outside capture
            for a in self.inputs:
                cholesky(a, mode=mode)
        torch.cuda.current_stream(dev).wait_stream(side)
        torch.cuda.synchronize(dev)
        self.graph = torch.cuda.CUDAGraph()
        self._streams = [torch.cuda.Stream(device=dev) for _ in self.inputs] if branches else []
        with torch.cuda.graph(self.graph):
            if branches:
                cap = torch.cuda.current_stream(dev)
                self.outputs = []
                for a, s in zip(self.inputs, self._streams):
                    s.wait_stream(cap)
                    with torch.cuda.stream(s):
                        self.outputs.append(cholesky(a, mode=mode))
                for s in self._streams:
                    cap.wait_stream(s)
            else:
                self.outputs = [cholesky(a, mode=mode) for a in self.inputs]

    def replay(self):
        """Replay the sequence on the current stream; returns the output
        buffers (one per input, in order)."""
        self.graph.replay()
        return self.outputs

    def __len__(self):
        return len(self.inputs)
