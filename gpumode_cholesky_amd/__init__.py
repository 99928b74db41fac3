"""MI355X (gfx950) batched Cholesky engine.

``custom_kernel(data) -> L`` is the drop-in for the GPU MODE ``cholesky``
entry point; ``cholesky(A, mode=...)`` exposes the precision mode and info.
All arithmetic runs in hand-written HIP kernels of ``libchol_amd.so``.
"""
from .cholesky import cholesky, custom_kernel, workspace_bytes  # noqa: F401
from .task import TestSpec, input_t, output_t  # noqa: F401

__all__ = ["cholesky", "custom_kernel", "workspace_bytes", "TestSpec",
           "input_t", "output_t"]
