"""``from gpumode_cholesky_amd.submission import custom_kernel`` mirrors the
reference harness's ``from submission import custom_kernel``
(reference/eval.py:107)."""
from .cholesky import custom_kernel  # noqa: F401
