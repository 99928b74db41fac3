"""Boundary types of the path, as in reference/task.py:6-15."""
from typing import TypeVar, TypedDict

try:  # Python >= 3.11
    from typing import NotRequired
except ImportError:  # Python 3.10
    from typing_extensions import NotRequired

import torch

input_t = TypeVar("input_t", bound=torch.Tensor)
output_t = TypeVar("output_t", bound=torch.Tensor)


class TestSpec(TypedDict):
    batch: int
    n: int
    cond: int
    seed: int
    case: NotRequired[str]
