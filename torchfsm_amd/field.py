"""Forcing fields of the reference that the engine integrates (field.py:128-148)."""

import torch

from .operator import ExplicitSource, ImplicitSource, Operator


def kolm_force(x: torch.Tensor, drag_coef: float = -0.1, k: float = 4.0, length_scale: float = 1.0) -> Operator:
    """Kolmogorov forcing in vorticity form: a w - k cos(k l x)  (field.py:128-148)."""
    return drag_coef * ImplicitSource() - ExplicitSource(k * torch.cos(k * length_scale * x))
