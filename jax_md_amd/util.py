"""Scalar conventions of the reference (jax_md/util.py) on the torch side."""

from __future__ import annotations

import numpy as onp
import torch

f16 = onp.float16
f32 = onp.float32
f64 = onp.float64
i32 = onp.int32


def static_cast(*xs):
  """Casts each Python scalar to the smallest NumPy dtype that holds it, as
  ``util.static_cast`` does on CPU/GPU (jax_md/util.py:58-64): 0.1 becomes
  float16, 5 becomes uint8."""
  return tuple(onp.asarray(x, dtype=onp.min_scalar_type(x))[()] for x in xs)


def maybe_downcast(x):
  """float32 unless already a float64 tensor (jax_md/util.py:109-112)."""
  if isinstance(x, torch.Tensor) and x.dtype == torch.float64:
    return x
  if isinstance(x, torch.Tensor):
    return x.to(torch.float32)
  return onp.asarray(x, onp.float32)


def as_f32_scalar(x) -> float:
  """Python float holding the float32 value of a scalar (tensor, array, float)."""
  if isinstance(x, torch.Tensor):
    x = x.detach().cpu().numpy()
  return float(onp.float32(onp.asarray(x, onp.float32)))


def merge_dicts(a, b, ignore_unused_parameters=False):
  """jax_md/util.py:73-82."""
  if not ignore_unused_parameters:
    return {**a, **b}
  merged = dict(a)
  for key in merged.keys():
    b_val = b.get(key)
    if b_val is not None:
      merged[key] = b_val
  return merged
