"""Spaces (jax_md/space.py), carrying the geometry the native kernels need.

``periodic(side)`` and ``free()`` return ``(displacement_fn, shift_fn)``
callables with the reference semantics on torch tensors.  The callables also
describe their geometry (``side`` as float32, ``periodic``), which is how
``partition.neighbor_list``, ``smap.pair_neighbor_list`` and the integrators
route to the gfx950 kernels.  Other spaces (``periodic_general``) are outside
this package's scope.
"""

from __future__ import annotations

from typing import Optional

import numpy as onp
import torch


def _side_f32(side) -> onp.ndarray:
  if isinstance(side, torch.Tensor):
    side = side.detach().cpu().numpy()
  side = onp.asarray(side, onp.float32)
  if side.ndim > 1:
    raise NotImplementedError(
        'Box matrices (space.periodic_general) are outside this package; pass a '
        'scalar side or a vector of per-dimension sides.')
  return side


def _mod(x: torch.Tensor, y) -> torch.Tensor:
  """jnp.mod semantics: fmod, then add y where the sign differs."""
  r = torch.fmod(x, y)
  fix = (r != 0) & ((r < 0) != (torch.as_tensor(y, device=x.device) < 0))
  return torch.where(fix, r + y, r)


def periodic_displacement(side, dR: torch.Tensor) -> torch.Tensor:
  """mod(dR + side/2, side) - side/2 (jax_md/space.py:213-224)."""
  s = torch.as_tensor(side, dtype=dR.dtype, device=dR.device)
  half = s * 0.5
  return _mod(dR + half, s) - half


def square_distance(dR: torch.Tensor) -> torch.Tensor:
  """sum(dR**2, -1), accumulated left to right (jax_md/space.py:227-235)."""
  acc = dR[..., 0] * dR[..., 0]
  for k in range(1, dR.shape[-1]):
    acc = acc + dR[..., k] * dR[..., k]
  return acc


def distance(dR: torch.Tensor) -> torch.Tensor:
  """safe_mask(dr2 > 0, sqrt, dr2) (jax_md/space.py:238-247)."""
  dr2 = square_distance(dR)
  return torch.where(dr2 > 0, torch.sqrt(torch.where(dr2 > 0, dr2, torch.ones_like(dr2))),
                     torch.zeros_like(dr2))


def periodic_shift(side, R: torch.Tensor, dR: torch.Tensor) -> torch.Tensor:
  """mod(R + dR, side) (jax_md/space.py:250-252)."""
  s = torch.as_tensor(side, dtype=R.dtype, device=R.device)
  return _mod(R + dR, s)


class Displacement:
  """displacement_fn(Ra, Rb) = wrap(Ra - Rb) (or Ra - Rb in free space)."""

  def __init__(self, side: Optional[onp.ndarray], periodic: bool):
    self.side = side
    self.periodic = periodic

  def __call__(self, Ra, Rb, perturbation=None, **unused_kwargs):
    if 'box' in unused_kwargs:
      raise ValueError('`space.periodic` does not accept a box argument.')
    dR = Ra - Rb
    if self.periodic:
      dR = periodic_displacement(torch.as_tensor(self.side), dR)
    if perturbation is not None:
      dR = dR @ torch.as_tensor(perturbation, dtype=dR.dtype, device=dR.device).T
    return dR

  def side_values(self, dim: int):
    if not self.periodic:
      return [1.0] * dim
    s = onp.asarray(self.side, onp.float32)
    if s.ndim == 0:
      return [float(s)] * dim
    if s.shape[0] != dim:
      raise ValueError(f'side has {s.shape[0]} entries for dimension {dim}')
    return [float(v) for v in s]


class Shift:
  """shift_fn(R, dR): mod(R + dR, side) when wrapped, else R + dR."""

  def __init__(self, side: Optional[onp.ndarray], periodic: bool, wrapped: bool = True):
    self.side = side
    self.periodic = periodic
    self.wrapped = wrapped

  @property
  def wraps(self):
    return self.periodic and self.wrapped

  def __call__(self, R, dR, **unused_kwargs):
    if 'box' in unused_kwargs:
      raise ValueError('`space.periodic` does not accept a box argument.')
    if self.wraps:
      return periodic_shift(torch.as_tensor(self.side), R, dR)
    return R + dR


class Metric:
  """metric(displacement): |displacement(Ra, Rb)| (jax_md/space.py:475-478)."""

  def __init__(self, displacement: Displacement):
    self.displacement = displacement

  def __call__(self, Ra, Rb, **kwargs):
    return distance(self.displacement(Ra, Rb, **kwargs))


def periodic(side, wrapped: bool = True):
  """Periodic hypercube of side `side` (jax_md/space.py:275-329)."""
  s = _side_f32(side)
  return Displacement(s, True), Shift(s, True, wrapped)


def free():
  """Free boundary conditions (jax_md/space.py:258-272)."""
  return Displacement(None, False), Shift(None, False)


def metric(displacement):
  return Metric(displacement)


def canonicalize_displacement_or_metric(displacement_or_metric):
  """Returns a metric for a displacement (jax_md/space.py:505-522)."""
  if isinstance(displacement_or_metric, Displacement):
    return Metric(displacement_or_metric)
  return displacement_or_metric


def geometry_of(displacement_or_metric):
  """(periodic, side) of a space built here; raises for foreign callables."""
  d = displacement_or_metric
  if isinstance(d, Metric):
    d = d.displacement
  if isinstance(d, Displacement):
    return d
  raise NotImplementedError(
      'The native neighbour-list path needs a displacement from '
      'jax_md_amd.space.periodic or jax_md_amd.space.free; got '
      f'{type(displacement_or_metric).__name__}.')


def map_product(metric_or_displacement):
  """All pairs: out[i, j] = f(R_a[i], R_b[j]) (jax_md/space.py:481-485)."""
  return lambda Ra, Rb, **kw: metric_or_displacement(Ra[:, None, :], Rb[None, :, :], **kw)


def map_bond(metric_or_displacement):
  """Bonds: out[k] = f(R_a[k], R_b[k]) (jax_md/space.py:488-491)."""
  return lambda Ra, Rb, **kw: metric_or_displacement(Ra, Rb, **kw)


def map_neighbor(metric_or_displacement):
  """Neighbourhoods: out[i, k] = f(R_neigh[i, k], R[i]) (jax_md/space.py:494-502)."""
  return lambda Ra, Rb, **kw: metric_or_displacement(Rb, Ra[:, None, :], **kw)
