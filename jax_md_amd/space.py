"""Spaces (jax_md/space.py), carrying the geometry the native kernels need.

``periodic(side)``, ``periodic_general(box)`` (fractional coordinates) and
``free()`` return ``(displacement_fn, shift_fn)`` callables with the
reference semantics on torch tensors.  The callables also describe their
geometry (``side`` as float32, ``periodic``: 0 free, 1 periodic box, 2 box
matrix), which is how ``partition.neighbor_list``, ``smap.pair_neighbor_list``
and the integrators route to the gfx950 kernels.
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


def _box_as_matrix(box, dim: int) -> onp.ndarray:
  """[dim, dim] float32 matrix of a scalar, vector or matrix box."""
  b = onp.asarray(box.detach().cpu().numpy() if isinstance(box, torch.Tensor) else box,
                  onp.float32)
  if b.ndim == 0:
    return onp.eye(dim, dtype=onp.float32) * b
  if b.ndim == 1:
    if b.shape[0] != dim:
      raise ValueError(f'box has {b.shape[0]} entries for dimension {dim}')
    return onp.diag(b).astype(onp.float32)
  if b.shape != (dim, dim):
    raise ValueError(f'box of shape {b.shape} for dimension {dim}')
  return b


def transform(box, R: torch.Tensor) -> torch.Tensor:
  """box applied to R (jax_md/space.py:128-169): R * box, box * R per
  component, or the matrix product sum_j box[i, j] R[..., j]."""
  b = torch.as_tensor(onp.asarray(box, onp.float32) if not isinstance(box, torch.Tensor)
                      else box, dtype=R.dtype, device=R.device)
  if b.ndim <= 1:
    return R * b
  return torch.einsum('ij,...j->...i', b, R)


def inverse(box):
  """Inverse of an affine box (jax_md/space.py:110-125)."""
  b = onp.asarray(box, onp.float32)
  if b.ndim <= 1:
    return (onp.float32(1) / b).astype(onp.float32)
  return onp.linalg.inv(b.astype(onp.float64)).astype(onp.float32)


class GeneralDisplacement(Displacement):
  """periodic_general with fractional coordinates: d = T wrap(u_a - u_b),
  the wrap on the unit cube (jax_md/space.py:420-443)."""

  def __init__(self, box):
    super().__init__(onp.float32(1.0), True)
    self.box = box
    self.periodic = 2

  def __call__(self, Ra, Rb, perturbation=None, **kwargs):
    box = kwargs.get('new_box', kwargs.get('box', self.box))
    dR = periodic_displacement(1.0, Ra - Rb)
    dR = transform(box, dR)
    if perturbation is not None:
      dR = dR @ torch.as_tensor(perturbation, dtype=dR.dtype, device=dR.device).T
    return dR

  def side_values(self, dim: int):
    return [float(v) for v in _box_as_matrix(self.box, dim).reshape(-1)]


class GeneralShift(Shift):
  """shift_fn(u, dR) = mod(u + T^-1 dR, 1) (jax_md/space.py:450-470)."""

  def __init__(self, box, wrapped: bool = True):
    super().__init__(onp.float32(1.0), True, wrapped)
    self.box = box

  @property
  def wraps(self):
    if not self.wrapped:
      raise NotImplementedError('periodic_general(wrapped=False) is not supported by the '
                                'native integrators')
    return 2

  def side_values(self, dim: int):
    return [float(v) for v in _box_as_matrix(self.box, dim).reshape(-1)]

  def __call__(self, R, dR, **kwargs):
    box = kwargs.get('new_box', kwargs.get('box', self.box))
    dR = transform(inverse(box), dR)
    if self.wrapped:
      return periodic_shift(1.0, R, dR)
    return R + dR


def periodic_general(box, fractional_coordinates: bool = True, wrapped: bool = True):
  """Periodic parallelepiped X = T [0, 1]^d (jax_md/space.py:332-472), with
  positions stored as fractional coordinates u (x = T u).  Displacements and
  forces live in X.  ``fractional_coordinates=False`` is not supported."""
  if not fractional_coordinates:
    raise NotImplementedError('periodic_general with real-space positions '
                              '(fractional_coordinates=False) is not supported; store '
                              'fractional coordinates')
  b = onp.asarray(box.detach().cpu().numpy() if isinstance(box, torch.Tensor) else box,
                  onp.float32)
  if b.ndim > 2:
    raise ValueError(f'Box must be either: a scalar, a vector, or a matrix. Found {box}.')
  return GeneralDisplacement(b), GeneralShift(b, wrapped)


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
