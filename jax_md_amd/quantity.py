"""Forces and kinetic quantities (jax_md/quantity.py:58-199)."""

from __future__ import annotations

import torch

from . import smap


class NativeForce:
  """-dE/dR of a native pair energy, evaluated by the gfx950 kernel directly
  (the reference differentiates the energy with jax.grad, quantity.py:58-60)."""

  _jmd_kind = 'force'

  def __init__(self, energy_fn: smap.PairEnergy):
    self.energy_fn = energy_fn

  def __call__(self, R, neighbor=None, **kwargs):
    return self.energy_fn.force(R, neighbor, **kwargs)


def force(energy_fn):
  """F = -grad E.  Native energies return a NativeForce; other differentiable
  torch energies go through torch.autograd."""
  if isinstance(energy_fn, smap.PairEnergy):
    return NativeForce(energy_fn)

  def force_fn(R, *args, **kwargs):
    R = R.detach().requires_grad_(True)
    with torch.enable_grad():
      E = energy_fn(R, *args, **kwargs)
      (g,) = torch.autograd.grad(E, R)
    return -g

  return force_fn


def canonicalize_force(energy_or_force_fn):
  """Energy -> force, force -> itself (quantity.py:76-102).  The reference
  decides by output shape; here native energies and forces are recognised by
  type, and any other callable is treated as an energy."""
  kind = getattr(energy_or_force_fn, '_jmd_kind', None)
  if kind == 'force':
    return energy_or_force_fn
  return force(energy_or_force_fn)


def native_energy_of(energy_or_force_fn) -> smap.PairEnergy:
  """The native pair energy behind an energy or force function, for the
  integrators' fused kernels; raises for anything else."""
  if isinstance(energy_or_force_fn, smap.PairEnergy):
    return energy_or_force_fn
  if isinstance(energy_or_force_fn, NativeForce):
    return energy_or_force_fn.energy_fn
  raise NotImplementedError(
      'The native integrators run energies built by jax_md_amd.energy.*_neighbor_list '
      '(or their quantity.force); got ' f'{type(energy_or_force_fn).__name__}.')


def _mass_col(mass, like):
  if isinstance(mass, torch.Tensor) and mass.ndim > 0:
    return mass.reshape(-1, 1).to(like.device, torch.float64)
  return float(mass)


def kinetic_energy(*unused_args, momentum=None, velocity=None, mass=1.0):
  """0.5 sum p^2/m summed in float64 (quantity.py:124-159)."""
  if unused_args:
    raise ValueError('To use the kinetic energy function, you must explicitly pass either '
                     'momentum or velocity as a keyword argument.')
  if (momentum is None) == (velocity is None):
    raise ValueError('pass exactly one of momentum or velocity')
  q = velocity if momentum is None else momentum
  m = _mass_col(mass, q)
  q64 = q.to(torch.float64)
  k = q64 ** 2 * m if momentum is None else q64 ** 2 / m
  return (0.5 * k.sum()).to(q.dtype)


def temperature(*unused_args, momentum=None, velocity=None, mass=1.0):
  """sum p^2/m / dof (quantity.py:162-199)."""
  if unused_args:
    raise ValueError('pass momentum or velocity as a keyword argument')
  if (momentum is None) == (velocity is None):
    raise ValueError('pass exactly one of momentum or velocity')
  q = velocity if momentum is None else momentum
  m = _mass_col(mass, q)
  q64 = q.to(torch.float64)
  t = q64 ** 2 * m if momentum is None else q64 ** 2 / m
  return (t.sum() / q.numel()).to(q.dtype)


def _box_volume(box, dim):
  """volume(dim, box) (quantity.py:285-300): a scalar side, a [dim] vector of
  sides or a [dim, dim] box matrix."""
  import numpy as onp
  b = box.detach().cpu().numpy() if isinstance(box, torch.Tensor) else onp.asarray(box)
  b = b.astype(onp.float64)
  if b.ndim == 0:
    return float(b) ** dim
  if b.ndim == 1:
    return float(onp.prod(b))
  return float(abs(onp.linalg.det(b)))


def pressure(energy_fn, position, box, kinetic_energy=0.0, **kwargs):
  """(2 KE - dU/d eps) / (dim V) with every displacement scaled by (1 + eps)
  (quantity.py:202-236): the trace of the native virial kernel's tensor.
  ``neighbor=`` (and dynamic parameters) go in ``kwargs`` as for the energy
  function.  Returns a float64 0-d tensor."""
  energy = native_energy_of(energy_fn)
  neighbor = kwargs.pop('neighbor', None)
  dim = position.shape[1]
  W = energy.virial(position, neighbor, **kwargs)
  ke = kinetic_energy.to(torch.float64) if isinstance(kinetic_energy, torch.Tensor) \
      else float(kinetic_energy)
  return (2.0 * ke - torch.trace(W)) / (dim * _box_volume(box, dim))


def stress(energy_fn, position, box, mass=1.0, velocity=None, **kwargs):
  """(sum m v v^T - dU/d eps) / V with every displacement scaled by
  (I + eps) (quantity.py:239-282).  Returns float64 [dim, dim]."""
  energy = native_energy_of(energy_fn)
  neighbor = kwargs.pop('neighbor', None)
  dim = position.shape[1]
  W = energy.virial(position, neighbor, **kwargs)
  vv = torch.zeros_like(W)
  if velocity is not None:
    v = velocity.to(torch.float64)
    m = _mass_col(mass, v)
    vv = ((m * v)[:, :, None] * v[:, None, :]).sum(dim=0)
  return (vv - W) / _box_volume(box, dim)
