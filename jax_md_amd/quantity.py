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
