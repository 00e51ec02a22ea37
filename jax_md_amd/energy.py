"""Lennard-Jones and soft-sphere energies over neighbour lists
(jax_md/energy.py:125-343,534-580).

``lennard_jones_neighbor_list`` and ``soft_sphere_neighbor_list`` keep the
reference's signatures and defaults, including the injectable
``neighbor_list_fn=`` / ``pair_neighbor_list_fn=`` hooks, and return
``(neighbor_fn, energy_fn)``.  ``lennard_jones`` / ``soft_sphere`` /
``multiplicative_isotropic_cutoff`` are the elementwise pair functions of the
reference (usable directly on torch tensors of distances); passed to
``smap.pair_neighbor_list`` they select the matching gfx950 kernel.
"""

from __future__ import annotations

from typing import Callable

import numpy as onp
import torch

from . import partition, smap, space
from .util import maybe_downcast

f32 = onp.float32


def _max_f32(x) -> float:
  if isinstance(x, torch.Tensor):
    return float(f32(x.detach().max().cpu().item()))
  return float(f32(onp.max(onp.asarray(x, f32))))


def soft_sphere(dr, sigma=1, epsilon=1, alpha=2, **unused_kwargs):
  """eps/alpha (1 - r/sigma)^alpha for r < sigma, else 0 (energy.py:125-173)."""
  dr = dr / sigma
  base = torch.clamp(1.0 - dr, min=0.0)
  out = epsilon / alpha * base ** alpha
  return torch.where(dr < 1.0, out, torch.zeros_like(out))


def lennard_jones(dr, sigma=1, epsilon=1, **unused_kwargs):
  """4 eps ((sigma/r)^12 - (sigma/r)^6) (energy.py:246-272)."""
  idr = sigma / dr
  idr = idr * idr
  idr6 = idr * idr * idr
  idr12 = idr6 * idr6
  return torch.nan_to_num(4.0 * epsilon * (idr12 - idr6))


class IsotropicCutoff:
  """S(r) f(r) with the C1 switch of multiplicative_isotropic_cutoff
  (energy.py:534-580): S = 1 below r_onset, 0 beyond r_cutoff,
  (rc^2 - r^2)^2 (rc^2 + 2 r^2 - 3 ro^2) / (rc^2 - ro^2)^3 between."""

  def __init__(self, fn: Callable, r_onset, r_cutoff):
    self.fn = fn
    self.r_onset = float(f32(r_onset))
    self.r_cutoff = float(f32(r_cutoff))
    self.__name__ = getattr(fn, '__name__', 'cutoff_fn')

  def smooth(self, dr):
    r_c = self.r_cutoff ** 2
    r_o = self.r_onset ** 2
    r = dr ** 2
    if r_c > r_o:
      inner = torch.where(dr < self.r_cutoff,
                          (r_c - r) ** 2 * (r_c + 2 * r - 3 * r_o) / (r_c - r_o) ** 3,
                          torch.zeros_like(dr))
    else:
      inner = torch.zeros_like(dr)
    return torch.where(dr < self.r_onset, torch.ones_like(dr), inner)

  def __call__(self, dr, *args, **kwargs):
    return self.smooth(dr) * self.fn(dr, *args, **kwargs)


def multiplicative_isotropic_cutoff(fn: Callable, r_onset, r_cutoff) -> IsotropicCutoff:
  return IsotropicCutoff(fn, r_onset, r_cutoff)


def soft_sphere_neighbor_list(displacement_or_metric, box_size, species=None, sigma=1.0,
                              epsilon=1.0, alpha=2.0, dr_threshold: float = 0.2,
                              per_particle: bool = False,
                              fractional_coordinates: bool = False,
                              format: partition.NeighborListFormat = partition.OrderedSparse,
                              neighbor_list_fn: Callable = partition.neighbor_list,
                              pair_neighbor_list_fn: Callable = smap.pair_neighbor_list,
                              **neighbor_kwargs):
  """Soft spheres with a neighbour list (energy.py:200-243); the list cutoff
  is max(sigma)."""
  sigma = maybe_downcast(sigma)
  epsilon = maybe_downcast(epsilon)
  alpha = maybe_downcast(alpha)
  list_cutoff = _max_f32(sigma)
  dr_threshold = float(f32(dr_threshold))
  neighbor_fn = neighbor_list_fn(displacement_or_metric, box_size, list_cutoff, dr_threshold,
                                 fractional_coordinates=fractional_coordinates, format=format,
                                 **neighbor_kwargs)
  energy_fn = pair_neighbor_list_fn(
      soft_sphere, space.canonicalize_displacement_or_metric(displacement_or_metric),
      ignore_unused_parameters=True, species=species, sigma=sigma, epsilon=epsilon,
      alpha=alpha, reduce_axis=(1,) if per_particle else None,
      fractional_coordinates=fractional_coordinates)
  return neighbor_fn, energy_fn


def lennard_jones_neighbor_list(displacement_or_metric, box_size, species=None, sigma=1.0,
                                epsilon=1.0, r_onset: float = 2.0, r_cutoff: float = 2.5,
                                dr_threshold: float = 0.5, per_particle: bool = False,
                                fractional_coordinates: bool = False,
                                format: partition.NeighborListFormat = partition.OrderedSparse,
                                neighbor_list_fn: Callable = partition.neighbor_list,
                                pair_neighbor_list_fn: Callable = smap.pair_neighbor_list,
                                **neighbor_kwargs):
  """Lennard-Jones with a neighbour list (energy.py:300-343); r_onset and
  r_cutoff are scaled by max(sigma) once, at construction."""
  sigma = maybe_downcast(sigma)
  epsilon = maybe_downcast(epsilon)
  smax = f32(_max_f32(sigma))
  r_onset = float(f32(f32(r_onset) * smax))
  r_cutoff = float(f32(f32(r_cutoff) * smax))
  dr_threshold = float(f32(dr_threshold))
  neighbor_fn = neighbor_list_fn(displacement_or_metric, box_size, r_cutoff, dr_threshold,
                                 fractional_coordinates=fractional_coordinates, format=format,
                                 **neighbor_kwargs)
  energy_fn = pair_neighbor_list_fn(
      multiplicative_isotropic_cutoff(lennard_jones, r_onset, r_cutoff),
      space.canonicalize_displacement_or_metric(displacement_or_metric),
      ignore_unused_parameters=True, species=species, sigma=sigma, epsilon=epsilon,
      reduce_axis=(1,) if per_particle else None,
      fractional_coordinates=fractional_coordinates)
  return neighbor_fn, energy_fn
