"""Pair reduction over neighbour lists (jax_md/smap.py:856-979) on gfx950.

``pair_neighbor_list(fn, metric, species=..., reduce_axis=..., **params)``
accepts the pair functions of this package's ``energy`` module
(``lennard_jones`` wrapped by ``multiplicative_isotropic_cutoff`` and
``soft_sphere``) and returns an energy function with the reference's call
signature ``energy_fn(R, neighbor, **dynamic_kwargs)``.  Energy and force come
from one kernel pass; the energy is differentiable with respect to ``R``
through ``torch.autograd`` (the counterpart of a ``jax.custom_vjp``: the
backward pass returns ``-g * F`` from the same kernel).  Other pair functions
raise ``NotImplementedError``: the package implements this path only.
"""

from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import numpy as onp
import torch

from . import _native as nat
from . import partition
from . import space as space_mod
from . import util

_FMT = {partition.Dense: nat.DENSE, partition.Sparse: nat.SPARSE,
        partition.OrderedSparse: nat.ORDERED_SPARSE}


def _native_kind(fn):
  """(kind, r_onset, r_cutoff) for the pair functions the kernels implement."""
  from . import energy  # local import: energy imports smap
  if isinstance(fn, energy.IsotropicCutoff) and fn.fn is energy.lennard_jones:
    return nat.LENNARD_JONES, float(fn.r_onset), float(fn.r_cutoff)
  if fn is energy.soft_sphere:
    return nat.SOFT_SPHERE, 0.0, 0.0
  raise NotImplementedError(
      'smap.pair_neighbor_list runs natively for energy.lennard_jones (with '
      'multiplicative_isotropic_cutoff) and energy.soft_sphere only; got '
      f'{getattr(fn, "__name__", fn)}.')


class _Params:
  """Converts reference-style parameters to the C ABI struct and keeps the
  device arrays alive for the duration of a call."""

  def __init__(self, device, n):
    self.device = device
    self.n = n
    self.keep = []

  def param(self, value, species) -> nat.JmdParam:
    p = nat.JmdParam()
    if isinstance(value, torch.Tensor) and value.ndim > 0:
      t = value.to(device=self.device, dtype=torch.float32).contiguous()
    elif isinstance(value, onp.ndarray) and value.ndim > 0:
      t = torch.as_tensor(value, dtype=torch.float32, device=self.device).contiguous()
    else:
      p.mode = nat.PARAM_SCALAR
      p.value = util.as_f32_scalar(value)
      p.values = None
      return p
    if t.ndim == 1 and (species is None or t.shape[0] == self.n):
      if t.shape[0] != self.n:
        raise ValueError(f'per-particle parameter has {t.shape[0]} entries for N={self.n}')
      p.mode = nat.PARAM_PER_PARTICLE  # 0.5 (x + y) combinator (smap.py:836)
    elif t.ndim == 2 and species is not None:
      p.mode = nat.PARAM_SPECIES
    else:
      raise NotImplementedError('parameter arrays must be per-particle [N] vectors or '
                                'species matrices; [N, N] matrices are not supported')
    self.keep.append(t)
    p.values = ctypes.c_void_p(t.data_ptr())
    p.value = 0.0
    return p


class PairEnergy:
  """Energy function returned by ``pair_neighbor_list``."""

  _jmd_kind = 'energy'

  def __init__(self, kind, r_onset, r_cutoff, geometry, species, reduce_axis,
               ignore_unused_parameters, params):
    self.kind = kind
    self.r_onset = r_onset
    self.r_cutoff = r_cutoff
    self.geometry = geometry
    self.species = species
    self.reduce_axis = reduce_axis
    self.ignore_unused_parameters = ignore_unused_parameters
    self.params = params
    self._ws = {}

  # -- helpers ---------------------------------------------------------------
  def _sigma_bound(self, sig) -> float:
    if not (isinstance(sig, (torch.Tensor, onp.ndarray)) and sig.ndim > 0):
      return 0.0
    if sig is self.params.get('sigma'):
      if getattr(self, '_sig_max', None) is None:
        v = sig.detach().max().cpu().item() if isinstance(sig, torch.Tensor) else onp.max(sig)
        self._sig_max = float(onp.nextafter(onp.float32(v), onp.float32(onp.inf)))
      return self._sig_max
    return 0.0


  def reads_particle_data(self) -> bool:
    """True when the potential reads arrays indexed by particle id (species
    ids or per-particle [N] parameters): such arrays follow the caller's
    particle order, so a loop may not renumber the particles."""
    if self.species is not None:
      return True
    for k in ('sigma', 'epsilon', 'alpha'):
      v = self.params.get(k)
      if v is not None and hasattr(v, 'ndim') and v.ndim == 1:
        return True
    return False

  def _workspace(self, device, n):
    key = (str(device), int(n))
    ws = self._ws.get(key)
    if ws is None:
      lib = nat.load()
      ws = torch.empty((lib.jmd_pair_workspace_bytes(n),), dtype=torch.uint8, device=device)
      self._ws = {key: ws}
    return ws

  def _potential(self, R, dynamic_kwargs):
    merged = util.merge_dicts(self.params, dynamic_kwargs, self.ignore_unused_parameters)
    species = dynamic_kwargs.get('species', self.species)
    N = R.shape[0]
    P = _Params(R.device, N)
    pot = nat.JmdPairPotential()
    pot.kind = self.kind
    sp = None
    if species is not None:
      sp = torch.as_tensor(species, device=R.device).to(torch.int32).contiguous()
      P.keep.append(sp)
    pot.sigma = P.param(merged.get('sigma', 1.0), sp)
    pot.epsilon = P.param(merged.get('epsilon', 1.0), sp)
    pot.alpha = P.param(merged.get('alpha', 2.0), sp)
    pot.r_onset = self.r_onset
    pot.r_cutoff = self.r_cutoff
    if self.kind == nat.SOFT_SPHERE:
      # bound on sigma for the wrap-free interior path: known when sigma is
      # the construction-time value (its max is taken once, on the host)
      sig = merged.get('sigma', 1.0)
      pot.r_cutoff = self._sigma_bound(sig)
    ns = 0
    for k in ('sigma', 'epsilon', 'alpha'):
      v = merged.get(k)
      if v is not None and hasattr(v, 'ndim') and v.ndim == 2:
        ns = max(ns, int(v.shape[0]))
    pot.n_species = ns
    pot.species = ctypes.c_void_p(sp.data_ptr()) if sp is not None else None
    return pot, P

  def _run64(self, R, neighbor, want_force, want_energy_total, dynamic_kwargs,
             momentum=None, half_dt=0.0):
    """The float64 path (the reference's x64 run): R, the force and the
    momentum are float64; Sparse-format lists run as the Dense rows of the
    same pairs (the energy counts each pair once either way)."""
    nat.require_device_tensor(R, 'R', torch.float64)
    if int(self.geometry.periodic) == 2:
      raise NotImplementedError('float64 positions with a periodic_general box are not '
                                'supported natively')
    if momentum is not None:
      nat.require_device_tensor(momentum, 'momentum', torch.float64)
    lib = nat.load()
    N, dim = R.shape
    geo = self.geometry
    pot, keep = self._potential(R, dynamic_kwargs)
    dev = R.device
    F = torch.empty((N, dim), dtype=torch.float64, device=dev) if want_force else None
    e_tot = torch.empty((1,), dtype=torch.float64, device=dev) if want_energy_total else None
    ws = self._workspace64(dev, N) if want_energy_total else None
    idx = self._dense_rows_of(neighbor, N)
    nat.check(lib.jmd_pair_energy_force_f64(
        nat.ptr(R), N, dim, int(geo.periodic), nat.d3(self.call_side(dim, dynamic_kwargs)),
        ctypes.byref(pot),
        nat.ptr(idx), idx.shape[1], nat.DENSE, nat.ptr(F), nat.ptr(e_tot), nat.ptr(momentum),
        float(half_dt), nat.ptr(ws), nat.stream_handle(dev)), 'jmd_pair_energy_force_f64')
    del keep
    return e_tot, None, F

  def _dense_rows_of(self, neighbor, N):
    """The Dense rows the float64 kernel reads: a Dense list as it is; a
    Sparse-format list's pairs as rows (both orders, ids ascending), kept
    until the list is rebuilt (its build epoch) or rewritten (its version)."""
    if neighbor.format is partition.Dense:
      return neighbor.idx
    shadow = partition.current_shadow(neighbor)
    if shadow is not None:
      return shadow.idx
    idx = neighbor.idx
    key = (idx.data_ptr(), idx._version, tuple(idx.shape), id(neighbor.plan),
           neighbor.plan.epoch if neighbor.plan is not None else None, N)
    hit = self._ws.get('rows64')
    if hit is not None and hit[0] == key:
      return hit[1]
    rows = partition.dense_rows(idx, N, neighbor.format)
    self._ws['rows64'] = (key, rows)
    return rows

  def _workspace64(self, device, n):
    key = ('f64', str(device), int(n))
    ws = self._ws.get(key)
    if ws is None:
      lib = nat.load()
      ws = torch.empty((lib.jmd_pair_workspace_bytes_f64(n),), dtype=torch.uint8, device=device)
      self._ws[key] = ws
    return ws

  def call_side(self, dim: int, dynamic_kwargs) -> list:
    """The geometry's side values for one call.  ``box=`` / ``new_box=``
    (``new_box`` wins, as in the reference) replace the box matrix of a
    ``periodic_general`` space with fractional coordinates for this call
    (space.py:421-443: displacements are the new box applied to the
    fractional difference); ``space.periodic`` refuses them as the
    reference's displacement does, free space ignores them."""
    geo = self.geometry
    if 'new_box' in dynamic_kwargs:
      box = dynamic_kwargs['new_box']
    elif 'box' in dynamic_kwargs:
      box = dynamic_kwargs['box']
    else:
      return geo.side_values(dim)
    if int(geo.periodic) == 2:
      return [float(v) for v in space_mod._box_as_matrix(box, dim).reshape(-1)]
    if geo.periodic:
      raise ValueError('`space.periodic` does not accept a box argument. Perhaps you '
                       'meant to use `space.periodic_general`?')
    return geo.side_values(dim)

  def geometry_code(self, in_box: bool = False) -> int:
    """The C ABI's ``periodic`` argument; ``in_box`` asserts that every
    position lies in the box (written by a wrapping shift)."""
    code = int(self.geometry.periodic)
    return code | nat.POSITIONS_IN_BOX if (in_box and code) else code

  def _run(self, R, neighbor, want_force, want_energy_total, want_energy_atom,
           dynamic_kwargs, momentum=None, half_dt=0.0, dt_device=None, in_box=False):
    if not isinstance(neighbor, partition.NeighborList):
      raise TypeError('energy_fn needs a partition.NeighborList as `neighbor`')
    if isinstance(R, torch.Tensor) and R.dtype == torch.float64:
      if want_energy_atom or dt_device is not None:
        raise NotImplementedError('per-particle energies and device dt are float32-only')
      return self._run64(R, neighbor, want_force, want_energy_total, dynamic_kwargs,
                         momentum, half_dt)
    stride = nat.position_stride(R, 'R')
    lib = nat.load()
    N, dim = R.shape
    geo = self.geometry
    side = self.call_side(dim, dynamic_kwargs)
    pot, keep = self._potential(R, dynamic_kwargs)
    dev = R.device
    F = torch.empty((N, dim), dtype=torch.float32, device=dev) if want_force else None
    e_tot = torch.empty((1,), dtype=torch.float64, device=dev) if want_energy_total else None
    e_atom = torch.empty((N,), dtype=torch.float32, device=dev) if want_energy_atom else None
    idx = neighbor.idx
    fmt = _FMT[neighbor.format]
    # a Sparse-format list from partition's allocate / update carries the
    # Dense rows of its pairs: the row kernel sums each particle's force in
    # a fixed order (no float atomics); energies and per-particle energies
    # count the same pairs
    shadow = partition.current_shadow(neighbor)
    if shadow is not None:
      idx, fmt = shadow.idx, nat.DENSE
    width = idx.shape[1]
    ws = self._workspace(dev, N)
    nat.check(lib.jmd_pair_energy_force(
        nat.ptr(R), N, dim, stride, self.geometry_code(in_box), nat.f3(side), ctypes.byref(pot),
        nat.ptr(idx),
        width, fmt, nat.ptr(F), nat.ptr(e_atom), nat.ptr(e_tot),
        nat.ptr(momentum), float(half_dt), nat.ptr(dt_device), nat.ptr(ws),
        nat.stream_handle(dev)), 'jmd_pair_energy_force')
    del keep
    return e_tot, e_atom, F

  # -- public ----------------------------------------------------------------
  def __call__(self, R, neighbor=None, **dynamic_kwargs):
    if 'fractional_coordinates' in dynamic_kwargs:
      # an unused pair-function argument in the reference as well; the
      # geometry comes from the displacement
      dynamic_kwargs = dict(dynamic_kwargs)
      dynamic_kwargs.pop('fractional_coordinates')
    if self.reduce_axis is not None and len(self.reduce_axis) == 0:
      return self.edge_energies(R, neighbor, **dynamic_kwargs)
    per_particle = self.reduce_axis is not None
    if per_particle and tuple(self.reduce_axis) != (1,):
      raise NotImplementedError(f'reduce_axis={self.reduce_axis} is not supported natively')
    if per_particle:
      if neighbor.format is partition.OrderedSparse:
        raise ValueError('Cannot report per-particle values with a neighbor list whose '
                         'format is `OrderedSparse`. Please use either `Dense` or `Sparse`.')
      _, e_atom, _ = self._run(R, neighbor, False, False, True, dynamic_kwargs)
      return e_atom
    if R.requires_grad and torch.is_grad_enabled():
      return _EnergyWithForce.apply(R, self, neighbor, dynamic_kwargs)
    e_tot, _, _ = self._run(R, neighbor, False, True, False, dynamic_kwargs)
    return e_tot.reshape(()).to(R.dtype)

  def energy_and_force(self, R, neighbor=None, **dynamic_kwargs):
    """(E float64 0-d, F [N, dim] of R's dtype) from one kernel pass."""
    e_tot, _, F = self._run(R, neighbor, True, True, False, dynamic_kwargs)
    return e_tot.reshape(()), F

  def force(self, R, neighbor=None, in_box=False, **dynamic_kwargs):
    """-dE/dR from the native kernel (quantity.force routes here).
    ``in_box``: the caller guarantees every position lies in the box."""
    _, _, F = self._run(R, neighbor, True, False, False, dynamic_kwargs, in_box=in_box)
    return F

  def edge_energies(self, R, neighbor=None, **dynamic_kwargs):
    """The energy of every list entry over the reference's normalization
    (reduce_axis=(), smap.py:940-953): [N, max_occupancy] for Dense lists,
    [capacity] for the Sparse formats, float32, 0 at padding entries."""
    if not isinstance(neighbor, partition.NeighborList):
      raise TypeError('energy_fn needs a partition.NeighborList as `neighbor`')
    stride = nat.position_stride(R, 'R')
    lib = nat.load()
    N, dim = R.shape
    pot, keep = self._potential(R, dynamic_kwargs)
    idx = neighbor.idx
    dense = not partition.is_sparse(neighbor.format)
    width = idx.shape[1]
    out = torch.empty((N, width) if dense else (width,), dtype=torch.float32, device=R.device)
    nat.check(lib.jmd_pair_edge_energies(
        nat.ptr(R), N, dim, stride, int(self.geometry.periodic),
        nat.f3(self.call_side(dim, dynamic_kwargs)), ctypes.byref(pot), nat.ptr(idx), width,
        _FMT[neighbor.format], nat.ptr(out), nat.stream_handle(R.device)),
        'jmd_pair_edge_energies')
    del keep
    return out

  def virial(self, R, neighbor=None, **dynamic_kwargs):
    """dU/d eps [dim, dim] float64 for every displacement perturbed as
    (I + eps) d: weight * sum over the list's pairs of (U'(r)/r) d d^T."""
    if not isinstance(neighbor, partition.NeighborList):
      raise TypeError('virial needs a partition.NeighborList as `neighbor`')
    stride = nat.position_stride(R, 'R')
    lib = nat.load()
    N, dim = R.shape
    geo = self.geometry
    pot, keep = self._potential(R, dynamic_kwargs)
    W = torch.empty((dim, dim), dtype=torch.float64, device=R.device)
    ws = self._ws.get(('virial', str(R.device)))
    if ws is None:
      ws = torch.empty((lib.jmd_pair_virial_workspace_bytes(),), dtype=torch.uint8,
                       device=R.device)
      self._ws[('virial', str(R.device))] = ws
    idx = neighbor.idx
    width = idx.shape[1]  # Dense: row width; Sparse: capacity of the [2, capacity] list
    nat.check(lib.jmd_pair_virial(
        nat.ptr(R), N, dim, stride, int(geo.periodic), nat.f3(self.call_side(dim, dynamic_kwargs)),
        ctypes.byref(pot), nat.ptr(idx), width, _FMT[neighbor.format], nat.ptr(W), nat.ptr(ws),
        nat.stream_handle(R.device)), 'jmd_pair_virial')
    del keep
    return W

  def force_and_kick(self, R, neighbor, momentum, half_dt=0.0, dt_device=None, in_box=False,
                     **dynamic_kwargs):
    """Force, plus the closing half kick p += half_dt F fused in the same pass."""
    _, _, F = self._run(R, neighbor, True, False, False, dynamic_kwargs, momentum=momentum,
                        half_dt=half_dt, dt_device=dt_device, in_box=in_box)
    return F


class _EnergyWithForce(torch.autograd.Function):
  """E(R) whose vector-Jacobian product is -g F, both from one kernel pass."""

  @staticmethod
  def forward(ctx, R, fn, neighbor, dynamic_kwargs):
    e_tot, _, F = fn._run(R.detach(), neighbor, True, True, False, dynamic_kwargs)
    ctx.save_for_backward(F)
    return e_tot.reshape(()).to(torch.float32)

  @staticmethod
  def backward(ctx, g):
    (F,) = ctx.saved_tensors
    return -g * F, None, None, None


def pair_neighbor_list(fn, displacement_or_metric, species=None,
                       reduce_axis: Optional[Tuple[int, ...]] = None,
                       ignore_unused_parameters: bool = False, **kwargs) -> PairEnergy:
  """Promotes a pair function to a neighbour-list energy
  (jax_md/smap.py:856-979)."""
  kind, r_on, r_c = _native_kind(fn)
  kwargs = dict(kwargs)
  # the reference passes it on to the pair function, which ignores it
  kwargs.pop('fractional_coordinates', None)
  geo = space_mod.geometry_of(displacement_or_metric)
  return PairEnergy(kind, r_on, r_c, geo, species, reduce_axis, ignore_unused_parameters,
                    kwargs)
