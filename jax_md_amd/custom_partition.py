"""Multi-image neighbour lists for small periodic cells
(jax_md/custom_partition.py:17-1122).

``neighbor_list_multi_image`` keeps every periodic image within the cutoff,
not only the nearest one, so it serves cells smaller than twice the cutoff.
Each edge stores its integer lattice shift (real-space shift ``s T^T``).  The
build runs in libjaxmd_native.so (``jmd_multi_image_build``); the shift
stencil, the capacities and the rebuild decision follow the reference.
"""

from __future__ import annotations

import dataclasses
import math
from typing import Any, Callable, Optional

import numpy as onp
import torch

from . import _native as nat
from .partition import Dense, NeighborListFormat, OrderedSparse, Sparse

f32 = onp.float32


@dataclasses.dataclass
class NeighborListMultiImage:
  """State of a multi-image list (custom_partition.py:17-122).  Sparse
  formats: ``idx = (receivers, senders)``, ``shifts`` [capacity, dim]; Dense:
  ``idx`` [N, max_neighbors], ``shifts`` [N, max_neighbors, dim].  Padding is
  index N with a zero shift."""
  idx: Any
  shifts: torch.Tensor
  reference_position: torch.Tensor
  reference_box: onp.ndarray
  box: onp.ndarray
  search_shifts: torch.Tensor
  format: NeighborListFormat
  max_occupancy: int
  update_fn: Callable[..., 'NeighborListMultiImage']
  did_buffer_overflow: torch.Tensor
  shift_max: Optional[onp.ndarray] = None  # host copy of the stencil's extent

  def update(self, position, **kwargs) -> 'NeighborListMultiImage':
    return self.update_fn(position, self, **kwargs)

  @property
  def senders(self):
    if self.format is Dense:
      raise ValueError('senders property not available for Dense format. Use idx directly.')
    return self.idx[1]

  @property
  def receivers(self):
    if self.format is Dense:
      raise ValueError('receivers property not available for Dense format. Use idx directly.')
    return self.idx[0]

  @property
  def n_edges(self) -> int:
    N = len(self.reference_position)
    if self.format is Dense:
      return int((self.idx < N).sum().item())
    return int((self.idx[0] < N).sum().item())

  @property
  def max_neighbors(self) -> int:
    if self.format is not Dense:
      raise ValueError('max_neighbors property only available for Dense format.')
    return self.idx.shape[1]

  @property
  def n_node(self) -> int:
    return len(self.reference_position)


@dataclasses.dataclass
class NeighborListMultiImageFns:
  allocate: Callable[..., NeighborListMultiImage]
  update: Callable[..., NeighborListMultiImage]

  def __iter__(self):
    return iter((self.allocate, self.update))


def _matrix(box) -> onp.ndarray:
  b = onp.asarray(box.detach().cpu().numpy() if isinstance(box, torch.Tensor) else box, f32)
  if b.ndim != 2 or b.shape[0] != b.shape[1]:
    raise ValueError(f'the multi-image list takes a [dim, dim] box matrix, got {b.shape}')
  return b


def _shift_n_max(box: onp.ndarray, cutoff: float, pbc) -> onp.ndarray:
  """Images per lattice direction: ceil(c / h_d) with the perpendicular
  heights h_d = 1 / |column d of inv(T)^T| (custom_partition.py:147-209)."""
  inv_t = onp.linalg.inv(box.astype(onp.float64)).T
  heights = 1.0 / onp.linalg.norm(inv_t, axis=0)
  n = onp.ceil(f32(cutoff) / heights.astype(f32)).astype(onp.int32)
  return onp.where(onp.asarray(pbc, bool), n, 0).astype(onp.int32)


def _search_shifts(n_max: onp.ndarray) -> onp.ndarray:
  grids = onp.meshgrid(*[onp.arange(-n, n + 1) for n in n_max], indexing='ij')
  return onp.stack([g.ravel() for g in grids], axis=-1).astype(onp.int32)


def estimate_max_neighbors(r_cutoff, atomic_density=0.1, safety_factor=2.0, dim=3) -> int:
  """safety * density * cutoff-sphere volume (custom_partition.py:627-686)."""
  if r_cutoff <= 0:
    return 0
  if dim > 3:
    raise ValueError(f'dim must be 1, 2, or 3, got {dim}')
  vol = {3: 4.0 / 3.0 * math.pi * r_cutoff ** 3, 2: math.pi * r_cutoff ** 2,
         1: 2.0 * r_cutoff}[dim]
  return max(int(math.ceil(safety_factor * atomic_density * vol)), 1)


def estimate_max_neighbors_from_box(box, r_cutoff, n_atoms, safety_factor=2.0,
                                    pbc=None) -> int:
  """safety * N / V * volume of the capped cutoff sphere
  (custom_partition.py:688-797)."""
  if r_cutoff <= 0:
    return 0
  b = onp.asarray(box, onp.float64)
  dim = b.shape[0]
  if dim > 3:
    raise ValueError(f'dim must be 1, 2, or 3, got {dim}')
  pbc = onp.ones(dim, bool) if pbc is None else onp.asarray(pbc, bool)
  density = n_atoms / abs(onp.linalg.det(b))
  lengths = onp.linalg.norm(b, axis=0)
  r_eff = float(onp.min(onp.where(pbc, r_cutoff, onp.minimum(r_cutoff, lengths))))
  vol = {3: 4.0 / 3.0 * math.pi * r_eff ** 3, 2: math.pi * r_eff ** 2, 1: 2.0 * r_eff}[dim]
  return max(int(math.ceil(safety_factor * density * vol)), 1)


def neighbor_list_multi_image(displacement_or_metric, box, r_cutoff, dr_threshold=0.0,
                              capacity_multiplier: float = 1.25, pbc=None,
                              fractional_coordinates: bool = True, ordered: bool = False,
                              format: NeighborListFormat = Sparse,
                              **kwargs) -> NeighborListMultiImageFns:
  """Multi-image lists (custom_partition.py:800-1104); the displacement
  argument is ignored as in the reference."""
  del displacement_or_metric, kwargs
  default_box = _matrix(box)
  dim = default_box.shape[0]
  use_dense = format is Dense
  use_ordered = ordered or format is OrderedSparse
  fmt = Dense if use_dense else (OrderedSparse if use_ordered else Sparse)
  code = nat.DENSE if use_dense else (nat.ORDERED_SPARSE if use_ordered else nat.SPARSE)
  pbc = onp.ones(dim, bool) if pbc is None else onp.asarray(pbc, bool)
  search_cutoff = float(r_cutoff) + float(dr_threshold)
  cutoff_sq = float(f32(search_cutoff ** 2))
  use_threshold = float(dr_threshold) > 0

  cache = {}

  def _workspace(lib, N, ns, dev):
    key = (N, ns, str(dev))
    if key not in cache:
      cache.clear()
      cache[key] = torch.zeros((int(lib.jmd_multi_image_workspace_bytes(N, ns)),),
                               dtype=torch.uint8, device=dev)
    return cache[key]

  def _build(position, cur_box, n_max, capacity, idx, shifts, overflow):
    lib = nat.load()
    N = position.shape[0]
    dev = position.device
    ns = int(onp.prod(2 * n_max.astype(onp.int64) + 1))
    ws = _workspace(lib, N, ns, dev)
    occ = torch.zeros((1,), dtype=torch.int64, device=dev)
    nat.check(lib.jmd_multi_image_build(
        nat.ptr(position), N, dim, int(bool(fractional_coordinates)),
        nat.f3(cur_box.reshape(-1)), (nat.c_int32 * 3)(*[int(v) for v in n_max] + [0] * (3 - dim)),
        cutoff_sq, code, nat.ptr(idx) if idx is not None else None,
        nat.ptr(shifts) if shifts is not None else None, int(capacity), nat.ptr(occ),
        nat.ptr(overflow) if overflow is not None else None, nat.ptr(ws),
        nat.stream_handle(dev)), 'jmd_multi_image_build')
    return occ

  def _make(position, cur_box, n_max, capacity):
    N = position.shape[0]
    dev = position.device
    if use_dense:
      idx = torch.empty((N, capacity), dtype=torch.int32, device=dev)
      shifts = torch.empty((N, capacity, dim), dtype=torch.int32, device=dev)
    else:
      idx = torch.empty((2, capacity), dtype=torch.int32, device=dev)
      shifts = torch.empty((capacity, dim), dtype=torch.int32, device=dev)
    overflow = torch.zeros((1,), dtype=torch.uint8, device=dev)
    _build(position, cur_box, n_max, capacity, idx, shifts, overflow)
    return NeighborListMultiImage(
        idx=idx if use_dense else (idx[0], idx[1]), shifts=shifts,
        reference_position=position.clone(), reference_box=cur_box.copy(), box=cur_box.copy(),
        search_shifts=torch.as_tensor(_search_shifts(n_max), device=dev), format=fmt,
        max_occupancy=capacity, update_fn=update_fn, did_buffer_overflow=overflow[0] != 0,
        shift_max=n_max.copy())

  def allocate_fn(position, extra_capacity: int = 0, box=None, **kw) -> NeighborListMultiImage:
    nat.require_device_tensor(position, 'position', torch.float32)
    position = position.contiguous()
    N = position.shape[0]
    cur_box = _matrix(box) if box is not None else default_box
    n_max = _shift_n_max(cur_box, search_cutoff, pbc)
    extra = extra_capacity if use_dense else N * extra_capacity
    actual = int(_build(position, cur_box, n_max, 0, None, None, None).item())
    capacity = max(int(actual * capacity_multiplier) + extra, 1)
    return _make(position, cur_box, n_max, capacity)

  def update_fn(position, neighbors: NeighborListMultiImage, **kw) -> NeighborListMultiImage:
    nat.require_device_tensor(position, 'position', torch.float32)
    position = position.contiguous()
    cur_box = _matrix(kw['box']) if 'box' in kw else neighbors.box
    need = _shift_n_max(cur_box, search_cutoff, pbc)
    have = (neighbors.shift_max if neighbors.shift_max is not None else
            neighbors.search_shifts.abs().amax(0).cpu().numpy().astype(onp.int32))
    range_overflow = bool(onp.any(need > have))
    if use_threshold and not range_overflow:
      lib = nat.load()
      out = torch.empty((1,), dtype=torch.float32, device=position.device)
      nat.check(lib.jmd_multi_image_max_displacement_sq(
          nat.ptr(position), nat.ptr(neighbors.reference_position), position.shape[0], dim,
          int(bool(fractional_coordinates)), nat.f3(cur_box.reshape(-1)),
          nat.f3(onp.asarray(neighbors.reference_box, f32).reshape(-1)), nat.ptr(out),
          nat.stream_handle(position.device)), 'jmd_multi_image_max_displacement_sq')
      max_atom = float(onp.sqrt(f32(out.item())))
      sh = _search_shifts(have).astype(f32)
      delta = sh @ (cur_box - onp.asarray(neighbors.reference_box, f32)).T
      max_shift = float(onp.sqrt(onp.max(onp.sum(delta * delta, axis=-1))))
      if not (2.0 * max_atom + max_shift >= float(dr_threshold)):
        return dataclasses.replace(neighbors, box=cur_box.copy())
    n_max = have
    nb = _make(position, cur_box, n_max, neighbors.max_occupancy)
    if range_overflow:
      nb.did_buffer_overflow = torch.ones((), dtype=torch.bool, device=position.device)
    return nb

  return NeighborListMultiImageFns(allocate_fn, update_fn)


def neighbor_list_multi_image_mask(neighbors: NeighborListMultiImage) -> torch.Tensor:
  """Valid-edge mask: receiver < N (custom_partition.py:1106-1122)."""
  return neighbors.idx[0] < len(neighbors.reference_position)
