"""Cell lists and neighbour lists (jax_md/partition.py) on gfx950.

Same surface as the reference: ``neighbor_list(...)`` returns
``NeighborListFns(allocate, update)``; a ``NeighborList`` exposes ``idx``,
``reference_position``, ``error``, ``did_buffer_overflow``, ``max_occupancy``,
``format`` and ``update``.  The host logic that the reference runs outside
jit (capacity estimates, ``int(occupancy * multiplier + extra)``) is restated
here in float32 NumPy; every per-particle computation runs in
libjaxmd_native.so.

Storage semantics.  ``update`` returns a new ``NeighborList`` object, but the
device buffers (``idx``, ``reference_position``, the error word) are updated
in place: the rebuild decision is taken on the device (the reference's
``lax.cond``) so no host round trip is needed, and an unchanged list costs
nothing.  The reference's own loops never read a list after updating it, and
re-allocate from positions on overflow, which this supports unchanged.
"""

from __future__ import annotations

import dataclasses
from enum import Enum, IntEnum
from functools import reduce
from operator import mul
from typing import NamedTuple, Any, Callable, Dict, Optional

import numpy as onp
import torch

from . import _native as nat
from . import space as space_mod

f32 = onp.float32


# ---------------------------------------------------------------------------
# Formats and error codes (partition.py:494-681)
# ---------------------------------------------------------------------------


class NeighborListFormat(Enum):
  """Dense [N, max_occupancy]; Sparse [2, capacity] = (receiver, sender);
  OrderedSparse like Sparse with only j < i (partition.py:641-661)."""
  Dense = 0
  Sparse = 1
  OrderedSparse = 2


Dense = NeighborListFormat.Dense
Sparse = NeighborListFormat.Sparse
OrderedSparse = NeighborListFormat.OrderedSparse


def is_sparse(fmt: NeighborListFormat) -> bool:
  return fmt is NeighborListFormat.Sparse or fmt is NeighborListFormat.OrderedSparse


def is_format_valid(fmt: NeighborListFormat):
  if fmt not in list(NeighborListFormat):
    raise ValueError(
        f'Neighbor list format must be a member of NeighborListFormat found {fmt}.')


class PartitionErrorCode(IntEnum):
  NONE = 0
  NEIGHBOR_LIST_OVERFLOW = 1 << 0
  CELL_LIST_OVERFLOW = 1 << 1
  CELL_SIZE_TOO_SMALL = 1 << 2
  MALFORMED_BOX = 1 << 3


PEC = PartitionErrorCode


@dataclasses.dataclass
class PartitionError:
  """Error word (device uint8 0-d tensor), sticky across updates."""
  code: torch.Tensor

  def update(self, bit, pred) -> 'PartitionError':
    pred = torch.as_tensor(pred, device=self.code.device)
    return PartitionError(self.code | torch.where(pred, int(bit), 0).to(torch.uint8))

  def __str__(self) -> str:
    c = int(self.code)
    if not c:
      return ''
    if c & PEC.NEIGHBOR_LIST_OVERFLOW:
      return 'Partition Error: Neighbor list buffer overflow.'
    if c & PEC.CELL_LIST_OVERFLOW:
      return 'Partition Error: Cell list buffer overflow'
    if c & PEC.CELL_SIZE_TOO_SMALL:
      return 'Partition Error: Cell size too small'
    if c & PEC.MALFORMED_BOX:
      return 'Partition Error: Incorrect box format. Expecting upper triangular.'
    raise ValueError(f'Unexpected Error Code {c}.')

  __repr__ = __str__


# ---------------------------------------------------------------------------
# Cell grid arithmetic (partition.py:146-188), float32 as in the reference
# ---------------------------------------------------------------------------


def _box_f32(box) -> onp.ndarray:
  if isinstance(box, torch.Tensor):
    box = box.detach().cpu().numpy()
  b = onp.asarray(box, f32)
  if b.ndim > 1:
    raise NotImplementedError(
        'Box matrices (periodic_general / fractional coordinates) are outside '
        'this package.')
  return b


def _fractional_cell_size(box, cutoff) -> f32:
  """Cell size on the unit cube for a box (partition.py:595-633), float32:
  the cutoff over the smallest perpendicular width of the cell."""
  if isinstance(box, torch.Tensor):
    box = box.detach().cpu().numpy()
  b = onp.asarray(box, f32)
  c = f32(cutoff)
  if b.ndim == 0:
    return f32(c / b)
  if b.ndim == 1:
    return f32(c / onp.min(b))
  one = f32(1)
  if b.shape[0] == 1:
    return f32(one / onp.floor(f32(b[0, 0] / c)))
  if b.shape[0] == 2:
    xx, yy = b[0, 0], b[1, 1]
    xy = f32(b[0, 1] / yy)
    nx = f32(xx / onp.sqrt(f32(one + xy * xy)))
    ny = yy
  elif b.shape[0] == 3:
    xx, yy, zz = b[0, 0], b[1, 1], b[2, 2]
    xy = f32(b[0, 1] / yy)
    xz = f32(b[0, 2] / zz)
    yz = f32(b[1, 2] / zz)
    t = f32(xy * yz - xz)
    nx = f32(xx / onp.sqrt(f32(f32(one + f32(xy * xy)) + f32(t * t))))
    ny = f32(yy / onp.sqrt(f32(one + f32(yz * yz))))
    nz = zz
  else:
    raise ValueError(f'Expected box to be either 1-, 2-, or 3-dimensional found {b.shape[0]}')
  widths = [nx, ny] if b.shape[0] == 2 else [nx, ny, nz]
  nmin = onp.floor(f32(f32(min(widths)) / c))
  nmin = f32(1) if nmin == 0 else f32(nmin)
  return f32(one / nmin)


def is_box_valid(box) -> bool:
  """Scalar / vector boxes, or upper-triangular matrices (partition.py:676-681)."""
  if isinstance(box, torch.Tensor):
    box = box.detach().cpu().numpy()
  b = onp.asarray(box)
  if b.ndim <= 1:
    return True
  if b.ndim == 2:
    return bool(onp.all(onp.triu(b) == b))
  return False


def cell_dimensions(dim: int, box, minimum_cell_size):
  """(box, cell_size, cells_per_side, cell_count) (partition.py:146-188)."""
  box = _box_f32(box)
  mcs = onp.asarray(minimum_cell_size, f32)
  cps = onp.floor(box / mcs).astype(f32)
  cell_size = (box / cps).astype(f32)
  cps_i = cps.astype(onp.int32)
  if box.ndim == 1:
    if box.shape[0] != dim:
      raise ValueError(f'box has {box.shape[0]} sides for dimension {dim}')
    if onp.any(cps_i < 3):
      raise ValueError(
          'Box must be at least 3x the size of the grid spacing in each dimension.')
    cell_count = reduce(mul, [int(c) for c in cps_i], 1)
  else:
    cell_count = int(cps_i) ** dim
  return box, cell_size, cps_i, cell_count


def _per_dim(x, dim):
  x = onp.asarray(x)
  return [x.item()] * dim if x.ndim == 0 else [v.item() for v in x.reshape(dim)]


class _Workspace:
  """Zero-initialised device scratch for the native list builds.  The native
  calls leave it zeroed at rest (their contract), so it is allocated once."""

  def __init__(self, device, nbytes):
    self.buf = torch.zeros(max(int(nbytes), 256), dtype=torch.uint8, device=device)

  @property
  def ptr(self):
    return nat.ptr(self.buf)


def _check_position(position) -> int:
  """Validates a position array and returns its row stride."""
  return nat.position_stride(position)


def _filling_max(position, cps, cell_size):
  """Largest cell occupancy on a grid (count_cell_filling, partition.py:191-208)."""
  lib = nat.load()
  N, dim = position.shape
  ncell = reduce(mul, [int(c) for c in cps], 1)
  ws = _Workspace(position.device, lib.jmd_neighbor_workspace_bytes(N, dim, ncell, 0, 0))
  out = torch.zeros((1,), dtype=torch.int32, device=position.device)
  nat.check(lib.jmd_cell_filling_max(nat.ptr(position), N, dim, _check_position(position),
                                     nat.i3(cps), nat.f3(cell_size),
                                     nat.ptr(out), ws.ptr, nat.stream_handle(position.device)),
            'jmd_cell_filling_max')
  return int(out.item())


# ---------------------------------------------------------------------------
# Cell list (partition.py:78-488)
# ---------------------------------------------------------------------------


@dataclasses.dataclass
class CellList:
  position_buffer: torch.Tensor
  id_buffer: torch.Tensor
  particle_cell_id: torch.Tensor
  named_buffer: Dict[str, torch.Tensor]
  did_buffer_overflow: torch.Tensor
  cell_capacity: int
  cell_size: Any
  update_fn: Callable[..., 'CellList']

  def update(self, position, **kwargs) -> 'CellList':
    return self.update_fn(position, self.cell_capacity, **kwargs)

  @property
  def kwarg_buffers(self):
    return self.named_buffer


@dataclasses.dataclass
class CellListFns:
  allocate: Callable[..., CellList]
  update: Callable[..., CellList]

  def __iter__(self):
    return iter((self.allocate, self.update))


def cell_list(box_size, minimum_cell_size, buffer_size_multiplier: float = 1.25) -> CellListFns:
  """Uniform-grid spatial partition (partition.py:296-488).

  Buffers have the reference's shapes: ``position_buffer``
  [..cells.., capacity, dim] (cells in reversed-dimension order),
  ``id_buffer`` [..cells.., capacity, 1] with N in empty slots,
  ``particle_cell_id`` [N, dim].  Side data (``**kwargs`` buffers) is not
  supported natively.
  """
  box_size = _box_f32(box_size)
  mcs = onp.asarray(minimum_cell_size, f32)

  def cell_list_fn(position, capacity=None, overflow=None, extra_capacity=0, **kwargs):
    if kwargs:
      raise NotImplementedError('cell-list side data (named buffers) is not supported natively')
    stride = _check_position(position)
    lib = nat.load()
    N, dim = position.shape
    _, cell_size, cps, ncell = cell_dimensions(dim, box_size, mcs)
    cs3, cps3 = _per_dim(cell_size, dim), _per_dim(cps, dim)
    if capacity is None:
      # _estimate_cell_capacity counts on the grid derived again from cell_size
      _, cs2, cps2, _ = cell_dimensions(dim, box_size, cell_size)
      fill = _filling_max(position, _per_dim(cps2, dim), _per_dim(cs2, dim))
      capacity = int(onp.int32(fill) * buffer_size_multiplier) + extra_capacity
    capacity = int(capacity)
    dev = position.device
    pos_buf = torch.empty((ncell * capacity, dim), dtype=torch.float32, device=dev)
    id_buf = torch.empty((ncell * capacity,), dtype=torch.int32, device=dev)
    pcid = torch.empty((N, dim), dtype=torch.int32, device=dev)
    ovf = torch.zeros((), dtype=torch.uint8, device=dev)
    ws = _Workspace(dev, lib.jmd_neighbor_workspace_bytes(N, dim, ncell, 0, 0))
    nat.check(lib.jmd_cell_list(nat.ptr(position), N, dim, stride, nat.i3(cps3), nat.f3(cs3),
                                capacity,
                                nat.ptr(pos_buf), nat.ptr(id_buf), nat.ptr(pcid), None,
                                nat.ptr(ovf), ws.ptr, nat.stream_handle(dev)), 'jmd_cell_list')
    cells = tuple(int(c) for c in reversed(cps3))
    pos_buf = pos_buf.reshape(cells + (capacity, dim))
    id_buf = id_buf.reshape(cells + (capacity, 1))
    did = ovf != 0
    if overflow is not None:
      did = did | torch.as_tensor(overflow, device=dev)
    return CellList(pos_buf, id_buf, pcid, {}, did, capacity, cell_size, update_fn)

  def allocate_fn(position, extra_capacity: int = 0, **kwargs) -> CellList:
    return cell_list_fn(position, extra_capacity=extra_capacity, **kwargs)

  def update_fn(position, cl_or_capacity, **kwargs) -> CellList:
    if isinstance(cl_or_capacity, int):
      return cell_list_fn(position, capacity=cl_or_capacity, **kwargs)
    cl = cl_or_capacity
    return cell_list_fn(position, capacity=cl.cell_capacity, overflow=cl.did_buffer_overflow,
                        **kwargs)

  return CellListFns(allocate_fn, update_fn)


# ---------------------------------------------------------------------------
# Neighbour list (partition.py:684-1164)
# ---------------------------------------------------------------------------


@dataclasses.dataclass
class _Plan:
  """Device-side state shared by a list and its updates."""
  workspace: _Workspace
  gate: torch.Tensor          # int32 [1]
  occupancy: torch.Tensor     # int32 [1]
  use_cell_list: bool
  cells_per_side: list
  cell_size3: list
  side: list
  periodic: int               # 0 free, 1 periodic box, 2 box matrix (fractional)
  dim: int
  spec: Any = None
  # bumped by every build this module enqueues (allocate / update), which
  # rewrite the list's buffers in place without touching their tensors'
  # version counters
  epoch: int = 0
  # (unmasked list, mask function) of a list with a custom mask
  custom_mask: Any = None
  # Sparse formats: the Dense rows of the same pairs, rebuilt by every
  # allocate / update beside the Sparse list (see _shadow_build), and the
  # (epoch, idx version) at which it matched the Sparse list
  shadow: Any = None
  shadow_mark: Any = None

  @property
  def builds(self) -> int:
    """Completed list builds since allocation (device counter; syncs)."""
    return int(self.workspace.buf[16:20].view(torch.int32).item())


@dataclasses.dataclass
class NeighborList:
  """State of a neighbour list (partition.py:684-737)."""
  idx: torch.Tensor
  reference_position: torch.Tensor
  error: PartitionError
  cell_list_capacity: Optional[int]
  max_occupancy: int
  format: NeighborListFormat
  cell_size: Any
  cell_list_fn: Any
  update_fn: Callable[..., 'NeighborList']
  plan: Optional[_Plan] = None

  def update(self, position, **kwargs) -> 'NeighborList':
    return self.update_fn(position, self, **kwargs)

  @property
  def did_buffer_overflow(self) -> torch.Tensor:
    return self.error.code & (PEC.NEIGHBOR_LIST_OVERFLOW | PEC.CELL_LIST_OVERFLOW)

  @property
  def cell_size_too_small(self) -> torch.Tensor:
    return self.error.code & PEC.CELL_SIZE_TOO_SMALL

  @property
  def malformed_box(self) -> torch.Tensor:
    return self.error.code & PEC.MALFORMED_BOX


@dataclasses.dataclass
class NeighborListFns:
  allocate: Callable[..., NeighborList]
  update: Callable[..., NeighborList]

  def __call__(self, position, neighbors=None, extra_capacity: int = 0, **kwargs):
    if neighbors is None:
      return self.allocate(position, extra_capacity, **kwargs)
    return self.update(position, neighbors, **kwargs)

  def __iter__(self):
    return iter((self.allocate, self.update))


class NeighborListSpec:
  """Static configuration of ``neighbor_list`` with the reference's float32
  derived constants (partition.py:887-895)."""

  def __init__(self, displacement_or_metric, box, r_cutoff, dr_threshold,
               capacity_multiplier, disable_cell_list, mask_self, fmt,
               fractional_coordinates=False, cell_size_factor=1.0):
    self.geometry = space_mod.geometry_of(displacement_or_metric)
    self.fractional = bool(fractional_coordinates)
    rc = f32(onp.asarray(r_cutoff, f32))
    dr = f32(onp.asarray(dr_threshold, f32))
    self.cutoff = f32(rc + dr)
    # cells at least cell_size_factor x the list cutoff (1: the reference's
    # grid and candidate order; larger: same neighbour sets, rows in that
    # grid's candidate order)
    self.cell_size_factor = float(cell_size_factor)
    if not self.cell_size_factor >= 1.0:
      raise ValueError(f'cell_size_factor must be >= 1, got {cell_size_factor}')
    min_size = self.cutoff if self.cell_size_factor == 1.0 else f32(
        self.cutoff * f32(self.cell_size_factor))
    self._min_size = min_size
    if self.fractional:
      # cells on the unit cube, sized from the box (partition.py:1045-1050)
      self.user_box = box
      self.box = onp.asarray(f32(1.0))
      self.min_cell_size = _fractional_cell_size(box, min_size)
    else:
      self.box = _box_f32(box)
      self.min_cell_size = min_size
    self.cutoff_sq = f32(self.cutoff * self.cutoff)
    self.threshold_sq = f32((dr / f32(2)) * (dr / f32(2)))
    self.always_rebuild = bool(dr == 0)
    self.capacity_multiplier = capacity_multiplier
    self.disable_cell_list = disable_cell_list
    self.mask_self = mask_self
    self.format = fmt

  def use_cell_list(self, min_cell_size=None) -> bool:
    if self.disable_cell_list:
      return False
    mcs = self.min_cell_size if min_cell_size is None else min_cell_size
    return bool(onp.all(mcs < (self.box / f32(3.0)).astype(f32)))

  def fractional_cell_size(self, box) -> f32:
    """Cell size on the unit cube for a box given at call time."""
    return _fractional_cell_size(box, self._min_size)


_FORMAT_CODE = {Dense: nat.DENSE, Sparse: nat.SPARSE, OrderedSparse: nat.ORDERED_SPARSE}


def neighbor_list(displacement_or_metric, box, r_cutoff, dr_threshold=0.0,
                  capacity_multiplier: float = 1.25, disable_cell_list: bool = False,
                  mask_self: bool = True, custom_mask_function=None,
                  fractional_coordinates: bool = False,
                  format: NeighborListFormat = NeighborListFormat.Dense,
                  cell_size_factor: float = 1.0,
                  **static_kwargs) -> NeighborListFns:
  """Neighbour lists for a periodic or free space (partition.py:801-1164).

  Supported: ``space.periodic`` (scalar or per-dimension side),
  ``space.periodic_general`` with ``fractional_coordinates=True`` (cells on
  the unit cube, distances through the box matrix) and ``space.free``
  displacements, Dense / Sparse / OrderedSparse, with or without the cell
  list, ``mask_self``.  ``custom_mask_function`` (Dense lists) takes the
  rows [N, W] (padding N) and returns them with masked entries set to N, as
  the reference's does with its candidate array (partition.py:1079-1080);
  the masked rows are compacted in order and sized with the reference's
  capacity rule.  ``cell_size_factor`` (no counterpart in the
  reference, where cells are the list cutoff wide, partition.py:146-188)
  makes cells at least that multiple of the cutoff: the same neighbour sets,
  rows in that grid's candidate order; ``autotune`` chooses it.
  """
  is_format_valid(format)
  if custom_mask_function is not None:
    if format is not Dense:
      raise NotImplementedError('custom_mask_function is supported for Dense lists')
    inner = neighbor_list(displacement_or_metric, box, r_cutoff, dr_threshold,
                          capacity_multiplier, disable_cell_list, mask_self, None,
                          fractional_coordinates, format, cell_size_factor, **static_kwargs)
    return _masked_list_fns(inner, custom_mask_function, capacity_multiplier, mask_self)
  general = isinstance(space_mod.geometry_of(displacement_or_metric),
                       space_mod.GeneralDisplacement)
  if fractional_coordinates and not general:
    raise NotImplementedError('fractional coordinates need a space.periodic_general '
                              'displacement')
  if general and not fractional_coordinates:
    raise NotImplementedError('a space.periodic_general displacement stores fractional '
                              'coordinates: pass fractional_coordinates=True')
  spec = NeighborListSpec(displacement_or_metric, box, r_cutoff, dr_threshold,
                          capacity_multiplier, disable_cell_list, mask_self, format,
                          fractional_coordinates, cell_size_factor)
  fcode = _FORMAT_CODE[format]

  def _build(position, nbrs: Optional[NeighborList], plan: _Plan, gate, side=None):
    lib = nat.load()
    plan.epoch += 1
    N, dim = position.shape
    st = nat.stream_handle(position.device)
    nat.check(lib.jmd_neighbor_list_build(
        nat.ptr(position), N, dim, _check_position(position), int(plan.periodic),
        nat.f3(plan.side if side is None else side), float(spec.cutoff_sq),
        int(plan.use_cell_list), nat.i3(plan.cells_per_side),
        nat.f3(plan.cell_size3), int(nbrs.cell_list_capacity or 0), fcode,
        int(spec.mask_self), nat.ptr(nbrs.idx), int(nbrs.max_occupancy),
        nat.ptr(plan.occupancy), nat.ptr(nbrs.error.code), nat.ptr(nbrs.reference_position),
        gate, plan.workspace.ptr, st), 'jmd_neighbor_list_build')

  spec.build = _build
  spec.format_code = fcode

  def _shadow_build(position, shadow, gate, side):
    """The Dense rows of the Sparse list's pairs: the same grid, skin decision
    and reference positions, each candidate kept by the Sparse pair test
    (wrap(R_i - R_j), both orders stored); a row longer than the shadow's
    width raises the list's overflow bit."""
    lib = nat.load()
    sp = shadow.plan
    N, dim = position.shape
    nat.check(lib.jmd_neighbor_list_build(
        nat.ptr(position), N, dim, _check_position(position), int(sp.periodic), nat.f3(side),
        float(spec.cutoff_sq), int(sp.use_cell_list), nat.i3(sp.cells_per_side),
        nat.f3(sp.cell_size3), int(shadow.cell_list_capacity or 0),
        nat.DENSE | nat.DENSE_SPARSE_TEST, int(spec.mask_self), nat.ptr(shadow.idx),
        int(shadow.max_occupancy), nat.ptr(sp.occupancy), nat.ptr(shadow.error.code),
        nat.ptr(shadow.reference_position), gate, sp.workspace.ptr,
        nat.stream_handle(position.device)), 'jmd_neighbor_list_build(shadow)')

  def _shadow_allocate(position, nbrs, plan, side):
    N, dim = position.shape
    dev = position.device
    recv, send = nbrs.idx[0].long(), nbrs.idx[1].long()
    valid = (recv < N) & (send < N)
    rows = torch.bincount(send[valid], minlength=N)
    if format is OrderedSparse:
      rows = rows + torch.bincount(recv[valid], minlength=N)
    longest = int(rows.max().item()) if N else 0
    limit = N - 1 if mask_self else N
    cand = (3 ** dim) * int(nbrs.cell_list_capacity or 0) if plan.use_cell_list else N
    width = max(min(max(2 * longest, longest + 32), limit, cand), 1)
    ncell = 1
    for c in plan.cells_per_side:
      ncell *= int(c)
    lib = nat.load()
    ws = _Workspace(dev, lib.jmd_neighbor_workspace_bytes(
        N, dim, ncell if plan.use_cell_list else 0, int(nbrs.cell_list_capacity or 0),
        nat.DENSE))
    plan2 = _Plan(ws, torch.zeros((1,), dtype=torch.int32, device=dev),
                  torch.zeros((1,), dtype=torch.int32, device=dev), plan.use_cell_list,
                  plan.cells_per_side, plan.cell_size3, plan.side, plan.periodic, plan.dim,
                  None)
    shadow = NeighborList(
        idx=torch.empty((N, width), dtype=torch.int32, device=dev),
        reference_position=nbrs.reference_position, error=nbrs.error,
        cell_list_capacity=nbrs.cell_list_capacity, max_occupancy=width, format=Dense,
        cell_size=nbrs.cell_size, cell_list_fn=None, update_fn=None, plan=plan2)
    _shadow_build(position, shadow, None, side)
    plan.shadow = shadow
    plan.shadow_mark = (plan.epoch, nbrs.idx._version)

  def allocate_fn(position, extra_capacity: int = 0, **kwargs) -> NeighborList:
    # box= at allocation (fractional coordinates): the cell grid on the unit
    # cube and the distances come from that box (partition.py:1045-1050)
    box_kw = kwargs.get('box')
    if box_kw is not None and not spec.fractional:
      raise ValueError('Neighbor list cannot accept a box keyword argument '
                       'if fractional_coordinates is not enabled.')
    stride = _check_position(position)
    lib = nat.load()
    N, dim = position.shape
    dev = position.device
    geo = spec.geometry
    side = geo.side_values(dim) if geo.periodic else [1.0] * dim
    min_cs = spec.min_cell_size
    if box_kw is not None:
      side = [float(v) for v in space_mod._box_as_matrix(box_kw, dim).reshape(-1)]
      min_cs = spec.fractional_cell_size(box_kw)
    use_cells = spec.use_cell_list(min_cs)
    cps3, cs3, ncell, cap = [1] * dim, [1.0] * dim, 1, None
    cell_size = None
    if use_cells:
      _, cell_size, cps, ncell = cell_dimensions(dim, spec.box, min_cs)
      cps3, cs3 = _per_dim(cps, dim), _per_dim(cell_size, dim)
      _, cs2, cps2, _ = cell_dimensions(dim, spec.box, cell_size)
      fill = _filling_max(position, _per_dim(cps2, dim), _per_dim(cs2, dim))
      cap = int(onp.int32(fill) * capacity_multiplier) + extra_capacity
      cell_size = min_cs
    ws = _Workspace(dev, lib.jmd_neighbor_workspace_bytes(N, dim, ncell if use_cells else 0,
                                                          int(cap or 0), fcode))
    plan = _Plan(ws, torch.zeros((1,), dtype=torch.int32, device=dev),
                 torch.zeros((1,), dtype=torch.int32, device=dev), use_cells, cps3, cs3, side,
                 int(geo.periodic), dim, spec)
    st = nat.stream_handle(dev)
    # 1) measure the occupancy (allocate is not jittable in the reference either)
    nat.check(lib.jmd_neighbor_list_build(
        nat.ptr(position), N, dim, stride, int(plan.periodic), nat.f3(side),
        float(spec.cutoff_sq),
        int(use_cells), nat.i3(cps3), nat.f3(cs3), int(cap or 0), fcode, int(spec.mask_self),
        None, 0, nat.ptr(plan.occupancy), None, None, None, ws.ptr, st),
        'jmd_neighbor_list_build(count)')
    occupancy = int(plan.occupancy.item())
    # 2) capacity (partition.py:1092-1104), float32 arithmetic like the reference
    sparse = is_sparse(format)
    extra = extra_capacity if not sparse else N * extra_capacity
    max_occ = int(f32(occupancy) * f32(capacity_multiplier) + f32(extra))
    n_cand = N if not use_cells else (3 ** dim) * cap
    total = n_cand if not sparse else N * n_cand
    max_occ = min(max_occ, total)
    if not sparse:
      limit = N - 1 if mask_self else N
    elif format is Sparse:
      limit = N * (N - 1) if mask_self else N * N
    else:
      limit = N * (N - 1) // 2
    max_occ = max(min(max_occ, limit), 0)
    shape = (N, max_occ) if not sparse else (2, max_occ)
    nbrs = NeighborList(
        idx=torch.empty(shape, dtype=torch.int32, device=dev),
        reference_position=torch.empty((N, dim), dtype=torch.float32, device=dev),
        error=PartitionError(torch.zeros((), dtype=torch.uint8, device=dev)),
        cell_list_capacity=cap, max_occupancy=max_occ, format=format, cell_size=cell_size,
        cell_list_fn=None, update_fn=update_fn, plan=plan)
    _build(position, nbrs, plan, None)
    if sparse:
      _shadow_allocate(position, nbrs, plan, side)
    if spec.fractional and not spec.disable_cell_list and is_box_valid(
        spec.user_box if box_kw is None else box_kw):
      # the reference raises this bit when is_box_valid holds (its predicate,
      # partition.py:1049, as written)
      nbrs.error.code.bitwise_or_(int(PEC.MALFORMED_BOX))
    return nbrs

  def update_fn(position, neighbors: NeighborList, **kwargs) -> NeighborList:
    stride = _check_position(position)
    plan = neighbors.plan
    if plan is None or tuple(position.shape) != tuple(neighbors.reference_position.shape):
      raise ValueError('update needs a list allocated for the same particle count')
    lib = nat.load()
    N, dim = position.shape
    side = None
    if 'box' in kwargs:
      # a new box for this update (partition.py:1123-1139): the cells stay
      # those of the allocation; CELL_SIZE_TOO_SMALL when the new box needs
      # larger cells on the unit cube, MALFORMED_BOX with the reference's
      # predicate; the skin test and the build measure in the new box
      new_box = kwargs['box']
      if not spec.disable_cell_list:
        if not spec.fractional:
          raise ValueError('Neighbor list cannot accept a box keyword argument '
                           'if fractional_coordinates is not enabled.')
        bits = 0
        if plan.use_cell_list:
          cur = f32(1.0) / onp.floor(f32(1.0) / f32(neighbors.cell_size)).astype(f32)
          new = f32(1.0) / onp.floor(
              f32(1.0) / _fractional_cell_size(new_box, spec.cutoff)).astype(f32)
          if bool(onp.any(new > cur)):
            bits |= int(PEC.CELL_SIZE_TOO_SMALL)
        if is_box_valid(new_box):
          bits |= int(PEC.MALFORMED_BOX)
        if bits:
          neighbors.error.code.bitwise_or_(bits)
      if spec.fractional:
        side = [float(v) for v in space_mod._box_as_matrix(new_box, dim).reshape(-1)]
      elif spec.geometry.periodic:
        raise ValueError('`space.periodic` does not accept a box argument. Perhaps you '
                         'meant to use `space.periodic_general`?')
    side_now = plan.side if side is None else side
    shadow = plan.shadow if plan.shadow is not None and _shadow_intact(neighbors) else None
    plan.shadow_mark = None
    if spec.always_rebuild:
      _build(position, neighbors, plan, None, side_now)
      if shadow is not None:
        _shadow_build(position, shadow, None, side_now)
    else:
      st = nat.stream_handle(position.device)
      nat.check(lib.jmd_neighbor_list_check(
          nat.ptr(position), stride, nat.ptr(neighbors.reference_position), N, dim,
          int(plan.periodic), nat.f3(side_now), float(spec.threshold_sq),
          nat.ptr(plan.gate), st), 'jmd_neighbor_list_check')
      if shadow is not None:
        shadow.plan.gate.copy_(plan.gate)  # the Sparse build resets its gate word
      _build(position, neighbors, plan, nat.ptr(plan.gate), side_now)
      if shadow is not None:
        _shadow_build(position, shadow, nat.ptr(shadow.plan.gate), side_now)
    if shadow is not None:
      plan.shadow_mark = (plan.epoch, neighbors.idx._version)
    return dataclasses.replace(neighbors)

  return NeighborListFns(allocate_fn, update_fn)


def _mask_compact(raw_idx, mask_fn, N, out):
  """Applies a custom mask to Dense rows and compacts the kept entries to
  the left of `out` in order (padding N); returns the kept count per row."""
  m = mask_fn(raw_idx)
  if not isinstance(m, torch.Tensor) or tuple(m.shape) != tuple(raw_idx.shape):
    raise ValueError('custom_mask_function must return an array of the shape it was given')
  m = m.to(device=raw_idx.device, dtype=torch.int32)
  keep = m < N
  cnt = keep.sum(1)
  pos = torch.cumsum(keep.to(torch.int32), 1) - 1
  out.fill_(N)
  ok = keep & (pos < out.shape[1])
  rows = torch.arange(N, device=raw_idx.device)[:, None].expand_as(m)
  out[rows[ok], pos[ok].long()] = m[ok]
  return cnt


def _masked_list_fns(inner: NeighborListFns, mask_fn, capacity_multiplier, mask_self):
  """Dense lists with a custom mask: the unmasked list is built and updated
  natively (gated on the skin criterion as usual); every allocate / update
  then masks its rows and compacts them into the list the caller sees, whose
  width follows the reference's capacity rule on the masked occupancy
  (partition.py:1092-1104).  Both share the error word and reference
  positions; a masked row longer than the width sets the overflow bit."""

  def finish(raw, out_idx):
    N = raw.reference_position.shape[0]
    cnt = _mask_compact(raw.idx, mask_fn, N, out_idx)
    over = (cnt.max() > out_idx.shape[1]) if N else torch.zeros((), dtype=torch.bool,
                                                                device=out_idx.device)
    raw.error.code.bitwise_or_(over.to(torch.uint8) * int(PEC.NEIGHBOR_LIST_OVERFLOW))
    return cnt

  def allocate_fn(position, extra_capacity: int = 0, **kwargs) -> NeighborList:
    raw = inner.allocate(position, extra_capacity, **kwargs)
    N = position.shape[0]
    tmp = torch.empty_like(raw.idx)
    cnt = _mask_compact(raw.idx, mask_fn, N, tmp)
    occupancy = int(cnt.max().item()) if N else 0
    max_occ = int(f32(occupancy) * f32(capacity_multiplier) + f32(extra_capacity))
    max_occ = max(min(max_occ, raw.idx.shape[1], N - 1 if mask_self else N), 0)
    idx = torch.empty((N, max_occ), dtype=torch.int32, device=position.device)
    nbrs = dataclasses.replace(raw, idx=idx, max_occupancy=max_occ, update_fn=update_fn)
    nbrs.plan.custom_mask = (raw, mask_fn)
    finish(raw, idx)
    return nbrs

  def update_fn(position, neighbors: NeighborList, **kwargs) -> NeighborList:
    raw, _ = neighbors.plan.custom_mask
    raw = inner.update(position, raw, **kwargs)
    neighbors.plan.custom_mask = (raw, mask_fn)
    finish(raw, neighbors.idx)
    return dataclasses.replace(neighbors)

  return NeighborListFns(allocate_fn, update_fn)


def neighbor_list_mask(neighbor: NeighborList, mask_self: bool = False) -> torch.Tensor:
  """Validity mask of a list (partition.py:1167-1182)."""
  if is_sparse(neighbor.format):
    mask = neighbor.idx[0] < len(neighbor.reference_position)
    if mask_self:
      mask = mask & (neighbor.idx[0] != neighbor.idx[1])
    return mask
  mask = neighbor.idx < len(neighbor.idx)
  if mask_self:
    N = len(neighbor.reference_position)
    self_mask = neighbor.idx != torch.arange(N, dtype=torch.int32,
                                             device=neighbor.idx.device).reshape(N, 1)
    mask = mask & self_mask
  return mask


class GraphsTuple(NamedTuple):
  """The fields of jraph.GraphsTuple (jraph is not a dependency here)."""
  nodes: Any
  edges: Any
  receivers: torch.Tensor
  senders: torch.Tensor
  globals: Any
  n_node: torch.Tensor
  n_edge: torch.Tensor


def _tree_map(fn, tree):
  if tree is None:
    return None
  if isinstance(tree, dict):
    return {k: _tree_map(fn, v) for k, v in tree.items()}
  if isinstance(tree, (list, tuple)) and not hasattr(tree, '_fields'):
    return type(tree)(_tree_map(fn, v) for v in tree)
  return fn(tree)


def to_jraph(neighbor: NeighborList, mask: Optional[torch.Tensor] = None, nodes=None,
             edges=None, globals=None) -> GraphsTuple:
  """A Sparse-format list as a graph (partition.py:1185-1250): padding is one
  fictitious graph of one node; with ``mask`` the kept edges move to the
  front in their order (edge features with them) and the rest become padding
  edges (receiver and sender N, features zero)."""
  if not is_sparse(neighbor.format):
    raise ValueError('Cannot convert a dense neighbor list to jraph format. '
                     'Please use either NeighborListFormat.Sparse or '
                     'NeighborListFormat.OrderedSparse.')
  receivers, senders = neighbor.idx
  N = len(neighbor.reference_position)
  _mask = neighbor_list_mask(neighbor)

  def pad(x):
    return torch.cat((x, torch.zeros((1,) + tuple(x.shape[1:]), dtype=x.dtype,
                                     device=x.device)), 0)

  nodes = _tree_map(pad, nodes)
  globals = _tree_map(pad, globals)
  if mask is not None:
    _mask = _mask & mask.to(device=_mask.device, dtype=torch.bool)
    E = receivers.shape[0]
    index = torch.where(_mask, torch.cumsum(_mask.to(torch.int64), 0) - 1,
                        torch.full_like(receivers, E, dtype=torch.int64))
    keep = index < E

    def reorder(x, fill):
      out = torch.full((E,) + tuple(x.shape[1:]), fill, dtype=x.dtype, device=x.device)
      out[index[keep]] = x[keep]
      return out

    receivers = reorder(receivers, N)
    senders = reorder(senders, N)
    edges = _tree_map(lambda x: reorder(x, 0), edges)
  dev = receivers.device
  n_valid = _mask.sum().to(torch.int32)
  return GraphsTuple(nodes=nodes, edges=edges, receivers=receivers, senders=senders,
                     globals=globals,
                     n_node=torch.tensor([N, 1], dtype=torch.int32, device=dev),
                     n_edge=torch.stack((n_valid, (~_mask).sum().to(torch.int32))))


def _shadow_intact(nbrs) -> bool:
  """The Sparse list was last written by this module's allocate / update
  (not by a loop, by hand or by another build)."""
  plan = nbrs.plan
  return plan.shadow_mark == (plan.epoch, nbrs.idx._version)


def current_shadow(nbrs) -> Optional[NeighborList]:
  """The Dense rows of a Sparse-format list's pairs when they match the
  list as it stands, else None."""
  plan = nbrs.plan
  if plan is None or plan.shadow is None or not is_sparse(nbrs.format):
    return None
  return plan.shadow if _shadow_intact(nbrs) else None


def dense_rows(idx: torch.Tensor, N: int, fmt) -> torch.Tensor:
  """Dense rows [N, W] (padding N) of a list: Dense as it is; Sparse
  (receiver, sender) pairs grouped by sender; OrderedSparse pairs in both
  orders.  Each row lists its neighbours in increasing id order."""
  if fmt is Dense:
    return idx
  recv, send = idx[0].long(), idx[1].long()
  valid = (recv < N) & (send < N)
  recv, send = recv[valid], send[valid]
  if fmt is OrderedSparse:
    rows, cols = torch.cat([send, recv]), torch.cat([recv, send])
  else:
    rows, cols = send, recv
  order = torch.argsort(rows * (N + 1) + cols)
  rows, cols = rows[order], cols[order]
  counts = torch.bincount(rows, minlength=N)
  W = max(int(counts.max().item()) if rows.numel() else 0, 1)
  start = torch.cumsum(counts, 0) - counts
  slot = torch.arange(rows.numel(), device=idx.device) - start[rows]
  out = torch.full((N, W), N, dtype=torch.int32, device=idx.device)
  out[rows, slot] = cols.to(torch.int32)
  return out


def to_dense(neighbor: NeighborList) -> torch.Tensor:
  """Sparse -> Dense ids (partition.py:1253-1271); host-side like the
  reference (not jittable there either)."""
  if neighbor.format is not Sparse:
    raise ValueError('Can only convert sparse neighbor lists to dense ones.')
  receivers, senders = neighbor.idx
  mask = neighbor_list_mask(neighbor)
  receivers = receivers[mask].long()
  senders = senders[mask].long()
  N = len(neighbor.reference_position)
  count = torch.bincount(receivers, minlength=N)
  max_count = int(count.max().item()) if receivers.numel() else 0
  offset = torch.arange(max_count, device=receivers.device).repeat(N)[:len(senders)]
  hashes = senders * max_count + offset
  dense = torch.full((N * max_count,), N, dtype=torch.int32, device=receivers.device)
  dense[hashes] = receivers.to(torch.int32)
  return dense.reshape(N, max_count)
