"""CPU oracle for the neighbour-list molecular-dynamics hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this module;
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker.

It restates, in NumPy, the reference algorithm of jax-md 0.2.29 for:

* periodic / free displacement, square distance and shift
  (jax_md/space.py:213-252,258-329,494-502), and periodic_general with
  fractional coordinates (jax_md/space.py:110-169,332-472) with the
  fractional cell size of its neighbour lists (jax_md/partition.py:595-633);
* cell-list binning with the reference's slot layout
  (jax_md/partition.py:146-293,349-471);
* the Dense / Sparse / OrderedSparse neighbour-list build, capacities, error
  bits and the skin criterion (jax_md/partition.py:494-561,891-1164);
* the Lennard-Jones and soft-sphere pair reductions over a neighbour list
  (jax_md/smap.py:697-979, jax_md/energy.py:125-173,246-343,534-580);
* the velocity-Verlet NVE step (jax_md/simulate.py:168-317);
* the FIRE descent step with its float16 scalar arithmetic
  (jax_md/minimize.py:124-226, jax_md/util.py:58-64).

Precision model.  ``dtype=np.float32`` emulates the reference run with x64
off operation by operation (NumPy float32 arithmetic rounds every operation to
nearest and never fuses a multiply-add, which is what XLA's CPU backend does
without fast-math).  ``dtype=np.float64`` emulates the reference's
``JAX_ENABLE_X64`` run: positions and arithmetic in float64, while the values
the reference casts to float32 regardless of x64 (the box inside
``neighbor_list``, ``maybe_downcast`` parameters, the NVE time step) stay
float32-valued.

Pinning.  Every function here is checked in ``tests/test_oracle.py`` against
the reference's golden fixtures (``tests/golden``), closed forms, brute force
or invariants; the header of that file lists which check pins which function.
"""

from __future__ import annotations

import itertools
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

F32 = np.float32
F64 = np.float64

# PartitionErrorCode bits (jax_md/partition.py:514-518).
NONE = 0
NEIGHBOR_LIST_OVERFLOW = 1 << 0
CELL_LIST_OVERFLOW = 1 << 1
CELL_SIZE_TOO_SMALL = 1 << 2
MALFORMED_BOX = 1 << 3

# NeighborListFormat (jax_md/partition.py:659-661).
DENSE, SPARSE, ORDERED_SPARSE = 0, 1, 2


def is_sparse(fmt: int) -> bool:
  return fmt in (SPARSE, ORDERED_SPARSE)


# ---------------------------------------------------------------------------
# Geometry (jax_md/space.py)
# ---------------------------------------------------------------------------


def jnp_mod(x, y):
  """Floor modulus as ``jnp.mod`` computes it: ``rem`` then a sign fix.

  ``lax.rem`` is C ``fmod`` (exact); the fix adds ``y`` when the remainder is
  non-zero and its sign differs from ``y``'s.  The addition rounds in the
  operands' dtype.
  """
  x = np.asarray(x)
  y = np.asarray(y, dtype=x.dtype)
  r = np.fmod(x, y)
  fix = (r != 0) & ((r < 0) != (y < 0))
  return np.where(fix, r + y, r).astype(x.dtype)


def periodic_displacement(side, dR):
  """``mod(dR + side/2, side) - side/2`` (jax_md/space.py:213-224)."""
  dR = np.asarray(dR)
  side = np.asarray(side, dtype=dR.dtype)
  half = side * dR.dtype.type(0.5)
  return (jnp_mod(dR + half, side) - half).astype(dR.dtype)


def square_distance(dR):
  """``sum(dR**2, -1)`` accumulated left to right (jax_md/space.py:227-235)."""
  dR = np.asarray(dR)
  acc = dR[..., 0] * dR[..., 0]
  for k in range(1, dR.shape[-1]):
    acc = acc + dR[..., k] * dR[..., k]
  return acc


def distance(dR):
  """``safe_mask(dr2 > 0, sqrt, dr2)`` (jax_md/space.py:238-247)."""
  dr2 = square_distance(dR)
  return np.where(dr2 > 0, np.sqrt(np.where(dr2 > 0, dr2, 0)), 0).astype(dr2.dtype)


def periodic_shift(side, R, dR):
  """``mod(R + dR, side)`` (jax_md/space.py:250-252)."""
  R = np.asarray(R)
  return jnp_mod(R + dR, np.asarray(side, R.dtype))


@dataclass
class Space:
  """A periodic hypercube (``side`` scalar or per-dimension) or free space.

  ``side`` is kept in the precision the reference's displacement function sees
  it: a Python float is weakly typed, so it acts as float32 in an x64-off run
  and as float64 in an x64 run (jax_md/space.py:275-329).
  """
  side: Optional[np.ndarray]
  periodic: bool = True

  def displacement(self, Ra, Rb):
    d = np.asarray(Ra) - np.asarray(Rb)
    if not self.periodic:
      return d
    return periodic_displacement(np.asarray(self.side, d.dtype), d)

  def metric_sq(self, Ra, Rb):
    return square_distance(self.displacement(Ra, Rb))

  def shift(self, R, dR):
    R = np.asarray(R)
    if not self.periodic:
      return R + dR
    return periodic_shift(np.asarray(self.side, R.dtype), R, dR)


def box_matrix(box, dim, dtype=F32):
  """[dim, dim] matrix of a scalar, vector or matrix box."""
  b = np.asarray(box, dtype)
  if b.ndim == 0:
    return np.eye(dim, dtype=dtype) * b
  if b.ndim == 1:
    return np.diag(b).astype(dtype)
  return b


def transform(T, dR):
  """T applied to the last axis of dR (raw_transform, space.py:128-153): each
  component sum_j T[i, j] dR[..., j], products rounded, summed left to right
  in the arrays' dtype (NumPy never fuses a multiply-add)."""
  dR = np.asarray(dR)
  T = np.asarray(T, dR.dtype)
  out = np.empty_like(dR)
  d = dR.shape[-1]
  for i in range(d):
    acc = T[i, 0] * dR[..., 0]
    for j in range(1, d):
      acc = acc + T[i, j] * dR[..., j]
    out[..., i] = acc
  return out


@dataclass
class GeneralSpace:
  """periodic_general with fractional coordinates (jax_md/space.py:332-472):
  positions u in the unit cube, x = T u.  displacement = T wrap_1(u_a - u_b),
  shift = mod(u + T^-1 dR, 1).  ``box`` is a scalar, vector or matrix, kept
  float32-valued as the reference's box; the inverse is inv(T) in float64
  rounded to the arrays' dtype."""
  box: np.ndarray
  periodic: bool = True

  def matrix(self, dim, dtype):
    return box_matrix(np.asarray(self.box, F32), dim).astype(dtype)

  def displacement(self, Ra, Rb):
    d = np.asarray(Ra) - np.asarray(Rb)
    d = periodic_displacement(np.asarray(1.0, d.dtype), d)
    return transform(self.matrix(d.shape[-1], d.dtype), d)

  def metric_sq(self, Ra, Rb):
    return square_distance(self.displacement(Ra, Rb))

  def shift(self, R, dR):
    R = np.asarray(R)
    T = self.matrix(R.shape[-1], F64)
    Ti = np.linalg.inv(T).astype(R.dtype)
    return periodic_shift(np.asarray(1.0, R.dtype), R, transform(Ti, np.asarray(dR, R.dtype)))


def fractional_cell_size(box, cutoff):
  """Cell size on the unit cube (jax_md/partition.py:595-633), float32: the
  cutoff over the smallest perpendicular width of the box."""
  b = np.asarray(box, F32)
  c = F32(cutoff)
  if b.ndim == 0:
    return F32(c / b)
  if b.ndim == 1:
    return F32(c / np.min(b))
  one = F32(1)
  if b.shape[0] == 1:
    return F32(one / np.floor(F32(b[0, 0] / c)))
  if b.shape[0] == 2:
    xy = F32(b[0, 1] / b[1, 1])
    widths = [F32(b[0, 0] / np.sqrt(F32(one + F32(xy * xy)))), b[1, 1]]
  else:
    xy = F32(b[0, 1] / b[1, 1])
    xz = F32(b[0, 2] / b[2, 2])
    yz = F32(b[1, 2] / b[2, 2])
    t = F32(F32(xy * yz) - xz)
    widths = [F32(b[0, 0] / np.sqrt(F32(F32(one + F32(xy * xy)) + F32(t * t)))),
              F32(b[1, 1] / np.sqrt(F32(one + F32(yz * yz)))), b[2, 2]]
  nmin = np.floor(F32(F32(min(widths)) / c))
  nmin = F32(1) if nmin == 0 else F32(nmin)
  return F32(one / nmin)


# ---------------------------------------------------------------------------
# Cell list (jax_md/partition.py:146-471)
# ---------------------------------------------------------------------------


def cell_dimensions(dim, box_size, minimum_cell_size):
  """cells per side, cell size and cell count (jax_md/partition.py:146-188).

  ``box_size`` is a float32 scalar or a float32 vector; the arithmetic is
  float32 NumPy as in the reference.
  """
  box = np.asarray(box_size, F32)
  mcs = np.asarray(minimum_cell_size, F32)
  cps = np.floor(box / mcs).astype(F32)
  cell_size = (box / cps).astype(F32)
  cps_i = cps.astype(np.int32)
  if box.ndim == 1:
    if np.any(cps_i < 3):
      raise ValueError(
        'Box must be at least 3x the size of the grid spacing in each dimension.')
    cell_count = int(np.prod(cps_i))
  else:
    cell_count = int(cps_i) ** dim
  return box, cell_size, cps_i, cell_count


def per_dim(x, dim):
  x = np.asarray(x)
  if x.ndim == 0:
    return np.full((dim,), x, dtype=x.dtype)
  return x.reshape(dim)


def hash_constants(dim, cps):
  """[1, cx, cx*cy] (jax_md/partition.py:211-223)."""
  c = per_dim(cps, dim).astype(np.int64)
  out = [1]
  for d in range(1, dim):
    out.append(out[-1] * int(c[d - 1]))
  return np.array(out, np.int64)


def neighboring_cells(dim):
  """Stencil offsets in iota order, first component slowest
  (jax_md/partition.py:231-240)."""
  return np.array(list(itertools.product((-1, 0, 1), repeat=dim)), np.int64)


def particle_cell_index(position, cell_size, cps):
  """``mod(int32(position / cell_size), cells_per_side)``
  (jax_md/partition.py:421-422).  Division in the positions' dtype, truncation
  toward zero, integer floor modulus."""
  position = np.asarray(position)
  dim = position.shape[1]
  cs = per_dim(np.asarray(cell_size, F32), dim).astype(position.dtype)
  q = position / cs
  idx = np.trunc(q).astype(np.int64)
  return np.mod(idx, per_dim(cps, dim).astype(np.int64))


def cell_hash(position, cell_size, cps):
  dim = position.shape[1]
  ci = particle_cell_index(position, cell_size, cps)
  return ci, ci @ hash_constants(dim, cps)


def estimate_cell_capacity(position, box, cell_size, mult):
  """int(max filling * multiplier) (jax_md/partition.py:191-208,243-250).

  ``cell_size`` is the already-derived cell size; ``count_cell_filling``
  derives the grid from it a second time (floor(box / cell_size) can come out
  one below the first cells-per-side in float32), and counts on that grid.
  """
  dim = position.shape[1]
  _, cs2, cps2, count2 = cell_dimensions(dim, box, cell_size)
  _, h = cell_hash(position, cs2, cps2)
  filling = np.bincount(h, minlength=count2)
  return int(np.int32(filling.max()) * mult)


@dataclass
class CellList:
  position_buffer: np.ndarray   # [ncell*cap, dim]
  id_buffer: np.ndarray         # [ncell*cap] int32, N for empty
  particle_cell_id: np.ndarray  # [N, dim]
  did_buffer_overflow: bool
  cell_capacity: int
  cells_per_side: np.ndarray
  cell_count: int
  hash: np.ndarray              # [N]
  filling: np.ndarray           # [ncell]


def cell_list(position, box_size, minimum_cell_size, buffer_size_multiplier=1.25,
              capacity=None, extra_capacity=0):
  """Bins particles into a uniform grid (jax_md/partition.py:349-471).

  Slot of the k-th particle in hash-sorted order (stable, so ids ascend inside
  a cell) is ``hash * cap + (k mod cap)`` (partition.py:432-442).
  """
  position = np.asarray(position)
  N, dim = position.shape
  box, cell_size, cps, cell_count = cell_dimensions(dim, box_size, minimum_cell_size)
  if capacity is None:
    capacity = estimate_cell_capacity(position, box, cell_size,
                                      buffer_size_multiplier) + extra_capacity
  ci, h = cell_hash(position, cell_size, cps)
  order = np.argsort(h, kind='stable')
  sorted_hash = h[order]
  slot = sorted_hash * capacity + np.arange(N) % capacity
  pos_buf = np.zeros((cell_count * capacity, dim), position.dtype)
  id_buf = np.full((cell_count * capacity,), N, np.int32)
  # With overflow the reference scatter has colliding writes; the last one
  # wins here, which the reference leaves unspecified.
  pos_buf[slot] = position[order]
  id_buf[slot] = order.astype(np.int32)
  filling = np.bincount(h, minlength=cell_count)
  overflow = bool(filling.max(initial=0) > capacity)
  return CellList(pos_buf, id_buf, ci, overflow, int(capacity), cps, cell_count, h,
                  filling)


# ---------------------------------------------------------------------------
# Neighbour list (jax_md/partition.py:801-1164)
# ---------------------------------------------------------------------------


@dataclass
class NeighborList:
  idx: np.ndarray
  reference_position: np.ndarray
  error: int
  cell_list_capacity: Optional[int]
  max_occupancy: int
  format: int
  occupancy: int = 0
  use_cell_list: bool = False


@dataclass
class NeighborListSpec:
  """Static configuration captured by ``partition.neighbor_list``
  (jax_md/partition.py:887-895)."""
  space: Space
  box: np.ndarray               # float32 scalar or vector
  r_cutoff: float
  dr_threshold: float = 0.0
  capacity_multiplier: float = 1.25
  disable_cell_list: bool = False
  mask_self: bool = True
  format: int = DENSE
  custom_mask_function: Optional[object] = None
  fractional_coordinates: bool = False
  cutoff: np.ndarray = field(init=False)
  cutoff_sq: np.ndarray = field(init=False)
  threshold_sq: np.ndarray = field(init=False)
  cell_box: np.ndarray = field(init=False)
  min_cell_size: np.ndarray = field(init=False)

  def __post_init__(self):
    # r_cutoff and dr_threshold arrive as float32 (maybe_downcast in the
    # energy wrappers, energy.py:318-321); box is cast with f32(box).
    self.box = np.asarray(self.box, F32)
    rc = F32(self.r_cutoff)
    dr = F32(self.dr_threshold)
    self.cutoff = F32(rc + dr)
    self.cutoff_sq = F32(self.cutoff * self.cutoff)
    self.threshold_sq = F32((dr / F32(2)) * (dr / F32(2)))
    if self.fractional_coordinates:
      # cells on the unit cube (partition.py:1045-1050)
      self.cell_box = np.asarray(F32(1.0))
      self.min_cell_size = fractional_cell_size(self.box, self.cutoff)
    else:
      self.cell_box = self.box
      self.min_cell_size = self.cutoff

  @property
  def always_rebuild(self):
    return F32(self.dr_threshold) == 0

  def uses_cell_list(self, dim):
    if self.disable_cell_list:
      return False
    box3 = (self.cell_box / F32(3.0)).astype(F32)
    return bool(np.all(self.min_cell_size < box3))


def _candidates_cell(spec, position, cl, rows):
  """Candidate ids for the given rows: 3^dim stencil cells x cap slots, ids
  failing ``wrap(Ri - Rj)^2 < c^2`` replaced by N (partition.py:911-951)."""
  N, dim = position.shape
  cap = cl.cell_capacity
  cps = per_dim(cl.cells_per_side, dim).astype(np.int64)
  shifts = neighboring_cells(dim)
  hm = hash_constants(dim, cps)
  shifted = np.mod(cl.particle_cell_id[rows][:, None, :] + shifts[None], cps)
  cell_1d = shifted @ hm                                  # [R, 3^dim]
  slots = cell_1d[:, :, None] * cap + np.arange(cap)[None, None, :]
  slots = slots.reshape(len(rows), -1)
  cand = cl.id_buffer[slots]
  nbr_pos = cl.position_buffer[slots]
  d2 = spec.space.metric_sq(position[rows][:, None, :], nbr_pos)
  return np.where(d2 < spec.cutoff_sq, cand, N).astype(np.int64)


def _prune_dense(spec, position, rows, idx):
  """Keep ``wrap(Rj - Ri)^2 < c^2 & idx < N`` in candidate order
  (partition.py:960-980).  Returns per-row valid lists."""
  N = position.shape[0]
  safe = np.minimum(idx, N - 1)
  d2 = spec.space.metric_sq(position[safe], position[rows][:, None, :])
  mask = (d2 < spec.cutoff_sq) & (idx < N)
  return [idx[r][mask[r]] for r in range(len(rows))]


def _rows_for(spec, position, cl, rows):
  N = position.shape[0]
  if cl is None:
    cand = np.broadcast_to(np.arange(N)[None, :], (len(rows), N)).copy()
  else:
    cand = _candidates_cell(spec, position, cl, rows)
  if spec.mask_self:
    cand = np.where(cand == np.asarray(rows)[:, None], N, cand)
  if spec.custom_mask_function is not None:
    cand = spec.custom_mask_function(cand, np.asarray(rows))
  return cand


def neighbor_rows(spec, position, cl, rows=None, chunk=4096):
  """Valid neighbour ids per row, in the reference's candidate order, for the
  Dense format (both-direction test with a cell list, one direction without)."""
  position = np.asarray(position)
  N = position.shape[0]
  rows = np.arange(N) if rows is None else np.asarray(rows)
  out = []
  for s in range(0, len(rows), chunk):
    r = rows[s:s + chunk]
    cand = _rows_for(spec, position, cl, r)
    out.extend(_prune_dense(spec, position, r, cand))
  return out


def sparse_pairs(spec, position, cl, chunk=4096):
  """(receiver j, sender i) pairs in flattened row-major candidate order
  (partition.py:982-1032)."""
  position = np.asarray(position)
  N = position.shape[0]
  recv, send = [], []
  for s in range(0, N, chunk):
    r = np.arange(s, min(N, s + chunk))
    cand = _rows_for(spec, position, cl, r)
    sender = np.broadcast_to(r[:, None], cand.shape)
    if cl is None:
      safe = np.minimum(cand, N - 1)
      d2 = spec.space.metric_sq(position[sender], position[safe])
      mask = (d2 < spec.cutoff_sq) & (cand < N)
    else:
      mask = cand < N
    if spec.format == ORDERED_SPARSE:
      mask = mask & (cand < sender)
    recv.append(cand[mask])
    send.append(sender[mask])
  return np.concatenate(recv), np.concatenate(send)


def build_cell_list_for(spec, position, neighbors=None, extra_capacity=0):
  if not spec.uses_cell_list(position.shape[1]):
    return None
  if neighbors is None:
    return cell_list(position, spec.cell_box, spec.min_cell_size, spec.capacity_multiplier,
                     extra_capacity=extra_capacity)
  if neighbors.cell_list_capacity is None:
    return None
  return cell_list(position, spec.cell_box, spec.min_cell_size,
                   capacity=neighbors.cell_list_capacity)


def neighbor_list_build(spec, position, neighbors=None, extra_capacity=0):
  """``neighbor_fn`` of partition.py:1037-1117: allocate when ``neighbors`` is
  None, otherwise rebuild at the stored capacities."""
  position = np.asarray(position)
  N, dim = position.shape
  err = 0 if neighbors is None else neighbors.error
  if neighbors is None and spec.fractional_coordinates and not spec.disable_cell_list:
    # raised when the box IS upper triangular: the reference's predicate as
    # written (partition.py:1049 with is_box_valid, partition.py:676-681)
    b = np.asarray(spec.box)
    if b.ndim <= 1 or bool(np.all(np.triu(b) == b)):
      err |= MALFORMED_BOX
  cl = build_cell_list_for(spec, position, neighbors, extra_capacity)
  if cl is not None and cl.did_buffer_overflow:
    err |= CELL_LIST_OVERFLOW
  n_cand = N if cl is None else (3 ** dim) * cl.cell_capacity
  if is_sparse(spec.format):
    recv, send = sparse_pairs(spec, position, cl)
    occupancy = len(recv)
    total = N * n_cand
  else:
    rows = neighbor_rows(spec, position, cl)
    occupancy = max((len(r) for r in rows), default=0)
    total = n_cand
  if neighbors is None:
    extra = extra_capacity if not is_sparse(spec.format) else N * extra_capacity
    max_occ = int(F32(occupancy) * F32(spec.capacity_multiplier) + F32(extra))
    max_occ = min(max_occ, total)
    if not is_sparse(spec.format):
      limit = N - 1 if spec.mask_self else N
    elif spec.format == SPARSE:
      limit = N * (N - 1) if spec.mask_self else N * N
    else:
      limit = N * (N - 1) // 2
    max_occ = min(max_occ, limit)
  else:
    max_occ = neighbors.max_occupancy
  if is_sparse(spec.format):
    idx = np.full((2, max_occ), N, np.int32)
    n = min(max_occ, occupancy)
    idx[0, :n] = recv[:n]
    idx[1, :n] = send[:n]
  else:
    idx = np.full((N, max_occ), N, np.int32)
    for i, r in enumerate(rows):
      n = min(max_occ, len(r))
      idx[i, :n] = r[:n]
  if occupancy > max_occ:
    err |= NEIGHBOR_LIST_OVERFLOW
  return NeighborList(idx, position.copy(), err,
                      None if cl is None else cl.cell_capacity, max_occ, spec.format,
                      occupancy, cl is not None)


def needs_rebuild(spec, position, reference_position):
  """``any(metric_sq(R, R_ref) > (dr/2)^2)`` (partition.py:1141-1154)."""
  if spec.always_rebuild:
    return True
  d2 = spec.space.metric_sq(np.asarray(position), np.asarray(reference_position))
  return bool(np.any(d2 > spec.threshold_sq))


def neighbor_list_update(spec, position, neighbors):
  if needs_rebuild(spec, position, neighbors.reference_position):
    return neighbor_list_build(spec, position, neighbors)
  return neighbors


def sorted_rows(nl):
  """Per-row sorted valid ids of a Dense list (the set the parity tests
  compare)."""
  N = nl.idx.shape[0]
  return [np.sort(r[r < N]) for r in nl.idx]


# ---------------------------------------------------------------------------
# Pair potentials (jax_md/energy.py)
# ---------------------------------------------------------------------------


@dataclass
class LJParams:
  """Lennard-Jones with the multiplicative C1 cutoff; r_onset and r_cutoff are
  scaled by max(sigma) at construction (energy.py:318-320)."""
  sigma: float = 1.0
  epsilon: float = 1.0
  r_onset: float = 2.0
  r_cutoff: float = 2.5

  @property
  def r_on(self):
    return F64(F32(F32(self.r_onset) * F32(np.max(self.sigma))))

  @property
  def r_c(self):
    return F64(F32(F32(self.r_cutoff) * F32(np.max(self.sigma))))


def lj_energy_pair(r, sigma=1.0, epsilon=1.0):
  """4 eps ((s/r)^12 - (s/r)^6) (energy.py:246-272)."""
  idr = (sigma / r) ** 2
  idr6 = idr * idr * idr
  return 4.0 * epsilon * (idr6 * idr6 - idr6)


def smooth(r, r_on, r_c):
  """S(r) of multiplicative_isotropic_cutoff (energy.py:534-580): 1 below
  r_on, (rc^2-r^2)^2 (rc^2+2r^2-3ro^2)/(rc^2-ro^2)^3 up to r_c, 0 beyond."""
  r = np.asarray(r, F64)
  rc2, ro2 = r_c * r_c, r_on * r_on
  r2 = r * r
  with np.errstate(divide='ignore', invalid='ignore'):
    inner = np.where(r < r_c, (rc2 - r2) ** 2 * (rc2 + 2 * r2 - 3 * ro2) / (rc2 - ro2) ** 3, 0.0)
  return np.where(r < r_on, 1.0, inner)


def lj_energy_smoothed(r, sigma, epsilon, r_on, r_c):
  r = np.asarray(r, F64)
  with np.errstate(divide='ignore', invalid='ignore'):
    e = np.nan_to_num(lj_energy_pair(r, sigma, epsilon))
  return smooth(r, r_on, r_c) * e


def lj_dE_over_r(r2, sigma, epsilon, r_on, r_c):
  """(dE/dr)/r of the smoothed LJ pair energy, as a function of r^2.

  E = S V;  V'/r = -24 eps (2 s^12 - s^6) / r^2;
  S'/r = 12 (rc^2 - r^2)(ro^2 - r^2) / (rc^2 - ro^2)^3 on [r_on, r_c).
  """
  r2 = np.asarray(r2, F64)
  s2 = sigma * sigma / r2
  s6 = s2 * s2 * s2
  V = 4.0 * epsilon * (s6 * s6 - s6)
  dV = -24.0 * epsilon * (2.0 * s6 * s6 - s6) / r2
  rc2, ro2 = r_c * r_c, r_on * r_on
  if rc2 > ro2:
    D = (rc2 - ro2) ** 3
    S = np.where(r2 < ro2, 1.0, np.where(r2 < rc2, (rc2 - r2) ** 2 * (rc2 + 2 * r2 - 3 * ro2) / D, 0.0))
    dS = np.where((r2 >= ro2) & (r2 < rc2), 12.0 * (rc2 - r2) * (ro2 - r2) / D, 0.0)
  else:
    S = np.where(r2 < rc2, 1.0, 0.0)
    dS = np.zeros_like(r2)
  return S * dV + dS * V, S * V


def soft_sphere_energy(r, sigma=1.0, epsilon=1.0, alpha=2.0):
  """eps/alpha (1 - r/sigma)^alpha for r < sigma (energy.py:125-173)."""
  r = np.asarray(r, F64)
  x = r / sigma
  return np.where(x < 1.0, epsilon / alpha * np.abs(1.0 - np.minimum(x, 1.0)) ** alpha, 0.0)


def soft_sphere_dE_over_r(r2, sigma, epsilon, alpha):
  r2 = np.asarray(r2, F64)
  r = np.sqrt(r2)
  x = r / sigma
  inside = x < 1.0
  base = np.where(inside, 1.0 - x, 0.0)
  E = np.where(inside, epsilon / alpha * base ** alpha, 0.0)
  with np.errstate(divide='ignore', invalid='ignore'):
    dEdr = np.where(inside, -epsilon / sigma * base ** (alpha - 1.0), 0.0)
    dEr = np.where(inside & (r > 0), dEdr / r, 0.0)
  return dEr, E


# ---------------------------------------------------------------------------
# Pair reduction over a neighbour list (jax_md/smap.py:856-979)
# ---------------------------------------------------------------------------


@dataclass
class PairModel:
  """kind 'lj' or 'soft_sphere'; parameters may be scalars, per-particle
  vectors (combined as 0.5 (x + y), smap.py:836) or species matrices
  (smap.py:787-797)."""
  kind: str
  sigma: object = 1.0
  epsilon: object = 1.0
  alpha: object = 2.0
  species: Optional[np.ndarray] = None
  r_onset: float = 2.0
  r_cutoff: float = 2.5
  fixed_smoothing: Optional[tuple] = None

  def pair_params(self, i, j, name):
    p = np.asarray(getattr(self, name), F64)
    if p.ndim == 0:
      return np.broadcast_to(p, np.broadcast(i, j).shape)
    if p.ndim == 1:
      return 0.5 * (p[i] + p[j])
    if self.species is not None:
      s = np.asarray(self.species)
      return p[s[i], s[j]]
    return p[i, j]

  def smoothing(self):
    if self.fixed_smoothing is not None:  # absolute (r_on, r_c), e.g. for d/dsigma
      return F64(self.fixed_smoothing[0]), F64(self.fixed_smoothing[1])
    smax = F32(np.max(np.asarray(self.sigma, F32)))
    r_on = F64(F32(F32(self.r_onset) * smax))
    r_c = F64(F32(F32(self.r_cutoff) * smax))
    return r_on, r_c


def pairs_from_dense(idx):
  N = idx.shape[0]
  i = np.broadcast_to(np.arange(N)[:, None], idx.shape)
  m = idx < N
  return i[m], idx[m].astype(np.int64)


def pair_energy_force(model, space, R, i, j, weight=0.5):
  """Energy and force over the ordered pairs (i, j), each counted with
  ``weight`` (0.5 for Dense/Sparse lists holding both orders, 1 for
  OrderedSparse).  FP64.  Displacement dR = disp(R_j, R_i) as in map_neighbor
  (space.py:494-502).  Returns (E_total, per-particle E, F)."""
  R = np.asarray(R, F64)
  N = R.shape[0]
  dR = space.displacement(R[j], R[i]).astype(F64)
  r2 = np.sum(dR * dR, axis=-1)
  sig = model.pair_params(i, j, 'sigma')
  eps = model.pair_params(i, j, 'epsilon')
  if model.kind == 'lj':
    r_on, r_c = model.smoothing()
    dEr, E = lj_dE_over_r(r2, sig, eps, r_on, r_c)
  else:
    alp = model.pair_params(i, j, 'alpha')
    dEr, E = soft_sphere_dE_over_r(r2, sig, eps, alp)
  E = weight * E
  f = (weight * dEr)[:, None] * dR       # force on i from pair (i, j)
  F = np.zeros_like(R)
  np.add.at(F, i, f)
  np.add.at(F, j, -f)
  e_atom = np.zeros(N)
  np.add.at(e_atom, i, E)
  return float(np.sum(E)), e_atom, F


def pair_energies(model, space, R, i, j, weight=0.5):
  """The energy of each ordered pair (i, j) times ``weight``, FP64: the
  reduce_axis=() output of smap.pair_neighbor_list (smap.py:940-953), whose
  ``out / normalization`` is weight = 1/2 for lists holding both orders and
  1 for OrderedSparse.  Same pair terms as ``pair_energy_force``."""
  R = np.asarray(R, F64)
  dR = space.displacement(R[j], R[i]).astype(F64)
  r2 = np.sum(dR * dR, axis=-1)
  sig = model.pair_params(i, j, 'sigma')
  eps = model.pair_params(i, j, 'epsilon')
  if model.kind == 'lj':
    r_on, r_c = model.smoothing()
    E = lj_dE_over_r(r2, sig, eps, r_on, r_c)[1]
  else:
    E = soft_sphere_dE_over_r(r2, sig, eps, model.pair_params(i, j, 'alpha'))[1]
  return weight * E


def pair_virial(model, space, R, i, j, weight=0.5):
  """W_ab = weight sum over the ordered pairs (i, j) of (U'(r)/r) d_a d_b,
  d = disp(R_j, R_i): the derivative dU/d eps of the energy with every
  displacement perturbed as (I + eps) d, which quantity.stress / pressure
  take with jax.grad (quantity.py:202-282, space.py:299-300).  FP64."""
  R = np.asarray(R, F64)
  dR = space.displacement(R[j], R[i]).astype(F64)
  r2 = np.sum(dR * dR, axis=-1)
  sig = model.pair_params(i, j, 'sigma')
  eps = model.pair_params(i, j, 'epsilon')
  if model.kind == 'lj':
    r_on, r_c = model.smoothing()
    dEr, _ = lj_dE_over_r(r2, sig, eps, r_on, r_c)
  else:
    dEr, _ = soft_sphere_dE_over_r(r2, sig, eps, model.pair_params(i, j, 'alpha'))
  return weight * np.einsum('p,pa,pb->ab', dEr, dR, dR)


def stress(W, volume, velocity=None, mass=1.0):
  """(sum m v v^T - W) / V (quantity.py:239-282)."""
  W = np.asarray(W, F64)
  vv = np.zeros_like(W)
  if velocity is not None:
    v = np.asarray(velocity, F64)
    m = mass if np.ndim(mass) == 0 else np.reshape(mass, (-1, 1, 1))
    vv = np.sum(m * v[:, :, None] * v[:, None, :], axis=0)
  return (vv - W) / volume


def pressure(W, volume, kinetic_energy=0.0):
  """(2 KE - tr W) / (dim V) (quantity.py:202-236)."""
  W = np.asarray(W, F64)
  return (2.0 * kinetic_energy - np.trace(W)) / (W.shape[0] * volume)


def pair_force_scale(model, space, R):
  """Per-particle sum of pair-force magnitudes sum_j |f_ij| (FP64): the scale
  against which a net force is judged when it nearly cancels (a packing at
  mechanical equilibrium)."""
  R = np.asarray(R, F64)
  if model.kind == 'lj':
    cutoff = model.smoothing()[1]
  else:
    cutoff = float(np.max(np.asarray(model.sigma, F64)))
  i, j = all_pairs(space, R, cutoff)
  dR = space.displacement(R[j], R[i]).astype(F64)
  r2 = np.sum(dR * dR, axis=-1)
  sig = model.pair_params(i, j, 'sigma')
  eps = model.pair_params(i, j, 'epsilon')
  if model.kind == 'lj':
    r_on, r_c = model.smoothing()
    dEr, _ = lj_dE_over_r(r2, sig, eps, r_on, r_c)
  else:
    dEr, _ = soft_sphere_dE_over_r(r2, sig, eps, model.pair_params(i, j, 'alpha'))
  out = np.zeros(R.shape[0])
  np.add.at(out, i, np.abs(dEr) * np.sqrt(r2))
  return out


def all_pairs(space, R, cutoff):
  """Brute-force ordered pairs with r < cutoff, i != j (FP64)."""
  R = np.asarray(R, F64)
  N = R.shape[0]
  ii, jj = [], []
  for s in range(0, N, 1024):
    rows = np.arange(s, min(N, s + 1024))
    dR = space.displacement(R[None, :, :], R[rows][:, None, :])
    r2 = np.sum(dR * dR, -1)
    m = (r2 < cutoff * cutoff)
    m[np.arange(len(rows)), rows] = False
    a, b = np.nonzero(m)
    ii.append(rows[a])
    jj.append(b)
  return np.concatenate(ii), np.concatenate(jj)


def energy_force_brute(model, space, R):
  """FP64 energy/force over every pair, the counterpart of smap.pair
  (smap.py:548-691) used as the pin for the neighbour-list reduction."""
  if model.kind == 'lj':
    cutoff = model.smoothing()[1]
  else:
    cutoff = float(np.max(np.asarray(model.sigma, F64)))
  i, j = all_pairs(space, R, cutoff)
  return pair_energy_force(model, space, R, i, j, 0.5)


# ---------------------------------------------------------------------------
# Second derivatives: the gradient of the force, which the reference takes
# with jax.grad through smap.pair_neighbor_list (quantity.py:58-60).  Pinned by
# central differences of pair_energy_force (tests/test_oracle.py).
# ---------------------------------------------------------------------------


def pair_d2(model, r2, sig, eps, alp=2.0):
  """FP64 (g, g', dg/dsigma, dg/depsilon, dU/dsigma, dU/depsilon) of one pair
  as functions of s = r^2, where g = U'(r)/r (the force is sum_j g d_ij) and
  g' = dg/ds.  LJ: U = S(s) V(s) with the smoothing constants fixed."""
  s = np.asarray(r2, F64)
  sig = np.asarray(sig, F64)
  eps = np.asarray(eps, F64)
  if model.kind == 'lj':
    r_on, r_c = model.smoothing()
    q = sig * sig / s
    q3 = q ** 3
    q6 = q3 * q3
    V = 4 * eps * (q6 - q3)
    gV = -24 * eps * (2 * q6 - q3) / s
    gVp = 48 * eps * (7 * q6 - 2 * q3) / s ** 2
    V_s = 24 * eps * (2 * q6 - q3) / sig
    gV_s = -144 * eps * (4 * q6 - q3) / (sig * s)
    V_e = 4 * (q6 - q3)
    gV_e = -24 * (2 * q6 - q3) / s
    rc2, ro2 = r_c * r_c, r_on * r_on
    if rc2 > ro2:
      D = (rc2 - ro2) ** 3
      inner = (s >= ro2) & (s < rc2)
      u = rc2 - s
      S = np.where(s < ro2, 1.0, np.where(s < rc2, u * u * (rc2 + 2 * s - 3 * ro2) / D, 0.0))
      Ss = np.where(inner, 6 * u * (ro2 - s) / D, 0.0)
      Sss = np.where(inner, 6 * (2 * s - ro2 - rc2) / D, 0.0)
    else:
      S = np.where(s < rc2, 1.0, 0.0)
      Ss = Sss = np.zeros_like(s)
    g = S * gV + 2 * Ss * V
    gp = gVp * S + 2 * gV * Ss + 2 * V * Sss
    gs = S * gV_s + 2 * Ss * V_s
    ge = S * gV_e + 2 * Ss * V_e
    us = S * V_s
    ue = S * V_e
    zero = ~(s < rc2)
    return tuple(np.where(zero, 0.0, x) for x in (g, gp, gs, ge, us, ue))
  alp = np.asarray(alp, F64)
  r = np.sqrt(s)
  b = 1 - r / sig
  inside = (r < sig) & (s > 0)
  with np.errstate(divide='ignore', invalid='ignore'):
    bm1 = np.where(inside, np.abs(b) ** (alp - 1), 0.0)
    bm2 = np.where(inside, np.abs(b) ** (alp - 2), 0.0)
    g = -eps * bm1 / (sig * r)
    gp = eps / (2 * sig * r) * ((alp - 1) * bm2 / (sig * r) + bm1 / r ** 2)
    ge = -bm1 / (sig * r)
    ue = bm1 * b / alp
    us = eps * bm1 * r / sig ** 2
    gs = -eps / (r * sig ** 2) * ((alp - 1) * bm2 * r / sig - bm1)
  return tuple(np.where(inside, x, 0.0) for x in (g, gp, gs, ge, us, ue))


def force_vjp(model, space, R, v, i, j, dsigma=None):
  """FP64 VJP of F = -dE/dR over the ordered pairs (i, j) of a list holding
  both orders: (J v)_i = sum_j g w + 2 g' (d.w) d with d = disp(R_j, R_i),
  w = v_j - v_i; v^T dF/dsigma, v^T dF/depsilon (scalar parameters); and
  dE/dsigma, dE/depsilon.  With ``dsigma`` (a species matrix of coefficients
  d sigma_ab / d theta) the sigma terms are taken with respect to theta.
  Returns (Jv, (ts, te), (Es, Ee))."""
  R = np.asarray(R, F64)
  v = np.asarray(v, F64)
  d = space.displacement(R[j], R[i]).astype(F64)
  s = np.sum(d * d, -1)
  sig = model.pair_params(i, j, 'sigma')
  eps = model.pair_params(i, j, 'epsilon')
  alp = model.pair_params(i, j, 'alpha') if model.kind != 'lj' else 2.0
  g, gp, gs, ge, us, ue = pair_d2(model, s, sig, eps, alp)
  w = v[j] - v[i]
  dw = np.sum(d * w, -1)
  h = g[:, None] * w + (2 * gp * dw)[:, None] * d
  Jv = np.zeros_like(R)
  np.add.at(Jv, i, h)
  lam = np.sum(v[i] * d, -1)
  if dsigma is not None:
    sp = np.asarray(model.species)
    w_ab = np.asarray(dsigma, F64)[sp[i], sp[j]]
    gs, us = gs * w_ab, us * w_ab
  return Jv, (float(np.sum(gs * lam)), float(np.sum(ge * lam))), \
      (0.5 * float(np.sum(us)), 0.5 * float(np.sum(ue)))


def force_vjp_per_particle(model, space, R, v, i, j):
  """FP64 parameter cotangents for per-particle sigma and epsilon (combined
  as p_ij = (p_i + p_j) / 2, smap.py:836) over the ordered pairs (i, j) of a
  list holding both orders: with F_i = sum_j g_ij d_ij and E = 1/2 sum U_ij,
    v^T dF/dp_k = sum_(i,j) (dg_ij/dp) (v_i . d_ij) (delta_ik + delta_jk) / 2,
    dE/dp_k     = 1/2 sum_(i,j) (dU_ij/dp) (delta_ik + delta_jk) / 2,
  each pair's term scattered to both of its particles.  Returns
  (cot_sigma, cot_eps, dE_dsigma, dE_deps), each float64 [N]."""
  R = np.asarray(R, F64)
  v = np.asarray(v, F64)
  N = R.shape[0]
  d = space.displacement(R[j], R[i]).astype(F64)
  s = np.sum(d * d, -1)
  sig = model.pair_params(i, j, 'sigma')
  eps = model.pair_params(i, j, 'epsilon')
  alp = model.pair_params(i, j, 'alpha') if model.kind != 'lj' else 2.0
  _, _, gs, ge, us, ue = pair_d2(model, s, sig, eps, alp)
  lam = np.sum(v[i] * d, -1)
  out = []
  for term in (gs * lam, ge * lam, 0.5 * us, 0.5 * ue):
    acc = np.zeros(N)
    np.add.at(acc, i, 0.5 * term)
    np.add.at(acc, j, 0.5 * term)
    out.append(acc)
  return tuple(out)


def box_vjp(model, space, R, v, i, j):
  """FP64 derivatives with respect to the side L of a periodic hypercube whose
  sides all equal L, positions held in real coordinates (periodic_general(L I,
  fractional_coordinates=False), space.py:332-472): L enters the pair terms
  only through the minimum image d = R_j - R_i - L n, n the integer image
  vector of the pair, so dd/dL = -n and over the ordered pairs (i, j)
    v^T dF/dL = -sum v_i . (g n + 2 g' (d.n) d),   dE/dL = -1/2 sum g (d.n).
  Returns (v^T dF/dL, dE/dL)."""
  R = np.asarray(R, F64)
  v = np.asarray(v, F64)
  L = F64(np.asarray(space.side).reshape(-1)[0])
  raw = R[j] - R[i]
  d = space.displacement(R[j], R[i]).astype(F64)
  n = np.rint((raw - d) / L)
  s = np.sum(d * d, -1)
  sig = model.pair_params(i, j, 'sigma')
  eps = model.pair_params(i, j, 'epsilon')
  alp = model.pair_params(i, j, 'alpha') if model.kind != 'lj' else 2.0
  g, gp = pair_d2(model, s, sig, eps, alp)[:2]
  dn = np.sum(d * n, -1)
  vF = -np.sum(g * np.sum(v[i] * n, -1) + 2 * gp * dn * np.sum(v[i] * d, -1))
  return float(vF), float(-0.5 * np.sum(g * dn))


def nve_energy_after(model, space, R, P, dt, n_steps):
  """E(x_T) after n_steps of FP64 velocity Verlet with brute-force forces:
  the function whose finite differences pin the trajectory adjoint."""
  def force(X):
    return energy_force_brute(model, space, X)[2]
  st = NVEState(np.asarray(R, F64), np.asarray(P, F64), force(R))
  for _ in range(n_steps):
    st = nve_step(st, force, space, dt)
  return energy_force_brute(model, space, st.position)[0]


# ---------------------------------------------------------------------------
# Dynamics (jax_md/simulate.py, jax_md/minimize.py)
# ---------------------------------------------------------------------------


@dataclass
class NVEState:
  position: np.ndarray
  momentum: np.ndarray
  force: np.ndarray
  mass: object = 1.0


def nve_step(state, force_fn, space, dt):
  """Velocity Verlet with dt cast to float32 (simulate.py:227-243)."""
  dt = F64(F32(dt))
  dt_2 = F64(F32(F32(dt) / F32(2)))
  m = state.mass if np.ndim(state.mass) == 0 else np.reshape(state.mass, (-1, 1))
  p = state.momentum + dt_2 * state.force
  R = space.shift(state.position, dt * p / m)
  F = force_fn(R)
  p = p + dt_2 * F
  return NVEState(R, p, F, state.mass)


@dataclass
class FireState:
  position: np.ndarray
  momentum: np.ndarray
  force: np.ndarray
  mass: object
  dt: np.float16
  alpha: np.float16
  n_pos: int


def static_cast(x):
  """onp.min_scalar_type cast (util.py:58-64)."""
  return np.asarray(x, dtype=np.min_scalar_type(x))[()]


@dataclass
class FireParams:
  dt_start: float = 0.1
  dt_max: float = 0.4
  n_min: float = 5
  f_inc: float = 1.1
  f_dec: float = 0.5
  alpha_start: float = 0.1
  f_alpha: float = 0.99

  def cast(self):
    return tuple(static_cast(x) for x in (self.dt_start, self.dt_max, self.n_min, self.f_inc,
                                           self.f_dec, self.alpha_start, self.f_alpha))


def fire_init(R, force_fn, params, mass=1.0):
  dt_start, _, _, _, _, alpha_start, _ = params.cast()
  R = np.asarray(R, F64)
  return FireState(R, np.zeros_like(R), force_fn(R), mass, dt_start, alpha_start, 0)


def fire_step(state, force_fn, space, params, tape=None):
  """FIRE (minimize.py:184-224): NVE step at dt=state.dt, then the mixing and
  the float16 dt/alpha bookkeeping.  ``tape`` (a list) receives what the
  backward recursion of ``fire_energy_gradient`` needs."""
  dt_start, dt_max, n_min, f_inc, f_dec, alpha_start, f_alpha = params.cast()
  s = nve_step(NVEState(state.position, state.momentum, state.force, state.mass),
               force_fn, space, F32(state.dt))
  R, P, F = s.position, s.momentum, s.force
  F_norm = np.sqrt(np.sum(F * F) + 1e-6)
  P_norm = np.sqrt(np.sum(P * P))
  F_dot_P = np.sum(F * P)
  alpha = state.alpha
  if tape is not None:
    tape.append(dict(x=R, x_prev=state.position, q=P, F=F, dt=F64(F32(state.dt)),
                     alpha=F64(alpha), keep=1.0 if F_dot_P >= 0 else 0.0, pn=P_norm, fn=F_norm))
  P = P + F64(alpha) * (F * P_norm / F_norm - P)
  n_pos = state.n_pos + 1 if F_dot_P >= 0 else 0
  dt = state.dt
  if F_dot_P > 0 and n_pos > n_min:
    dt = min(np.float16(dt * f_inc), dt_max)
  if F_dot_P < 0:
    dt = np.float16(dt * f_dec)
  if F_dot_P > 0 and n_pos > n_min:
    alpha = np.float16(alpha * f_alpha)
  if F_dot_P < 0:
    alpha = alpha_start
  P = P * (1.0 if F_dot_P >= 0 else 0.0)
  return FireState(R, P, F, state.mass, np.float16(dt), np.float16(alpha), n_pos)


def fire_energy_after(model, space, R, n_steps, params):
  """E(x_T) after n_steps of FP64 FIRE from rest with brute-force forces: the
  function whose finite differences pin ``fire_energy_gradient``."""
  def force(X):
    return energy_force_brute(model, space, X)[2]
  st = fire_init(R, force, params)
  for _ in range(n_steps):
    st = fire_step(st, force, space, params)
  return energy_force_brute(model, space, st.position)[0]


def fire_energy_gradient(model, space, R, n_steps, params, dsigma, box=False):
  """(E(x_T), dE(x_T)/dtheta, dE(x_T)/dx_0) through n_steps of FP64 FIRE from
  rest, where the species matrix sigma depends on theta with coefficients
  ``dsigma``; with ``box`` a fourth entry, dE(x_T)/dL for the side L of a
  periodic square or cube (x_0 held fixed in real coordinates, see
  ``box_vjp``).  The reference would take this with jax.grad through the
  minimiser's loop (minimize.py:184-224); the step-size / alpha decisions and
  the F.P >= 0 reset are piecewise constant and carry no gradient.

  Backward recursion per step k (h = dt_k / 2, q = momentum before mixing,
  G = keep * P, s1 = F.G):
    Q  = (1 - a) G + a s1 q / (|q| |F|~)
    Fb = a (|q| / |F|~) G - a s1 |q| F / |F|~^3 + h Q + W
    X += J(x_{k+1}) Fb;  theta += Fb . dF/dtheta (x_{k+1})
    P  = Q + dt X;  W = h P
  and finally X += J(x_0) W, theta += W . dF/dtheta (x_0).  The box adjoint
  collects dE/dL at x_T, Fb . dF/dL at every node, and -X . m_k for the wrap
  x_{k+1} = x_k + dt ph - L m_k of the shift (m_k the integer fold of step k,
  rint((x_k - x_{k+1}) / L) while a step moves less than L / 2)."""
  cutoff = float(np.max(np.asarray(model.sigma, F64)))

  def force(X):
    return energy_force_brute(model, space, X)[2]

  def jvp(X, v):
    i, j = all_pairs(space, X, cutoff)
    Jv, (ts, _), (es, _) = force_vjp(model, space, X, v, i, j, dsigma)
    return Jv, ts, es

  def lvp(X, v):
    i, j = all_pairs(space, X, cutoff)
    return box_vjp(model, space, X, v, i, j)

  st = fire_init(R, force, params)
  x0 = st.position
  tape = []
  for _ in range(n_steps):
    st = fire_step(st, force, space, params, tape)
  E_T, _, F_T = energy_force_brute(model, space, st.position)
  _, _, theta = jvp(st.position, np.zeros_like(F_T))
  lbar = lvp(st.position, np.zeros_like(F_T))[1] if box else 0.0
  L = F64(np.asarray(space.side).reshape(-1)[0]) if box else 1.0
  X = -F_T
  P = np.zeros_like(X)
  W = np.zeros_like(X)
  for rec in reversed(tape):
    a, h, dt = rec['alpha'], 0.5 * rec['dt'], rec['dt']
    G = rec['keep'] * P
    s1 = float(np.sum(rec['F'] * G))
    pn, fn = rec['pn'], rec['fn']
    cq = a * s1 / (pn * fn) if pn > 0 else 0.0
    Q = (1 - a) * G + cq * rec['q']
    Fb = a * (pn / fn) * G - a * s1 * pn / fn ** 3 * rec['F'] + h * Q + W
    Jv, ts, _ = jvp(rec['x'], Fb)
    X = X + Jv
    theta += ts
    if box:
      m = np.rint((rec['x_prev'] - rec['x']) / L)
      lbar += lvp(rec['x'], Fb)[0] - float(np.sum(X * m))
    P = Q + dt * X
    W = h * P
  Jv, ts, _ = jvp(x0, W)
  if box:
    return E_T, theta + ts, X + Jv, lbar + lvp(x0, W)[0]
  return E_T, theta + ts, X + Jv


# Nose-Hoover chain thermostat (simulate.py:328-347,384-672; Martyna et al.
# 1992/1996 with Suzuki-Yoshida substeps).  The weights are the published
# constants of the scheme.
SY_WEIGHTS = {
    1: [1.0],
    3: [0.828981543588751, -0.657963087177502, 0.828981543588751],
    5: [0.2967324292201065, 0.2967324292201065, -0.186929716880426, 0.2967324292201065,
        0.2967324292201065],
    7: [0.784513610477560, 0.235573213359357, -1.17767998417887, 1.31518632068391,
        -1.17767998417887, 0.235573213359357, 0.784513610477560],
}


@dataclass
class NHCState:
  position: np.ndarray
  momentum: np.ndarray
  force: np.ndarray
  mass: object
  xi: np.ndarray        # chain positions [L]
  p_xi: np.ndarray      # chain momenta [L]
  Q: np.ndarray         # chain masses [L]
  ke: float             # kinetic energy the chain last saw
  dof: int


def nhc_masses(L, kT, tau, dof):
  Q = F64(F32(kT)) * F64(F32(tau)) ** 2 * np.ones(L)
  Q[0] *= dof
  return Q


def nhc_substep(delta, xi, p_xi, Q, ke, dof, kT):
  """One chain substep of length delta; returns (xi, p_xi, ke, momentum scale)."""
  d2, d4, d8 = delta / 2, delta / 4, delta / 8
  L = len(p_xi)
  M = L - 1
  p = p_xi.copy()
  p[M] += d4 * (p_xi[M - 1] ** 2 / Q[M - 1] - kT)
  carry = p[M]
  for m in range(M - 1, 0, -1):   # backward sweep, old neighbours below
    G = p_xi[m - 1] ** 2 / Q[m - 1] - kT
    s = np.exp(-d8 * carry / Q[m + 1])
    carry = s * (s * p_xi[m] + d4 * G)
    p[m] = carry
  s = np.exp(-d8 * p[1] / Q[1])
  p[0] = s * (s * p[0] + d4 * (2.0 * ke - dof * kT))
  scale = np.exp(-d2 * p[0] / Q[0])
  ke = ke * scale ** 2
  xi = xi + d2 * p / Q
  G = 2.0 * ke - dof * kT
  q = p.copy()
  for m in range(M):               # forward sweep, neighbours above as left
    s = np.exp(-d8 * p[m + 1] / Q[m + 1])
    q[m] = s * (s * p[m] + d4 * G)
    G = q[m] ** 2 / Q[m] - kT
  q[M] += d4 * G
  return xi, q, ke, scale


def nhc_half_step(dt, chain_steps, sy_steps, xi, p_xi, Q, ke, dof, kT):
  """Half a step of the chain (simulate.py:494-507); the momentum scale is
  the product over the substeps."""
  total = 1.0
  if chain_steps == 1 and sy_steps == 1:
    subs = [F64(F32(dt))]
  else:
    delta = F64(F32(dt)) / chain_steps
    ws = SY_WEIGHTS[sy_steps]
    subs = [F64(F32(delta * ws[i % sy_steps])) for i in range(chain_steps * sy_steps)]
  for d in subs:
    xi, p_xi, ke, sc = nhc_substep(d, xi, p_xi, Q, ke, dof, kT)
    total *= sc
  return xi, p_xi, ke, total


def nvt_init(R, P, force_fn, kT, dt, mass=1.0, chain_length=5, tau=None):
  R = np.asarray(R, F64)
  P = np.asarray(P, F64)
  dof = R.size
  tau = F32(dt) * 100 if tau is None else tau
  m = mass if np.ndim(mass) == 0 else np.reshape(mass, (-1, 1))
  ke = 0.5 * float(np.sum(P ** 2 / m))
  L = chain_length
  return NHCState(R, P, force_fn(R), mass, np.zeros(L), np.zeros(L),
                  nhc_masses(L, kT, tau, dof), ke, dof), F64(F32(tau))


def nvt_step(st, force_fn, space, dt, kT, tau, chain_steps=2, sy_steps=3):
  """NVT Nose-Hoover step (simulate.py:644-660): chain masses, chain half
  step, velocity Verlet, kinetic energy, chain half step."""
  Q = nhc_masses(len(st.Q), kT, tau, st.dof)
  xi, p_xi, ke, s1 = nhc_half_step(dt, chain_steps, sy_steps, st.xi, st.p_xi, Q, st.ke,
                                   st.dof, kT)
  nv = nve_step(NVEState(st.position, st.momentum * s1, st.force, st.mass), force_fn, space,
                dt)
  m = st.mass if np.ndim(st.mass) == 0 else np.reshape(st.mass, (-1, 1))
  ke = 0.5 * float(np.sum(nv.momentum ** 2 / m))
  xi, p_xi, ke, s2 = nhc_half_step(dt, chain_steps, sy_steps, xi, p_xi, Q, ke, st.dof, kT)
  return NHCState(nv.position, nv.momentum * s2, nv.force, st.mass, xi, p_xi, Q, ke, st.dof)


def nvt_invariant(st, energy, kT):
  """Extended-system energy (simulate.py:675-705)."""
  m = st.mass if np.ndim(st.mass) == 0 else np.reshape(st.mass, (-1, 1))
  E = energy + 0.5 * float(np.sum(st.momentum ** 2 / m))
  E += st.p_xi[0] ** 2 / (2 * st.Q[0]) + st.dof * kT * st.xi[0]
  for r, p, q in zip(st.xi[1:], st.p_xi[1:], st.Q[1:]):
    E += p ** 2 / (2 * q) + kT * r
  return E


def kinetic_energy(momentum, mass=1.0):
  """0.5 sum p^2/m (quantity.py:124-159)."""
  m = mass if np.ndim(mass) == 0 else np.reshape(mass, (-1, 1))
  return 0.5 * float(np.sum(np.asarray(momentum, F64) ** 2 / m))


# ---------------------------------------------------------------------------
# Synthetic inputs used by the tests and the benchmark
# ---------------------------------------------------------------------------


def lj_lattice(n_side, rho=0.8442, jitter=0.05, kT=1.0, seed=0, dim=3):
  """Jittered simple-cubic lattice in a periodic cube, L = (N/rho)^(1/dim)
  rounded to float32; momenta N(0, kT) with the mean removed."""
  rng = np.random.default_rng(seed)
  N = n_side ** dim
  L = F32((N / rho) ** (1.0 / dim))
  a = F64(L) / n_side
  grid = np.stack(np.meshgrid(*([np.arange(n_side)] * dim), indexing='ij'), -1)
  R = (grid.reshape(-1, dim)[:, ::-1] + 0.5) * a
  R = R + rng.normal(0.0, jitter, R.shape)
  R = np.mod(R, F64(L))
  R = R.astype(F32)
  R = np.where(R >= L, F32(0), R).astype(F32)
  P = rng.normal(0.0, np.sqrt(kT), R.shape)
  P = (P - P.mean(0, keepdims=True)).astype(F32)
  return R, P, L


def fcc_cubic(n_cells, a=1.5496):
  """FCC crystal, n_cells^3 conventional cubic cells of side a (4 atoms
  each) in a periodic cube.  Returns (R f32, L f32)."""
  basis = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]], F64)
  g = np.stack(np.meshgrid(*([np.arange(n_cells)] * 3), indexing='ij'), -1).reshape(-1, 3)
  R = ((g[:, None, :] + basis[None]) * a).reshape(-1, 3)
  return R.astype(F32), F32(n_cells * a)


def fcc_triclinic(n, a=1.5496, jitter=0.0, kT=0.0, seed=0):
  """The same FCC crystal as an n^3 supercell of its primitive cell (one
  atom each), rotated so that the box matrix is upper triangular: the
  primitive vectors M = a/2 [(0,1,1), (1,0,1), (1,1,0)] (columns) factor as
  M = Q R, and the box is T = n R with positive diagonal.  Optional Gaussian
  jitter in real space.  Returns (fractional positions f32 in [0, 1), T f32
  [3, 3], momenta f32 with the mean removed)."""
  M = 0.5 * a * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0]], F64)
  Q, Rm = np.linalg.qr(M)
  D = np.diag(np.sign(np.diag(Rm)))
  Rm = D @ Rm
  T = (n * Rm).astype(F32)
  g = np.stack(np.meshgrid(*([np.arange(n)] * 3), indexing='ij'), -1).reshape(-1, 3)
  u = g.astype(F64) / n
  rng = np.random.default_rng(seed)
  if jitter > 0:
    x = u @ T.astype(F64).T + rng.normal(0.0, jitter, u.shape)
    u = x @ np.linalg.inv(T.astype(F64)).T
  u = np.mod(u, 1.0).astype(F32)
  u = np.where(u >= F32(1), F32(0), u).astype(F32)
  P = rng.normal(0.0, np.sqrt(kT), u.shape) if kT > 0 else np.zeros(u.shape)
  P = (P - P.mean(0, keepdims=True)).astype(F32)
  return u, T, P


# ---------------------------------------------------------------------------
# Multi-image lists (jax_md/custom_partition.py:147-626)
# ---------------------------------------------------------------------------


def multi_image_shift_n_max(box, cutoff, pbc=None):
  """ceil(c / h_d), h_d = 1 / |column d of inv(T)^T| (custom_partition.py:
  147-209); zero for non-periodic directions."""
  b = np.asarray(box, F64)
  inv_t = np.linalg.inv(b).T
  heights = (1.0 / np.linalg.norm(inv_t, axis=0)).astype(F32)
  n = np.ceil(F32(cutoff) / heights).astype(np.int32)
  pbc = np.ones(b.shape[0], bool) if pbc is None else np.asarray(pbc, bool)
  return np.where(pbc, n, 0).astype(np.int32)


def multi_image_shifts(n_max):
  grids = np.meshgrid(*[np.arange(-n, n + 1) for n in n_max], indexing='ij')
  return np.stack([g.ravel() for g in grids], axis=-1).astype(np.int32)


def multi_image_candidates(position, box, cutoff_sq, shifts, fractional=True):
  """Kept (s, i, j, d^2) over every shift s and pair (i, j): d = (x_j + s T^T)
  - x_i in float32 (x = T u when fractional; products rounded, sums left to
  right), d^2 < cutoff_sq, minus i = j at s = 0."""
  T = np.asarray(box, F32)
  u = np.asarray(position, F32)
  x = transform(T, u) if fractional else u
  sr = transform(T, shifts.astype(F32))
  N = len(x)
  zero = int(np.argmin(np.sum(shifts.astype(np.int64) ** 2, axis=1)))
  out = []
  for k in range(len(shifts)):
    d = (x[None, :, :] + sr[k][None, None, :]) - x[:, None, :]   # [i, j, dim]
    d2 = square_distance(d)
    keep = d2 < F32(cutoff_sq)
    if k == zero:
      keep &= ~np.eye(N, dtype=bool)
    ii, jj = np.nonzero(keep)
    out.append((np.full(len(ii), k), ii, jj, d2[ii, jj]))
  ks, ii, jj, dd = (np.concatenate(v) for v in zip(*out))
  return ks, ii, jj, dd, zero


def multi_image_list(position, box, cutoff_sq, n_max, fmt, capacity, fractional=True):
  """The reference's multi-image build (custom_partition.py:376-626): Sparse
  entries in (s, i, j) order (receiver i, sender j, integer shift);
  OrderedSparse keeps i < j at s = 0 and canonical shifts (first nonzero
  component positive); Dense rows the capacity nearest (s, j) per atom,
  ordered by (d^2, s N + j).  Returns (idx, shifts, occupancy)."""
  shifts = multi_image_shifts(n_max)
  ks, ii, jj, dd, zero = multi_image_candidates(position, box, cutoff_sq, shifts, fractional)
  N = len(position)
  dim = shifts.shape[1]
  if fmt == DENSE:
    idx = np.full((N, capacity), N, np.int32)
    sh = np.zeros((N, capacity, dim), np.int32)
    occ = 0
    for i in range(N):
      m = ii == i
      flat = ks[m].astype(np.int64) * N + jj[m]
      order = np.lexsort((flat, dd[m]))
      occ = max(occ, int(m.sum()))
      for r, o in enumerate(order[:capacity]):
        idx[i, r] = jj[m][o]
        sh[i, r] = shifts[ks[m][o]]
    return idx, sh, occ
  if fmt == ORDERED_SPARSE:
    canon = np.array([next((v > 0 for v in s if v != 0), True) for s in shifts])
    keep = np.where(ks == zero, ii < jj, canon[ks])
    ks, ii, jj = ks[keep], ii[keep], jj[keep]
  order = np.lexsort((jj, ii, ks))            # (s, i, j), shift slowest
  ks, ii, jj = ks[order], ii[order], jj[order]
  occ = len(ks)
  idx = np.full((2, capacity), N, np.int32)
  sh = np.zeros((capacity, dim), np.int32)
  n = min(occ, capacity)
  idx[0, :n], idx[1, :n] = ii[:n], jj[:n]
  sh[:n] = shifts[ks[:n]]
  return idx, sh, occ


def bidisperse_2d(N, phi=0.84, d_small=1.0, d_large=1.4, seed=0):
  """50:50 bidisperse soft-sphere disks at packing fraction phi in a periodic
  square; uniform random positions.  Returns (R f32, species int32, L f32)."""
  rng = np.random.default_rng(seed)
  species = np.zeros(N, np.int32)
  species[N // 2:] = 1
  radii = np.where(species == 0, d_small, d_large) / 2.0
  L = F32(np.sqrt(np.sum(np.pi * radii ** 2) / phi))
  R = (rng.uniform(0.0, 1.0, (N, 2)) * F64(L)).astype(F32)
  R = np.where(R >= L, F32(0), R).astype(F32)
  return R, species, L
