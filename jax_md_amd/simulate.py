"""Velocity-Verlet NVE (jax_md/simulate.py:227-317) on gfx950.

``nve(energy_or_force_fn, shift_fn, dt)`` returns ``(init_fn, step_fn)`` with
the reference's signatures.  A step is two kernels: the half kick + drift
(positions accumulated in float64, mirrored to float32) and the force pass
with the closing half kick fused in.  ``step_fn`` is functional: it returns
new arrays and leaves the input state untouched, so the reference's
speculative-block pattern (discard the block on overflow) works unchanged.
"""

from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as onp
import torch

from . import _native as nat
from . import quantity
from . import space as space_mod

f32 = onp.float32


@dataclasses.dataclass
class NVEState:
  """position, momentum, force [N, dim] float32; mass scalar or [N];
  position_hi: the float64 master copy of the positions.  In the x64 form
  (float64 R given to ``init_fn``, the reference's JAX_ENABLE_X64 run)
  momentum and force are float64 and ``position`` is the float32 mirror of
  ``position_hi`` that the neighbour list is built from."""
  position: torch.Tensor
  momentum: torch.Tensor
  force: torch.Tensor
  mass: object
  position_hi: Optional[torch.Tensor] = None

  @property
  def x64(self) -> bool:
    return self.momentum.dtype == torch.float64

  @property
  def velocity(self):
    m = self.mass.reshape(-1, 1) if isinstance(self.mass, torch.Tensor) and self.mass.ndim else self.mass
    return self.momentum / m

  def set(self, **kw):
    return dataclasses.replace(self, **kw)


def canonicalize_mass(mass, device):
  """float, [N] or [N, 1] (simulate.py:121-140) -> (scalar, device [N] or None)."""
  if isinstance(mass, torch.Tensor) and mass.ndim > 0:
    m = mass.to(device=device, dtype=torch.float32).reshape(-1).contiguous()
    return 0.0, m
  if isinstance(mass, onp.ndarray) and mass.ndim > 0:
    return 0.0, torch.as_tensor(mass.reshape(-1), dtype=torch.float32, device=device)
  return float(f32(float(mass))), None


def native_shift(shift_fn) -> space_mod.Shift:
  if isinstance(shift_fn, space_mod.Shift):
    return shift_fn
  raise NotImplementedError(
      'The native integrators need a shift_fn from jax_md_amd.space.periodic, '
      f'periodic_general or free; got {type(shift_fn).__name__}.')


def _side(shift: space_mod.Shift, dim):
  if isinstance(shift, space_mod.GeneralShift):
    return shift.side_values(dim)  # the box matrix, row-major
  if not shift.periodic:
    return [1.0] * dim
  s = onp.asarray(shift.side, f32)
  return [float(s)] * dim if s.ndim == 0 else [float(v) for v in s]


def positions_in_box(shift: space_mod.Shift, energy, dim: int) -> bool:
  """True when ``shift`` wraps into the energy's own box, so every position it
  writes lies in that box (the force kernels may then skip folds)."""
  if not shift.wraps:
    return False
  geo = energy.geometry
  if not geo.periodic:
    return False
  return [float(v) for v in _side(shift, dim)] == [float(v) for v in geo.side_values(dim)]


def refuse_box(kwargs, what: str):
  """A step that changes the box (``box=`` / ``new_box=`` through the
  integrator, as NPT drives it) is outside the native path: the drift's
  box is the shift function's."""
  if 'box' in kwargs or 'new_box' in kwargs:
    raise NotImplementedError(f'{what}: a changing box (box= / new_box=) is not supported by '
                              'the native integrators; pass it to the energy and the '
                              'neighbour list directly')


def velocity_verlet_native(energy, shift, state: NVEState, dt, neighbor, dt_device=None,
                           **kwargs) -> NVEState:
  """One step: p += dt/2 F; R = shift(R, dt p/m); F = F(R); p += dt/2 F
  (simulate.py:227-243), dt cast to float32 (F5)."""
  refuse_box(kwargs, 'velocity Verlet')
  lib = nat.load()
  R = state.position
  N, dim = R.shape
  dev = R.device
  if state.x64:
    return _velocity_verlet64(energy, shift, state, dt, neighbor, dt_device, **kwargs)
  dt32 = f32(dt)
  half = f32(dt32 / f32(2))
  mass_s, mass_v = canonicalize_mass(state.mass, dev)
  x_hi = state.position_hi if state.position_hi is not None else R.to(torch.float64)
  x_hi_new = torch.empty_like(x_hi)
  x_new = new_positions(N, dim, dev)
  p_new = torch.empty_like(state.momentum)
  wraps = shift.wraps
  nat.check(lib.jmd_nve_kick_drift(
      nat.ptr(x_hi), nat.ptr(state.momentum), nat.ptr(state.force), nat.ptr(mass_v),
      mass_s if mass_v is None else 0.0, N, dim, int(wraps), nat.f3(_side(shift, dim)),
      float(dt32), float(half), nat.ptr(dt_device), nat.ptr(x_hi_new), nat.ptr(x_new),
      nat.position_stride(x_new), nat.ptr(p_new), None, 0.0, None, nat.stream_handle(dev)),
      'jmd_nve_kick_drift')
  F_new = energy.force_and_kick(x_new, neighbor, p_new, half_dt=float(half),
                                dt_device=dt_device, in_box=positions_in_box(shift, energy, dim),
                                **kwargs)
  return NVEState(x_new, p_new, F_new, state.mass, x_hi_new)


def _mass64(mass, device):
  """The x64 mass: a float or a device float64 [N]."""
  if isinstance(mass, (torch.Tensor, onp.ndarray)) and onp.ndim(mass) > 0:
    return 0.0, torch.as_tensor(mass, device=device).to(torch.float64).reshape(-1).contiguous()
  return float(mass), None


def _velocity_verlet64(energy, shift, state: NVEState, dt, neighbor, dt_device=None,
                       **kwargs) -> NVEState:
  """velocity Verlet in float64 (the x64 run; dt stays the float32 value the
  reference casts it to, simulate.py:236)."""
  if dt_device is not None:
    raise NotImplementedError('a device dt is float32-only')
  lib = nat.load()
  x_hi = state.position_hi.clone()
  N, dim = x_hi.shape
  dev = x_hi.device
  dt32 = float(f32(dt))
  half = float(f32(f32(dt) / f32(2)))
  mass_s, mass_v = _mass64(state.mass, dev)
  x_new = new_positions(N, dim, dev)
  p_new = state.momentum.clone()
  nat.check(lib.jmd_nve_kick_drift_f64(
      nat.ptr(x_hi), nat.ptr(p_new), nat.ptr(state.force), nat.ptr(mass_v), mass_s, N, dim,
      int(shift.wraps), nat.f3(_side(shift, dim)), dt32, half, nat.ptr(x_new),
      nat.position_stride(x_new), None, 0.0, None, nat.stream_handle(dev)),
      'jmd_nve_kick_drift_f64')
  F_new = energy.force_and_kick(x_hi, neighbor, p_new, half_dt=half, **kwargs)
  return NVEState(x_new, p_new, F_new, state.mass, x_hi)


def new_positions(N, dim, device, like=None):
  """Float32 positions in the layout the kernels gather fastest: 3-D rows
  padded to 16 bytes (an [N, 3] view of an [N, 4] buffer), 2-D contiguous."""
  if dim == 3:
    return nat.padded_positions(N, device, like)
  out = torch.empty((N, dim), dtype=torch.float32, device=device)
  if like is not None:
    out.copy_(like)
  return out


def _momenta(key, R, kT, mass):
  g = key if isinstance(key, torch.Generator) else None
  if g is None:
    g = torch.Generator(device=R.device)
    g.manual_seed(int(key) if key is not None else 0)
  P = torch.randn(R.shape, generator=g, device=R.device, dtype=torch.float32)
  m = mass.reshape(-1, 1) if isinstance(mass, torch.Tensor) and mass.ndim else mass
  P = torch.sqrt(torch.as_tensor(m * kT, dtype=torch.float32, device=R.device)) * P
  if R.shape[0] > 1:
    P = P - P.mean(dim=0, keepdim=True)
  return P.contiguous()


def nve(energy_or_force_fn, shift_fn, dt=1e-3, **sim_kwargs):
  """NVE ensemble by velocity Verlet (simulate.py:279-317)."""
  energy = quantity.native_energy_of(energy_or_force_fn)
  shift = native_shift(shift_fn)

  def init_fn(key, R, kT, mass=1.0, momenta=None, **kwargs):
    if isinstance(R, torch.Tensor) and R.dtype == torch.float64:
      return _init64(key, R, kT, mass, momenta, **kwargs)
    nat.position_stride(R, 'R')
    R = new_positions(R.shape[0], R.shape[1], R.device, R)
    neighbor = kwargs.pop('neighbor', None)
    F = energy.force(R, neighbor, **kwargs)
    if momenta is None:
      P = _momenta(key, R, kT, mass)
    else:
      # a private copy: NVELoop advances the state in place
      P = momenta.to(device=R.device, dtype=torch.float32).clone().contiguous()
    if isinstance(mass, torch.Tensor) and mass.ndim == 2:
      mass = mass.reshape(-1)
    return NVEState(R, P, F, mass, R.to(torch.float64))

  def _init64(key, R, kT, mass, momenta, **kwargs):
    nat.require_device_tensor(R, 'R', torch.float64)
    x_hi = R.clone().contiguous()
    neighbor = kwargs.pop('neighbor', None)
    F = energy.force(x_hi, neighbor, **kwargs)
    if momenta is None:
      P = _momenta(key, R, kT, mass).to(torch.float64)
    else:
      P = momenta.to(device=R.device, dtype=torch.float64).clone().contiguous()
    if isinstance(mass, torch.Tensor) and mass.ndim == 2:
      mass = mass.reshape(-1)
    x32 = new_positions(R.shape[0], R.shape[1], R.device, R.to(torch.float32))
    return NVEState(x32, P, F, mass, x_hi)

  def step_fn(state: NVEState, **kwargs):
    _dt = kwargs.pop('dt', dt)
    neighbor = kwargs.pop('neighbor', None)
    if isinstance(_dt, torch.Tensor):
      return velocity_verlet_native(energy, shift, state, 0.0, neighbor,
                                    dt_device=_dt.to(torch.float32).reshape(1), **kwargs)
    return velocity_verlet_native(energy, shift, state, _dt, neighbor, **kwargs)

  return init_fn, step_fn


def _dense_spec(spec):
  """A copy of a list's static configuration with the Dense format."""
  import copy
  from . import partition
  d = copy.copy(spec)
  d.format = partition.Dense
  # rows kept by the Sparse test, read off as the Sparse list after a block
  d.format_code = nat.DENSE | nat.DENSE_SPARSE_TEST
  return d


def _sparse_pairs(nbrs, N):
  """(rows, cols) of every ordered pair of a Sparse-format list, in entry
  order (OrderedSparse pairs in both orders, the stored one first)."""
  from . import partition
  recv, send = nbrs.idx[0].long(), nbrs.idx[1].long()
  valid = (recv < N) & (send < N)
  recv, send = recv[valid], send[valid]
  if nbrs.format is partition.OrderedSparse:
    return torch.cat([send, recv]), torch.cat([recv, send])
  return send, recv


def _sparse_row_counts(nbrs, N):
  rows, _ = _sparse_pairs(nbrs, N)
  counts = torch.zeros((N,), dtype=torch.int64, device=rows.device)
  counts.scatter_add_(0, rows, torch.ones_like(rows))
  return counts


def _sparse_to_dense_rows(nbrs, out):
  """Dense rows [N, W] of a Sparse-format list's pairs: row i holds the
  partners of i in entry order (deterministic); a row longer than W sets the
  list's overflow bit."""
  from . import partition
  N, W = out.shape
  rows, cols = _sparse_pairs(nbrs, N)
  M = rows.numel()
  out.fill_(N)
  if M == 0:
    return
  order = torch.argsort(rows * M + torch.arange(M, device=rows.device))
  rows, cols = rows[order], cols[order]
  counts = torch.zeros((N,), dtype=torch.int64, device=rows.device)
  counts.scatter_add_(0, rows, torch.ones_like(rows))
  start = torch.cumsum(counts, 0) - counts
  slot = torch.arange(M, device=rows.device) - start[rows]
  ok = slot < W
  out[rows[ok], slot[ok]] = cols[ok].to(torch.int32)
  over = (counts.max() > W).to(torch.uint8) * int(partition.PEC.NEIGHBOR_LIST_OVERFLOW)
  nbrs.error.code.bitwise_or_(over)


def _dense_rows_to_sparse(dense, nbrs):
  """A Sparse-format list from Dense rows: (receiver j, sender i) for each
  row i's entries in order (OrderedSparse: j < i), rows in order, padded
  with N; more pairs than the list's capacity set the overflow bit."""
  from . import partition
  d = dense.idx
  N = d.shape[0]
  cap = nbrs.idx.shape[1]
  keep = d < N
  if nbrs.format is partition.OrderedSparse:
    keep = keep & (d < torch.arange(N, device=d.device, dtype=d.dtype)[:, None])
  flat = keep.reshape(-1)
  pos = torch.cumsum(flat.to(torch.int64), 0) - 1
  ok = flat & (pos < cap)
  senders = torch.arange(N, device=d.device, dtype=torch.int32)[:, None].expand_as(d).reshape(-1)
  nbrs.idx.fill_(N)
  nbrs.idx[0, pos[ok]] = d.reshape(-1)[ok]
  nbrs.idx[1, pos[ok]] = senders[ok]
  over = (flat.sum() > cap).to(torch.uint8) * int(partition.PEC.NEIGHBOR_LIST_OVERFLOW)
  nbrs.error.code.bitwise_or_(over)


class StepEvents:
  """Per-step timing of an eager block: HIP events around each step's list
  update (with the half step when fused) and around its force launch, an
  empty event pair after each (what a pair of markers itself adds), and the
  list's build counter after each update (copied on the device, read after
  the block), so that steps that rebuilt can be told from steps that did not.
  Events are recorded on the current stream, which is the one the kernels
  are launched on."""

  def __init__(self):
    self.update = []   # (start, end, mark_start, mark_end) per step
    self.force = []
    self._builds = []  # device int32 [1] per step

  @staticmethod
  def _ev():
    e = torch.cuda.Event(enable_timing=True)
    e.record()
    return e

  def start(self):
    return self._ev()

  def stop(self, kind, e0):
    e1, m0, m1 = self._ev(), self._ev(), self._ev()
    getattr(self, kind).append((e0, e1, m0, m1))

  def builds_after_update(self, plan):
    self._builds.append(plan.workspace.buf[16:20].view(torch.int32).clone())

  @staticmethod
  def _us(pairs):
    return [1e3 * (a.elapsed_time(b) - c.elapsed_time(d)) for a, b, c, d in pairs]

  def summary(self):
    """Durations in us (marker cost removed); call after a synchronize."""
    upd, frc = self._us(self.update), self._us(self.force)
    out = {'force_us': frc, 'update_us': upd}
    if self._builds:
      b = [int(t.item()) for t in self._builds]
      out['rebuilt'] = [b[0] > 0 if k == 0 else b[k] > b[k - 1] for k in range(len(b))]
    return out


class NVELoop:
  """``n_steps`` of velocity Verlet with the neighbour list maintained on the
  device: the counterpart of the reference's
  ``lax.fori_loop(0, n, lambda i, (s, nbrs): (apply_fn(s, neighbor=nbrs.update(s.position)), ...))``
  (examples/nve_neighbor_list.py:186-201), run in place on the state's
  buffers with no host synchronisation.

  Per step: (1) the list rebuild gated on the skin criterion of the current
  positions, (2) half kick + drift, which also evaluates the skin criterion
  on the new positions for the next step, (3) force pass with the closing
  half kick.  The rebuild decisions are those of the reference's
  ``update`` before ``apply_fn``; overflow is reported in the list's error
  word and checked by the caller once per block, as the reference does.
  ``fused`` (default, Dense lists): (2) runs inside the previous step's
  force launch, from the force and momentum in its registers, writing the
  new positions into a second buffer while the next list update still reads
  the current ones (the two buffers and two gate words alternate by step);
  the block's first half step is one launch before the loop.
  ``fused='update'`` (Dense, cell list): (2) runs in the binning launch of
  (1) instead.  Every form gives the same bits.

  With ``graph`` (default) a block is one HIP graph launch: the first call for
  a given set of buffers and block length runs eagerly and records the
  block's kernels into a graph; later calls on the same buffers replay it.
  Every rebuild decision stays on the device (gated kernels), so the recorded
  sequence is valid for any trajectory.
  """

  def __init__(self, energy_or_force_fn, shift_fn, dt, reorder: bool = True,
               graph: bool = True, fused: bool = True, reorder_every: int = 1000,
               sparse_as_dense: bool = True):
    self.energy = quantity.native_energy_of(energy_or_force_fn)
    self.shift = native_shift(shift_fn)
    self.dt = f32(dt)
    self.half = f32(self.dt / f32(2))
    self.reorder = reorder
    self.graph = graph
    self.fused = fused
    self._alt = None   # second position buffer and gate word of the fused step
    self._bufs = None
    self._bufs_key = None
    self._graphs = {}  # block key -> (graph, the objects the key's addresses belong to)
    self._keep = None
    # Consecutive blocks keep the cell-order working copies: when the state
    # and list a block starts from are exactly what the previous block left
    # (same buffers, untouched since), the block starts from those copies
    # instead of ordering, permuting and relabelling again; the cell order is
    # recomputed once it is reorder_every steps old.
    self.reorder_every = int(reorder_every)
    # Sparse-format lists run as a Dense copy (deterministic row sums, no
    # float atomics), converted back to the Sparse format after each block
    self.sparse_as_dense = bool(sparse_as_dense)
    self._shadow = None  # [key, Dense list, Sparse list signature at the last sync]
    self._cont = None  # (signature after the last block, steps since the order was taken)
    # blocks run eagerly (host-launched), graphs recorded, blocks replayed;
    # continued: blocks that started from the previous block's cell order
    self.stats = {'eager': 0, 'recorded': 0, 'replayed': 0, 'continued': 0}

  def prime(self, state: NVEState, nbrs) -> None:
    """Sets the gate from the skin criterion of the current positions."""
    lib = nat.load()
    plan = nbrs.plan
    spec = plan.spec
    N, dim = state.position.shape
    if spec.always_rebuild:
      return
    nat.check(lib.jmd_neighbor_list_check(
        nat.ptr(state.position), nat.position_stride(state.position),
        nat.ptr(nbrs.reference_position), N, dim,
        int(plan.periodic), nat.f3(plan.side), float(spec.threshold_sq), nat.ptr(plan.gate),
        nat.stream_handle(state.position.device)), 'jmd_neighbor_list_check')

  def _use_reorder(self, state, nbrs) -> bool:
    if state.x64:
      return False
    # per-particle parameter arrays are indexed by the caller's ids
    return bool(self.reorder and nbrs.plan.use_cell_list and state.position.shape[0] > 1
                and not self.energy.reads_particle_data())

  @staticmethod
  def _signature(state, nbrs):
    """Buffers and in-place modification counters of the state and list: a
    write through torch bumps a tensor's version, a list build through
    ``partition`` bumps the plan's epoch."""
    def v(t):
      return (t.data_ptr(), t._version) if isinstance(t, torch.Tensor) else t
    return (v(state.position), v(state.position_hi), v(state.momentum), v(state.force),
            v(state.mass), v(nbrs.idx), v(nbrs.reference_position), v(nbrs.error.code),
            id(nbrs.plan), nbrs.plan.epoch)

  def _continues(self, state, nbrs, reorder) -> bool:
    """Whether a block may start from the previous block's cell-order copies."""
    return bool(reorder and self._cont is not None and self._bufs is not None
                and self._bufs_key == self._buffers_key(state, nbrs)
                and self._cont[1] < self.reorder_every
                and self._cont[0] == self._signature(state, nbrs))

  def _after_block(self, state, nbrs, n_steps, reorder, cont):
    if not reorder:
      self._cont = None
      return
    age = (self._cont[1] if cont else 0) + int(n_steps)
    self._cont = (self._signature(state, nbrs), age)
    if cont:
      self.stats['continued'] += 1

  @staticmethod
  def _key(state, nbrs, n_steps, reorder):
    """What a recorded block depends on: buffer addresses and shapes, the
    list's plan (grid, capacities, workspace) and the block length (and,
    appended by run, whether it starts from the previous block's copies)."""
    def p(t):
      return (t.data_ptr(), tuple(t.shape)) if isinstance(t, torch.Tensor) else t
    plan = nbrs.plan
    return (p(state.position), p(state.position_hi), p(state.momentum), p(state.force),
            p(state.mass), p(nbrs.idx), p(nbrs.reference_position), p(nbrs.error.code),
            id(plan), plan.workspace.buf.data_ptr(), int(nbrs.cell_list_capacity or 0),
            int(nbrs.max_occupancy), int(n_steps), bool(reorder))

  def _key_refs(self, state, nbrs):
    """The objects whose addresses a recorded block holds: kept alive with
    the graph so that no freed buffer's address can match its key again."""
    return (state.position, state.position_hi, state.momentum, state.force, state.mass,
            nbrs.idx, nbrs.reference_position, nbrs.error.code, nbrs.plan)

  def run(self, state: NVEState, nbrs, n_steps: int, events=None) -> NVEState:
    """Advances ``state`` and ``nbrs`` in place by ``n_steps``.  With
    ``reorder`` (lists on a cell grid) the block runs with particles in
    cell order, so that a wavefront's neighbour gathers share cache lines;
    user order, and the list in user ids, are restored at the end (Dense rows
    hold the reference's neighbour sets and Sparse lists its pairs, in the
    cell order of the block; OrderedSparse pairs keep receiver < sender).
    ``events`` (a ``StepEvents``) times each step's list update and force
    launch; it forces the eager path."""
    if getattr(nbrs.plan, 'custom_mask', None) is not None:
      raise NotImplementedError('NVELoop maintains lists without a custom mask; step a '
                                'masked list with nve and neighbor_fn.update')
    if self._via_dense(nbrs):
      shadow = self._shadow_of(nbrs)
      self._run_block(state, shadow, n_steps, events)
      self._shadow_to_list(state, shadow, nbrs)
      return state
    return self._run_block(state, nbrs, n_steps, events)

  # -- Sparse-format lists: the block runs over a Dense copy ------------------
  def _via_dense(self, nbrs) -> bool:
    from . import partition
    spec = nbrs.plan.spec
    return bool(self.sparse_as_dense and partition.is_sparse(nbrs.format)
                and nbrs.plan.use_cell_list and not spec.always_rebuild)

  @staticmethod
  def _list_signature(nbrs):
    def v(t):
      return (t.data_ptr(), t._version)
    return (v(nbrs.idx), v(nbrs.reference_position), v(nbrs.error.code), id(nbrs.plan),
            nbrs.plan.epoch)

  def _shadow_of(self, nbrs):
    """The Dense list a Sparse-format list's block runs over: the same pairs
    in both orders (rows built on the same grid, in the reference's candidate
    order), sharing the list's reference positions, error word and gate.  It
    is filled from the Sparse list when that list changed since the last
    block, and converted back after each block."""
    from . import partition
    plan = nbrs.plan
    key = (id(plan), nbrs.idx.data_ptr(), tuple(nbrs.idx.shape))
    if self._shadow is None or self._shadow[0] != key:
      N, dim = nbrs.reference_position.shape
      dev = nbrs.idx.device
      rows = _sparse_row_counts(nbrs, N)
      longest = int(rows.max().item()) if N else 0
      limit = N - 1 if plan.spec.mask_self else N
      cand = (3 ** dim) * int(nbrs.cell_list_capacity or 0)
      width = max(min(max(2 * longest, longest + 32), limit, cand), 1)
      spec = _dense_spec(plan.spec)
      ncell = 1
      for c in plan.cells_per_side:
        ncell *= int(c)
      lib = nat.load()
      ws = partition._Workspace(dev, lib.jmd_neighbor_workspace_bytes(
          N, dim, ncell, int(nbrs.cell_list_capacity or 0), nat.DENSE))
      plan2 = partition._Plan(ws, plan.gate, torch.zeros((1,), dtype=torch.int32, device=dev),
                              plan.use_cell_list, plan.cells_per_side, plan.cell_size3,
                              plan.side, plan.periodic, plan.dim, spec)
      shadow = partition.NeighborList(
          idx=torch.empty((N, width), dtype=torch.int32, device=dev),
          reference_position=nbrs.reference_position, error=nbrs.error,
          cell_list_capacity=nbrs.cell_list_capacity, max_occupancy=width,
          format=partition.Dense, cell_size=nbrs.cell_size, cell_list_fn=None, update_fn=None,
          plan=plan2)
      self._shadow = [key, shadow, None]
    shadow = self._shadow[1]
    if self._shadow[2] != self._list_signature(nbrs):
      _sparse_to_dense_rows(nbrs, shadow.idx)
      shadow.plan.epoch += 1
    return shadow

  def _shadow_to_list(self, state, shadow, nbrs):
    """Writes the block's list back in the Sparse format: each row's entries
    in row order (OrderedSparse: those below the row's id), the reference's
    sender-major layout, overflow when the pairs exceed the capacity."""
    _dense_rows_to_sparse(shadow, nbrs)
    self._shadow[2] = self._list_signature(nbrs)
    if self._cont is not None:  # the conversion leaves the working copies valid
      self._cont = (self._signature(state, shadow), self._cont[1])

  def _run_block(self, state: NVEState, nbrs, n_steps: int, events=None) -> NVEState:
    if state.position_hi is None:
      state.position_hi = state.position.to(torch.float64)
    reorder = self._use_reorder(state, nbrs)
    cont = self._continues(state, nbrs, reorder)
    if events is not None or not self.graph:
      self._enqueue(state, nbrs, n_steps, reorder, events, cont)
      self.stats['eager'] += 1
      self._after_block(state, nbrs, n_steps, reorder, cont)
      return state
    key = self._key(state, nbrs, n_steps, reorder) + (cont,)
    g = self._graphs.get(key)
    if g is not None:
      g[0].replay()
      self.stats['replayed'] += 1
      self._after_block(state, nbrs, n_steps, reorder, cont)
      return state
    # eager first block: it also loads every kernel before the capture
    self._enqueue(state, nbrs, n_steps, reorder, None, cont)
    self.stats['eager'] += 1
    self._after_block(state, nbrs, n_steps, reorder, cont)
    # the recorded block starts as the next block will (from these copies)
    cont = self._continues(state, nbrs, reorder)
    self._record(state, nbrs, n_steps, reorder, self._key(state, nbrs, n_steps, reorder) + (cont,),
                 cont)
    return state

  def capture(self, state, nbrs, n_steps: int) -> None:
    """Records the block graph for ``n_steps`` on these buffers without
    running it (the state is not advanced), so that a later ``run`` of that
    length is a replay.  Every kernel must have been loaded by an earlier
    eager call (a ``run`` of any length does that)."""
    if state.position_hi is None:
      state.position_hi = state.position.to(torch.float64)
    if not self.graph:
      return
    if self._via_dense(nbrs):
      nbrs = self._shadow_of(nbrs)
    reorder = self._use_reorder(state, nbrs)
    # a block that continues the last one, and (cell order only) one that
    # starts afresh: the order is recomputed every reorder_every steps
    variants = [self._continues(state, nbrs, reorder)]
    if reorder and variants[0]:
      variants.append(False)
    for cont in variants:
      key = self._key(state, nbrs, n_steps, reorder) + (cont,)
      if key not in self._graphs:
        self._record(state, nbrs, n_steps, reorder, key, cont)

  def _record(self, state, nbrs, n_steps, reorder, key, cont=False):
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
      self._enqueue(state, nbrs, n_steps, reorder, None, cont)  # recorded, not run
    # the oldest graph goes first (each entry keeps its buffers alive)
    if len(self._graphs) >= 6:
      self._graphs.pop(next(iter(self._graphs)))
    self._graphs[key] = (g, self._key_refs(state, nbrs))
    self.stats['recorded'] += 1

  def _enqueue(self, state, nbrs, n_steps, reorder, events, cont=False):
    if reorder:
      self._run_reordered(state, nbrs, n_steps, events, cont)
    else:
      self._run_steps(state, nbrs, n_steps, events)

  @staticmethod
  def _buffers_key(state, nbrs):
    R = state.position
    N, dim = R.shape
    return (N, dim, str(R.device), tuple(nbrs.idx.shape),
            isinstance(state.mass, torch.Tensor) and state.mass.ndim > 0)

  def _buffers(self, state, nbrs):
    """Cell-order working copies of the state and list, kept across blocks
    (fixed addresses for graph replay)."""
    R = state.position
    N, dim = R.shape
    dev = R.device
    key = self._buffers_key(state, nbrs)
    if self._bufs_key != key:
      b = dict(
          order=torch.empty((N,), dtype=torch.int32, device=dev),
          inv=torch.empty((N,), dtype=torch.int32, device=dev),
          xh=torch.empty((N, dim), dtype=torch.float64, device=dev),
          x=new_positions(N, dim, dev),
          p=torch.empty((N, dim), dtype=torch.float32, device=dev),
          F=torch.empty((N, dim), dtype=torch.float32, device=dev),
          ref=torch.empty((N, dim), dtype=torch.float32, device=dev),
          idx=torch.empty(tuple(nbrs.idx.shape), dtype=torch.int32, device=dev),
          mass=(torch.empty((N,), dtype=torch.float32, device=dev) if key[-1] else None),
          m32=(torch.empty((N,), dtype=torch.float32, device=dev) if key[-1] else None))
      self._bufs, self._bufs_key = b, key
    return self._bufs

  def _run_steps64(self, state, nbrs, n_steps, events, hook, gate, thr, cps, cs, psides, side,
                   wraps, pot, idx, width, err, ref, occ, ws):
    """The x64 step: gated build on the float32 mirror, float64 half kick and
    drift (mirror and skin test), float64 force with the closing half kick."""
    lib = nat.load()
    plan = nbrs.plan
    spec = plan.spec
    if (spec.format_code & 0xF) != nat.DENSE:
      raise NotImplementedError('float64 states run over Dense neighbour lists')
    R = state.position
    N, dim = R.shape
    dev = R.device
    st = nat.stream_handle(dev)
    stride = nat.position_stride(R)
    mass_s, mass_v = _mass64(state.mass, dev)
    geo = self.energy.geometry
    dsides = nat.d3(geo.side_values(dim))
    dt32 = float(self.dt)
    half = float(self.half)
    for k in range(n_steps):
      if events is not None:
        e0 = events.start()
      nat.check(lib.jmd_neighbor_list_build(
          nat.ptr(R), N, dim, stride, int(plan.periodic), psides, float(spec.cutoff_sq),
          int(plan.use_cell_list), cps, cs, int(nbrs.cell_list_capacity or 0),
          spec.format_code, int(spec.mask_self), idx, int(nbrs.max_occupancy), occ, err, ref,
          gate, ws, st), 'jmd_neighbor_list_build')
      if events is not None:
        events.stop('update', e0)
        events.builds_after_update(plan)
      if hook is not None:
        hook(k, state.position)
      nat.check(lib.jmd_nve_kick_drift_f64(
          nat.ptr(state.position_hi), nat.ptr(state.momentum), nat.ptr(state.force),
          nat.ptr(mass_v), mass_s, N, dim, wraps, side, dt32, half, nat.ptr(R), stride, ref,
          thr, gate, st), 'jmd_nve_kick_drift_f64')
      if events is not None:
        e0 = events.start()
      nat.check(lib.jmd_pair_energy_force_f64(
          nat.ptr(state.position_hi), N, dim, int(geo.periodic), dsides, nat.ctypes.byref(pot),
          idx, width, spec.format_code & 0xF, nat.ptr(state.force), None, nat.ptr(state.momentum), half,
          None, st), 'jmd_pair_energy_force_f64')
      if events is not None:
        events.stop('force', e0)
    return state

  def _alternates(self, R, gate):
    """[R, R2] and [gate, gate2]: the second position buffer (R's layout) and
    the second gate word (zero) of the fused step."""
    key = (tuple(R.shape), str(R.device), nat.position_stride(R), R.data_ptr())
    if self._alt is None or self._alt[0] != key:
      N, dim = R.shape
      if nat.position_stride(R) == 4:
        R2 = nat.padded_positions(N, R.device)
      else:
        R2 = torch.empty_like(R)
      g2 = torch.zeros_like(gate)
      self._alt = (key, R2, g2)
    return [R, self._alt[1]], [gate, self._alt[2]]

  def _run_reordered(self, state, nbrs, n_steps, events, cont=False):
    """One block in cell order.  ``cont``: the working copies already hold
    this state and list in the previous block's cell order (nothing touched
    them since it ended), so the block starts from them."""
    lib = nat.load()
    plan = nbrs.plan
    R = state.position
    N, dim = R.shape
    dev = R.device
    st = nat.stream_handle(dev)
    stride = nat.position_stride(R)
    b = self._buffers(state, nbrs)
    order, inv = b['order'], b['inv']

    def permute(src, dst, row_bytes, src_stride, dst_stride, inverse):
      nat.check(lib.jmd_permute_rows(nat.ptr(src), nat.ptr(dst), nat.ptr(order), N, row_bytes,
                                     src_stride, dst_stride, inverse, st), 'jmd_permute_rows')

    rb = 4 * dim
    xh, x, p, F, ref, idx = b['xh'], b['x'], b['p'], b['F'], b['ref'], b['idx']
    xs = nat.position_stride(x)
    mass = state.mass if b['mass'] is None else b['mass']
    width = nbrs.idx.shape[1]
    fmt = plan.spec.format_code & 0xF
    if not cont:
      nat.check(lib.jmd_cell_order(nat.ptr(R), N, dim, stride, nat.i3(plan.cells_per_side),
                                   nat.f3(plan.cell_size3), nat.ptr(order), nat.ptr(inv),
                                   plan.workspace.ptr, st), 'jmd_cell_order')
      permute(state.position_hi, xh, 8 * dim, 8 * dim, 8 * dim, 0)
      permute(R, x, rb, 4 * stride, 4 * xs, 0)
      permute(state.momentum, p, rb, rb, rb, 0)
      permute(state.force, F, rb, rb, rb, 0)
      if b['mass'] is not None:
        b['m32'].copy_(state.mass.to(device=dev, dtype=torch.float32).reshape(-1))
        permute(b['m32'], b['mass'], 4, 4, 4, 0)
      permute(nbrs.reference_position, ref, rb, rb, rb, 0)
      nat.check(lib.jmd_remap_neighbor_ids(nat.ptr(nbrs.idx), nat.ptr(idx), nat.ptr(inv), N,
                                           fmt, width, st), 'jmd_remap_neighbor_ids')
    nb = dataclasses.replace(nbrs, idx=idx, reference_position=ref)
    sp = NVEState(x, p, F, mass, xh)
    self._run_steps(sp, nb, n_steps, events)
    # back to user order
    permute(xh, state.position_hi, 8 * dim, 8 * dim, 8 * dim, 1)
    permute(x, R, rb, 4 * xs, 4 * stride, 1)
    permute(p, state.momentum, rb, rb, rb, 1)
    permute(F, state.force, rb, rb, rb, 1)
    permute(ref, nbrs.reference_position, rb, rb, rb, 1)
    nat.check(lib.jmd_remap_neighbor_ids(nat.ptr(idx), nat.ptr(nbrs.idx), nat.ptr(order), N,
                                         fmt, width, st), 'jmd_remap_neighbor_ids')
    return state

  def _run_steps(self, state: NVEState, nbrs, n_steps: int, events=None,
                 hook=None) -> NVEState:
    """The per-step kernel sequence; ``hook(k, x)``, when given, runs after
    step k's (gated) list update, when ``x`` (float32, the state's position
    or the step's position buffer) holds x_k and the list is valid for it."""
    lib = nat.load()
    plan = nbrs.plan
    spec = plan.spec
    R = state.position
    N, dim = R.shape
    dev = R.device
    st = nat.stream_handle(dev)
    stride = nat.position_stride(R)
    if state.position_hi is None:
      state.position_hi = R.to(torch.float64)
    mass_s, mass_v = canonicalize_mass(state.mass, dev)
    side = nat.f3(_side(self.shift, dim))
    wraps = int(self.shift.wraps)
    gate = None if spec.always_rebuild else nat.ptr(plan.gate)
    thr = float(spec.threshold_sq)
    cps, cs = nat.i3(plan.cells_per_side), nat.f3(plan.cell_size3)
    psides = nat.f3(plan.side)
    x_hi, x, p, F = (nat.ptr(state.position_hi), nat.ptr(R), nat.ptr(state.momentum),
                     nat.ptr(state.force))
    mv = nat.ptr(mass_v)
    idx, err, ref = nat.ptr(nbrs.idx), nat.ptr(nbrs.error.code), nat.ptr(nbrs.reference_position)
    occ, ws = nat.ptr(plan.occupancy), plan.workspace.ptr
    pot, keep = self.energy._potential(R, {})
    self._keep = (pot, keep)  # device parameter arrays stay alive for graph replays
    pws = nat.ptr(self.energy._workspace(dev, N))
    bfmt = spec.format_code  # the build's format (Sparse-test flag included)
    fmt = bfmt & 0xF
    width = nbrs.idx.shape[1]
    geo = self.energy.geometry
    esides = nat.f3(geo.side_values(dim))
    gcode = self.energy.geometry_code(positions_in_box(self.shift, self.energy, dim))
    if state.x64:
      return self._run_steps64(state, nbrs, n_steps, events, hook, gate, thr, cps, cs, psides,
                               side, wraps, pot, idx, width, err, ref, occ, ws)
    if (self.fused and self.fused != 'update' and gate is not None and fmt == nat.DENSE):
      # The force launch also starts the next step: skin test of the
      # positions it used (the next update's gate word), opening half kick
      # and drift into the other position buffer; the block's first opening
      # half step is one launch before the loop.  Per step: the gated list
      # update on x_k, then the force at x_{k+1}.
      xs, gs = self._alternates(R, plan.gate)
      if n_steps > 0:
        nat.check(lib.jmd_nve_kick_drift(
            x_hi, p, F, mv, mass_s if mass_v is None else 0.0, N, dim, wraps, side,
            float(self.dt), float(self.half), None, x_hi, nat.ptr(xs[1]), stride, p, None, thr,
            None, st), 'jmd_nve_kick_drift')
      # the two calls' arguments for even and odd steps, converted once (the
      # host issues them faster than a small system's steps run)
      xp, gp = [nat.ptr(t) for t in xs], [nat.ptr(t) for t in gs]
      build_args = [(xp[e], N, dim, stride, int(plan.periodic), psides,
                     float(spec.cutoff_sq), int(plan.use_cell_list), cps, cs,
                     int(nbrs.cell_list_capacity or 0), bfmt, int(spec.mask_self), idx,
                     int(nbrs.max_occupancy), occ, err, ref, gp[e], ws, st) for e in (0, 1)]
      pot_ref = nat.ctypes.byref(pot)
      ms = mass_s if mass_v is None else 0.0
      step_args = [[(xp[1 - e], N, dim, stride, gcode, esides, pot_ref, idx, width, F, p,
                     float(self.half), None, x_hi, mv, ms, wraps, side, float(self.dt), ref, thr,
                     gp[1 - e], drift, xp[e], st) for drift in (0, 1)] for e in (0, 1)]
      build, step_fn = lib.jmd_neighbor_list_build, lib.jmd_nve_force_kick_drift
      for k in range(n_steps):
        e = k % 2
        if events is not None:
          e0 = events.start()
        rc = build(*build_args[e])
        if rc:
          nat.check(rc, 'jmd_neighbor_list_build')
        if events is not None:
          events.stop('update', e0)
          events.builds_after_update(plan)
          e0 = events.start()
        if hook is not None:
          hook(k, xs[e])
        rc = step_fn(*step_args[e][int(k + 1 < n_steps)])
        if rc:
          nat.check(rc, 'jmd_nve_force_kick_drift')
        if events is not None:
          events.stop('force', e0)
      if n_steps % 2:  # the state's buffer and gate word hold step n's values
        R.copy_(xs[1])
        gs[0].copy_(gs[1])
        gs[1].zero_()
      return state
    if (self.fused == 'update' and hook is None and gate is not None and plan.use_cell_list
        and fmt == nat.DENSE):
      xs, gs = self._alternates(R, plan.gate)
      for k in range(n_steps):
        xin, xout = xs[k % 2], xs[(k + 1) % 2]
        if events is not None:
          e0 = events.start()
        nat.check(lib.jmd_nve_list_update_kick_drift(
            nat.ptr(xin), N, dim, stride, int(plan.periodic), psides, float(spec.cutoff_sq), 1,
            cps, cs, int(nbrs.cell_list_capacity or 0), bfmt, int(spec.mask_self), idx,
            int(nbrs.max_occupancy), occ, err, ref, nat.ptr(gs[k % 2]), ws, x_hi, p, F, mv,
            mass_s if mass_v is None else 0.0, wraps, side, float(self.dt), float(self.half),
            None, nat.ptr(xout), thr, nat.ptr(gs[(k + 1) % 2]), st),
            'jmd_nve_list_update_kick_drift')
        if events is not None:
          events.stop('update', e0)
          events.builds_after_update(plan)
          e0 = events.start()
        nat.check(lib.jmd_pair_energy_force(
            nat.ptr(xout), N, dim, stride, gcode, esides, nat.ctypes.byref(pot), idx,
            width, fmt, F, None, None, p, float(self.half), None, pws, st),
            'jmd_pair_energy_force')
        if events is not None:
          events.stop('force', e0)
      if n_steps % 2:  # the state's buffer and gate word hold step n's values
        R.copy_(xs[1])
        gs[0].copy_(gs[1])
        gs[1].zero_()
      return state
    for k in range(n_steps):
      if events is not None:
        e0 = events.start()
      nat.check(lib.jmd_neighbor_list_build(
          x, N, dim, stride, int(plan.periodic), psides, float(spec.cutoff_sq), int(plan.use_cell_list),
          cps, cs, int(nbrs.cell_list_capacity or 0), bfmt, int(spec.mask_self), idx,
          int(nbrs.max_occupancy), occ, err, ref, gate, ws, st), 'jmd_neighbor_list_build')
      if events is not None:
        events.stop('update', e0)
        events.builds_after_update(plan)
      if hook is not None:
        hook(k, state.position)
      nat.check(lib.jmd_nve_kick_drift(
          x_hi, p, F, mv, mass_s if mass_v is None else 0.0, N, dim, wraps, side,
          float(self.dt), float(self.half), None, x_hi, x, stride, p, ref, thr, gate, st),
          'jmd_nve_kick_drift')
      if events is not None:
        e0 = events.start()
      nat.check(lib.jmd_pair_energy_force(
          x, N, dim, stride, gcode, esides, nat.ctypes.byref(pot), idx, width, fmt, F,
          None, None, p, float(self.half), None, pws, st), 'jmd_pair_energy_force')
      if events is not None:
        events.stop('force', e0)
    return state


# ---------------------------------------------------------------------------
# NVT with a Nose-Hoover chain (simulate.py:384-705)
# ---------------------------------------------------------------------------


@dataclasses.dataclass
class NVTNoseHooverState:
  """position, momentum, force [N, dim] float32; mass; chain float64 [3, L]
  (positions, momenta, masses of the chain); kinetic_energy float64 [1] (the
  value the chain last saw); position_hi the float64 master positions."""
  position: torch.Tensor
  momentum: torch.Tensor
  force: torch.Tensor
  mass: object
  chain: torch.Tensor
  kinetic_energy: torch.Tensor
  position_hi: Optional[torch.Tensor] = None

  @property
  def velocity(self):
    m = self.mass.reshape(-1, 1) if isinstance(self.mass, torch.Tensor) and self.mass.ndim else self.mass
    return self.momentum / m

  @property
  def x64(self) -> bool:
    return False

  def set(self, **kw):
    return dataclasses.replace(self, **kw)


class _Thermostat:
  """The chain's constants and device scratch (partials, factor)."""

  def __init__(self, dt, kT, chain_length, chain_steps, sy_steps, tau):
    if sy_steps not in (1, 3, 5, 7):
      raise ValueError('sy_steps must be 1, 3, 5 or 7')
    if not 2 <= chain_length <= 16:
      raise ValueError('chain_length must be between 2 and 16')
    self.dt = f32(dt)
    self.kT = f32(kT)
    self.tau = f32(self.dt * 100) if tau is None else f32(tau)
    self.L, self.chain_steps, self.sy_steps = int(chain_length), int(chain_steps), int(sy_steps)
    self._bufs = {}

  def buffers(self, N, device):
    key = (int(N), str(device))
    b = self._bufs.get(key)
    if b is None:
      lib = nat.load()
      b = dict(part=torch.zeros((lib.jmd_kick_kinetic_partials(N),), dtype=torch.float64,
                                device=device),
               scale=torch.ones((1,), dtype=torch.float64, device=device),
               ticket=torch.zeros((1,), dtype=torch.int32, device=device))
      self._bufs = {key: b}
    return b

  def step(self, mode, state, N, dim, kT=None, st=None):
    lib = nat.load()
    b = self.buffers(N, state.momentum.device)
    kT = self.kT if kT is None else f32(kT)
    nat.check(lib.jmd_nhc_step(
        int(mode), nat.ptr(state.chain), self.L, self.chain_steps, self.sy_steps,
        float(self.dt), float(kT), float(self.tau), float(N * dim), nat.ptr(state.kinetic_energy),
        nat.ptr(b['part']), int(b['part'].numel()), nat.ptr(b['scale']),
        st if st is not None else nat.stream_handle(state.momentum.device)), 'jmd_nhc_step')
    return b['scale']

  def kinetic_step(self, mode, state, F, half, st=None, kT=None):
    """kinetic() then step(mode, ...) in one launch (same results)."""
    lib = nat.load()
    N, dim = state.momentum.shape
    dev = state.momentum.device
    b = self.buffers(N, dev)
    kT = self.kT if kT is None else f32(kT)
    mass_s, mass_v = canonicalize_mass(state.mass, dev)
    nat.check(lib.jmd_kick_kinetic_nhc(
        nat.ptr(state.momentum), nat.ptr(F), nat.ptr(mass_v), mass_s if mass_v is None else 0.0,
        N, dim, float(half), nat.ptr(b['part']), nat.ptr(b['ticket']), int(mode),
        nat.ptr(state.chain), self.L, self.chain_steps, self.sy_steps, float(self.dt), float(kT),
        float(self.tau), float(N * dim), nat.ptr(state.kinetic_energy), nat.ptr(b['scale']),
        st if st is not None else nat.stream_handle(dev)), 'jmd_kick_kinetic_nhc')
    return b['scale']

  def kinetic(self, state, F, half, st=None):
    """p += half F (F None: no kick) with the kinetic-energy partials."""
    lib = nat.load()
    N, dim = state.momentum.shape
    dev = state.momentum.device
    b = self.buffers(N, dev)
    mass_s, mass_v = canonicalize_mass(state.mass, dev)
    nat.check(lib.jmd_kick_kinetic(
        nat.ptr(state.momentum), nat.ptr(F), nat.ptr(mass_v), mass_s if mass_v is None else 0.0,
        N, dim, float(half), nat.ptr(b['part']),
        st if st is not None else nat.stream_handle(dev)), 'jmd_kick_kinetic')


_NHC_FINISH, _NHC_START, _NHC_INIT = 1, 2, 4


def nvt_nose_hoover(energy_or_force_fn, shift_fn, dt, kT, chain_length=5, chain_steps=2,
                    sy_steps=3, tau=None, **sim_kwargs):
  """NVT with a Nose-Hoover chain thermostat (simulate.py:562-660):
  ``(init_fn(key, R, mass=1.0, momenta=None, **kw), apply_fn(state, **kw))``.
  Per step: chain masses and a chain half step (one momentum factor),
  velocity Verlet, the kinetic energy, a second chain half step.  The chain
  runs on the device in float64; ``kT=`` in ``apply_fn`` overrides the bath
  temperature as in the reference."""
  energy = quantity.native_energy_of(energy_or_force_fn)
  shift = native_shift(shift_fn)
  th = _Thermostat(dt, kT, chain_length, chain_steps, sy_steps, tau)

  def init_fn(key, R, mass=1.0, momenta=None, **kwargs):
    _kT = kwargs.pop('kT', kT)
    nat.position_stride(R, 'R')
    R = new_positions(R.shape[0], R.shape[1], R.device, R)
    neighbor = kwargs.pop('neighbor', None)
    F = energy.force(R, neighbor, **kwargs)
    if momenta is None:
      P = _momenta(key, R, float(f32(_kT)), mass)
    else:
      P = momenta.to(device=R.device, dtype=torch.float32).clone().contiguous()
    if isinstance(mass, torch.Tensor) and mass.ndim == 2:
      mass = mass.reshape(-1)
    N, dim = R.shape
    state = NVTNoseHooverState(R, P, F, mass,
                               torch.zeros((3, th.L), dtype=torch.float64, device=R.device),
                               torch.zeros((1,), dtype=torch.float64, device=R.device),
                               R.to(torch.float64))
    th.kinetic(state, None, 0.0)
    th.step(_NHC_INIT, state, N, dim, kT=_kT)
    return state

  def apply_fn(state: NVTNoseHooverState, **kwargs):
    refuse_box(kwargs, 'nvt_nose_hoover')
    _kT = kwargs.pop('kT', kT)
    neighbor = kwargs.pop('neighbor', None)
    lib = nat.load()
    R = state.position
    N, dim = R.shape
    dev = R.device
    st = nat.stream_handle(dev)
    new = state.set(momentum=state.momentum.clone(), chain=state.chain.clone(),
                    kinetic_energy=state.kinetic_energy.clone())
    scale = th.step(_NHC_START, new, N, dim, kT=_kT)
    mass_s, mass_v = canonicalize_mass(state.mass, dev)
    x_hi = state.position_hi if state.position_hi is not None else R.to(torch.float64)
    x_hi_new = torch.empty_like(x_hi)
    x_new = new_positions(N, dim, dev)
    half = f32(th.dt / f32(2))
    nat.check(lib.jmd_nvt_kick_drift(
        nat.ptr(x_hi), nat.ptr(new.momentum), nat.ptr(state.force), nat.ptr(mass_v),
        mass_s if mass_v is None else 0.0, N, dim, int(shift.wraps), nat.f3(_side(shift, dim)),
        float(th.dt), float(half), nat.ptr(x_hi_new), nat.ptr(x_new), nat.position_stride(x_new),
        nat.ptr(new.momentum), None, 0.0, None, nat.ptr(scale), st), 'jmd_nvt_kick_drift')
    F = energy.force(x_new, neighbor, in_box=positions_in_box(shift, energy, dim), **kwargs)
    new = new.set(position=x_new, force=F, position_hi=x_hi_new)
    th.kinetic(new, F, half)
    scale = th.step(_NHC_FINISH, new, N, dim, kT=_kT)
    nat.check(lib.jmd_scale_momentum(nat.ptr(new.momentum), N, dim, nat.ptr(scale), st),
              'jmd_scale_momentum')
    return new

  init_fn.thermostat = th
  return init_fn, apply_fn


def nvt_nose_hoover_invariant(energy_fn, state: NVTNoseHooverState, kT, **kwargs) -> float:
  """Extended-system energy of the NVT chain (simulate.py:675-705), float64."""
  energy = quantity.native_energy_of(energy_fn)
  neighbor = kwargs.pop('neighbor', None)
  E, _ = energy.energy_and_force(state.position, neighbor, **kwargs)
  KE = float(quantity.kinetic_energy(momentum=state.momentum.to(torch.float64),
                                     mass=state.mass))
  c = state.chain.cpu().numpy()
  L = c.shape[1]
  xi, p, Q = c[0], c[1], c[2]
  dof = state.position.numel()
  kT = float(f32(kT))
  H = float(E) + KE + p[0] ** 2 / (2 * Q[0]) + dof * kT * xi[0]
  for m in range(1, L):
    H += p[m] ** 2 / (2 * Q[m]) + kT * xi[m]
  return H


class NVTLoop(NVELoop):
  """``n_steps`` of nvt_nose_hoover's apply_fn with the neighbour list
  maintained on the device (the reference's ``fori_loop`` of ``update`` and
  ``apply_fn``, notebooks/lj_benchmark.ipynb), one HIP graph per block like
  ``NVELoop``.  Per step: the gated list update, the half kick and drift
  with the chain's factor, the force, the closing kick with the kinetic
  energy partials, and the chain (closing half step of this step and
  opening half step of the next in one launch)."""

  def __init__(self, energy_or_force_fn, shift_fn, dt, kT, chain_length=5, chain_steps=2,
               sy_steps=3, tau=None, reorder: bool = True, graph: bool = True):
    super().__init__(energy_or_force_fn, shift_fn, dt, reorder=reorder, graph=graph,
                     fused=False)
    self.thermostat = _Thermostat(dt, kT, chain_length, chain_steps, sy_steps, tau)
    self._nvt_state = None

  def run(self, state: NVTNoseHooverState, nbrs, n_steps: int, events=None):
    self._nvt_state = state
    return super().run(state, nbrs, n_steps, events)

  def capture(self, state: NVTNoseHooverState, nbrs, n_steps: int) -> None:
    self._nvt_state = state
    super().capture(state, nbrs, n_steps)

  def _key(self, state, nbrs, n_steps, reorder):
    return super()._key(state, nbrs, n_steps, reorder) + (state.chain.data_ptr(),
                                                          state.kinetic_energy.data_ptr())

  def _key_refs(self, state, nbrs):
    return super()._key_refs(state, nbrs) + (state.chain, state.kinetic_energy)

  def _run_steps(self, state, nbrs, n_steps: int, events=None, hook=None):
    if hook is not None:
      raise NotImplementedError('NVTLoop runs without a step hook')
    lib = nat.load()
    th = self.thermostat
    chain_state = self._nvt_state
    plan = nbrs.plan
    spec = plan.spec
    R = state.position
    N, dim = R.shape
    dev = R.device
    st = nat.stream_handle(dev)
    stride = nat.position_stride(R)
    if state.position_hi is None:
      state.position_hi = R.to(torch.float64)
    mass_s, mass_v = canonicalize_mass(state.mass, dev)
    side = nat.f3(_side(self.shift, dim))
    gate = None if spec.always_rebuild else nat.ptr(plan.gate)
    thr = float(spec.threshold_sq)
    cps, cs = nat.i3(plan.cells_per_side), nat.f3(plan.cell_size3)
    psides = nat.f3(plan.side)
    idx, err, ref = nat.ptr(nbrs.idx), nat.ptr(nbrs.error.code), nat.ptr(nbrs.reference_position)
    occ, ws = nat.ptr(plan.occupancy), plan.workspace.ptr
    pot, keep = self.energy._potential(R, {})
    self._keep = (pot, keep)
    pws = nat.ptr(self.energy._workspace(dev, N))
    bfmt = spec.format_code  # the build's format (Sparse-test flag included)
    fmt = bfmt & 0xF
    width = nbrs.idx.shape[1]
    geo = self.energy.geometry
    esides = nat.f3(geo.side_values(dim))
    gcode = self.energy.geometry_code(positions_in_box(self.shift, self.energy, dim))
    b = th.buffers(N, dev)
    kin = dataclasses.replace(chain_state, momentum=state.momentum, mass=state.mass)
    if n_steps <= 0:
      return state
    th.step(_NHC_START, kin, N, dim, st=st)
    for k in range(n_steps):
      if events is not None:
        e0 = events.start()
      nat.check(lib.jmd_neighbor_list_build(
          nat.ptr(R), N, dim, stride, int(plan.periodic), psides, float(spec.cutoff_sq),
          int(plan.use_cell_list), cps, cs, int(nbrs.cell_list_capacity or 0), bfmt,
          int(spec.mask_self), idx, int(nbrs.max_occupancy), occ, err, ref, gate, ws, st),
          'jmd_neighbor_list_build')
      if events is not None:
        events.stop('update', e0)
        events.builds_after_update(plan)
      nat.check(lib.jmd_nvt_kick_drift(
          nat.ptr(state.position_hi), nat.ptr(state.momentum), nat.ptr(state.force),
          nat.ptr(mass_v), mass_s if mass_v is None else 0.0, N, dim, int(self.shift.wraps),
          side, float(self.dt), float(self.half), nat.ptr(state.position_hi), nat.ptr(R), stride,
          nat.ptr(state.momentum), ref, thr, gate, nat.ptr(b['scale']), st),
          'jmd_nvt_kick_drift')
      if events is not None:
        e0 = events.start()
      nat.check(lib.jmd_pair_energy_force(
          nat.ptr(R), N, dim, stride, gcode, esides, nat.ctypes.byref(pot), idx, width,
          fmt, nat.ptr(state.force), None, None, None, 0.0, None, pws, st),
          'jmd_pair_energy_force')
      if events is not None:
        events.stop('force', e0)
      # closing kick, kinetic energy and the chain in one launch
      th.kinetic_step(_NHC_FINISH | _NHC_START if k + 1 < n_steps else _NHC_FINISH, kin,
                      state.force, self.half, st=st)
    nat.check(lib.jmd_scale_momentum(nat.ptr(state.momentum), N, dim, nat.ptr(b['scale']), st),
              'jmd_scale_momentum')
    return state
