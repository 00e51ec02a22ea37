"""Ragged ensembles of independent 2-D soft-sphere systems, minimised by FIRE in
one persistent launch.

The reference runs such an ensemble one system at a time: per system
``neighbor_fn, energy_fn = energy.soft_sphere_neighbor_list(...)``,
``init_fn, apply_fn = minimize.fire_descent(energy_fn, shift_fn, ...)`` and a
loop ``nbrs = nbrs.update(R); state = apply_fn(state, neighbor=nbrs)``
(energy.py:200-243, minimize.py:124-226, partition.py:801-1164).  Systems of
uneven size cannot share one jit-compiled vmap.  Here the systems are one CSR
(concatenated particles, per-system offsets) and ``jmd_fire_ensemble_2d`` runs
every system's loop in its own workgroup of one launch.  Per system the
neighbour list follows the reference's Dense rules (capacities from
``allocate``, candidate order, overflow bits), so each system is what the
plain path computes for it.
"""

from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as onp
import torch

from . import _native as nat
from . import partition
from . import space as space_mod
from . import util

f32 = onp.float32

_FILL, _OCCUPANCY, _INIT, _RUN = 0, 1, 2, 3


class _System2D(ctypes.Structure):
  _fields_ = [('offset', ctypes.c_int32), ('n', ctypes.c_int32),
              ('cells_per_side', ctypes.c_int32 * 2), ('cell_size', ctypes.c_float * 2),
              ('fill_cells_per_side', ctypes.c_int32 * 2),
              ('fill_cell_size', ctypes.c_float * 2), ('side', ctypes.c_float * 2),
              ('cell_capacity', ctypes.c_int32), ('max_occupancy', ctypes.c_int32),
              ('cutoff_sq', ctypes.c_float), ('threshold_sq', ctypes.c_float),
              ('param_offset', ctypes.c_int32), ('reserved', ctypes.c_int32)]


class SoftSphereFireEnsemble:
  """FIRE minimisation of many periodic 2-D soft-sphere systems.

  positions: sequence of [N_s, 2] arrays; boxes: per-system side lengths;
  species: sequence of [N_s] integer arrays; sigma: a scalar, a species
  matrix shared by all systems, or one species matrix per system
  ([S, n_species, n_species]); epsilon, alpha: scalars.  Each system's list
  cutoff is its largest sigma plus dr_threshold (energy.py:207-219).  FIRE
  constants as in ``minimize.fire_descent`` (static_cast to float16).
  """

  def __init__(self, positions: Sequence, boxes: Sequence, species: Optional[Sequence] = None,
               sigma=1.0, epsilon=1.0, alpha=2.0, dr_threshold=0.2,
               capacity_multiplier: float = 1.25, dt_start=0.1, dt_max=0.4, n_min=5,
               f_inc=1.1, f_dec=0.5, alpha_start=0.1, f_alpha=0.99, device='cuda'):
    lib = nat.load()
    self.device = torch.device(device)
    S = len(positions)
    if len(boxes) != S or (species is not None and len(species) != S):
      raise ValueError('positions, boxes and species need one entry per system')
    ns = [int(onp.asarray(p).shape[0]) for p in positions]
    self.sizes = ns
    self.offsets = onp.concatenate([[0], onp.cumsum(ns)]).astype(onp.int64)
    self.n_total = int(self.offsets[-1])
    if self.n_total >= 2 ** 31 - 1:
      raise ValueError('too many particles for int32 ids')
    R = onp.concatenate([onp.asarray(p, f32).reshape(-1, 2) for p in positions])
    sig = onp.asarray(sigma, f32)
    per_system = sig.ndim == 3
    if per_system and sig.shape[0] != S:
      raise ValueError('per-system sigma needs one species matrix per system')
    sig_max = [float(f32(onp.max(sig[s] if per_system else sig))) for s in range(S)]
    self.cutoff_sigma = max(sig_max)
    specs = []
    for s, L in enumerate(boxes):
      disp, _ = space_mod.periodic(f32(L))
      specs.append(partition.NeighborListSpec(disp, f32(L), sig_max[s], dr_threshold,
                                              capacity_multiplier, False, True,
                                              partition.Dense))
    self.specs = specs
    self.capacity_multiplier = capacity_multiplier
    recs = (_System2D * S)()
    max_cells = lib.jmd_fire_ensemble_max_cells()
    block = int(sig[0].size) if per_system else 0
    for s, (L, sp) in enumerate(zip(boxes, specs)):
      if not sp.use_cell_list():
        raise NotImplementedError('every system needs a cell list (cutoff < box / 3)')
      _, cs, cps, nc = partition.cell_dimensions(2, f32(L), sp.cutoff)
      _, cs2, cps2, nc2 = partition.cell_dimensions(2, f32(L), cs)
      if max(nc, nc2) > max_cells:
        raise NotImplementedError(f'system {s}: {max(nc, nc2)} cells exceed {max_cells}')
      r = recs[s]
      r.offset, r.n = int(self.offsets[s]), ns[s]
      cps_l, cs_l = partition._per_dim(cps, 2), partition._per_dim(cs, 2)
      cps2_l, cs2_l = partition._per_dim(cps2, 2), partition._per_dim(cs2, 2)
      for d in range(2):
        r.cells_per_side[d] = int(cps_l[d])
        r.cell_size[d] = float(cs_l[d])
        r.fill_cells_per_side[d] = int(cps2_l[d])
        r.fill_cell_size[d] = float(cs2_l[d])
        r.side[d] = float(f32(L))
      r.cutoff_sq = float(sp.cutoff_sq)
      r.threshold_sq = float(sp.threshold_sq)
      r.param_offset = s * block
    self._recs = recs
    dev = self.device
    order = onp.argsort(-onp.asarray(ns), kind='stable').astype(onp.int32)  # largest first
    self.order = torch.as_tensor(order, device=dev)
    self.position = torch.as_tensor(R, device=dev).contiguous()
    self.position_hi = self.position.to(torch.float64)
    self.momentum = torch.zeros_like(self.position)
    self.force = torch.zeros_like(self.position)
    self.reference_position = torch.empty_like(self.position)
    self.scratch = torch.empty((3 * self.n_total,), dtype=torch.int32, device=dev)
    self.error = torch.zeros((S,), dtype=torch.uint8, device=dev)
    self.builds = torch.zeros((S,), dtype=torch.int32, device=dev)
    self.energy = torch.zeros((S,), dtype=torch.float64, device=dev)
    self.fire_state = torch.zeros((S, 2), dtype=torch.float32, device=dev)
    self.fire_n_pos = torch.zeros((S,), dtype=torch.int32, device=dev)
    self._stat = torch.zeros((S,), dtype=torch.int32, device=dev)
    cast = util.static_cast(dt_start, dt_max, n_min, f_inc, f_dec, alpha_start, f_alpha)
    self._fire = (ctypes.c_float * 7)(*[float(f32(c)) for c in cast])
    if species is None:
      species = [onp.zeros(n, onp.int32) for n in ns]
    sp_all = onp.concatenate([onp.asarray(x, onp.int32).reshape(-1) for x in species])
    self.species = torch.as_tensor(sp_all, device=dev)
    self._sig = torch.as_tensor(onp.ascontiguousarray(sig).reshape(-1), device=dev)
    pot = nat.JmdPairPotential()
    pot.kind = nat.SOFT_SPHERE
    if sig.ndim == 0:
      pot.sigma = nat.JmdParam(nat.PARAM_SCALAR, float(sig), None)
      pot.n_species = 0
    else:
      pot.sigma = nat.JmdParam(nat.PARAM_SPECIES, 0.0, ctypes.c_void_p(self._sig.data_ptr()))
      pot.n_species = int(sig.shape[-1])
    pot.epsilon = nat.JmdParam(nat.PARAM_SCALAR, float(f32(epsilon)), None)
    pot.alpha = nat.JmdParam(nat.PARAM_SCALAR, float(f32(alpha)), None)
    pot.r_onset = 0.0
    pot.r_cutoff = self.cutoff_sigma
    pot.species = ctypes.c_void_p(self.species.data_ptr())
    self._pot = pot
    self._systems = torch.empty((ctypes.sizeof(recs),), dtype=torch.uint8, device=dev)
    self.idx = None
    self.width = 0
    self.box_adjoint = None

  # -- native call -------------------------------------------------------------
  def _launch(self, mode, n_steps=0, tape=None):
    lib = nat.load()
    tp = tape or {}
    nat.check(lib.jmd_fire_ensemble_2d(
        nat.ptr(self._systems), nat.ptr(self.order), len(self.sizes), self.n_total, mode,
        n_steps, nat.ptr(self.position_hi), nat.ptr(self.position), nat.ptr(self.momentum),
        nat.ptr(self.force), nat.ptr(self.reference_position), nat.ptr(self.idx),
        self.width, nat.ptr(self.scratch), ctypes.byref(self._pot), self._fire,
        nat.ptr(self.fire_state), nat.ptr(self.fire_n_pos), nat.ptr(self.error),
        nat.ptr(self._stat), nat.ptr(self.energy), nat.ptr(self.builds),
        nat.ptr(tp.get('x')), nat.ptr(tp.get('F')), nat.ptr(tp.get('q')),
        nat.ptr(tp.get('idx')), nat.ptr(tp.get('sc')), nat.stream_handle(self.device)),
        'jmd_fire_ensemble_2d')

  def _upload(self):
    host = torch.frombuffer(bytearray(bytes(self._recs)), dtype=torch.uint8)
    self._systems.copy_(host)

  # -- public --------------------------------------------------------------------
  def allocate(self, extra_capacity: int = 0):
    """Capacities as ``neighbor_fn.allocate`` computes them per system
    (partition.py:243-250,1092-1104), then the list, F(R), P = 0 and the FIRE
    scalars (minimize.py:124-170).  Reads the measured sizes on the host, as
    the reference's allocate does."""
    self._upload()
    self._launch(_FILL)
    fill = self._stat.cpu().numpy()
    cm = self.capacity_multiplier
    for s in range(len(self.sizes)):
      self._recs[s].cell_capacity = int(onp.int32(fill[s]) * cm) + extra_capacity
    self._upload()
    self._launch(_OCCUPANCY)
    occ = self._stat.cpu().numpy()
    width = 1
    for s in range(len(self.sizes)):
      m = int(f32(occ[s]) * f32(cm) + f32(extra_capacity))
      m = min(m, 9 * int(self._recs[s].cell_capacity), max(self.sizes[s] - 1, 0))
      self._recs[s].max_occupancy = max(m, 0)
      width = max(width, self._recs[s].max_occupancy)
    self._upload()
    self.width = -(-width // 4) * 4  # row stride: whole 16-byte groups
    self.idx = torch.empty((self.n_total, self.width), dtype=torch.int32, device=self.device)
    self.error.zero_()
    self._launch(_INIT)
    return self

  def run(self, n_steps: int):
    """n_steps of FIRE for every system (one launch)."""
    if self.idx is None:
      raise RuntimeError('allocate() first')
    self._launch(_RUN, int(n_steps))
    return self

  @property
  def did_buffer_overflow(self) -> torch.Tensor:
    return self.error & (partition.PEC.NEIGHBOR_LIST_OVERFLOW | partition.PEC.CELL_LIST_OVERFLOW)

  def system_slice(self, s: int) -> slice:
    return slice(int(self.offsets[s]), int(self.offsets[s + 1]))

  def system_neighbors(self, s: int) -> torch.Tensor:
    """System s's Dense list in local ids (padding N_s), max_occupancy columns."""
    sl = self.system_slice(s)
    rows = self.idx[sl, :self._recs[s].max_occupancy]
    off = int(self.offsets[s])
    n = self.sizes[s]
    return torch.where(rows >= self.n_total, torch.full_like(rows, n), rows - off)

  def capacities(self, s: int):
    r = self._recs[s]
    return int(r.cell_capacity), int(r.max_occupancy)

  # -- gradients -------------------------------------------------------------
  _STATE = ('position_hi', 'position', 'momentum', 'force', 'reference_position', 'fire_state',
            'fire_n_pos', 'idx', 'error')

  def _snapshot(self):
    return {k: getattr(self, k).clone() for k in self._STATE}

  def _restore(self, snap):
    for k in self._STATE:
      getattr(self, k).copy_(snap[k])

  def energy_gradient(self, n_steps: int, dsigma, checkpoint_every: int = 25,
                      scale: Optional[float] = None, dbox=None):
    """FIRE for n_steps from the allocated state, then the gradient of
    scale * sum_s E_s(x_T) (default scale 1/S: the mean final energy) with
    respect to each system's parameter theta_s, where the species matrix
    depends on theta_s with coefficients dsigma (a matrix shared by all
    systems, or one per system).  The state and list are kept every
    ``checkpoint_every`` steps; each segment is replayed once with a tape and
    run backward (one force-Jacobian product per step).  Returns
    (E_T float64 [S], dL/dtheta float64 [S], dL/dx_0 float64 [N_total, 2]).
    The ensemble is left in its final state.

    ``dbox`` (None, a scalar or [S]): d side / d theta_s of each box.  With
    it the box follows theta_s as in the meta-optimisation notebook
    (``periodic_general(L(d) * eye(2), fractional_coordinates=False)`` with x_0
    held in real coordinates): theta's derivative gains
    dbox_s * d loss / d side_s, taken through the minimum images of every
    force and of E(x_T) and through the folds of the shift.
    ``self.box_adjoint`` then holds d loss / d side_s [S].  None: the boxes
    are constants (the derivative at fixed box)."""
    if self.idx is None:
      raise RuntimeError('allocate() first')
    lib = nat.load()
    S, dev = len(self.sizes), self.device
    scale = 1.0 / S if scale is None else float(scale)
    ds = onp.asarray(dsigma, f32)
    if ds.ndim == 2:
      ds = onp.broadcast_to(ds, (S,) + ds.shape) if self._sig.numel() > ds.size else ds
    dsig = torch.as_tensor(onp.ascontiguousarray(ds).reshape(-1), device=dev)
    if dsig.numel() != self._sig.numel():
      raise ValueError('dsigma must be laid out like sigma')
    box_adj = None
    if dbox is not None:
      db = onp.broadcast_to(onp.asarray(dbox, onp.float64), (S,))
      box_adj = torch.zeros((S,), dtype=torch.float64, device=dev)
    K = max(1, int(checkpoint_every))
    snaps = []
    done = 0
    while done < n_steps:
      k = min(K, n_steps - done)
      snaps.append((k, self._snapshot()))
      self.run(k)
      done += k
    E_T = self.energy.clone()
    final = self._snapshot()
    final_builds = self.builds.clone()
    n2 = (self.n_total, 2)
    X = torch.zeros(n2, dtype=torch.float64, device=dev)
    P = torch.zeros(n2, dtype=torch.float64, device=dev)
    W = torch.zeros(n2, dtype=torch.float32, device=dev)
    V = torch.zeros(n2, dtype=torch.float32, device=dev)
    theta = torch.zeros((S,), dtype=torch.float64, device=dev)

    def new_tape():
      return dict(x=torch.empty((K + 1,) + n2, dtype=torch.float32, device=dev),
                  F=torch.empty((K + 1,) + n2, dtype=torch.float32, device=dev),
                  q=torch.empty((K,) + n2, dtype=torch.float32, device=dev),
                  idx=torch.empty((K + 1, self.n_total, self.width), dtype=torch.int32,
                                  device=dev),
                  sc=torch.empty((K, S, 6), dtype=torch.float32, device=dev))

    # Segment replays (the taped forward runs) go to a second stream, so the
    # replay of the segment before runs beside the backward kernel of the
    # segment after (each launch fills one CU per system); two tapes
    # alternate, events order their reuse.
    overlap = len(snaps) > 1
    tapes = [new_tape() for _ in range(2 if overlap else 1)]
    tape = tapes[0]

    def back(mode, k):
      nat.check(lib.jmd_fire_ensemble_adjoint_2d(
          nat.ptr(self._systems), S, self.n_total, mode, k, nat.ptr(self.position),
          nat.ptr(self.force), nat.ptr(self.idx), self.width, ctypes.byref(self._pot),
          nat.ptr(dsig), nat.ptr(tape['x']), nat.ptr(tape['F']), nat.ptr(tape['q']),
          nat.ptr(tape['idx']), nat.ptr(tape['sc']), nat.ptr(X), nat.ptr(P), nat.ptr(W),
          nat.ptr(V), nat.ptr(theta), nat.ptr(box_adj), float(scale), nat.stream_handle(dev)),
          'jmd_fire_ensemble_adjoint_2d')

    back(0, 0)
    back_stream = torch.cuda.current_stream(dev)
    replay_stream = nat.replay_stream(dev) if overlap else back_stream
    if overlap:
      replay_stream.wait_stream(back_stream)  # the state's last readers above
    back_done = []
    for i, (k, snap) in enumerate(reversed(snaps)):
      tape = tapes[i % len(tapes)]
      with torch.cuda.stream(replay_stream):
        if overlap and i >= len(tapes):  # this tape's previous segment is done
          replay_stream.wait_event(back_done[i - len(tapes)])
        self._restore(snap)
        self._launch(_RUN, k, tape)
        ready = torch.cuda.Event()
        ready.record(replay_stream)
      back_stream.wait_event(ready)
      back(1, k)
      done_ev = torch.cuda.Event()
      done_ev.record(back_stream)
      back_done.append(done_ev)
    back_stream.wait_stream(replay_stream)  # the state, restored below
    if snaps:
      back(2, 0)
    self._restore(final)
    self.builds.copy_(final_builds)
    self.energy.copy_(E_T)
    self.box_adjoint = box_adj
    if box_adj is not None:
      theta = theta + torch.as_tensor(db.copy(), device=dev) * box_adj
    return E_T, theta, X
