"""Gradients through the pair force and through NVE trajectories on gfx950.

The reference obtains every gradient from ``jax.grad``: of the energy for the
force (quantity.py:58-60), of that force for the Hessian, and through
``lax.fori_loop`` over ``simulate.nve``'s step for a trajectory
(simulate.py:227-317).  Here each is a kernel of its own:

* ``force_vjp(energy_fn, R, neighbor, v)``: the VJP of F(R; sigma, epsilon) =
  -dE/dR, that is J v with J = dF/dR (symmetric), v^T dF/dsigma and
  v^T dF/depsilon, and dE/dsigma, dE/depsilon, from one pass over the list;
* ``ForceFunction``: the force as a ``torch.autograd.Function`` whose backward
  is that kernel (the counterpart of a ``jax.custom_vjp`` on the force);
* ``nve_energy_gradient``: d E(x_T) / d(x_0, p_0, sigma, epsilon) after T
  velocity-Verlet steps, by checkpointed recomputation: the forward run keeps
  the state and list every ``checkpoint_every`` steps; the backward pass
  replays one segment at a time (same kernels, same list decisions, so the
  replay is bit-identical) and runs one fused adjoint kernel per step.

Parameters are differentiated when scalar or per-particle (diameters,
combined as (p_i + p_j) / 2, smap.py:836): a per-particle parameter gets one
cotangent per particle, each a row sum of its own.  Species matrices are not
differentiated.  Sparse and OrderedSparse lists are differentiated through
their Dense rows (``dense_rows``: every neighbour of a particle, in id order),
which hold the same pairs.  r_onset / r_cutoff stay fixed as the reference
fixes them at construction (energy.py:319-320).
"""

from __future__ import annotations

import dataclasses
import time

import numpy as onp
import torch

from . import _native as nat
from . import partition
from . import quantity
from . import simulate

f32 = onp.float32


def _energy(energy_fn):
  e = quantity.native_energy_of(energy_fn)
  return e


def _workspace(device, n, dim):
  lib = nat.load()
  return torch.empty((lib.jmd_adjoint_workspace_bytes(n, dim),), dtype=torch.uint8,
                     device=device)


dense_rows = partition.dense_rows


def _param_count(pot, N) -> int:
  c = int(nat.load().jmd_adjoint_param_count(nat.ctypes.byref(pot), N))
  if c < 0:
    raise NotImplementedError('parameter cotangents are computed for scalar and per-particle '
                              'sigma / epsilon; species matrices are not differentiated')
  return c


def split_params(vec: torch.Tensor, pot, N):
  """(sigma part, epsilon part) of a parameter output: a 0-d tensor for a
  scalar parameter, [N] for a per-particle one."""
  ns = N if pot.sigma.mode == nat.PARAM_PER_PARTICLE else 1
  s, e = vec[:ns], vec[ns:]
  return (s[0] if ns == 1 else s), (e[0] if e.numel() == 1 else e)


def force_vjp(energy_fn, R, neighbor, cotangent, params: bool = True, **dynamic_kwargs):
  """(J v [N, dim] float32, v.dF/d(sigma, eps), dE/d(sigma, eps)) for the
  pair energy ``energy_fn`` over ``neighbor`` (any format).  The parameter
  outputs are float64 vectors [sigma block, epsilon block] (1 value for a
  scalar parameter, N for a per-particle one; ``split_params`` splits them),
  or None without ``params``."""
  energy = _energy(energy_fn)
  lib = nat.load()
  N, dim = R.shape
  dev = R.device
  rows = dense_rows(neighbor.idx, N, neighbor.format)
  v = cotangent.to(device=dev, dtype=torch.float32).contiguous()
  out = torch.empty((N, dim), dtype=torch.float32, device=dev)
  pot, keep = energy._potential(R, dynamic_kwargs)
  count = _param_count(pot, N) if params else 0
  pc = torch.empty((count,), dtype=torch.float64, device=dev) if params else None
  eg = torch.empty((count,), dtype=torch.float64, device=dev) if params else None
  geo = energy.geometry
  if int(geo.periodic) == 2:
    raise NotImplementedError('the adjoint kernels take periodic boxes and free space; '
                              'periodic_general boxes are not supported')
  ws = _workspace(dev, N, dim)
  # float64 positions (an x64 trajectory): row stride code 0
  stride = 0 if R.dtype == torch.float64 else nat.position_stride(R, 'R')
  if stride == 0:
    nat.require_device_tensor(R, 'R', torch.float64)
  nat.check(lib.jmd_pair_force_vjp(
      nat.ptr(R), N, dim, stride, int(geo.periodic),
      nat.f3(geo.side_values(dim)), nat.ctypes.byref(pot), nat.ptr(rows),
      rows.shape[1], nat.DENSE, nat.ptr(v), nat.ptr(out), nat.ptr(pc), nat.ptr(eg),
      nat.ptr(ws), nat.stream_handle(dev)), 'jmd_pair_force_vjp')
  del keep
  return out, pc, eg


class ForceFunction(torch.autograd.Function):
  """F = -dE/dR as a differentiable function of (R, sigma, epsilon); the
  backward pass is ``force_vjp``.  Use through ``differentiable_force``."""

  @staticmethod
  def forward(ctx, R, sigma, epsilon, energy_fn, neighbor):
    kw = {}
    for name, p in (('sigma', sigma), ('epsilon', epsilon)):
      if p is None:
        continue
      if isinstance(p, torch.Tensor) and p.ndim > 0:
        kw[name] = p.detach().to(device=R.device, dtype=torch.float32).contiguous()
      else:
        kw[name] = float(p)
    F = _energy(energy_fn).force(R.detach(), neighbor, **kw)
    ctx.save_for_backward(R.detach())
    ctx.energy_fn, ctx.neighbor, ctx.kw = energy_fn, neighbor, kw
    ctx.has = (sigma is not None, epsilon is not None)
    return F

  @staticmethod
  def backward(ctx, g):
    (R,) = ctx.saved_tensors
    want = ctx.needs_input_grad[1] or ctx.needs_input_grad[2]
    Jv, pc, _ = force_vjp(ctx.energy_fn, R, ctx.neighbor, g, params=want, **ctx.kw)
    gs = ge = None
    if want:
      pot, keep = _energy(ctx.energy_fn)._potential(R, ctx.kw)
      ps, pe = split_params(pc, pot, R.shape[0])
      del keep
      gs = ps.to(torch.float32) if ctx.has[0] else None
      ge = pe.to(torch.float32) if ctx.has[1] else None
    return Jv, gs, ge, None, None


def differentiable_force(energy_fn):
  """force(R, neighbor, sigma=None, epsilon=None) differentiable in R and in
  the parameters passed (0-d tensors, or [N] per-particle tensors)."""
  def force(R, neighbor, sigma=None, epsilon=None):
    return ForceFunction.apply(R, sigma, epsilon, energy_fn, neighbor)
  return force


@dataclasses.dataclass
class TrajectoryGradient:
  energy: float                 # E(x_T), float64
  position: torch.Tensor        # dE/dx_0 [N, dim] float64
  momentum: torch.Tensor        # dE/dp_0 [N, dim] float32
  sigma: object                 # dE/dsigma (through the trajectory and directly):
  epsilon: object               # a float, or float64 [N] for a per-particle parameter
  seconds_forward: float        # forward run with checkpointing
  seconds_backward: float       # replay + adjoint
  checkpoint_bytes: int         # device memory held by checkpoints and a segment


def _snapshot(state, nbrs):
  return dict(x_hi=state.position_hi.clone(), x=state.position.clone(),
              p=state.momentum.clone(), F=state.force.clone(), idx=nbrs.idx.clone(),
              ref=nbrs.reference_position.clone(), err=nbrs.error.code.clone(),
              gate=nbrs.plan.gate.clone())


def _restore(snap, state, nbrs):
  state.position_hi.copy_(snap['x_hi'])
  state.position.copy_(snap['x'])
  state.momentum.copy_(snap['p'])
  state.force.copy_(snap['F'])
  nbrs.idx.copy_(snap['idx'])
  nbrs.reference_position.copy_(snap['ref'])
  nbrs.error.code.copy_(snap['err'])
  nbrs.plan.gate.copy_(snap['gate'])


def nve_energy_gradient(energy_fn, shift_fn, neighbor_fn, R, P, dt, n_steps: int,
                        checkpoint_every: int = 20, mass=1.0,
                        neighbor=None, graph: bool = False) -> TrajectoryGradient:
  """Gradient of E(x_T) after ``n_steps`` velocity-Verlet steps from (R, P)
  with respect to R, P, sigma and epsilon (scalar or per-particle).  A
  float64 R runs the x64 form: float64 trajectory (positions, momenta,
  forces), adjoint products formed from float64 displacements.  Sparse and
  OrderedSparse lists run the trajectory as given (their forces sum with
  float atomics, so a segment's replay matches the forward run to rounding,
  not bit for bit) and each step's adjoint over the list's Dense rows.

  Backward recursion (one fused kernel per step): with adjoints X of x and PH
  of the half-step momentum, X_T = -F(x_T), PH = (dt/m) X_T, and for
  k = T-1 .. 0:  X += c_k J(x_k) PH;  theta += c_k PH^T dF(x_k)/dtheta;
  PH <- PH + (dt/m) X, where c_k = dt for k >= 1 and dt/2 for k = 0 (the
  product at x_k enters steps k and k-1); dE/dp_0 is the last PH.

  With ``graph`` the forward segments and the recording replays run as HIP
  graph replays (one graph per segment length and kind, recorded after the
  first eager segment of that kind); the steps are the same kernels either
  way.  The graphs live for one call, so their recording pays off only for
  long runs of short segments (at 200 steps in segments of 20 it does not).
  """
  energy = _energy(energy_fn)
  if not isinstance(mass, (int, float)):
    raise NotImplementedError('the trajectory adjoint takes a scalar mass')
  lib = nat.load()
  dev = R.device
  N, dim = R.shape
  dt32 = f32(dt)
  m = float(f32(mass))
  nbrs = neighbor if neighbor is not None else neighbor_fn.allocate(
      R.to(torch.float32) if R.dtype == torch.float64 else R)
  x64 = R.dtype == torch.float64
  if x64 and nbrs.format is not partition.Dense:
    raise NotImplementedError('the x64 trajectory adjoint runs over Dense neighbour lists')
  init_fn, _ = simulate.nve(energy, shift_fn, dt)
  state = init_fn(0, R, kT=0.0, mass=mass, momenta=P, neighbor=nbrs)
  loop = simulate.NVELoop(energy, shift_fn, dt, reorder=False, graph=False)
  loop.prime(state, nbrs)
  K = max(1, int(checkpoint_every))
  graphs = {}
  use_graph = bool(graph) and state.position.is_cuda

  def segment(k, hook=None):
    """k steps from the current state (eager the first time a kind runs,
    then recorded; later segments of that kind replay the graph)."""
    key = (k, hook is not None)
    g = graphs.get(key)
    if g is not None:
      g[0].replay()
      return
    loop._run_steps(state, nbrs, k, hook=hook)
    if use_graph:
      g = torch.cuda.CUDAGraph()
      with torch.cuda.graph(g):
        loop._run_steps(state, nbrs, k, hook=hook)  # recorded, not run
      # the parameter arrays the recorded kernels read stay alive with it
      graphs[key] = (g, loop._keep)

  torch.cuda.synchronize(dev)
  t0 = time.perf_counter()
  ckpts = []
  done = 0
  while done < n_steps:
    k = min(K, n_steps - done)
    ckpts.append((done, k, _snapshot(state, nbrs)))
    segment(k)
    done += k
  torch.cuda.synchronize(dev)
  t1 = time.perf_counter()

  # final energy, X_T = -F(x_T), direct parameter gradient
  xT = state.position_hi if x64 else state.position
  E_T, F_T = energy.energy_and_force(xT, nbrs)
  zero = torch.zeros((N, dim), dtype=torch.float32, device=dev)
  _, _, direct = force_vjp(energy, xT, nbrs, zero, params=True)
  X = (-F_T).to(torch.float64).contiguous()
  PH = ((dt32 / m) * X).to(torch.float32).contiguous()
  PH2 = torch.empty_like(PH)
  acc = torch.zeros((lib.jmd_adjoint_param_acc_doubles(N, dim),), dtype=torch.float64,
                    device=dev)
  pot, keep = energy._potential(state.position, {})
  count = _param_count(pot, N)
  geo = energy.geometry
  if int(geo.periodic) == 2:
    raise NotImplementedError('the adjoint kernels take periodic boxes and free space; '
                              'periodic_general boxes are not supported')
  side = nat.f3(geo.side_values(dim))
  st = nat.stream_handle(dev)
  if n_steps == 0:
    PH.zero_()  # dE/dp_0 = 0 when x_T = x_0
  # Segment replays and the adjoint steps run on two streams: the replay of
  # the segment before (in time) the one being differentiated fills the
  # other of two recording sets meanwhile (it depends only on its
  # checkpoint).  Graph replays keep one stream.
  overlap = state.position.is_cuda and not use_graph and len(ckpts) > 1
  n_sets = 2 if overlap else 1

  def new_set():
    if x64:
      xs_ = [torch.empty((N, dim), dtype=torch.float64, device=dev) for _ in range(K)]
    else:
      xs_ = [simulate.new_positions(N, dim, dev) for _ in range(K)]
    return xs_, [torch.empty_like(nbrs.idx) for _ in range(K)]

  sets = [new_set() for _ in range(n_sets)]
  # the backward loop's call arguments, converted once
  dense = nbrs.format is partition.Dense
  set_ptrs = [([nat.ptr(t) for t in xs_], [nat.ptr(t) for t in idxs_]) for xs_, idxs_ in sets]
  dense_width = nbrs.idx.shape[1]
  x_stride = 0 if x64 else nat.position_stride(sets[0][0][0])
  periodic = int(geo.periodic)
  pot_ref = nat.ctypes.byref(pot)
  ph_a, ph_b = nat.ptr(PH), nat.ptr(PH2)
  flip = 0  # PH is ph_a while flip is 0
  X_p, acc_p = nat.ptr(X), nat.ptr(acc)
  c_step, c_first = float(dt32), float(f32(dt32 / f32(2)))
  dt_m = float(dt32) / m
  ckpt_bytes = sum(sum(t.numel() * t.element_size() for t in c[2].values()) for c in ckpts)
  ckpt_bytes += sum(t.numel() * t.element_size() for xs_, idxs_ in sets for t in xs_ + idxs_)

  back_stream = torch.cuda.current_stream(dev) if state.position.is_cuda else None
  replay_stream = nat.replay_stream(dev) if overlap else None
  if overlap:
    replay_stream.wait_stream(back_stream)  # the state's last readers above
  back_done = []

  def replay(i, snap, k):
    """Replays segment i (in backward order) into set i % n_sets."""
    xs_, idxs_ = sets[i % n_sets]

    def hook(q, x):
      xs_[q].copy_(state.position_hi if x64 else x)
      idxs_[q].copy_(nbrs.idx)

    if not overlap:
      _restore(snap, state, nbrs)
      segment(k, hook)
      return None
    with torch.cuda.stream(replay_stream):
      if i >= n_sets:  # the set's previous segment is differentiated
        replay_stream.wait_event(back_done[i - n_sets])
      _restore(snap, state, nbrs)
      segment(k, hook)
      ev = torch.cuda.Event()
      ev.record(replay_stream)
    return ev

  for i, (start, k, snap) in enumerate(reversed(ckpts)):
    ready = replay(i, snap, k)
    if ready is not None:
      back_stream.wait_event(ready)
    x_ptrs, idx_ptrs = set_ptrs[i % n_sets]
    idxs = sets[i % n_sets][1]
    for q in range(k - 1, -1, -1):
      step = start + q
      c = c_step if step >= 1 else c_first
      if dense:
        rows_p, width = idx_ptrs[q], dense_width
      else:
        rows = dense_rows(idxs[q], N, nbrs.format)
        rows_p, width = nat.ptr(rows), rows.shape[1]
      ph, ph2 = (ph_a, ph_b) if flip == 0 else (ph_b, ph_a)
      if x64:
        rc = lib.jmd_nve_adjoint_step_x64(x_ptrs[q], N, dim, periodic, side, pot_ref, rows_p,
                                          width, ph, ph2, X_p, c, dt_m, acc_p, c, st)
      else:
        rc = lib.jmd_nve_adjoint_step(x_ptrs[q], N, dim, x_stride, periodic, side, pot_ref,
                                      rows_p, width, ph, ph2, X_p, c, dt_m, acc_p, c, st)
      if rc:
        nat.check(rc, 'jmd_nve_adjoint_step')
      if step >= 1:
        PH, PH2 = PH2, PH
        flip ^= 1
    if overlap:
      ev = torch.cuda.Event()
      ev.record(back_stream)
      back_done.append(ev)
  if overlap:
    back_stream.wait_stream(replay_stream)
  out = torch.empty((count,), dtype=torch.float64, device=dev)
  nat.check(lib.jmd_adjoint_param_gather(nat.ptr(acc), N, dim, nat.ctypes.byref(pot),
                                         nat.ptr(out), st), 'jmd_adjoint_param_gather')
  torch.cuda.synchronize(dev)
  t2 = time.perf_counter()
  gs, ge = split_params(out + direct, pot, N)
  del keep
  as_out = lambda t: float(t) if t.ndim == 0 else t
  return TrajectoryGradient(float(E_T), X, PH.clone(), as_out(gs), as_out(ge),
                            t1 - t0, t2 - t1, int(ckpt_bytes))
