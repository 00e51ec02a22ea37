#!/usr/bin/env python
"""Benchmark: LJ NVE at N = 262,144 on one MI355X (atom-steps/s).

Workload: 64^3 simple-cubic lattice at rho = 0.8442 (L rounded to float32),
jitter N(0, 0.05), momenta N(0, kT=1) with the mean removed; LJ sigma = eps =
1 with the C1 switch between r_on = 2.0 and r_c = 2.5, Dense neighbour list
with skin 0.5 rebuilt on the skin criterion, dt = 0.005, float32 pair
arithmetic, float64 master positions and energy reductions.  One step = skin
rebuild (gated on the device) + force pass with the closing half kick, which
also runs the next step's skin test, half kick and drift.  Steps run in
blocks of 100, each block one HIP graph replay, with a host overflow check per
block; an overflowing block is redone from its saved start state after
re-allocating the list (the reference's protocol,
examples/nve_neighbor_list.py:186-201).

Setup, before the warm-up: the autotuner (jax_md_amd.autotune) times the
block loop on copies of the initial state for threads per row, cell size
and row capacity; the line records its choice.  Before the timed region a
HIP graph is recorded for every block the timed loop uses, so the timed
blocks are graph replays (the line reports the blocks' kinds).  The force
kernel's duration, the list update with and without a rebuild, and the
block-boundary cost are measured with HIP events in one more (eager)
100-step block after the timed region, on the stream the kernels run on,
with the device kept behind the host.  HBM traffic per launch comes from
the committed counter summary of the same kernel and workload
(profiles/k_pair_dense_step_lj_n262144_pmc.json), scaled to this run's mean
row length.

Prints one JSON line on stdout; progress goes to stderr.
"""

import argparse
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'atom-steps/s LJ NVE N=262k, plain vs batched/adjoint; error vs FP64 oracle'
HBM_PEAK_GBS = 8000.0


def without_collector(fn):
  """fn() with Python's cyclic garbage collector paused (after one pass):
  a collection over the garbage of earlier sections landing inside a
  host-driven timed loop would dominate it (timeit does the same)."""
  gc.collect()
  gc.disable()
  try:
    return fn()
  finally:
    gc.enable()


def log(*a):
  print(*a, file=sys.stderr, flush=True)


def step_breakdown(jmd, loop, state, nbrs, steps=100):
  """Per-step kernel times inside the step loop: one eager block of `steps`
  steps (the kernels of the timed blocks, launched from the host instead of a
  graph replay) with HIP events around each step's list update (binning +
  half step, cell sort, list build; gated) and around each force launch, on
  the stream the kernels run on, less what an empty event pair adds.  A spin
  kernel queued first keeps the device behind the host, so no interval holds
  a host gap.  Steps that rebuilt the list are told apart by the list's build
  counter."""
  import torch
  ev = jmd.simulate.StepEvents()
  loop.prime(state, nbrs)
  try:
    torch.cuda._sleep(int(5e8))  # ~0.2 s of device time queued ahead
  except Exception:
    pass
  b0 = torch.cuda.Event(enable_timing=True)
  b1 = torch.cuda.Event(enable_timing=True)
  b0.record()
  loop.run(state, nbrs, steps, ev)
  b1.record()
  torch.cuda.synchronize()
  t = ev.summary()
  rebuilt = t['rebuilt']
  n_open = sum(rebuilt)
  closed_upd = [u for u, r in zip(t['update_us'], rebuilt) if not r]
  open_upd = [u for u, r in zip(t['update_us'], rebuilt) if r]

  def mean(v):
    return sum(v) / len(v) if v else None

  upd_closed, upd_open = mean(closed_upd), mean(open_upd)
  block_us = 1e3 * b0.elapsed_time(b1)
  in_steps = sum(t['update_us']) + sum(t['force_us'])
  return {
      'steps': steps, 'rebuilds': n_open,
      'force_us': mean(t['force_us']),
      'update_closed_us': upd_closed,
      'update_open_us': upd_open,
      'rebuild_us': (upd_open - upd_closed) if (upd_open is not None and upd_closed) else None,
      # what the block costs outside the steps' kernels: the cell-order
      # permutation of state and list at its start and end, event markers
      'block_boundary_us': block_us - in_steps,
      'block_us': block_us,
      'note': 'eager block with HIP events; update = binning, cell sort and list build, '
              'gated on the skin test (the force launch runs the half step and the test); '
              'rebuild_us = open update less closed update',
  }


# The force launch of the block loop (jmd_nve_force_kick_drift): force, the
# closing half kick and the next step's start.  Algorithmic bytes per atom:
# the Dense row, 4 B per valid entry plus the padding entry that ends it (4);
# float32 position read 16 (padded row); force write 12; momentum read+write
# 24; parameters 4; float64 position read+write 48; next float32 position
# write 12; reference position read 12 (skin test) -> 4 Kbar + 132 B, the
# whole step's per-atom bytes apart from the gated list update.
FORCE_KERNEL = ('k_pair_dense STEP (force + closing half kick + skin test + next half kick '
                'and drift)')
FORCE_STATE_BYTES = 132.0


def counter_traffic(N, kbar):
  """HBM bytes per force launch from the committed counter summary, scaled to
  this run's mean row length: the summary's measured traffic over its own
  algorithmic bytes, times this run's algorithmic bytes.  None without it."""
  path = os.path.join(ROOT, 'profiles', f'k_pair_dense_step_lj_n{N}_pmc.json')
  try:
    with open(path) as f:
      d = json.load(f)
    t = float(d['traffic_bytes_per_launch'])
    k_prof = float(d['kbar'])
  except (OSError, KeyError, ValueError):
    return None, None
  ratio = t / (N * (4.0 * k_prof + FORCE_STATE_BYTES))
  return ratio * N * (4.0 * kbar + FORCE_STATE_BYTES), {
      'measured_bytes': t, 'measured_kbar': k_prof, 'ratio_to_algorithmic': ratio,
      'source': os.path.relpath(path, ROOT),
      'note': 'FETCH_SIZE + WRITE_SIZE of the force kernel (copy-calibrated), measured at '
              'measured_kbar; roofline.traffic = ratio_to_algorithmic x this run\'s '
              'algorithmic bytes'}


def c1_drift(jmd, O, n_steps=1000, block=100):
  """Total-energy drift of the small LJ NVE workload (16^3 lattice, N = 4096,
  the same density, jitter, kT, cutoff, skin and dt) over 1000 steps,
  through the same step loop; energies in float64 at every block end."""
  import numpy as np
  import torch
  dev = 'cuda:0'
  R, P, L = O.lj_lattice(16, seed=0)
  disp, shift = jmd.space.periodic(L)
  nfn, efn = jmd.energy.lennard_jones_neighbor_list(disp, L, format=jmd.partition.Dense)
  Rt = torch.as_tensor(R, device=dev)
  nb = nfn.allocate(Rt, extra_capacity=16)
  init_fn, _ = jmd.simulate.nve(efn, shift, 0.005)
  st = init_fn(0, Rt, kT=1.0, momenta=torch.as_tensor(P, device=dev), neighbor=nb)
  loop = jmd.simulate.NVELoop(efn, shift, 0.005)

  def etot():
    E, _ = efn.energy_and_force(st.position, nb)
    return float(E) + float(jmd.quantity.kinetic_energy(momentum=st.momentum))

  e0 = etot()
  series = [e0]
  loop.prime(st, nb)
  done = 0
  while done < n_steps:
    loop.run(st, nb, block)
    if bool(nb.did_buffer_overflow):
      raise RuntimeError('C1 drift run overflowed its list')
    done += block
    series.append(etot())
  dev_max = max(abs(e - e0) for e in series) / abs(e0)
  return {'workload': 'LJ NVE N=4096 (16^3 lattice, rho=0.8442, jitter 0.05, kT=1), Dense, '
                      'r_c=2.5 skin=0.5, dt=0.005, 1000 steps from t=0',
          'E_total_0': e0, 'E_total_1000': series[-1],
          'drift_rel': abs(series[-1] - e0) / abs(e0), 'max_dev_rel': dev_max,
          'every_100_steps': series}


def adjoint_section(jmd, O, check=True):
  """Adjoint workload: LJ N = 32,768 (32^3 lattice), gradient of E(x_T) after
  200 NVE steps (dt = 0.005) with respect to x_0, sigma and epsilon, with
  checkpoints every 20 steps; time against the same forward run.  Run in
  float32 (the workload's precision) and in the x64 form; both are compared
  with central differences of the C FP64 restatement."""
  import numpy as np
  import torch
  dev = 'cuda:0'
  R, P, L = O.lj_lattice(32, seed=1)
  disp, shift = jmd.space.periodic(L)
  nfn, efn = jmd.energy.lennard_jones_neighbor_list(disp, L, format=jmd.partition.Dense)
  T, K, dt = 200, 20, 0.005

  def run(dtype):
    Rt = torch.as_tensor(R, device=dev).to(dtype)
    Pt = torch.as_tensor(P, device=dev).to(dtype)
    R32 = Rt.to(torch.float32)
    init_fn, _ = jmd.simulate.nve(efn, shift, dt)
    loop = jmd.simulate.NVELoop(efn, shift, dt, reorder=False, graph=False)
    nb = nfn.allocate(R32)
    st = init_fn(0, Rt, kT=0.0, momenta=Pt, neighbor=nb)
    loop.prime(st, nb)
    loop._run_steps(st, nb, 20)  # warm
    nb = nfn.allocate(R32)
    st = init_fn(0, Rt, kT=0.0, momenta=Pt, neighbor=nb)
    loop.prime(st, nb)
    torch.cuda.synchronize()
    t = time.perf_counter()
    loop._run_steps(st, nb, T)
    efn.energy_and_force(st.position_hi if st.x64 else st.position, nb)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t
    # warm: two segments, so the second recording set and the replay stream
    # exist before the timed call (allocations the caching allocator keeps)
    jmd.adjoint.nve_energy_gradient(efn, shift, nfn, Rt, Pt, dt, 2 * K, checkpoint_every=K)
    torch.cuda.synchronize()
    t = time.perf_counter()
    g = jmd.adjoint.nve_energy_gradient(efn, shift, nfn, Rt, Pt, dt, T, checkpoint_every=K)
    torch.cuda.synchronize()
    return g, t_fwd, time.perf_counter() - t

  g32, f32_fwd, f32_grad = run(torch.float32)
  g64, f64_fwd, f64_grad = run(torch.float64)
  out = {'workload': 'LJ N=32768 (32^3 lattice, rho=0.8442, jitter 0.05, kT=1), 200 NVE '
                     'steps dt=0.005, dE(x_T)/d(x_0, sigma, epsilon), checkpoints every 20',
         'forward_s': f32_fwd, 'gradient_s': f32_grad, 'ratio_vs_forward': f32_grad / f32_fwd,
         'checkpoint_bytes': g32.checkpoint_bytes, 'energy_T': g32.energy,
         'x64': {'forward_s': f64_fwd, 'gradient_s': f64_grad,
                 'ratio_vs_forward': f64_grad / f64_fwd, 'energy_T': g64.energy}}
  if check:
    try:
      from oracle import c_oracle
      # E(x_T) is C1 in the parameters (the smoothing is C1 at r_on, r_c), so
      # central differences converge at first order: on this run h = 1e-4 is
      # 7% off the limit, h = 1e-7 within 3e-5
      h = 1e-7
      rng = np.random.default_rng(5)
      u = rng.normal(size=R.shape)
      u /= np.linalg.norm(u)
      dt32 = float(np.float32(dt))

      def e_after(R0, sigma=1.0, eps=1.0):
        sysc = c_oracle.LJSystem(R0, float(L), sigma=sigma, eps=eps)
        sysc.nve(P.astype(np.float64), dt32, T)
        return sysc.forces()[0]

      R64 = R.astype(np.float64)
      fd_x = (e_after(R64 + h * u) - e_after(R64 - h * u)) / (2 * h)
      fd_s = (e_after(R64, sigma=1 + h) - e_after(R64, sigma=1 - h)) / (2 * h)
      fd_e = (e_after(R64, eps=1 + h) - e_after(R64, eps=1 - h)) / (2 * h)

      def compare(g):
        ad_x = float(np.sum(g.position.cpu().numpy() * u))
        gn = float(np.linalg.norm(g.position.cpu().numpy()))
        return {'direction_x': {'adjoint': ad_x, 'fd': fd_x,
                                'rel_err': abs(ad_x - fd_x) / max(abs(fd_x), 1e-30),
                                'err_over_grad_norm': abs(ad_x - fd_x) / gn},
                'sigma': {'adjoint': g.sigma, 'fd': fd_s,
                          'rel_err': abs(g.sigma - fd_s) / abs(fd_s)},
                'epsilon': {'adjoint': g.epsilon, 'fd': fd_e,
                            'rel_err': abs(g.epsilon - fd_e) / abs(fd_e)}}

      out['fd_check'] = dict(compare(g32), h=h, reference='C FP64 restatement '
                             '(oracle/md_oracle.c), central differences')
      out['x64']['fd_check'] = compare(g64)
    except Exception as e:  # the checker needs gcc-built oracle/libmd_oracle.so
      out['fd_check'] = f'skipped: {e}'
  log(f'adjoint: forward {f32_fwd * 1e3:.1f} ms, gradient {f32_grad * 1e3:.1f} ms, '
      f'ratio {f32_grad / f32_fwd:.2f}; x64 forward {f64_fwd * 1e3:.1f} ms, gradient '
      f'{f64_grad * 1e3:.1f} ms, ratio {f64_grad / f64_fwd:.2f}')
  return out


def plain_fire_rate(jmd, R, sp, L, sigma, fire, n_steps, block=50):
  """The plain per-system path for one 2-D soft-sphere system (Dense list,
  ``nbrs.update`` + ``fire_descent``'s apply_fn per step) with its steps
  recorded in a HIP graph of ``block`` steps and replayed: seconds for
  ``n_steps`` steps (a multiple of block), excluding allocation, warm-up and
  capture."""
  import torch
  dev = 'cuda:0'
  disp, shift = jmd.space.periodic(L)
  nfn, efn = jmd.energy.soft_sphere_neighbor_list(
      disp, L, species=torch.as_tensor(sp, device=dev),
      sigma=torch.as_tensor(sigma, device=dev), format=jmd.partition.Dense)
  init_fn, apply_fn = jmd.minimize.fire_descent(efn, shift, **fire)
  Rt = torch.as_tensor(R, device=dev)
  nb = nfn.allocate(Rt)
  st = init_fn(Rt, neighbor=nb)
  for _ in range(3):  # loads the kernels, fills the workspace caches
    nb = nb.update(st.position)
    st = apply_fn(st, neighbor=nb)
  torch.cuda.synchronize()
  fields = ('position', 'momentum', 'force', 'fire_scalars', 'fire_n_pos', 'position_hi')
  g = torch.cuda.CUDAGraph()
  with torch.cuda.graph(g):
    cur, nbc = st, nb
    for _ in range(block):
      nbc = nbc.update(cur.position)
      cur = apply_fn(cur, neighbor=nbc)
    for f in fields:
      getattr(st, f).copy_(getattr(cur, f))
  g.replay()  # one untimed replay
  torch.cuda.synchronize()
  t = time.perf_counter()
  for _ in range(max(1, n_steps // block)):
    g.replay()
  torch.cuda.synchronize()
  return time.perf_counter() - t


def batched_section(jmd, O, n_systems=512, n_steps=500, plain_sample=8, plain_steps=50):
  """Batched workload: 512 bidisperse 2-D soft-sphere systems (N log-uniform in
  [64, 4096], seed 0), 500 FIRE steps (dt_start 0.01, dt_max 0.1) in one
  launch, against per-system plain launches on a sample of the systems."""
  import numpy as np
  import torch
  dev = 'cuda:0'
  rng = np.random.default_rng(0)
  sizes = np.floor(np.exp(rng.uniform(np.log(64), np.log(4096), n_systems))).astype(int)
  systems = [O.bidisperse_2d(int(n), seed=1000 + k) for k, n in enumerate(sizes)]
  sigma = np.array([[1.0, 1.2], [1.2, 1.4]], np.float32)
  fire = dict(dt_start=0.01, dt_max=0.1)

  def make():
    return jmd.ensemble.SoftSphereFireEnsemble([s[0] for s in systems], [s[2] for s in systems],
                                               [s[1] for s in systems], sigma=sigma, **fire)

  ens = make().allocate()
  ens.run(5)  # warm
  ens = make().allocate()
  torch.cuda.synchronize()
  t = time.perf_counter()
  ens.run(n_steps)
  torch.cuda.synchronize()
  t_ens = time.perf_counter() - t
  total = int(sizes.sum())
  value = total * n_steps / t_ens
  overflow = int((ens.did_buffer_overflow != 0).sum().item())
  # plain per-system path on a sample spread over the sizes: every system's
  # steps as graph replays (the fair comparator), and host-launched
  pick = np.argsort(sizes)[np.linspace(0, n_systems - 1, plain_sample).astype(int)]
  t_graph, graph_atoms = 0.0, 0
  for k in pick:
    R, sp, L = systems[k]
    t_graph += plain_fire_rate(jmd, R, sp, L, sigma, fire, n_steps)
    graph_atoms += len(R) * n_steps
  plain_graph = graph_atoms / t_graph
  t_plain, steps_atoms = 0.0, 0
  for k in pick:
    R, sp, L = systems[k]
    disp, shift = jmd.space.periodic(L)
    nfn, efn = jmd.energy.soft_sphere_neighbor_list(
        disp, L, species=torch.as_tensor(sp, device=dev),
        sigma=torch.as_tensor(sigma, device=dev), format=jmd.partition.Dense)
    init_fn, apply_fn = jmd.minimize.fire_descent(efn, shift, **fire)
    Rt = torch.as_tensor(R, device=dev)
    nb = nfn.allocate(Rt)
    st = init_fn(Rt, neighbor=nb)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(plain_steps):
      nb = nb.update(st.position)
      st = apply_fn(st, neighbor=nb)
    torch.cuda.synchronize()
    t_plain += time.perf_counter() - t
    steps_atoms += len(R) * plain_steps
  plain = steps_atoms / t_plain
  # energies at the final positions against FP64 on a few systems
  E = ens.energy.cpu().numpy()
  errs = []
  for k in pick[:4]:
    R, sp, L = systems[k]
    model = O.PairModel('soft_sphere', sigma=sigma.astype(np.float64), species=sp)
    pos = ens.position[ens.system_slice(int(k))].cpu().numpy().astype(np.float64)
    E64 = O.energy_force_brute(model, O.Space(np.float64(L)), pos)[0]
    errs.append(abs(E[k] - E64) / max(abs(E64), 1e-30))
  log(f'batched: {value:.3e} atom-steps/s in {t_ens * 1e3:.1f} ms, plain graph '
      f'{plain_graph:.3e} (x{value / plain_graph:.1f}), plain host loop {plain:.3e}')
  return {'workload': f'{n_systems} bidisperse 2-D soft-sphere systems (N log-uniform '
                      f'64..4096, total {total}), {n_steps} FIRE steps, one launch',
          'value': value, 'unit': 'atom-steps/s', 'seconds': t_ens,
          'plain_graph_value': plain_graph,
          'plain_graph_sample': f'{plain_sample} systems spread over the sizes, {n_steps} '
                                f'steps each as replays of a 50-step HIP graph, one system '
                                f'after another',
          'speedup_vs_plain_graph': value / plain_graph,
          'plain_per_system_value': plain,
          'plain_sample': f'{plain_sample} systems spread over the sizes, {plain_steps} steps '
                          f'each, host-launched',
          'speedup_vs_plain': value / plain, 'systems_overflowed': overflow,
          'energy_rel_err_vs_fp64_max': float(max(errs)) if errs else None}


def ensemble_gradient_section(jmd, O, n_rep=64, n=8192, n_steps=300, every=25, phi=0.84,
                              plain_sample=8):
  """Ensemble-gradient workload: 64 replicas of a 2-D bidisperse soft-sphere
  packing (N = 8192 each, phi = 0.84), replica k with diameter ratio
  r_k = 1.2 + 0.4 k / 63 (sigma_ab = [[1, (1+r)/2], [(1+r)/2, r]]) in a
  square of side L(r_k) = sqrt(N pi (1 + r_k^2) / (8 phi)), 300 FIRE steps in
  one batched launch, then d mean_k E_k(x_T) / d r_k for every k through the
  checkpointed FIRE adjoint (checkpoints every 25 steps).  The reported
  derivative lets the box follow the ratio, as the reference's
  meta-optimisation notebook does (periodic_general(L(d) I) with positions
  in real coordinates, x_0 held fixed); the fixed-box derivative is reported
  beside it.  The plain per-replica path (graph-replayed FIRE steps of one
  system at a time) is timed on a sample of the replicas."""
  import numpy as np
  import torch
  r = [float(np.float32(1.2 + 0.4 * k / (n_rep - 1))) for k in range(n_rep)]
  systems = [O.bidisperse_2d(n, d_large=r[k], seed=2000 + k) for k in range(n_rep)]
  fire = dict(dt_start=0.01, dt_max=0.1)
  dsig = np.array([[0.0, 0.5], [0.5, 1.0]], np.float32)
  side = [float(np.sqrt(n * np.pi * (1 + x * x) / (8 * phi))) for x in r]
  dLdr = np.array([L * x / (1 + x * x) for L, x in zip(side, r)])

  def sig_of(x):
    return np.array([[1.0, (1 + x) / 2], [(1 + x) / 2, x]], np.float32)

  def make():
    sig = np.stack([sig_of(x) for x in r])
    return jmd.ensemble.SoftSphereFireEnsemble([s[0] for s in systems], [s[2] for s in systems],
                                               [s[1] for s in systems], sigma=sig,
                                               **fire).allocate()

  make().run(5)  # warm
  ens = make()
  torch.cuda.synchronize()
  t = time.perf_counter()
  ens.run(n_steps)
  torch.cuda.synchronize()
  t_fwd = time.perf_counter() - t
  # warm: two segments, so both tapes and the replay stream exist before the
  # timed call (allocations the caching allocator keeps)
  make().energy_gradient(2 * every, dsig, checkpoint_every=every, dbox=dLdr)
  ens = make()
  torch.cuda.synchronize()
  t = time.perf_counter()
  E, th, _ = ens.energy_gradient(n_steps, dsig, checkpoint_every=every, dbox=dLdr)
  torch.cuda.synchronize()
  t_grad = time.perf_counter() - t
  overflow = int((ens.did_buffer_overflow != 0).sum().item())
  box_adj = ens.box_adjoint.cpu().numpy()
  th = th.cpu().numpy()
  ens = make()
  torch.cuda.synchronize()
  t = time.perf_counter()
  _, th_fixed, _ = ens.energy_gradient(n_steps, dsig, checkpoint_every=every)
  torch.cuda.synchronize()
  t_grad_fixed = time.perf_counter() - t
  th_fixed = th_fixed.cpu().numpy()
  # plain per-replica forward on a sample spread over the ratios
  pick = np.linspace(0, n_rep - 1, plain_sample).astype(int)
  t_plain = 0.0
  for k in pick:
    R, sp, L = systems[k]
    t_plain += plain_fire_rate(jmd, R, sp, L, sig_of(r[k]), fire, n_steps)
  plain = plain_sample * n * n_steps / t_plain
  fwd = n_rep * n * n_steps / t_fwd
  out = {'workload': f'{n_rep} replicas of a 2-D bidisperse soft-sphere packing (N={n} each, '
                     f'phi {phi}), diameter ratio r_k = 1.2 + 0.4 k/{n_rep - 1}, side L(r_k) = '
                     f'sqrt(N pi (1 + r_k^2) / (8 phi)), {n_steps} FIRE steps batched, '
                     f'd mean E(x_T) / d r_k by the checkpointed adjoint (checkpoints every '
                     f'{every})',
         'derivative': 'dmeanE_dr: the box follows r_k (through the minimum images and the '
                       'shift folds, x_0 held in real coordinates), as the reference notebook '
                       'differentiates periodic_general(L(d) I); dmeanE_dr_fixed_box: boxes '
                       'held constant',
         'forward_s': t_fwd, 'forward_atom_steps_per_s': fwd,
         'gradient_s': t_grad, 'ratio_vs_forward': t_grad / t_fwd,
         'gradient_fixed_box_s': t_grad_fixed,
         'plain_per_replica_atom_steps_per_s': plain,
         'plain_sample': f'{plain_sample} replicas spread over r, {n_steps} steps each as '
                         f'replays of a 50-step HIP graph, one replica after another',
         'forward_speedup_vs_plain': fwd / plain,
         'mean_energy_T': float(E.mean().item()), 'systems_overflowed': overflow,
         'dmeanE_dr': [float(x) for x in th],
         'dmeanE_dr_fixed_box': [float(x) for x in th_fixed],
         'dmeanE_dside': [float(x) for x in box_adj]}
  # FIRE's dt / alpha / reset decisions make E(x_T) piecewise smooth in r: in
  # FP64, central differences over 300 steps reach the derivative only for
  # h <~ 1e-7, out of float32's reach, so parity is checked on small replicas
  # against the FP64 adjoint (itself pinned by such differences) in the tests
  out['parity'] = ('tests/test_gpu_ensemble.py: FP64 FIRE adjoint on four replicas '
                   '(N 64..256, 30 steps), fixed box and box following r, 1e-5 relative')
  log(f'ensemble gradient: forward {t_fwd * 1e3:.1f} ms ({fwd:.3e} atom-steps/s, plain '
      f'per replica {plain:.3e}), gradient {t_grad * 1e3:.1f} ms, ratio {t_grad / t_fwd:.2f}')
  return out


def published_section(jmd, O, n_steps=2000, warmup=200, block=100):
  """The reference's only published run (notebooks/lj_benchmark.ipynb): 64,000
  LJ atoms on a 40^3 simple-cubic lattice of constant 1.3782 (number density
  0.382), r_cutoff 3.0 sigma, skin 1.0, NVT Nose-Hoover at kT = 1.2, dt =
  0.005, OrderedSparse lists, float32; published 2.24 ms/step (2.86e7
  atom-steps/s) on an A100.  Momenta are drawn here (the notebook's JAX key
  cannot be reproduced).  Also run with Dense lists."""
  import numpy as np
  import torch
  dev = 'cuda:0'
  a, nrep = 1.37820, 40
  g = np.arange(nrep)
  R = (np.stack(np.meshgrid(g, g, g, indexing='ij'), -1).reshape(-1, 3) * a).astype(np.float32)
  L = np.float32(nrep * a)
  N = R.shape[0]
  rng = np.random.default_rng(7)
  P = (rng.normal(size=R.shape) * np.sqrt(1.2)).astype(np.float32)
  P -= P.mean(0)
  disp, shift = jmd.space.periodic(L)
  out = {}
  for name in ('OrderedSparse', 'Dense'):
    nfn, efn = jmd.energy.lennard_jones_neighbor_list(disp, L, r_cutoff=3.0, dr_threshold=1.0,
                                                      format=getattr(jmd.partition, name))
    Rt = torch.as_tensor(R, device=dev)
    init_fn, _ = jmd.simulate.nvt_nose_hoover(efn, shift, 5e-3, kT=1.2)
    nb = nfn.allocate(Rt, extra_capacity=55 if name == 'Dense' else 0)
    st = init_fn(0, Rt, momenta=torch.as_tensor(P, device=dev), neighbor=nb)
    loop = jmd.simulate.NVTLoop(efn, shift, 5e-3, kT=1.2)
    loop.prime(st, nb)
    reallocs = 0

    def run(n):
      nonlocal nb, reallocs
      done = 0
      while done < n:
        snap = (st.position_hi.clone(), st.position.clone(), st.momentum.clone(),
                st.force.clone(), st.chain.clone(), st.kinetic_energy.clone())
        loop.run(st, nb, block)
        if bool(nb.did_buffer_overflow):
          for t, v in zip((st.position_hi, st.position, st.momentum, st.force, st.chain,
                           st.kinetic_energy), snap):
            t.copy_(v)
          nb = nfn.allocate(st.position)
          loop.prime(st, nb)
          reallocs += 1
          continue
        done += block

    run(warmup)
    torch.cuda.synchronize()
    t = time.perf_counter()
    run(n_steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t
    T_end = float(jmd.quantity.temperature(momentum=st.momentum))
    out[name] = {'ms_per_step': 1e3 * el / n_steps, 'atom_steps_per_s': N * n_steps / el,
                 'vs_published_a100': (N * n_steps / el) / 2.86e7, 'reallocs': reallocs,
                 'temperature_end': T_end}
    log(f'published setup, {name}: {1e3 * el / n_steps:.3f} ms/step, '
        f'{N * n_steps / el:.3e} atom-steps/s, x{(N * n_steps / el) / 2.86e7:.1f} the A100 '
        f'figure, T_end {T_end:.3f}')
  # the drop-in per-step form with the notebook's format: nbrs.update and
  # apply_fn called from Python each step (the list's Dense rows carry the
  # force; host-launched, no graph)
  nfn, efn = jmd.energy.lennard_jones_neighbor_list(disp, L, r_cutoff=3.0, dr_threshold=1.0,
                                                    format=jmd.partition.OrderedSparse)
  Rt = torch.as_tensor(R, device=dev)
  init_fn, apply_fn = jmd.simulate.nvt_nose_hoover(efn, shift, 5e-3, kT=1.2)
  nb = nfn.allocate(Rt)
  st = init_fn(0, Rt, momenta=torch.as_tensor(P, device=dev), neighbor=nb)
  steps_api = 300
  for _ in range(20):
    nb = nb.update(st.position)
    st = apply_fn(st, neighbor=nb)
  torch.cuda.synchronize()
  t = time.perf_counter()
  for _ in range(steps_api):
    nb = nb.update(st.position)
    st = apply_fn(st, neighbor=nb)
  torch.cuda.synchronize()
  el = time.perf_counter() - t
  out['OrderedSparse_per_step_api'] = {
      'ms_per_step': 1e3 * el / steps_api, 'atom_steps_per_s': N * steps_api / el,
      'vs_published_a100': (N * steps_api / el) / 2.86e7,
      'overflow': bool(nb.did_buffer_overflow),
      'note': f'{steps_api} steps of nbrs.update + apply_fn from Python after 20, '
              'host-launched'}
  log(f'published setup, OrderedSparse per-step API: {1e3 * el / steps_api:.3f} ms/step')
  out['workload'] = ('64,000 LJ atoms, 40^3 lattice a = 1.3782 (rho 0.382), r_c 3.0, skin 1.0, '
                     f'NVT Nose-Hoover kT 1.2, dt 0.005, {n_steps} steps after {warmup}')
  out['published'] = {'ms_per_step': 2.24, 'atom_steps_per_s': 2.86e7,
                      'hardware': 'NVIDIA A100-SXM4-40GB', 'format': 'OrderedSparse',
                      'source': 'notebooks/lj_benchmark.ipynb'}
  return out


def general_section(jmd, O, n=64, n_cells=40, n_steps=1000, warmup=200, block=100):
  """periodic_general measurement: LJ NVE (r_c 2.5, skin 0.5, dt 0.005,
  Dense lists, kT 1) on an FCC crystal at rho 0.8442 stored as fractional
  coordinates in a triclinic cell (an n^3 supercell of the primitive cell,
  box matrix upper triangular: 262,144 atoms at n = 64), next to the same
  crystal in a cubic cell with space.periodic (4 n_cells^3 = 256,000 atoms).
  Also times the virial kernel behind quantity.pressure on the triclinic
  system."""
  import numpy as np
  import torch
  dev = 'cuda:0'
  a = float((4.0 / 0.8442) ** (1.0 / 3.0))
  out = {}

  def run(name, R, shift, nfn, efn, P):
    Rt = torch.as_tensor(R, device=dev)
    init_fn, _ = jmd.simulate.nve(efn, shift, 0.005)
    nb = nfn.allocate(Rt, extra_capacity=8)
    st = init_fn(0, Rt, 0.0, momenta=torch.as_tensor(P, device=dev), neighbor=nb)
    loop = jmd.simulate.NVELoop(efn, shift, 0.005)
    reallocs = 0

    def go(k_total):
      nonlocal nb, reallocs
      done = 0
      loop.prime(st, nb)
      while done < k_total:
        snap = (st.position_hi.clone(), st.position.clone(), st.momentum.clone(),
                st.force.clone())
        loop.run(st, nb, block)
        if bool(nb.did_buffer_overflow):
          for t, v in zip((st.position_hi, st.position, st.momentum, st.force), snap):
            t.copy_(v)
          nb = nfn.allocate(st.position, extra_capacity=8)
          loop.prime(st, nb)
          reallocs += 1
          continue
        done += block

    go(warmup)
    E0, _ = efn.energy_and_force(st.position, nb)
    K0 = float(jmd.quantity.kinetic_energy(momentum=st.momentum))
    torch.cuda.synchronize()
    t = time.perf_counter()
    go(n_steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t
    E1, _ = efn.energy_and_force(st.position, nb)
    K1 = float(jmd.quantity.kinetic_energy(momentum=st.momentum))
    N = R.shape[0]
    res = {'atoms': N, 'ms_per_step': 1e3 * el / n_steps, 'atom_steps_per_s': N * n_steps / el,
           'total_energy_drift_rel': abs((float(E1) + K1) - (float(E0) + K0)) /
                                     abs(float(E0) + K0),
           'reallocs': reallocs}
    log(f'general section, {name}: {N} atoms, {1e3 * el / n_steps:.3f} ms/step, '
        f'{N * n_steps / el:.3e} atom-steps/s')
    return res, st, nb

  u, T, P = O.fcc_triclinic(n, a=a, jitter=0.03, kT=1.0, seed=21)
  disp, shift = jmd.space.periodic_general(T)
  nfn, efn = jmd.energy.lennard_jones_neighbor_list(disp, T, fractional_coordinates=True,
                                                    format=jmd.partition.Dense)
  out['triclinic'], st, nb = run('triclinic', u, shift, nfn, efn, P)
  # the virial kernel (quantity.pressure / stress) on the triclinic state
  for _ in range(3):
    efn.virial(st.position, nb)
  torch.cuda.synchronize()
  e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
  e0.record()
  for _ in range(20):
    W = efn.virial(st.position, nb)
  e1.record()
  torch.cuda.synchronize()
  out['virial_us_per_call'] = 1e3 * e0.elapsed_time(e1) / 20
  Rc, Lc = O.fcc_cubic(n_cells, a=a)
  rng = np.random.default_rng(22)
  Rc = np.mod(Rc + rng.normal(0, 0.03, Rc.shape), Lc).astype(np.float32)
  Rc = np.where(Rc >= Lc, np.float32(0), Rc).astype(np.float32)
  Pc = rng.normal(0.0, 1.0, Rc.shape)
  Pc = (Pc - Pc.mean(0)).astype(np.float32)
  dc, sc = jmd.space.periodic(Lc)
  nfc, efc = jmd.energy.lennard_jones_neighbor_list(dc, Lc, format=jmd.partition.Dense)
  out['cubic'], _, _ = run('cubic', Rc, sc, nfc, efc, Pc)
  out['workload'] = (f'LJ NVE, FCC at rho 0.8442 (a = {a:.4f}), jitter 0.03, kT 1, r_c 2.5, '
                     f'skin 0.5, dt 0.005, Dense lists, NVELoop; triclinic: {n}^3 primitive '
                     f'cells as fractional coordinates in an upper-triangular box '
                     f'(space.periodic_general); cubic: {n_cells}^3 conventional cells '
                     f'(space.periodic); {n_steps} steps after {warmup}')
  log(f'general section: virial {out["virial_us_per_call"]:.1f} us per call')
  # multi-image list of a small triclinic cell (cutoff beyond half the cell)
  rng = np.random.default_rng(23)
  Tm = np.array([[6.2, 1.4, -0.8], [0.0, 5.6, 1.0], [0.0, 0.0, 6.6]], np.float32)
  um = torch.as_tensor(rng.uniform(0, 1, (500, 3)).astype(np.float32), device=dev)
  mfn = jmd.custom_partition.neighbor_list_multi_image(None, Tm, 3.5, 0.0)
  mnb = mfn.allocate(um)
  for _ in range(3):
    mnb = mnb.update(um)
  torch.cuda.synchronize()
  gc.collect()  # the section's own garbage (the collector is paused)
  t = time.perf_counter()
  for _ in range(20):
    mnb = mnb.update(um)
  torch.cuda.synchronize()
  elapsed = time.perf_counter() - t
  out['multi_image'] = {'atoms': 500, 'edges': mnb.n_edges,
                        'shift_stencil': int(mnb.search_shifts.shape[0]),
                        'rebuild_us': 1e6 * elapsed / 20,
                        'workload': '500 atoms, triclinic cell of widths ~5.4-6.6, cutoff 3.5, '
                                    'Sparse, dr_threshold 0 (a rebuild per update, host '
                                    'overhead included)'}
  log(f'general section: multi-image rebuild {out["multi_image"]["rebuild_us"]:.1f} us, '
      f'{out["multi_image"]["edges"]} edges over {out["multi_image"]["shift_stencil"]} shifts')
  return out


def main():
  ap = argparse.ArgumentParser()
  ap.add_argument('--gpus', type=int, default=1)
  ap.add_argument('--steps', type=int, default=2000)
  ap.add_argument('--warmup', type=int, default=200)
  ap.add_argument('--n-side', type=int, default=64)
  ap.add_argument('--block', type=int, default=100)
  ap.add_argument('--cpu-steps', type=int, default=250)
  ap.add_argument('--no-cpu', action='store_true')
  ap.add_argument('--tpa', type=int, default=0,
                  help='threads per row of the pair kernel (skips the autotuner)')
  ap.add_argument('--no-tune', action='store_true',
                  help='the reference\'s cell size and capacity, the built-in threads per row')
  ap.add_argument('--eager', action='store_true', help='no HIP graphs (host-launched steps)')
  ap.add_argument('--no-reorder', action='store_true',
                  help='run the blocks in the caller\'s particle order (no cell-order layout)')
  ap.add_argument('--no-extra', action='store_true',
                  help='skip the batched-ensemble and adjoint measurements')
  ap.add_argument('--copy-calibration', action='store_true',
                  help='only run a plain 256 MiB device copy (the reference kernel that '
                       'scales the FETCH_SIZE / WRITE_SIZE counters) and exit')
  ap.add_argument('--section', choices=('batched', 'adjoint', 'ensemble_gradient', 'published',
                                        'general'),
                  help='only run this one extra measurement and print its JSON')
  args = ap.parse_args()
  if args.gpus != 1:
    print(f'bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU '
          'decomposition is described in DESIGN.md and not built.', file=sys.stderr)
    sys.exit(2)

  import numpy as np
  import torch

  import jax_md_amd as jmd
  from jax_md_amd import _native as nat
  from oracle import jaxmd_oracle as O

  dev = 'cuda:0'
  torch.cuda.set_device(0)
  if args.copy_calibration:
    nbytes = 256 << 20
    src = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(5):
      dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({'copy_calibration_bytes_read': nbytes, 'bytes_written': nbytes,
                      'dispatches': 5}), flush=True)
    return
  if args.section:
    fn = {'batched': lambda: batched_section(jmd, O),
          'adjoint': lambda: adjoint_section(jmd, O, check=not args.no_cpu),
          'ensemble_gradient': lambda: ensemble_gradient_section(jmd, O),
          'published': lambda: published_section(jmd, O),
          'general': lambda: general_section(jmd, O)}[args.section]
    print(json.dumps({args.section: without_collector(fn)}), flush=True)
    return
  R, P, L = O.lj_lattice(args.n_side, seed=0)
  N = R.shape[0]
  disp, shift = jmd.space.periodic(L)

  def make_fns(**kw):
    return jmd.energy.lennard_jones_neighbor_list(disp, L, format=jmd.partition.Dense, **kw)

  Rt = torch.as_tensor(R, device=dev)
  tuned = None
  if args.tpa:
    jmd.autotune.set_threads_per_atom(args.tpa)
    nfn, efn = make_fns()
  elif args.no_tune:
    nfn, efn = make_fns()
  else:
    # setup, before the warm-up: threads per row, cell size and capacity
    # timed on copies of the initial state (the user's state is untouched)
    t_tune = time.perf_counter()
    tuned = jmd.autotune.tune_nve_loop(
        make_fns, Rt, torch.as_tensor(P, device=dev), shift, 0.005,
        key=jmd.autotune.workload_key(Rt, L, 2.5, 0.5, jmd.partition.Dense, 'lj'),
        cache_path=os.environ.get('JAX_MD_AMD_TUNE_CACHE'))
    tuned.apply()
    nfn, efn = make_fns(**tuned.neighbor_kwargs)
    log(f'autotune ({time.perf_counter() - t_tune:.1f} s): {json.dumps(tuned.as_dict())}')
  nbrs = nfn.allocate(Rt)
  init_fn, _ = jmd.simulate.nve(efn, shift, 0.005)
  state = init_fn(0, Rt.clone(), kT=1.0, momenta=torch.as_tensor(P, device=dev), neighbor=nbrs)
  loop = jmd.simulate.NVELoop(efn, shift, 0.005, graph=not args.eager,
                              reorder=not args.no_reorder)
  log(f'N={N} L={L} max_occupancy={nbrs.max_occupancy} cell_capacity={nbrs.cell_list_capacity}')

  def snapshot(s):
    return (s.position_hi.clone(), s.position.clone(), s.momentum.clone(), s.force.clone())

  def restore(s, snap):
    s.position_hi.copy_(snap[0])
    s.position.copy_(snap[1])
    s.momentum.copy_(snap[2])
    s.force.copy_(snap[3])

  reallocs = 0
  builds_before = 0  # builds of lists replaced by a re-allocation

  def run_steps(n):
    nonlocal nbrs, reallocs, builds_before
    done = 0
    loop.prime(state, nbrs)
    while done < n:
      k = min(args.block, n - done)
      snap = snapshot(state)
      loop.run(state, nbrs, k)
      if bool(nbrs.did_buffer_overflow):
        restore(state, snap)
        builds_before += nbrs.plan.builds
        nbrs = nfn.allocate(state.position)
        loop.prime(state, nbrs)
        reallocs += 1
        log(f'  overflow: re-allocated, max_occupancy={nbrs.max_occupancy}')
        continue
      done += k
      log(f'  {done}/{n} steps')

  # warmup: W steps in all; before the last block the list is allocated again
  # for the equilibrated liquid (the lattice's capacities are tighter), as a
  # production run allocates after equilibration, and the last warmup block
  # records the block graph on the new list's buffers
  w1 = max(0, args.warmup - args.block)
  run_steps(w1)
  builds_before += nbrs.plan.builds
  nbrs = nfn.allocate(state.position)
  run_steps(args.warmup - w1)
  # the graphs of every block length the timed loop uses are recorded now,
  # so the timed blocks are replays (an eager warm-up block has loaded the
  # kernels; a block after a re-allocation in the timed region still records)
  lengths = {min(args.block, args.steps)}
  if args.steps > args.block and args.steps % args.block:
    lengths.add(args.steps % args.block)
  if loop.stats['eager'] > 0 and not args.eager:
    for k in sorted(lengths):
      loop.capture(state, nbrs, k)
  torch.cuda.synchronize()
  reallocs_warm = reallocs
  E_start, _ = efn.energy_and_force(state.position, nbrs)
  KE_start = jmd.quantity.kinetic_energy(momentum=state.momentum)
  builds0 = builds_before + nbrs.plan.builds
  stats0 = dict(loop.stats)
  torch.cuda.synchronize()

  def timed_run():
    t0 = time.perf_counter()
    run_steps(args.steps)
    torch.cuda.synchronize()
    return time.perf_counter() - t0

  elapsed = without_collector(timed_run)
  timed_blocks = {k: loop.stats[k] - stats0[k] for k in loop.stats}
  builds = builds_before + nbrs.plan.builds - builds0
  E_end, F_gpu = efn.energy_and_force(state.position, nbrs)
  KE_end = jmd.quantity.kinetic_energy(momentum=state.momentum)
  X_end = state.position.cpu().numpy().astype(np.float64)  # where F_gpu was taken
  value = N * args.steps / elapsed
  ms = 1e3 * elapsed / args.steps
  idx = nbrs.idx
  kbar = float((idx < N).sum().item()) / N
  brk = step_breakdown(jmd, loop, state, nbrs)
  force_us = brk['force_us']
  force_bytes = N * (4.0 * kbar + FORCE_STATE_BYTES)
  achieved = force_bytes / (force_us * 1e-6) / 1e9
  log(f'{args.steps} steps in {elapsed:.3f}s: {ms * 1e3:.1f} us/step, force kernel '
      f'{force_us:.1f} us, Kbar={kbar:.2f}, list builds={builds}, reallocs={reallocs}, '
      f'timed blocks {timed_blocks}')
  log(f'breakdown: {json.dumps({k: v for k, v in brk.items() if k != "note"})}')
  traffic, traffic_src = counter_traffic(N, kbar)
  drift = None
  try:
    drift = c1_drift(jmd, O)
    log(f'C1 drift over 1000 steps: {drift["drift_rel"]:.2e} (max {drift["max_dev_rel"]:.2e})')
  except Exception as e:
    drift = f'failed: {type(e).__name__}: {e}'
    log(f'C1 drift run failed: {e}')

  # force error against the float64 C oracle at the final float32 positions
  force_err = None
  try:
    from oracle import c_oracle
    sysc = c_oracle.LJSystem(X_end, float(L))
    E64, F64 = sysc.forces()
    Fg = F_gpu.cpu().numpy().astype(np.float64)
    rel = float(np.linalg.norm(Fg - F64) / np.linalg.norm(F64))
    worst = float(np.max(np.abs(Fg - F64)) / np.sqrt(np.mean(F64 ** 2)))
    force_err = {'force_rel_norm': rel, 'force_worst_over_rms': worst,
                 'energy_rel': float(abs(float(E_end) - E64) / abs(E64))}
  except Exception as e:  # the checker is optional on hosts without gcc
    log(f'force check skipped: {e}')

  cpu = None
  if not args.no_cpu:
    try:
      from oracle import c_oracle
      sysc = c_oracle.LJSystem(R.astype(np.float64), float(L))
      t = time.perf_counter()
      _, rb = sysc.nve(P.astype(np.float64), 0.005, args.cpu_steps)
      tc = time.perf_counter() - t
      cpu = {'value': N * args.cpu_steps / tc, 'unit': 'atom-steps/s',
             'cores': c_oracle.num_threads(), 'kind': 'port',
             'sample': f'C FP64 OpenMP restatement (oracle/md_oracle.c), same 64^3 LJ NVE '
                       f'workload, {args.cpu_steps} steps from t=0 incl. {rb} list rebuilds; '
                       'the reference (JAX CPU) cannot run here: jax is not installed'}
      log(f'cpu baseline: {cpu["value"]:.3e} atom-steps/s on {cpu["cores"]} threads')
    except Exception as e:
      log(f'cpu baseline skipped: {e}')

  out = {
      'metric': METRIC, 'value': value, 'unit': 'atom-steps/s', 'n_gpus': 1,
      'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': ms,
      'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32',
      'data': 'synthetic',
      'config': {'workload': f'LJ NVE N={N} ({args.n_side}^3 lattice, rho=0.8442, jitter 0.05, '
                             'kT=1), Dense list r_c=2.5 skin=0.5, dt=0.005, fp32 pair '
                             'arithmetic, fp64 positions/energies',
                 'timed_steps_from_lattice': f'{args.warmup}..{args.warmup + args.steps}',
                 'list_builds': builds, 'reallocs': reallocs - reallocs_warm,
                 'max_occupancy': nbrs.max_occupancy, 'kbar': kbar,
                 'timed_blocks': timed_blocks,
                 'graph_replays_only': timed_blocks.get('eager', 0) == 0,
                 'tuned': (tuned.as_dict() if tuned is not None else
                           {'threads_per_atom': args.tpa or 'built-in',
                            'cell_size_factor': 1.0, 'capacity_multiplier': 1.25})},
      'roofline': {'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                   'frac': achieved / HBM_PEAK_GBS, 'traffic': traffic,
                   'kernel': FORCE_KERNEL,
                   'kernel_us': force_us, 'algorithmic_bytes': force_bytes,
                   'traffic_source': traffic_src},
      'step_breakdown': brk,
      'rebuild_us': brk['rebuild_us'],
      'rebuilds_per_step_timed': builds / args.steps,
      'c1_nve_drift_1000': drift,
      # the whole step against the survey's per-atom-step bytes 4 K + 132
      # (x64 r+w 48, p r+w 24, F w+r 24, x32 w+r 24, x_ref 12): the roofline
      # the 50% target is stated on
      'roofline_step': {'bound': 'hbm', 'bytes_per_atom_step': 4.0 * kbar + 132.0,
                        'achieved': N * (4.0 * kbar + 132.0) / (ms * 1e-3) / 1e9,
                        'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                        'frac': N * (4.0 * kbar + 132.0) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
      'cpu_baseline': cpu,
      'error_vs_fp64': force_err,
      'energy': {'E_total_start': float(E_start) + float(KE_start),
                 'E_total_end': float(E_end) + float(KE_end),
                 'note': f'total energy at steps {args.warmup} and {args.warmup + args.steps} '
                         'from the lattice: early steps carry the melting transient of the '
                         'jittered lattice (see c1_nve_drift_1000 for a 1000-step drift)'},
  }
  if not args.no_extra:
    for key, fn in (('batched', lambda: batched_section(jmd, O)),
                    ('adjoint', lambda: adjoint_section(jmd, O, check=not args.no_cpu)),
                    ('ensemble_gradient', lambda: ensemble_gradient_section(jmd, O)),
                    ('published', lambda: published_section(jmd, O)),
                    ('general', lambda: general_section(jmd, O))):
      try:
        out[key] = without_collector(fn)
      except Exception as e:  # reported, never silently dropped
        out[key] = f'failed: {type(e).__name__}: {e}'
        log(f'{key} section failed: {e}')
  print(json.dumps(out), flush=True)


if __name__ == '__main__':
  main()
