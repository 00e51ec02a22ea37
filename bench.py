#!/usr/bin/env python
"""Benchmark: LJ NVE at N = 262,144 on one MI355X (atom-steps/s).

Workload: 64^3 simple-cubic lattice at rho = 0.8442 (L rounded to float32),
jitter N(0, 0.05), momenta N(0, kT=1) with the mean removed; LJ sigma = eps =
1 with the C1 switch between r_on = 2.0 and r_c = 2.5, Dense neighbour list
with skin 0.5 rebuilt on the skin criterion, dt = 0.005, float32 pair
arithmetic, float64 master positions and energy reductions.  One step = skin
rebuild (gated on the device) + half kick / drift + force pass with the
closing half kick.  Steps run in blocks of 100, each block one HIP graph
replay, with a host overflow check per block; an overflowing block is redone
from its saved start state after re-allocating the list (the reference's
protocol, examples/nve_neighbor_list.py:186-201).

The force kernel's duration is measured with HIP events around a graph of
20 back-to-back force launches on the loop's own cell-ordered buffers (the
state the timed region ended in), on the stream they run on.  HBM traffic per
launch comes from the committed counter summary of the same kernel and
workload (profiles/k_pair_dense_lj_n262144_pmc.json), when present.

Prints one JSON line on stdout; progress goes to stderr.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'atom-steps/s LJ NVE N=262k, plain vs batched/adjoint; error vs FP64 oracle'
HBM_PEAK_GBS = 8000.0


def log(*a):
  print(*a, file=sys.stderr, flush=True)


def force_kernel_us(jmd, loop, state, nbrs, efn, reps=20):
  """Average duration of the force + half-kick kernel, from HIP events around
  a graph of `reps` launches on the loop's cell-ordered buffers."""
  import torch
  b = loop._bufs
  if b is not None:
    import dataclasses
    x, idx = b['x'], b['idx']
    nb = dataclasses.replace(nbrs, idx=idx)
  else:
    x, nb = state.position, nbrs
  p = state.momentum.clone() if b is None else b['p'].clone()
  half = float(loop.half)
  efn.force_and_kick(x, nb, p, half_dt=half)  # warm
  g = torch.cuda.CUDAGraph()
  with torch.cuda.graph(g):
    for _ in range(reps):
      efn.force_and_kick(x, nb, p, half_dt=half)
  s = torch.cuda.current_stream()
  e0 = torch.cuda.Event(enable_timing=True)
  e1 = torch.cuda.Event(enable_timing=True)
  g.replay()
  e0.record(s)
  for _ in range(5):
    g.replay()
  e1.record(s)
  torch.cuda.synchronize()
  return e0.elapsed_time(e1) * 1e3 / (5 * reps)


def counter_traffic(N):
  """HBM bytes per force launch from the committed counter summary, or None."""
  path = os.path.join(ROOT, 'profiles', f'k_pair_dense_lj_n{N}_pmc.json')
  try:
    with open(path) as f:
      d = json.load(f)
    return float(d['traffic_bytes_per_launch'])
  except (OSError, KeyError, ValueError):
    return None


def main():
  ap = argparse.ArgumentParser()
  ap.add_argument('--gpus', type=int, default=1)
  ap.add_argument('--steps', type=int, default=2000)
  ap.add_argument('--warmup', type=int, default=200)
  ap.add_argument('--n-side', type=int, default=64)
  ap.add_argument('--block', type=int, default=100)
  ap.add_argument('--cpu-steps', type=int, default=250)
  ap.add_argument('--no-cpu', action='store_true')
  ap.add_argument('--tpa', type=int, default=0, help='threads per row of the pair kernel')
  ap.add_argument('--eager', action='store_true', help='no HIP graphs (host-launched steps)')
  ap.add_argument('--copy-calibration', action='store_true',
                  help='only run a plain 256 MiB device copy (the reference kernel that '
                       'scales the FETCH_SIZE / WRITE_SIZE counters) and exit')
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
  if args.tpa:
    nat.check(nat.load().jmd_set_threads_per_atom(args.tpa), 'jmd_set_threads_per_atom')
  R, P, L = O.lj_lattice(args.n_side, seed=0)
  N = R.shape[0]
  disp, shift = jmd.space.periodic(L)
  nfn, efn = jmd.energy.lennard_jones_neighbor_list(disp, L, format=jmd.partition.Dense)
  Rt = torch.as_tensor(R, device=dev)
  nbrs = nfn.allocate(Rt)
  init_fn, _ = jmd.simulate.nve(efn, shift, 0.005)
  state = init_fn(0, Rt.clone(), kT=1.0, momenta=torch.as_tensor(P, device=dev), neighbor=nbrs)
  loop = jmd.simulate.NVELoop(efn, shift, 0.005, graph=not args.eager)
  log(f'N={N} L={L} max_occupancy={nbrs.max_occupancy} cell_capacity={nbrs.cell_list_capacity}')

  def snapshot(s):
    return (s.position_hi.clone(), s.position.clone(), s.momentum.clone(), s.force.clone())

  def restore(s, snap):
    s.position_hi.copy_(snap[0])
    s.position.copy_(snap[1])
    s.momentum.copy_(snap[2])
    s.force.copy_(snap[3])

  reallocs = 0

  def run_steps(n):
    nonlocal nbrs, reallocs
    done = 0
    loop.prime(state, nbrs)
    while done < n:
      k = min(args.block, n - done)
      snap = snapshot(state)
      loop.run(state, nbrs, k)
      if bool(nbrs.did_buffer_overflow):
        restore(state, snap)
        nbrs = nfn.allocate(state.position)
        loop.prime(state, nbrs)
        reallocs += 1
        log(f'  overflow: re-allocated, max_occupancy={nbrs.max_occupancy}')
        continue
      done += k
      log(f'  {done}/{n} steps')

  run_steps(args.warmup)
  torch.cuda.synchronize()
  reallocs_warm = reallocs
  E_start, _ = efn.energy_and_force(state.position, nbrs)
  KE_start = jmd.quantity.kinetic_energy(momentum=state.momentum)
  builds0 = nbrs.plan.builds
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  run_steps(args.steps)
  torch.cuda.synchronize()
  elapsed = time.perf_counter() - t0
  builds = nbrs.plan.builds - builds0
  E_end, F_gpu = efn.energy_and_force(state.position, nbrs)
  KE_end = jmd.quantity.kinetic_energy(momentum=state.momentum)
  value = N * args.steps / elapsed
  ms = 1e3 * elapsed / args.steps
  idx = nbrs.idx
  kbar = float((idx < N).sum().item()) / N
  force_us = force_kernel_us(jmd, loop, state, nbrs, efn)
  # algorithmic bytes of one force launch: idx row (4 B per valid entry) plus
  # per-atom state: x read 16 (padded row), F write 12, p read+write 24,
  # row-end padding entry 4, energy/params 4 -> 4 Kbar + 60 B per atom
  force_bytes = N * (4.0 * kbar + 60.0)
  achieved = force_bytes / (force_us * 1e-6) / 1e9
  log(f'{args.steps} steps in {elapsed:.3f}s: {ms * 1e3:.1f} us/step, force kernel '
      f'{force_us:.1f} us, Kbar={kbar:.2f}, list builds={builds}, reallocs={reallocs}')
  traffic = counter_traffic(N)

  # force error against the float64 C oracle at the final float32 positions
  force_err = None
  try:
    from oracle import c_oracle
    sysc = c_oracle.LJSystem(state.position.cpu().numpy().astype(np.float64), float(L))
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
                 'list_builds': builds, 'reallocs': reallocs - reallocs_warm,
                 'max_occupancy': nbrs.max_occupancy, 'kbar': kbar},
      'roofline': {'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                   'frac': achieved / HBM_PEAK_GBS, 'traffic': traffic,
                   'kernel': 'k_pair_dense (force + closing half kick)',
                   'kernel_us': force_us, 'algorithmic_bytes': force_bytes},
      'cpu_baseline': cpu,
      'error_vs_fp64': force_err,
      'energy': {'E_total_start': float(E_start) + float(KE_start),
                 'E_total_end': float(E_end) + float(KE_end)},
  }
  print(json.dumps(out), flush=True)


if __name__ == '__main__':
  main()
