"""The large LJ NVE workload at full size on the GPU: N = 262,144 atoms
(64^3 simple-cubic lattice at rho = 0.8442, jitter 0.05, kT = 1), Dense list
r_c = 2.5, skin 0.5, dt = 0.005, through the step loop.

Bars (stated per check):
  * force at t = 0 and after 200 steps against the C FP64 restatement at the
    same positions: norm-wise relative error <= 1e-5, worst component
    <= 5e-5 x rms;
  * 200 steps of NVELoop (two graph-replayed blocks): no overflow, the Dense
    list symmetric (j in row i <=> i in row j) and free of duplicates and of
    self pairs, every pair within the list cutoff at the list's reference
    positions present (checked on a sample of rows against brute force);
  * total energy after 200 steps against a 200-step FP64 trajectory from the
    same start (oracle/md_oracle.c, velocity Verlet with the float32 dt):
    relative difference <= 2e-5; the float32 run's own total energy from step
    100 to 200 within 2e-4 relative.
"""

import numpy as np
import pytest
import torch

from oracle import jaxmd_oracle as O

pytestmark = pytest.mark.gpu

DEV = 'cuda:0'


def _force_err(F, F64):
  rel = float(np.linalg.norm(F - F64) / np.linalg.norm(F64))
  worst = float(np.max(np.abs(F - F64)) / np.sqrt(np.mean(F64 ** 2)))
  return rel, worst


def test_c2_full_size_force_loop_energy():
  import jax_md_amd as jmd
  from oracle import c_oracle
  R, P, L = O.lj_lattice(64, seed=0)
  N = R.shape[0]
  assert N == 262144
  disp, shift = jmd.space.periodic(L)
  nfn, efn = jmd.energy.lennard_jones_neighbor_list(disp, L, format=jmd.partition.Dense)
  Rt = torch.as_tensor(R, device=DEV)
  print('case c2: allocate + force at t=0', flush=True)
  nb = nfn.allocate(Rt, extra_capacity=16)
  init_fn, _ = jmd.simulate.nve(efn, shift, 0.005)
  st = init_fn(0, Rt.clone(), kT=1.0, momenta=torch.as_tensor(P, device=DEV), neighbor=nb)
  sysc = c_oracle.LJSystem(R.astype(np.float64), float(L))
  E64_0, F64_0 = sysc.forces()
  rel, worst = _force_err(st.force.cpu().numpy().astype(np.float64), F64_0)
  print(f'c2 t=0: force rel {rel:.2e} worst/rms {worst:.2e}', flush=True)
  assert rel <= 1e-5 and worst <= 5e-5, (rel, worst)

  loop = jmd.simulate.NVELoop(efn, shift, 0.005)
  loop.prime(st, nb)
  etot = []
  for blk in range(2):
    print(f'case c2: NVELoop block {blk} (100 steps)', flush=True)
    loop.run(st, nb, 100)
    assert not bool(nb.did_buffer_overflow), f'overflow in block {blk}'
    E, _ = efn.energy_and_force(st.position, nb)
    etot.append(float(E) + float(jmd.quantity.kinetic_energy(momentum=st.momentum)))
  assert loop.stats['replayed'] >= 1

  # list invariants at step 200
  print('case c2: list invariants', flush=True)
  idx = nb.idx
  rows = torch.arange(N, device=DEV, dtype=torch.int64).unsqueeze(1).expand_as(idx)
  valid = idx < N
  i = rows[valid]
  j = idx[valid].to(torch.int64)
  assert not bool((i == j).any()), 'self pair in the list'
  fwd = torch.sort(i * N + j).values
  bwd = torch.sort(j * N + i).values
  assert torch.equal(fwd, bwd), 'Dense list is not symmetric'
  assert not bool((fwd[1:] == fwd[:-1]).any()), 'duplicate entries'
  ref = nb.reference_position.cpu().numpy()
  rng = np.random.default_rng(0)
  sample = rng.choice(N, 64, replace=False)
  osp = O.Space(np.float32(L))
  c2 = np.float32(3.0) * np.float32(3.0)
  idx_np = idx.cpu().numpy()
  for r in sample:
    d2 = osp.metric_sq(ref, np.broadcast_to(ref[r], ref.shape))
    want = np.flatnonzero(d2 < c2)
    want = np.sort(want[want != r])
    have = np.sort(idx_np[r][idx_np[r] < N])
    assert np.array_equal(want, have), f'row {r}'

  # force at the step-200 positions against FP64
  print('case c2: force at step 200', flush=True)
  X = st.position.cpu().numpy().astype(np.float64)
  _, F64 = c_oracle.LJSystem(X, float(L)).forces()
  rel, worst = _force_err(st.force.cpu().numpy().astype(np.float64), F64)
  print(f'c2 t=200: force rel {rel:.2e} worst/rms {worst:.2e}', flush=True)
  assert rel <= 1e-5 and worst <= 5e-5, (rel, worst)

  # total energy against a 200-step FP64 trajectory from the same start
  print('case c2: FP64 trajectory (CPU)', flush=True)
  sys64 = c_oracle.LJSystem(R.astype(np.float64), float(L))
  P64, _ = sys64.nve(P.astype(np.float64), float(np.float32(0.005)), 200)
  E64, _ = sys64.forces()
  e64_tot = E64 + 0.5 * float(np.sum(P64 ** 2))
  d = abs(etot[-1] - e64_tot) / abs(e64_tot)
  own = abs(etot[1] - etot[0]) / abs(etot[0])
  print(f'c2 E_total(200) gpu {etot[-1]:.6f} fp64 {e64_tot:.6f} rel {d:.2e}; '
        f'own drift 100->200 {own:.2e}', flush=True)
  assert d <= 2e-5, d
  assert own <= 2e-4, own
