"""NVT Nose-Hoover (simulate.nvt_nose_hoover) on the GPU against the FP64
oracle's chain.

Bars: a 40-step trajectory of apply_fn agrees with the oracle's float64
Nose-Hoover step to 1e-5 in positions and 1e-4 (relative) in the chain
momenta; the extended-system energy on the jammed fixture is conserved to
1e-5 over 300 steps (the reference's float32 bar is 1e-3,
simulate_test.py:317-350);
the block loop (NVTLoop) matches apply_fn to 1e-5 and replays its graph bit
for bit; a liquid started hot relaxes to within 10% of the bath temperature
(simulate_test.py:264-304).
"""

import numpy as np
import pytest
import torch

from oracle import jaxmd_oracle as O
from tests.golden import fixtures

pytestmark = pytest.mark.gpu

DEV = 'cuda:0'
F32 = np.float32


def _jmd():
  import jax_md_amd
  return jax_md_amd


def _lj(n_side, seed):
  jmd = _jmd()
  R, P, L = O.lj_lattice(n_side, seed=seed)
  disp, shift = jmd.space.periodic(L)
  nfn, efn = jmd.energy.lennard_jones_neighbor_list(disp, L, format=jmd.partition.Dense)
  return jmd, R, P, L, shift, nfn, efn


def test_nvt_apply_fn_matches_fp64_oracle():
  print('case NVT apply_fn vs oracle', flush=True)
  jmd, R, P, L, shift, nfn, efn = _lj(6, seed=31)
  kT, dt, T = 1.5, 0.005, 40
  Rt = torch.as_tensor(R, device=DEV)
  nb = nfn.allocate(Rt)
  init_fn, apply_fn = jmd.simulate.nvt_nose_hoover(efn, shift, dt, kT)
  st = init_fn(0, Rt, momenta=torch.as_tensor(P, device=DEV), neighbor=nb)
  sp = O.Space(np.float64(L))
  model = O.PairModel('lj', fixed_smoothing=(2.0, 2.5))
  force = lambda X: O.energy_force_brute(model, sp, X)[2]
  o, tau = O.nvt_init(R, P, force, kT, dt)
  for _ in range(T):
    nb = nb.update(st.position)
    st = apply_fn(st, neighbor=nb)
    o = O.nvt_step(o, force, sp, dt, kT, tau)
  dx = np.abs(st.position_hi.cpu().numpy() - o.position)
  dx = np.minimum(dx, L - dx).max()
  c = st.chain.cpu().numpy()
  dp = np.abs(c[1] - o.p_xi).max() / np.abs(o.p_xi).max()
  print(f'  max |dx| {dx:.2e}, chain momenta rel {dp:.2e}, KE {float(st.kinetic_energy):.6f} '
        f'vs {o.ke:.6f}')
  assert dx <= 1e-5
  assert dp <= 1e-4
  assert abs(float(st.kinetic_energy) - o.ke) <= 1e-5 * o.ke


def test_nvt_invariant_jammed():
  print('case NVT invariant jammed', flush=True)
  jmd = _jmd()
  stj = fixtures.jammed_state()
  box = F32(stj['box'][0, 0])
  R = torch.as_tensor(stj['real_position'].astype(F32), device=DEV)
  disp, shift = jmd.space.periodic(box)
  nfn, efn = jmd.energy.soft_sphere_neighbor_list(
      disp, box, species=torch.as_tensor(stj['species'], device=DEV),
      sigma=torch.as_tensor(stj['sigma'], dtype=torch.float32, device=DEV))
  nb = nfn.allocate(R)
  kT = 1e-3
  init_fn, apply_fn = jmd.simulate.nvt_nose_hoover(efn, shift, 1e-3, kT)
  st = init_fn(0, R, neighbor=nb)
  H0 = jmd.simulate.nvt_nose_hoover_invariant(efn, st, kT, neighbor=nb)
  worst = 0.0
  for k in range(300):
    nb = nb.update(st.position)
    st = apply_fn(st, neighbor=nb)
    if k % 30 == 29:
      H = jmd.simulate.nvt_nose_hoover_invariant(efn, st, kT, neighbor=nb)
      worst = max(worst, abs(H - H0) / abs(H0))
  print(f'  invariant drift {worst:.2e}')
  assert worst <= 1e-5


@pytest.mark.parametrize('reorder', [False, True])
def test_nvt_loop_matches_apply_fn_and_replays(reorder):
  print(f'case NVTLoop reorder={reorder}', flush=True)
  jmd, R, P, L, shift, nfn, efn = _lj(12, seed=37)
  kT, dt = 1.2, 0.005
  Rt = torch.as_tensor(R, device=DEV)
  Pt = torch.as_tensor(P, device=DEV)
  init_fn, apply_fn = jmd.simulate.nvt_nose_hoover(efn, shift, dt, kT)
  nb = nfn.allocate(Rt)
  ref = init_fn(0, Rt, momenta=Pt, neighbor=nb)
  for _ in range(30):
    nb = nb.update(ref.position)
    ref = apply_fn(ref, neighbor=nb)
  runs = []
  for graph in (False, True):
    nb = nfn.allocate(Rt)
    st = init_fn(0, Rt, momenta=Pt, neighbor=nb)
    loop = jmd.simulate.NVTLoop(efn, shift, dt, kT, reorder=reorder, graph=graph)
    loop.prime(st, nb)
    snaps = []
    for _ in range(3):
      loop.run(st, nb, 10)
      torch.cuda.synchronize()
      snaps.append((st.position_hi.clone(), st.momentum.clone(), st.chain.clone()))
    runs.append(snaps)
  for blk, (a, b) in enumerate(zip(*runs)):
    for k in range(3):
      assert torch.equal(a[k], b[k]), (reorder, blk, k)
  xh, p, c = runs[1][-1]
  dx = (xh - ref.position_hi).abs()
  dx = torch.minimum(dx, float(L) - dx).max().item()
  dp = (p - ref.momentum).abs().max().item()
  print(f'  loop vs apply_fn: max |dx| {dx:.2e}, max |dp| {dp:.2e}')
  assert dx <= 1e-5 and dp <= 1e-4


def test_nvt_relaxes_to_bath_temperature():
  print('case NVT thermalisation', flush=True)
  jmd, R, P, L, shift, nfn, efn = _lj(16, seed=41)
  kT, dt = 0.8, 0.005
  Rt = torch.as_tensor(R, device=DEV)
  Pt = torch.as_tensor(P * np.sqrt(2.0), device=DEV)  # started at about 2 kT
  init_fn, _ = jmd.simulate.nvt_nose_hoover(efn, shift, dt, kT)
  nb = nfn.allocate(Rt, extra_capacity=8)
  st = init_fn(0, Rt, momenta=Pt, neighbor=nb)
  loop = jmd.simulate.NVTLoop(efn, shift, dt, kT)
  loop.prime(st, nb)
  temps = []
  for _ in range(30):
    loop.run(st, nb, 100)
    assert not bool(nb.did_buffer_overflow)
    temps.append(float(jmd.quantity.temperature(momentum=st.momentum)))
  T_end = float(np.mean(temps[-10:]))
  print(f'  T start {temps[0]:.3f}, mean of the last 1000 steps {T_end:.3f}, bath {kT}')
  assert abs(T_end - kT) / kT < 0.1
