"""NVT Nose-Hoover (simulate.nvt_nose_hoover) on the GPU against the FP64
oracle's chain.

Bars: a 40-step trajectory of apply_fn agrees with the oracle's float64
Nose-Hoover step to 1e-5 in positions and 1e-4 (relative) in the chain
momenta; the extended-system energy on the jammed fixture is conserved to
1e-5 over 300 steps (the reference's float32 bar is 1e-3,
simulate_test.py:317-350);
the block loop (NVTLoop) matches apply_fn to 1e-5 and replays its graph bit
for bit; a liquid started hot relaxes to within 10% of the bath temperature
(simulate_test.py:264-304); the fused kick + kinetic energy + chain launch
equals the two separate launches bit for bit; the chain's half steps alone
match the oracle's float64 chain to 1e-12 (polynomial range of the
exponentials) and 1e-10 (library range).
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


@pytest.mark.parametrize('n,mode', [(1000, 3), (70001, 3), (70001, 1), (300, 5)])
def test_fused_kick_kinetic_chain_equals_two_launches(n, mode):
  """jmd_kick_kinetic_nhc against jmd_kick_kinetic + jmd_nhc_step, bit for bit,
  on several consecutive calls sharing one ticket (several hundred workgroups
  at n = 70001)."""
  print(f'case fused kick+chain n={n} mode={mode}', flush=True)
  jmd = _jmd()
  nat = jmd._native
  lib = nat.load()
  rng = np.random.default_rng(n + mode)
  dim, L = 3, 5
  P0 = torch.as_tensor(rng.normal(size=(n, dim)).astype(F32), device=DEV)
  F = torch.as_tensor(rng.normal(size=(n, dim)).astype(F32), device=DEV)
  mass = torch.as_tensor(rng.uniform(0.5, 2.0, size=n).astype(F32), device=DEV)
  nparts = int(lib.jmd_kick_kinetic_partials(n))
  out = []
  for fused in (False, True):
    P = P0.clone()
    chain = torch.zeros((3, L), dtype=torch.float64, device=DEV)
    chain[1] = torch.linspace(-0.3, 0.4, L, dtype=torch.float64)
    chain[2] = 1.0
    ke = torch.zeros((1,), dtype=torch.float64, device=DEV)
    part = torch.zeros((nparts,), dtype=torch.float64, device=DEV)
    scale = torch.ones((1,), dtype=torch.float64, device=DEV)
    ticket = torch.zeros((1,), dtype=torch.int32, device=DEV)
    st = nat.stream_handle(P.device)
    scales = []
    for _ in range(3):
      args = (float(F32(0.005)), float(F32(1.2)), float(F32(0.5)), float(n * dim))
      if fused:
        nat.check(lib.jmd_kick_kinetic_nhc(
            nat.ptr(P), nat.ptr(F), nat.ptr(mass), 0.0, n, dim, float(F32(0.0025)),
            nat.ptr(part), nat.ptr(ticket), mode, nat.ptr(chain), L, 2, 3, *args[:3], args[3],
            nat.ptr(ke), nat.ptr(scale), st), 'jmd_kick_kinetic_nhc')
      else:
        nat.check(lib.jmd_kick_kinetic(nat.ptr(P), nat.ptr(F), nat.ptr(mass), 0.0, n, dim,
                                       float(F32(0.0025)), nat.ptr(part), st), 'kick')
        nat.check(lib.jmd_nhc_step(mode, nat.ptr(chain), L, 2, 3, *args[:3], args[3],
                                   nat.ptr(ke), nat.ptr(part), nparts, nat.ptr(scale), st),
                  'nhc')
      scales.append(scale.clone())
    torch.cuda.synchronize()
    out.append((P.cpu(), chain.cpu(), ke.cpu(), torch.cat(scales).cpu(), int(ticket.item())))
  a, b = out
  print(f'  scale {a[3].tolist()}, ke {float(a[2]):.10f}')
  assert b[4] == 0
  assert torch.equal(a[0], b[0])
  for x, y in zip(a[1:4], b[1:4]):
    assert torch.equal(x, y)


@pytest.mark.parametrize('rate,substeps,mode', [(0.3, (2, 3), 3), (200.0, (1, 1), 1),
                                                (400.0, (1, 1), 1)])
def test_chain_half_steps_vs_oracle(rate, substeps, mode):
  """jmd_nhc_step against the oracle's float64 chain (simulate.py:492-515).
  rate = p_xi / Q sets the exponentials' arguments: about 1e-4 (the Taylor
  polynomial's range) over two half steps of 2 x 3 substeps, and 0.12 / 0.25
  (the library exp; the sweeps then also reach arguments beyond 708, where
  the factor underflows to 0) in one substep.  Bar: 1e-12 relative in the
  polynomial range, 1e-10 beyond."""
  print(f'case chain half steps rate={rate} substeps={substeps} mode={mode}', flush=True)
  jmd = _jmd()
  nat = jmd._native
  lib = nat.load()
  L, n, dim = 5, 100, 3
  dof = float(n * dim)
  kT, dt, tau = F32(1.2), F32(0.005), F32(0.5)
  Q = O.nhc_masses(L, kT, tau, dof)
  p_xi = np.linspace(-1.0, 1.0, L) * rate * Q
  xi = 0.1 * np.arange(L, dtype=np.float64)
  ke = 0.5 * dof * float(kT) * 1.3
  chain = torch.as_tensor(np.stack([xi, p_xi, Q]), device=DEV)
  part = torch.as_tensor([ke], dtype=torch.float64, device=DEV)
  kin = torch.zeros((1,), dtype=torch.float64, device=DEV)
  scale = torch.zeros((1,), dtype=torch.float64, device=DEV)
  cs, sy = substeps
  nat.check(lib.jmd_nhc_step(mode, nat.ptr(chain), L, cs, sy, float(dt), float(kT), float(tau),
                             dof, nat.ptr(kin), nat.ptr(part), 1, nat.ptr(scale),
                             nat.stream_handle(chain.device)), 'jmd_nhc_step')
  torch.cuda.synchronize()
  k64 = np.float64(kT)
  x2, p2, ke2, s = O.nhc_half_step(dt, cs, sy, xi, p_xi, Q, ke, dof, k64)
  if mode & 2:
    x2, p2, ke2, s2 = O.nhc_half_step(dt, cs, sy, x2, p2, Q, ke2, dof, k64)
    s *= s2
  c = chain.cpu().numpy()
  tol = 1e-12 if rate < 1 else 1e-10
  errs = (np.abs(c[1] - p2).max() / np.abs(p2).max(), np.abs(c[0] - x2).max() / np.abs(x2).max(),
          abs(float(kin) - ke2) / ke2, abs(float(scale) - s) / s)
  print(f'  rel errors p_xi {errs[0]:.2e} xi {errs[1]:.2e} ke {errs[2]:.2e} scale {errs[3]:.2e}')
  assert all(np.isfinite(errs)) and max(errs) < tol
