"""Adjoint kernels on the GPU against the FP64 oracle and finite differences.

Bars: the force VJP (J v, the scalar-parameter cotangents and the
per-particle ones) within 1e-5 norm-wise of the FP64 oracle on the same
pairs; Sparse-format lists give the Dense list's VJP to 1e-6; the
checkpointed NVE trajectory gradient within 1e-4 of central finite
differences (h = 1e-4, FP64 oracle trajectory with brute-force forces), also
with respect to per-particle diameters over an OrderedSparse list.
"""

import dataclasses

import numpy as np
import pytest
import torch

from oracle import jaxmd_oracle as O

pytestmark = pytest.mark.gpu

DEV = 'cuda:0'
F32 = np.float32


def _jmd():
  import jax_md_amd
  return jax_md_amd


def _lj_system(n_side, seed, jitter=0.05):
  jmd = _jmd()
  R, P, L = O.lj_lattice(n_side, seed=seed, jitter=jitter)
  disp, shift = jmd.space.periodic(L)
  nfn, efn = jmd.energy.lennard_jones_neighbor_list(disp, L, format=jmd.partition.Dense)
  return jmd, R, P, L, shift, nfn, efn


@pytest.mark.parametrize('n_side', [8, 12])
def test_lj_force_vjp_matches_oracle(n_side):
  print(f'case LJ force vjp n_side={n_side}', flush=True)
  jmd, R, _, L, _, nfn, efn = _lj_system(n_side, seed=11)
  Rt = torch.as_tensor(R, device=DEV)
  nb = nfn.allocate(Rt)
  v = np.random.default_rng(3).normal(size=R.shape).astype(F32)
  Jv, pc, eg = jmd.adjoint.force_vjp(efn, Rt, nb, torch.as_tensor(v, device=DEV))
  i, j = O.pairs_from_dense(nb.idx.cpu().numpy())
  model = O.PairModel('lj', fixed_smoothing=(2.0, 2.5))
  want, (ts, te), (es, ee) = O.force_vjp(model, O.Space(np.float64(L)), R, v, i, j)
  got = Jv.cpu().numpy().astype(np.float64)
  rel = np.linalg.norm(got - want) / np.linalg.norm(want)
  print(f'  |Jv| rel err {rel:.2e}')
  assert rel <= 1e-5
  pc = pc.cpu().numpy()
  eg = eg.cpu().numpy()
  scale_t = np.sqrt(np.sum(v.astype(np.float64) ** 2)) * np.sqrt(len(i))
  assert abs(pc[0] - ts) <= 1e-5 * (abs(ts) + scale_t)
  assert abs(pc[1] - te) <= 1e-5 * (abs(te) + scale_t)
  assert abs(eg[0] - es) <= 1e-5 * abs(es) + 1e-6
  assert abs(eg[1] - ee) <= 1e-5 * abs(ee) + 1e-6


def test_soft_sphere_force_vjp_2d():
  """Soft spheres (alpha = 2) in 2-D: J v against the oracle; the parameter
  cotangents for a scalar sigma, and one per particle for per-particle
  diameters (against the oracle's pair-scattered form)."""
  jmd = _jmd()
  rng = np.random.default_rng(8)
  L = F32(14.0)
  R = (L * rng.uniform(size=(300, 2))).astype(F32)
  disp, _ = jmd.space.periodic(L)
  v = rng.normal(size=R.shape).astype(F32)
  sp = O.Space(np.float64(L))
  for sigma in (F32(1.2), (1.0 + 0.4 * (rng.uniform(size=300) < 0.5)).astype(F32)):
    scalar = np.ndim(sigma) == 0
    print(f'case SS vjp scalar_sigma={scalar}', flush=True)
    sig_t = float(sigma) if scalar else torch.as_tensor(sigma, device=DEV)
    nfn, efn = jmd.energy.soft_sphere_neighbor_list(disp, L, sigma=sig_t,
                                                      format=jmd.partition.Dense)
    Rt = torch.as_tensor(R, device=DEV)
    nb = nfn.allocate(Rt)
    i, j = O.pairs_from_dense(nb.idx.cpu().numpy())
    model = O.PairModel('soft_sphere', sigma=np.float64(sigma) if scalar
                        else sigma.astype(np.float64))
    want, (ts, te), _ = O.force_vjp(model, sp, R, v, i, j)
    if scalar:
      Jv, pc, _ = jmd.adjoint.force_vjp(efn, Rt, nb, torch.as_tensor(v, device=DEV))
      pc = pc.cpu().numpy()
      scale = np.sqrt(np.sum(v.astype(np.float64) ** 2)) * np.sqrt(len(i))
      assert abs(pc[0] - ts) <= 1e-5 * (abs(ts) + scale)
      assert abs(pc[1] - te) <= 1e-5 * (abs(te) + scale)
    else:
      Jv, pc, eg = jmd.adjoint.force_vjp(efn, Rt, nb, torch.as_tensor(v, device=DEV))
      pc, eg = pc.cpu().numpy(), eg.cpu().numpy()
      assert pc.shape == (301,) and eg.shape == (301,)  # [sigma_i] + [epsilon]
      cs, ce, es, ee = O.force_vjp_per_particle(model, sp, R, v, i, j)
      assert np.linalg.norm(pc[:300] - cs) <= 1e-5 * np.linalg.norm(cs)
      assert np.linalg.norm(eg[:300] - es) <= 1e-5 * np.linalg.norm(es)
      scale = np.sqrt(np.sum(v.astype(np.float64) ** 2)) * np.sqrt(len(i))
      assert abs(pc[300] - te) <= 1e-5 * (abs(te) + scale)  # the scalar epsilon
    got = Jv.cpu().numpy().astype(np.float64)
    assert np.linalg.norm(got - want) <= 1e-5 * np.linalg.norm(want)


def test_lj_force_vjp_per_particle_parameters():
  """Per-particle sigma and epsilon (LJ, 3-D): one cotangent per particle for
  each, against the oracle's pair-scattered FP64 form (1e-5 norm-wise), and
  through ForceFunction's backward."""
  print('case LJ per-particle vjp', flush=True)
  jmd = _jmd()
  R, _, L = O.lj_lattice(8, seed=14)
  N = R.shape[0]
  rng = np.random.default_rng(15)
  sig = (1.0 + 0.05 * rng.uniform(-1, 1, N)).astype(F32)
  eps = (1.0 + 0.2 * rng.uniform(-1, 1, N)).astype(F32)
  disp, _ = jmd.space.periodic(L)
  nfn, efn = jmd.energy.lennard_jones_neighbor_list(
      disp, L, sigma=torch.as_tensor(sig, device=DEV), epsilon=torch.as_tensor(eps, device=DEV),
      format=jmd.partition.Dense)
  Rt = torch.as_tensor(R, device=DEV)
  nb = nfn.allocate(Rt)
  v = rng.normal(size=R.shape).astype(F32)
  Jv, pc, eg = jmd.adjoint.force_vjp(efn, Rt, nb, torch.as_tensor(v, device=DEV))
  i, j = O.pairs_from_dense(nb.idx.cpu().numpy())
  smax = float(np.max(sig))
  model = O.PairModel('lj', sigma=sig.astype(np.float64), epsilon=eps.astype(np.float64),
                      fixed_smoothing=(float(F32(2.0 * smax)), float(F32(2.5 * smax))))
  sp = O.Space(np.float64(L))
  want, _, _ = O.force_vjp(model, sp, R, v, i, j)
  cs, ce, es, ee = O.force_vjp_per_particle(model, sp, R, v, i, j)
  pc, eg = pc.cpu().numpy(), eg.cpu().numpy()
  assert pc.shape == (2 * N,)
  for got, want_k, name in ((pc[:N], cs, 'sigma'), (pc[N:], ce, 'epsilon'),
                            (eg[:N], es, 'dE/dsigma'), (eg[N:], ee, 'dE/deps')):
    rel = np.linalg.norm(got - want_k) / np.linalg.norm(want_k)
    print(f'  {name}: rel err {rel:.2e}')
    assert rel <= 1e-5
  assert np.linalg.norm(Jv.cpu().numpy() - want) <= 1e-5 * np.linalg.norm(want)
  # ForceFunction with per-particle tensors
  sig_t = torch.as_tensor(sig, device=DEV).requires_grad_(True)
  force = jmd.adjoint.differentiable_force(efn)
  F = force(Rt, nb, sig_t, None)
  (F * torch.as_tensor(v, device=DEV)).sum().backward()
  assert sig_t.grad.shape == (N,)
  assert np.linalg.norm(sig_t.grad.cpu().numpy() - cs) <= 1e-5 * np.linalg.norm(cs)


@pytest.mark.parametrize('fmt', ['Sparse', 'OrderedSparse'])
def test_force_vjp_over_sparse_lists(fmt):
  """The VJP over a Sparse-format list equals the VJP over the Dense list of
  the same pairs (J v and the scalar parameter cotangents, 1e-6)."""
  print(f'case vjp {fmt}', flush=True)
  jmd, R, _, L, _, nfn, efn = _lj_system(8, seed=16)
  disp, _ = jmd.space.periodic(L)
  nfs, efs = jmd.energy.lennard_jones_neighbor_list(disp, L, format=getattr(jmd.partition, fmt))
  Rt = torch.as_tensor(R, device=DEV)
  v = torch.as_tensor(np.random.default_rng(17).normal(size=R.shape).astype(F32), device=DEV)
  Jd, pd, ed = jmd.adjoint.force_vjp(efn, Rt, nfn.allocate(Rt), v)
  Js, ps, es = jmd.adjoint.force_vjp(efs, Rt, nfs.allocate(Rt), v)
  assert torch.linalg.norm(Js - Jd) <= 1e-6 * torch.linalg.norm(Jd)
  assert torch.allclose(ps, pd, rtol=1e-6, atol=0) and torch.allclose(es, ed, rtol=1e-6, atol=0)


def test_trajectory_gradient_per_particle_diameters():
  """d E(x_T) / d(diameters) through 30 NVE steps of 2-D bidisperse soft
  spheres (N = 256, per-particle diameters, an OrderedSparse list) against
  central differences of the FP64 oracle trajectory along random directions
  of the diameter vector (h = 1e-4), 1e-4 relative."""
  print('case trajectory gradient diameters', flush=True)
  jmd = _jmd()
  R, species, L = O.bidisperse_2d(256, seed=18)
  N = R.shape[0]
  diam = np.where(species == 0, 1.0, 1.4).astype(F32)
  rng = np.random.default_rng(19)
  P = (0.3 * rng.normal(size=R.shape)).astype(F32)
  P -= P.mean(0, keepdims=True)
  disp, shift = jmd.space.periodic(L)
  nfn, efn = jmd.energy.soft_sphere_neighbor_list(
      disp, L, sigma=torch.as_tensor(diam, device=DEV), format=jmd.partition.OrderedSparse)
  dt, T = 0.01, 30
  res = jmd.adjoint.nve_energy_gradient(efn, shift, nfn, torch.as_tensor(R, device=DEV),
                                        torch.as_tensor(P, device=DEV), dt, T, checkpoint_every=10)
  g = res.sigma.cpu().numpy()
  assert g.shape == (N,)
  sp = O.Space(np.float64(L))
  R64, P64 = R.astype(np.float64), P.astype(np.float64)
  h = 1e-4
  for trial in range(3):
    u = rng.normal(size=N)
    u /= np.linalg.norm(u)
    mp = O.PairModel('soft_sphere', sigma=diam.astype(np.float64) + h * u)
    mm = O.PairModel('soft_sphere', sigma=diam.astype(np.float64) - h * u)
    fd = (O.nve_energy_after(mp, sp, R64, P64, dt, T)
          - O.nve_energy_after(mm, sp, R64, P64, dt, T)) / (2 * h)
    ad = float(np.sum(g * u))
    print(f'  direction {trial}: adjoint {ad:.6e} fd {fd:.6e} |grad| {np.linalg.norm(g):.3e}')
    assert abs(ad - fd) <= 1e-4 * max(abs(fd), np.linalg.norm(g))


def test_force_function_autograd():
  """ForceFunction (the custom-VJP force): backward of <g, F> gives J g and
  the parameter cotangents of the kernel."""
  jmd, R, _, L, _, nfn, efn = _lj_system(8, seed=12)
  Rt = torch.as_tensor(R, device=DEV).requires_grad_(True)
  nb = nfn.allocate(Rt.detach())
  sig = torch.tensor(1.0, device=DEV, requires_grad=True)
  eps = torch.tensor(1.0, device=DEV, requires_grad=True)
  force = jmd.adjoint.differentiable_force(efn)
  F = force(Rt, nb, sig, eps)
  g = torch.as_tensor(np.random.default_rng(4).normal(size=R.shape).astype(F32), device=DEV)
  (F * g).sum().backward()
  Jv, pc, _ = jmd.adjoint.force_vjp(efn, Rt.detach(), nb, g)
  assert torch.equal(Rt.grad, Jv)
  assert abs(sig.grad.item() - float(pc[0])) <= 1e-6 * abs(float(pc[0])) + 1e-6
  assert abs(eps.grad.item() - float(pc[1])) <= 1e-6 * abs(float(pc[1])) + 1e-6


@pytest.mark.parametrize('n_steps,every', [(30, 10), (37, 8)])
def test_nve_trajectory_gradient_vs_finite_differences(n_steps, every):
  """d E(x_T) / d(x_0, p_0, sigma, epsilon) against central differences of the
  FP64 oracle trajectory (h = 1e-4 along random unit directions)."""
  print(f'case trajectory gradient T={n_steps} K={every}', flush=True)
  jmd, R, P, L, shift, nfn, efn = _lj_system(6, seed=13)
  Rt = torch.as_tensor(R, device=DEV)
  Pt = torch.as_tensor(P, device=DEV)
  dt = 0.005
  res = jmd.adjoint.nve_energy_gradient(efn, shift, nfn, Rt, Pt, dt, n_steps,
                                        checkpoint_every=every)
  sp = O.Space(np.float64(L))
  model = O.PairModel('lj', fixed_smoothing=(2.0, 2.5))
  R64, P64 = R.astype(np.float64), P.astype(np.float64)
  E64 = O.nve_energy_after(model, sp, R64, P64, dt, n_steps)
  assert abs(res.energy - E64) <= 1e-5 * abs(E64)
  h = 1e-4
  rng = np.random.default_rng(21)
  for name in ('position', 'momentum'):
    u = rng.normal(size=R.shape)
    u /= np.linalg.norm(u)
    if name == 'position':
      ep = O.nve_energy_after(model, sp, R64 + h * u, P64, dt, n_steps)
      em = O.nve_energy_after(model, sp, R64 - h * u, P64, dt, n_steps)
    else:
      ep = O.nve_energy_after(model, sp, R64, P64 + h * u, dt, n_steps)
      em = O.nve_energy_after(model, sp, R64, P64 - h * u, dt, n_steps)
    fd = (ep - em) / (2 * h)
    g = getattr(res, name).cpu().numpy().astype(np.float64)
    ad = float(np.sum(g * u))
    scale = np.linalg.norm(g)
    print(f'  d/d{name}: adjoint {ad:.6e} fd {fd:.6e} |grad| {scale:.3e}')
    assert abs(ad - fd) <= 1e-4 * max(abs(fd), scale)
  for name in ('sigma', 'epsilon'):
    mp = dataclasses.replace(model, **{name: 1.0 + h})
    mm = dataclasses.replace(model, **{name: 1.0 - h})
    fd = (O.nve_energy_after(mp, sp, R64, P64, dt, n_steps)
          - O.nve_energy_after(mm, sp, R64, P64, dt, n_steps)) / (2 * h)
    ad = getattr(res, name)
    print(f'  d/d{name}: adjoint {ad:.6e} fd {fd:.6e}')
    assert abs(ad - fd) <= 1e-4 * abs(fd)


@pytest.mark.parametrize('x64', [False, True])
def test_trajectory_gradient_graph_replays_equal_eager(x64):
  """The graph-replayed segments (forward and recording replays) give the
  eager run's gradient bit for bit, over several segments of each kind and
  a shorter last one, for float32 and x64 trajectories."""
  print(f'case trajectory gradient graphs x64={x64}', flush=True)
  jmd, R, P, L, shift, nfn, efn = _lj_system(6, seed=19)
  dt = 0.005
  out = []
  for graph in (False, True):
    Rt = torch.as_tensor(R.astype(np.float64) if x64 else R, device=DEV)
    Pt = torch.as_tensor(P.astype(np.float64) if x64 else P, device=DEV)
    res = jmd.adjoint.nve_energy_gradient(efn, shift, nfn, Rt, Pt, dt, 47,
                                          checkpoint_every=10, graph=graph)
    out.append(res)
  a, b = out
  assert a.energy == b.energy
  assert torch.equal(a.position, b.position) and torch.equal(a.momentum, b.momentum)
  assert a.sigma == b.sigma and a.epsilon == b.epsilon


def test_x64_force_and_trajectory_match_fp64():
  """The float64 path: forces against the FP64 oracle to 1e-12 (norm-wise)
  and a 60-step trajectory against the oracle's float64 velocity Verlet to
  1e-9 in positions."""
  print('case x64 force and trajectory', flush=True)
  jmd, R, P, L, shift, nfn, efn = _lj_system(6, seed=17)
  R64, P64 = R.astype(np.float64), P.astype(np.float64)
  Rt = torch.as_tensor(R64, device=DEV)
  nb = nfn.allocate(Rt.to(torch.float32))
  E, F = efn.energy_and_force(Rt, nb)
  assert F.dtype == torch.float64
  sp = O.Space(np.float64(L))
  model = O.PairModel('lj', fixed_smoothing=(2.0, 2.5))
  E64, _, F64 = O.energy_force_brute(model, sp, R64)
  err = np.linalg.norm(F.cpu().numpy() - F64) / np.linalg.norm(F64)
  print(f'  force rel err {err:.2e}, energy {float(E):.12f} vs {E64:.12f}')
  assert err <= 1e-12
  assert abs(float(E) - E64) <= 1e-12 * abs(E64)
  init_fn, _ = jmd.simulate.nve(efn, shift, 0.005)
  st = init_fn(0, Rt, kT=0.0, momenta=torch.as_tensor(P64, device=DEV), neighbor=nb)
  assert st.x64
  loop = jmd.simulate.NVELoop(efn, shift, 0.005)
  loop.prime(st, nb)
  loop.run(st, nb, 60)
  o = O.NVEState(R64, P64, O.energy_force_brute(model, sp, R64)[2])
  for _ in range(60):
    o = O.nve_step(o, lambda X: O.energy_force_brute(model, sp, X)[2], sp, 0.005)
  dx = np.abs(st.position_hi.cpu().numpy() - o.position)
  dx = np.minimum(dx, L - dx).max()
  dp = np.abs(st.momentum.cpu().numpy() - o.momentum).max()
  print(f'  60 steps: max |dx| {dx:.2e}, max |dp| {dp:.2e}')
  assert dx <= 1e-9 and dp <= 1e-8


def test_x64_trajectory_gradient_long_horizon():
  """Over 200 steps the float32 trajectory's rounding is amplified past the
  1e-4 bar against FP64 finite differences (chaos); the x64 form keeps the
  gradient within 1e-4 of converged central differences (h = 1e-7)."""
  print('case x64 trajectory gradient T=200', flush=True)
  jmd, R, P, L, shift, nfn, efn = _lj_system(8, seed=23)
  R64, P64 = R.astype(np.float64), P.astype(np.float64)
  dt, T = 0.005, 200
  res = jmd.adjoint.nve_energy_gradient(efn, shift, nfn, torch.as_tensor(R64, device=DEV),
                                        torch.as_tensor(P64, device=DEV), dt, T,
                                        checkpoint_every=20)
  sp = O.Space(np.float64(L))
  model = O.PairModel('lj', fixed_smoothing=(2.0, 2.5))
  E64 = O.nve_energy_after(model, sp, R64, P64, dt, T)
  print(f'  E_T {res.energy:.10f} vs FP64 {E64:.10f}')
  assert abs(res.energy - E64) <= 1e-9 * abs(E64)
  h = 1e-7
  for name in ('sigma', 'epsilon'):
    mp = dataclasses.replace(model, **{name: 1.0 + h})
    mm = dataclasses.replace(model, **{name: 1.0 - h})
    fd = (O.nve_energy_after(mp, sp, R64, P64, dt, T)
          - O.nve_energy_after(mm, sp, R64, P64, dt, T)) / (2 * h)
    ad = getattr(res, name)
    print(f'  d/d{name}: adjoint {ad:.8e} fd {fd:.8e} rel {abs(ad - fd) / abs(fd):.2e}')
    assert abs(ad - fd) <= 1e-4 * abs(fd)
  u = np.random.default_rng(29).normal(size=R.shape)
  u /= np.linalg.norm(u)
  fd = (O.nve_energy_after(model, sp, R64 + h * u, P64, dt, T)
        - O.nve_energy_after(model, sp, R64 - h * u, P64, dt, T)) / (2 * h)
  g = res.position.cpu().numpy()
  ad = float(np.sum(g * u))
  print(f'  d/dx0 . u: adjoint {ad:.8e} fd {fd:.8e} |grad| {np.linalg.norm(g):.3e}')
  assert abs(ad - fd) <= 1e-4 * max(abs(fd), np.linalg.norm(g))
