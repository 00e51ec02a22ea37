"""Ragged FIRE ensembles (one launch) against the plain per-system path.

Bars: per system, the allocated capacities and the Dense list equal the
plain ``neighbor_fn.allocate`` (and the oracle) bit for bit; forces at
allocation equal the plain Dense kernel's bit for bit (same lane grouping);
after FIRE steps the FIRE scalars match the plain path and positions agree
to 1e-5 (the float64 FIRE reductions are summed in a different order);
final energies match the FP64 oracle at the final positions to 1e-6.
"""

import numpy as np
import pytest
import torch

from oracle import jaxmd_oracle as O

pytestmark = pytest.mark.gpu

DEV = 'cuda:0'
F32 = np.float32
SIGMA = np.array([[1.0, 1.2], [1.2, 1.4]], F32)


def _jmd():
  import jax_md_amd
  return jax_md_amd


def _systems(sizes, seed=0):
  out = []
  for k, n in enumerate(sizes):
    R, sp, L = O.bidisperse_2d(n, seed=seed + k)
    out.append((R, sp, L))
  return out


def _plain(jmd, R, sp, L, **fire):
  disp, shift = jmd.space.periodic(L)
  nfn, efn = jmd.energy.soft_sphere_neighbor_list(
      disp, L, species=torch.as_tensor(sp, device=DEV), sigma=torch.as_tensor(SIGMA, device=DEV),
      format=jmd.partition.Dense)
  init_fn, apply_fn = jmd.minimize.fire_descent(efn, shift, **fire)
  return nfn, efn, init_fn, apply_fn


FIRE = dict(dt_start=0.01, dt_max=0.1)


def test_ensemble_allocate_matches_plain_lists_and_forces():
  jmd = _jmd()
  sizes = [64, 300, 97, 1024, 150, 8192]  # 8192: 4356 cells
  syst = _systems(sizes, seed=3)
  print('case ensemble allocate', flush=True)
  ens = jmd.ensemble.SoftSphereFireEnsemble([s[0] for s in syst], [s[2] for s in syst],
                                            [s[1] for s in syst], sigma=SIGMA, **FIRE)
  ens.allocate()
  assert int(ens.did_buffer_overflow.max()) == 0
  for s, (R, sp, L) in enumerate(syst):
    nfn, efn, init_fn, _ = _plain(jmd, R, sp, L, **FIRE)
    Rt = torch.as_tensor(R, device=DEV)
    nb = nfn.allocate(Rt)
    assert ens.capacities(s) == (nb.cell_list_capacity, nb.max_occupancy), s
    assert torch.equal(ens.system_neighbors(s), nb.idx), s
    st = init_fn(Rt, neighbor=nb)
    assert torch.equal(ens.force[ens.system_slice(s)], st.force), s
    # oracle list
    spec = O.NeighborListSpec(O.Space(np.float32(L)), L, F32(1.4), 0.2, 1.25, False, True,
                              O.DENSE)
    o = O.neighbor_list_build(spec, R)
    assert np.array_equal(nb.idx.cpu().numpy(), o.idx), s


def test_ensemble_fire_matches_plain_fire():
  jmd = _jmd()
  sizes = [128, 257, 64, 512]
  syst = _systems(sizes, seed=11)
  n_steps = 25
  print('case ensemble fire', flush=True)
  ens = jmd.ensemble.SoftSphereFireEnsemble([s[0] for s in syst], [s[2] for s in syst],
                                            [s[1] for s in syst], sigma=SIGMA, **FIRE)
  ens.allocate().run(n_steps)
  assert int(ens.did_buffer_overflow.max()) == 0
  E = ens.energy.cpu().numpy()
  for s, (R, sp, L) in enumerate(syst):
    nfn, efn, init_fn, apply_fn = _plain(jmd, R, sp, L, **FIRE)
    Rt = torch.as_tensor(R, device=DEV)
    nb = nfn.allocate(Rt)
    st = init_fn(Rt, neighbor=nb)
    for _ in range(n_steps):
      nb = nb.update(st.position)
      st = apply_fn(st, neighbor=nb)
    sl = ens.system_slice(s)
    dx = (ens.position_hi[sl] - st.position_hi).abs().max().item()
    assert dx <= 1e-5, (s, dx)
    assert float(ens.fire_state[s, 0]) == float(st.fire_scalars[0]), s
    assert float(ens.fire_state[s, 1]) == float(st.fire_scalars[1]), s
    assert int(ens.fire_n_pos[s]) == int(st.fire_n_pos[0]), s
    # energy against FP64 at the ensemble's final positions
    model = O.PairModel('soft_sphere', sigma=SIGMA.astype(np.float64), species=sp)
    E64 = O.energy_force_brute(model, O.Space(np.float64(L)),
                               ens.position[sl].cpu().numpy().astype(np.float64))[0]
    assert abs(E[s] - E64) <= 1e-6 * abs(E64) + 1e-9, (s, E[s], E64)


def test_ensemble_rebuilds_and_overflow_flag():
  """Systems rebuild their lists independently (build counts differ), and a
  forced small width raises the overflow bit of the affected systems only."""
  jmd = _jmd()
  syst = _systems([200, 90, 400], seed=21)
  ens = jmd.ensemble.SoftSphereFireEnsemble([s[0] for s in syst], [s[2] for s in syst],
                                            [s[1] for s in syst], sigma=SIGMA, **FIRE)
  ens.allocate().run(60)
  b = ens.builds.cpu().numpy()
  assert np.all(b >= 2), b
  assert int(ens.did_buffer_overflow.max()) == 0


def _sigma_r(r):
  return np.array([[1.0, (1 + r) / 2], [(1 + r) / 2, r]], F32)


DSIGMA = np.array([[0.0, 0.5], [0.5, 1.0]], F32)


def test_ensemble_gradient_matches_fp64_fire_adjoint():
  """d E_s(x_T) / d r_s through T FIRE steps, per system with its own
  diameter ratio r_s (sigma_ab = [[1, (1+r)/2], [(1+r)/2, r]], box fixed),
  against the FP64 oracle's adjoint (itself pinned to central differences).
  The GPU trajectory is float32, so the bar is 1e-5 relative on dE/dr and on
  the projection of dE/dx_0; segment lengths change nothing (bit for bit)."""
  jmd = _jmd()
  sizes, ratios = [64, 150, 97, 256], [1.2, 1.3, 1.45, 1.6]
  syst = [O.bidisperse_2d(n, d_large=r, seed=40 + k) for k, (n, r) in
          enumerate(zip(sizes, ratios))]
  sig = np.stack([_sigma_r(r) for r in ratios])
  T = 30
  print('case ensemble gradient', flush=True)

  def make():
    return jmd.ensemble.SoftSphereFireEnsemble(
        [s[0] for s in syst], [s[2] for s in syst], [s[1] for s in syst], sigma=sig,
        **FIRE).allocate()

  ens = make()
  E, th, X = ens.energy_gradient(T, DSIGMA, checkpoint_every=7, scale=1.0)
  assert int(ens.did_buffer_overflow.max()) == 0
  # the ensemble is left at its final state
  ref = make().run(T)
  assert torch.equal(ens.position, ref.position)
  assert torch.equal(ens.energy, ref.energy)
  # segment length does not change the result
  E2, th2, X2 = make().energy_gradient(T, DSIGMA, checkpoint_every=T, scale=1.0)
  assert torch.equal(th, th2) and torch.equal(X, X2)
  th, X = th.cpu().numpy(), X.cpu().numpy()
  par = O.FireParams(dt_start=0.01, dt_max=0.1)
  v = np.random.default_rng(9).normal(size=X.shape)
  for s, ((R, sp, L), r) in enumerate(zip(syst, ratios)):
    model = O.PairModel('soft_sphere', sigma=_sigma_r(r).astype(np.float64), species=sp)
    Eo, tho, Xo = O.fire_energy_gradient(model, O.Space(np.float64(L)),
                                         R.astype(np.float64), T, par, DSIGMA)
    sl = ens.system_slice(s)
    print(s, th[s], tho, float(E[s]), Eo, flush=True)
    assert abs(th[s] - tho) <= 1e-5 * abs(tho), (s, th[s], tho)
    px, po = float(np.sum(X[sl] * v[sl])), float(np.sum(Xo * v[sl]))
    assert abs(px - po) <= 1e-5 * np.linalg.norm(Xo) * np.linalg.norm(v[sl]), (s, px, po)


def test_ensemble_gradient_with_box_following_ratio():
  """The meta-optimisation notebook's derivative: the box side follows the
  diameter ratio, L(r) = sqrt(N pi (1 + r^2) / (8 phi)), with x_0 held in
  real coordinates.  Per system, d E_s(x_T) / d side and the total
  d E_s / d r = (fixed-box part) + dL/dr * d E_s / d side against the FP64
  oracle's box adjoint (pinned to central differences over the side and over
  r with both sigma and L moving); 1e-5 relative, as the fixed-box test.  The
  oracle's runs fold particles across the boundary, so the shift's fold term
  is exercised."""
  jmd = _jmd()
  sizes, ratios = [64, 150, 97, 256], [1.2, 1.3, 1.45, 1.6]
  syst = [O.bidisperse_2d(n, d_large=r, seed=40 + k) for k, (n, r) in
          enumerate(zip(sizes, ratios))]
  sig = np.stack([_sigma_r(r) for r in ratios])
  T, phi = 30, 0.84
  dLdr = np.array([np.sqrt(n * np.pi * (1 + r * r) / (8 * phi)) * r / (1 + r * r)
                   for n, r in zip(sizes, ratios)])
  print('case ensemble gradient, box following the ratio', flush=True)

  def make():
    return jmd.ensemble.SoftSphereFireEnsemble(
        [s[0] for s in syst], [s[2] for s in syst], [s[1] for s in syst], sigma=sig,
        **FIRE).allocate()

  _, th_fixed, X_fixed = make().energy_gradient(T, DSIGMA, checkpoint_every=7, scale=1.0)
  ens = make()
  E, th, X = ens.energy_gradient(T, DSIGMA, checkpoint_every=7, scale=1.0, dbox=dLdr)
  assert int(ens.did_buffer_overflow.max()) == 0
  lb = ens.box_adjoint.cpu().numpy()
  # the box terms leave dE/dx_0 and the fixed-box part as they were
  assert torch.equal(X, X_fixed)
  assert np.array_equal(th.cpu().numpy(), th_fixed.cpu().numpy() + dLdr * lb)
  th = th.cpu().numpy()
  par = O.FireParams(dt_start=0.01, dt_max=0.1)
  folds = 0
  for s, ((R, sp, L), r) in enumerate(zip(syst, ratios)):
    model = O.PairModel('soft_sphere', sigma=_sigma_r(r).astype(np.float64), species=sp)
    space = O.Space(np.float64(L))
    R64 = R.astype(np.float64)
    _, tho, _, lbo = O.fire_energy_gradient(model, space, R64, T, par, DSIGMA, box=True)
    force = lambda x: O.energy_force_brute(model, space, x)[2]
    st, tape = O.fire_init(R64, force, par), []
    for _ in range(T):
      st = O.fire_step(st, force, space, par, tape)
    folds += sum(int(np.abs(np.rint((t['x_prev'] - t['x']) / np.float64(L))).sum())
                 for t in tape)
    print(s, lb[s], lbo, th[s], tho + dLdr[s] * lbo, flush=True)
    assert abs(lb[s] - lbo) <= 1e-5 * abs(lbo), (s, lb[s], lbo)
    total = tho + dLdr[s] * lbo
    assert abs(th[s] - total) <= 1e-5 * abs(total), (s, th[s], total)
  assert folds >= 3


def test_ensemble_gradient_mean_energy_scale():
  """scale defaults to 1 / S (the mean final energy)."""
  jmd = _jmd()
  syst = [O.bidisperse_2d(n, d_large=1.4, seed=60 + k) for k, n in enumerate([80, 120])]
  sig = np.stack([_sigma_r(1.4)] * 2)

  def make():
    return jmd.ensemble.SoftSphereFireEnsemble(
        [s[0] for s in syst], [s[2] for s in syst], [s[1] for s in syst], sigma=sig,
        **FIRE).allocate()

  _, th_mean, X_mean = make().energy_gradient(12, DSIGMA, checkpoint_every=5)
  _, th_sum, X_sum = make().energy_gradient(12, DSIGMA, checkpoint_every=5, scale=1.0)
  assert torch.allclose(th_mean * 2, th_sum, rtol=1e-12, atol=0)
  assert torch.allclose(X_mean * 2, X_sum, rtol=1e-6, atol=1e-12)


def test_ensemble_large_systems_thread_counts_agree():
  """Systems of 4096 particles or more run with 1024 threads per system (the
  size rule), smaller ones with 512: on two systems of 4200 and 4600
  particles, allocation (capacities, lists, forces) is bit-identical for 512
  and 1024 threads; 40 FIRE steps give the same float16 dt / alpha
  sequence, energies within 1e-7 and positions within 1e-5; the gradient
  with the box following the ratio agrees to 1e-6.  The threads knob only
  changes the order of the workgroup sums.  At 512 threads both systems
  exceed the LDS staging limit (8 particles per thread) and run the
  global-memory kernel instances; at 1024 they are staged in LDS, so this
  also compares the two instances (and runs the gradient's overlapped
  segment replays, three segments)."""
  jmd = _jmd()
  from jax_md_amd import _native as nat
  lib = nat.load()
  sizes, ratios = [4200, 4600], [1.3, 1.5]
  syst = [O.bidisperse_2d(n, d_large=r, seed=80 + k) for k, (n, r) in
          enumerate(zip(sizes, ratios))]
  sig = np.stack([_sigma_r(r) for r in ratios])
  dLdr = np.array([np.sqrt(n * np.pi * (1 + r * r) / (8 * 0.84)) * r / (1 + r * r)
                   for n, r in zip(sizes, ratios)])
  print('case ensemble large systems, 512 vs 1024 threads', flush=True)

  def make():
    return jmd.ensemble.SoftSphereFireEnsemble(
        [s[0] for s in syst], [s[2] for s in syst], [s[1] for s in syst], sigma=sig,
        **FIRE).allocate()

  out = {}
  try:
    for threads in (512, 1024):
      nat.check(lib.jmd_fire_ensemble_set_threads(threads), 'set_threads')
      a = make()
      alloc = (a.idx.clone(), a.force.clone(), [a.capacities(s) for s in range(2)])
      a.run(40)
      fwd = (a.position.clone(), a.energy.clone(), a.fire_state.clone())
      g = make()
      _, th, _ = g.energy_gradient(30, DSIGMA, checkpoint_every=10, scale=1.0, dbox=dLdr)
      out[threads] = (alloc, fwd, th.cpu().numpy(), g.box_adjoint.cpu().numpy())
  finally:
    nat.check(lib.jmd_fire_ensemble_set_threads(0), 'set_threads')
  (al5, fw5, th5, lb5), (al1, fw1, th1, lb1) = out[512], out[1024]
  assert al5[2] == al1[2]
  assert torch.equal(al5[0], al1[0]) and torch.equal(al5[1], al1[1])
  assert torch.equal(fw5[2], fw1[2])
  assert torch.allclose(fw5[1], fw1[1], rtol=1e-7, atol=0)
  assert float((fw5[0] - fw1[0]).abs().max()) <= 1e-5
  print(th5, th1, lb5, lb1, flush=True)
  assert np.allclose(th5, th1, rtol=1e-6, atol=0)
  assert np.allclose(lb5, lb1, rtol=1e-6, atol=1e-9)
