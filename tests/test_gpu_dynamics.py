"""NVE and FIRE on the GPU against the float64 oracle integrators.

Bars: after 100 NVE steps (dt = 0.005, LJ, N = 512) positions within 1e-4
and momenta within 1e-3 of the float64 trajectory started from the same
float32 state; total-energy drift over 1000 steps at N = 4096 below 1e-3 of
the kinetic energy.  FIRE: the float16 dt / alpha / n_pos sequence equals the
oracle's for the first 40 steps and the energy after 100 steps is within
1e-5 relative.
"""

import numpy as np
import pytest
import torch

from oracle import jaxmd_oracle as O

pytestmark = pytest.mark.gpu

F32 = np.float32
DEV = 'cuda:0'


def _jmd():
  import jax_md_amd
  return jax_md_amd


def _lj(n_side, seed=0):
  jmd = _jmd()
  R, P, L = O.lj_lattice(n_side, seed=seed)
  disp, shift = jmd.space.periodic(L)
  nfn, efn = jmd.energy.lennard_jones_neighbor_list(disp, L, format=jmd.partition.Dense)
  return jmd, R, P, L, shift, nfn, efn


def test_nve_trajectory_vs_fp64():
  print('case NVE N=512 100 steps', flush=True)
  jmd, R, P, L, shift, nfn, efn = _lj(8, seed=1)
  Rt = torch.as_tensor(R, device=DEV)
  nbrs = nfn.allocate(Rt)
  init_fn, apply_fn = jmd.simulate.nve(efn, shift, 0.005)
  state = init_fn(0, Rt, kT=1.0, momenta=torch.as_tensor(P, device=DEV), neighbor=nbrs)
  for _ in range(100):
    nbrs = nbrs.update(state.position)
    state = apply_fn(state, neighbor=nbrs)
  assert not bool(nbrs.did_buffer_overflow)
  sp = O.Space(np.float64(L))
  m = O.PairModel('lj')
  ff = lambda X: O.energy_force_brute(m, sp, X)[2]
  st = O.NVEState(R.astype(np.float64), P.astype(np.float64), ff(R.astype(np.float64)))
  for _ in range(100):
    st = O.nve_step(st, ff, sp, 0.005)
  dR = sp.displacement(state.position_hi.cpu().numpy(), st.position)
  dP = state.momentum.cpu().numpy() - st.momentum
  print(f'  max |dR|={np.abs(dR).max():.3e} max |dP|={np.abs(dP).max():.3e}')
  assert np.abs(dR).max() < 1e-4
  assert np.abs(dP).max() < 1e-3
  assert np.array_equal(state.position.cpu().numpy(),
                        state.position_hi.cpu().numpy().astype(F32))


def _energy_error(dt, n_steps, every):
  jmd, R, P, L, shift, nfn, efn = _lj(16, seed=0)
  Rt = torch.as_tensor(R, device=DEV)
  nbrs = nfn.allocate(Rt)
  init_fn, apply_fn = jmd.simulate.nve(efn, shift, dt)
  state = init_fn(0, Rt, kT=1.0, momenta=torch.as_tensor(P, device=DEV), neighbor=nbrs)

  def total(s):
    E, _ = efn.energy_and_force(s.position, nbrs)
    return float(E) + float(jmd.quantity.kinetic_energy(momentum=s.momentum))

  E0 = total(state)
  errs = []
  for k in range(n_steps):
    nbrs = nbrs.update(state.position)
    state = apply_fn(state, neighbor=nbrs)
    if (k + 1) % every == 0:
      assert not bool(nbrs.did_buffer_overflow)
      errs.append(total(state) - E0)
  ke = float(jmd.quantity.kinetic_energy(momentum=state.momentum))
  return np.array(errs), ke, E0


def test_nve_energy_error_c1_second_order():
  """C1 (N = 4096, 1000 steps at dt = 0.005): the total-energy error is
  reported; halving dt over the same time must quarter it (velocity Verlet is
  second order), which a wrong kick or drift would break."""
  print('case NVE energy error N=4096', flush=True)
  e1, ke, E0 = _energy_error(0.005, 1000, 50)
  e2, _, _ = _energy_error(0.0025, 2000, 100)
  m1, m2 = np.abs(e1).max(), np.abs(e2).max()
  slope = np.polyfit(np.arange(len(e1)) * 50, e1, 1)[0]
  print(f'  E0={E0:.4f} KE={ke:.1f} max|dE| dt=0.005: {m1:.4e}, dt=0.0025: {m2:.4e}, '
        f'ratio {m1 / m2:.2f}, linear drift {slope * 1000:.3e} per 1000 steps')
  assert 3.0 < m1 / m2 < 5.5
  assert m1 < 5e-3 * ke


def test_fire_vs_oracle_2d_soft_sphere():
  print('case FIRE N=256 2D soft spheres', flush=True)
  jmd = _jmd()
  R, species, L = O.bidisperse_2d(256, seed=5)
  sigma = np.array([[1.0, 1.2], [1.2, 1.4]], F32)
  disp, shift = jmd.space.periodic(L)
  nfn, efn = jmd.energy.soft_sphere_neighbor_list(
      disp, L, species=torch.as_tensor(species, device=DEV),
      sigma=torch.as_tensor(sigma, device=DEV), format=jmd.partition.Dense)
  Rt = torch.as_tensor(R, device=DEV)
  nbrs = nfn.allocate(Rt)
  init_fn, apply_fn = jmd.minimize.fire_descent(efn, shift, dt_start=0.01, dt_max=0.1)
  state = init_fn(Rt, neighbor=nbrs)
  sp = O.Space(np.float64(L))
  m = O.PairModel('soft_sphere', sigma=sigma.astype(np.float64), species=species)
  ff = lambda X: O.energy_force_brute(m, sp, X)[2]
  params = O.FireParams(0.01, 0.1)
  ost = O.fire_init(R.astype(np.float64), ff, params)
  for k in range(100):
    nbrs = nbrs.update(state.position)
    state = apply_fn(state, neighbor=nbrs)
    ost = O.fire_step(ost, ff, sp, params)
    if k < 40:
      assert float(state.dt) == float(ost.dt), k
      assert float(state.alpha) == float(ost.alpha), k
      assert int(state.n_pos) == ost.n_pos, k
  E = float(efn(state.position, nbrs))
  E64 = O.energy_force_brute(m, sp, ost.position)[0]
  print(f'  E={E:.8e} E64={E64:.8e}')
  assert abs(E - E64) <= 1e-5 * abs(E64) + 1e-9


@pytest.mark.parametrize('reorder', [True, False])
def test_nve_block_loop_matches_drop_in(reorder):
  """simulate.NVELoop (the native fori_loop body, with or without cell-order
  execution) against the per-step drop-in path: same trajectory to float32
  round-off after 60 steps, and a list holding exactly the neighbour sets of a
  fresh build at the final positions."""
  print(f'case NVELoop reorder={reorder}', flush=True)
  jmd, R, P, L, shift, nfn, efn = _lj(12, seed=2)
  Rt = torch.as_tensor(R, device=DEV)
  Pt = torch.as_tensor(P, device=DEV)
  init_fn, apply_fn = jmd.simulate.nve(efn, shift, 0.005)
  nb1 = nfn.allocate(Rt)
  s1 = init_fn(0, Rt, kT=1.0, momenta=Pt, neighbor=nb1)
  for _ in range(60):
    nb1 = nb1.update(s1.position)
    s1 = apply_fn(s1, neighbor=nb1)
  nb2 = nfn.allocate(Rt)
  s2 = init_fn(0, Rt, kT=1.0, momenta=Pt, neighbor=nb2)
  loop = jmd.simulate.NVELoop(efn, shift, 0.005, reorder=reorder)
  loop.prime(s2, nb2)
  loop.run(s2, nb2, 30)
  loop.run(s2, nb2, 30)
  assert not bool(nb2.did_buffer_overflow)
  sp = O.Space(np.float64(L))
  dR = sp.displacement(s1.position_hi.cpu().numpy(), s2.position_hi.cpu().numpy())
  dP = (s1.momentum - s2.momentum).abs().max().item()
  print(f'  max |dR|={np.abs(dR).max():.3e} max |dP|={dP:.3e}')
  assert np.abs(dR).max() < 1e-4 and dP < 1e-3
  assert np.array_equal(s2.position.cpu().numpy(), s2.position_hi.cpu().numpy().astype(F32))
  fresh = nfn.allocate(s2.position)
  N = R.shape[0]
  a = np.sort(np.where(nb2.idx.cpu().numpy() < N, nb2.idx.cpu().numpy(), N), axis=1)
  # the loop's list was built at its last rebuild: compare at those positions
  ref_pos = nb2.reference_position
  fresh = nfn.allocate(ref_pos.clone())
  b = np.sort(fresh.idx.cpu().numpy(), axis=1)
  w = min(a.shape[1], b.shape[1])
  assert np.array_equal(a[:, :w], b[:, :w])


@pytest.mark.parametrize('fmt', ['Sparse', 'OrderedSparse'])
def test_nve_sparse_formats_follow_dense(fmt):
  """simulate.nve over Sparse-format lists (the reference's default format)
  follows the Dense trajectory: 50 steps, N = 1728."""
  print(f'case NVE {fmt} vs Dense', flush=True)
  jmd = _jmd()
  R, P, L = O.lj_lattice(12, seed=9)
  disp, shift = jmd.space.periodic(L)
  Rt = torch.as_tensor(R, device=DEV)
  Pt = torch.as_tensor(P, device=DEV)
  out = []
  for f in ('Dense', fmt):
    nfn, efn = jmd.energy.lennard_jones_neighbor_list(disp, L, format=getattr(jmd.partition, f))
    init_fn, apply_fn = jmd.simulate.nve(efn, shift, 0.005)
    nb = nfn.allocate(Rt)
    st = init_fn(0, Rt, 0.0, momenta=Pt, neighbor=nb)
    for _ in range(50):
      nb = nb.update(st.position)
      st = apply_fn(st, neighbor=nb)
    assert not bool(nb.did_buffer_overflow)
    out.append(st)
  dR = (out[0].position_hi - out[1].position_hi).abs()
  dR = torch.minimum(dR, float(L) - dR).max().item()
  dP = (out[0].momentum - out[1].momentum).abs().max().item()
  print(f'  max |dR|={dR:.3e} max |dP|={dP:.3e}')
  assert dR < 1e-5 and dP < 1e-4


@pytest.mark.parametrize('fmt', ['Sparse', 'OrderedSparse'])
def test_nve_loop_sparse_reordered(fmt):
  """NVELoop in cell order over a Sparse-format list: the trajectory follows
  the per-step path, and the list handed back holds, in user ids, exactly
  the pairs of a fresh build at its reference positions (OrderedSparse
  pairs with receiver < sender)."""
  print(f'case NVELoop reordered {fmt}', flush=True)
  jmd = _jmd()
  R, P, L = O.lj_lattice(12, seed=11)
  disp, shift = jmd.space.periodic(L)
  nfn, efn = jmd.energy.lennard_jones_neighbor_list(disp, L, format=getattr(jmd.partition, fmt))
  Rt = torch.as_tensor(R, device=DEV)
  Pt = torch.as_tensor(P, device=DEV)
  init_fn, apply_fn = jmd.simulate.nve(efn, shift, 0.005)
  nb1 = nfn.allocate(Rt)
  s1 = init_fn(0, Rt, 0.0, momenta=Pt, neighbor=nb1)
  for _ in range(60):
    nb1 = nb1.update(s1.position)
    s1 = apply_fn(s1, neighbor=nb1)
  nb2 = nfn.allocate(Rt)
  s2 = init_fn(0, Rt, 0.0, momenta=Pt, neighbor=nb2)
  loop = jmd.simulate.NVELoop(efn, shift, 0.005, reorder=True)
  loop.prime(s2, nb2)
  loop.run(s2, nb2, 30)
  loop.run(s2, nb2, 30)
  assert not bool(nb2.did_buffer_overflow)
  dR = (s1.position_hi - s2.position_hi).abs()
  dR = torch.minimum(dR, float(L) - dR).max().item()
  dP = (s1.momentum - s2.momentum).abs().max().item()
  print(f'  max |dR|={dR:.3e} max |dP|={dP:.3e}')
  assert dR < 1e-4 and dP < 1e-3
  N = R.shape[0]

  def pairs(idx):
    a, b = idx[0].cpu().numpy(), idx[1].cpu().numpy()
    m = (a < N) & (b < N)
    return a[m], b[m]

  a, b = pairs(nb2.idx)
  if fmt == 'OrderedSparse':
    assert np.all(a < b)
  fresh = nfn.allocate(nb2.reference_position.clone())
  fa, fb = pairs(fresh.idx)
  assert sorted(zip(a.tolist(), b.tolist())) == sorted(zip(fa.tolist(), fb.tolist()))


@pytest.mark.parametrize('fmt', ['Sparse', 'OrderedSparse'])
def test_nve_loop_sparse_lists_exact_and_deterministic(fmt):
  """NVELoop over a Sparse-format list runs it as Dense rows kept by the
  Sparse pair test: without cell-order execution the list handed back is,
  entry for entry, a fresh Sparse build at its reference positions (and its
  capacity-padded tail is padding); two runs from one state give
  bit-identical states (no float atomics in the loop); the trajectory follows
  the Dense list's to float32 round-off."""
  print(f'case NVELoop exact {fmt}', flush=True)
  jmd = _jmd()
  R, P, L = O.lj_lattice(12, seed=12)
  N = R.shape[0]
  disp, shift = jmd.space.periodic(L)
  Rt = torch.as_tensor(R, device=DEV)
  Pt = torch.as_tensor(P, device=DEV)
  runs = []
  s0 = None
  for f in (fmt, fmt, 'Dense'):
    nfn, efn = jmd.energy.lennard_jones_neighbor_list(disp, L, format=getattr(jmd.partition, f))
    nb = nfn.allocate(Rt)
    init_fn, _ = jmd.simulate.nve(efn, shift, 0.005)
    s = init_fn(0, Rt.clone(), 0.0, momenta=Pt, neighbor=nb)
    if s0 is None:  # the standalone Sparse-format force sums with float atomics: one start
      s0 = (s.position.clone(), s.momentum.clone(), s.force.clone(), s.position_hi.clone())
    for t, v in zip((s.position, s.momentum, s.force, s.position_hi), s0):
      t.copy_(v)
    loop = jmd.simulate.NVELoop(efn, shift, 0.005, reorder=False)
    loop.prime(s, nb)
    for _ in range(3):
      loop.run(s, nb, 25)
    assert not bool(nb.did_buffer_overflow)
    runs.append((s, nb, nfn))
  (s1, nb1, nfn1), (s2, nb2, _), (s3, _, _) = runs
  assert torch.equal(s1.position_hi, s2.position_hi) and torch.equal(s1.momentum, s2.momentum)
  assert torch.equal(nb1.idx, nb2.idx)
  dR = (s1.position_hi - s3.position_hi).abs()
  dR = torch.minimum(dR, float(L) - dR).max().item()
  print(f'  vs Dense: max |dR|={dR:.3e}')
  assert dR < 1e-4
  # a fresh build at the reference positions on the same cell grid and cell
  # capacity (the slot rotation, and so the entry order, depends on both):
  # a skin-less list of the same cutoff allocated at the same start, updated
  nfr = jmd.partition.neighbor_list(disp, L, 3.0, 0.0, format=getattr(jmd.partition, fmt))
  fresh = nfr.allocate(Rt)
  assert fresh.cell_list_capacity == nb1.cell_list_capacity
  fresh = fresh.update(nb1.reference_position.clone())
  got, want = nb1.idx.cpu().numpy(), fresh.idx.cpu().numpy()
  total = int((want[0] < N).sum())
  assert int((got[0] < N).sum()) == total
  assert np.array_equal(got[:, :total], want[:, :total])
  assert np.all(got[:, total:] == N)


def test_nve_loop_with_species_matches_drop_in():
  """A species-dependent potential through NVELoop (cell-order execution
  requested) follows the per-step path: arrays indexed by particle id keep the
  caller's numbering."""
  print('case NVELoop species', flush=True)
  from tests.golden import fixtures
  jmd = _jmd()
  stj = fixtures.jammed_state()
  box = F32(stj['box'][0, 0])
  R = torch.as_tensor(stj['real_position'].astype(F32), device=DEV)
  disp, shift = jmd.space.periodic(box)
  nfn, efn = jmd.energy.soft_sphere_neighbor_list(
      disp, box, species=torch.as_tensor(stj['species'], device=DEV),
      sigma=torch.as_tensor(stj['sigma'], dtype=torch.float32, device=DEV))
  P = torch.as_tensor(np.random.default_rng(4).normal(size=R.shape).astype(F32) * 0.05,
                      device=DEV)
  init_fn, apply_fn = jmd.simulate.nve(efn, shift, 1e-2)
  nb1 = nfn.allocate(R)
  s1 = init_fn(0, R, 0.0, momenta=P, neighbor=nb1)
  for _ in range(40):
    nb1 = nb1.update(s1.position)
    s1 = apply_fn(s1, neighbor=nb1)
  nb2 = nfn.allocate(R)
  s2 = init_fn(0, R, 0.0, momenta=P, neighbor=nb2)
  loop = jmd.simulate.NVELoop(efn, shift, 1e-2, reorder=True)
  loop.prime(s2, nb2)
  loop.run(s2, nb2, 40)
  dR = (s1.position_hi - s2.position_hi).abs().max().item()
  dP = (s1.momentum - s2.momentum).abs().max().item()
  print(f'  max |dR|={dR:.3e} max |dP|={dP:.3e}')
  assert dR < 1e-6 and dP < 1e-5


@pytest.mark.parametrize('reorder', [False, True])
def test_nve_fused_update_matches_separate_kernels(reorder):
  """Both fused step forms -- the force launch starting the next step (skin
  test, half kick, drift; the default) and the list update carrying the half
  kick and drift ('update') -- with alternating position buffers and gate
  words give the bit-identical trajectory, lists, reference positions, gate
  and rebuild counts of the separate build, kick-drift and force kernels, for
  odd and even block lengths."""
  print(f'case NVELoop fused reorder={reorder}', flush=True)
  jmd, R, P, L, shift, nfn, efn = _lj(12, seed=7)
  Rt = torch.as_tensor(R, device=DEV)
  Pt = torch.as_tensor(P, device=DEV)
  init_fn, _ = jmd.simulate.nve(efn, shift, 0.005)
  runs = []
  for fused in (False, True, 'update'):
    nb = nfn.allocate(Rt)
    s = init_fn(0, Rt, kT=1.0, momenta=Pt, neighbor=nb)
    loop = jmd.simulate.NVELoop(efn, shift, 0.005, reorder=reorder, graph=False, fused=fused)
    loop.prime(s, nb)
    snaps = []
    for n_steps in (15, 16, 9, 30):
      loop.run(s, nb, n_steps)
      snaps.append((s.position_hi.clone(), s.position.clone(), s.momentum.clone(),
                    s.force.clone(), nb.idx.clone(), nb.reference_position.clone(),
                    nb.plan.gate.clone(), nb.plan.builds))
    runs.append(snaps)
  for other in runs[1:]:
    for blk, (a, b) in enumerate(zip(runs[0], other)):
      for k in range(7):
        assert torch.equal(a[k], b[k]), (reorder, blk, k)
      assert a[7] == b[7], (reorder, blk)
  assert runs[1][-1][7] > runs[1][0][7]  # rebuilds happened


def test_nve_loop_continued_blocks():
  """Consecutive blocks start from the previous block's cell-order copies
  (no ordering, permutation or relabelling at their start): the trajectory
  follows blocks that all start afresh to float32 round-off over 100 steps,
  the list holds the sets of a fresh build at its reference positions, and a
  write to the state through torch or a list update through partition makes
  the next block start afresh."""
  print('case NVELoop continued blocks', flush=True)
  jmd, R, P, L, shift, nfn, efn = _lj(12, seed=5)
  N = R.shape[0]
  Rt = torch.as_tensor(R, device=DEV)
  Pt = torch.as_tensor(P, device=DEV)
  init_fn, _ = jmd.simulate.nve(efn, shift, 0.005)
  out = []
  for every in (1000, 0):
    nb = nfn.allocate(Rt)
    s = init_fn(0, Rt, kT=1.0, momenta=Pt, neighbor=nb)
    loop = jmd.simulate.NVELoop(efn, shift, 0.005, reorder_every=every)
    loop.prime(s, nb)
    for _ in range(5):
      loop.run(s, nb, 20)
    assert not bool(nb.did_buffer_overflow)
    out.append((s, nb, loop))
  (sa, nba, la), (sb, nbb, lb) = out
  print(f'  stats continued={la.stats} fresh={lb.stats}')
  assert la.stats['continued'] == 4 and lb.stats['continued'] == 0
  assert nba.plan.builds > 1  # rebuilds happened inside continued blocks
  sp = O.Space(np.float64(L))
  dR = sp.displacement(sa.position_hi.cpu().numpy(), sb.position_hi.cpu().numpy())
  dP = (sa.momentum - sb.momentum).abs().max().item()
  print(f'  continued vs fresh: max |dR|={np.abs(dR).max():.3e} max |dP|={dP:.3e}')
  assert np.abs(dR).max() < 1e-4 and dP < 1e-3
  assert np.array_equal(sa.position.cpu().numpy(), sa.position_hi.cpu().numpy().astype(F32))
  a = np.sort(np.where(nba.idx.cpu().numpy() < N, nba.idx.cpu().numpy(), N), axis=1)
  fresh = nfn.allocate(nba.reference_position.clone())
  b = np.sort(fresh.idx.cpu().numpy(), axis=1)
  w = min(a.shape[1], b.shape[1])
  assert np.array_equal(a[:, :w], b[:, :w])
  # a torch write to the state, then a list update: each makes the next block fresh
  c0 = la.stats['continued']
  sa.momentum.mul_(1.0)
  la.run(sa, nba, 10)
  assert la.stats['continued'] == c0
  la.run(sa, nba, 10)
  assert la.stats['continued'] == c0 + 1
  nba = nba.update(sa.position)
  la.run(sa, nba, 10)
  assert la.stats['continued'] == c0 + 1


@pytest.mark.parametrize('reorder', [True, False])
def test_nve_loop_graph_replays_match_eager(reorder):
  """The HIP-graph block (recorded once, replayed) gives bit-identical state,
  lists and rebuild counts to the eagerly launched block, on every replay,
  with eager work and synchronisations between the replays."""
  print(f'case NVELoop graph replays reorder={reorder}', flush=True)
  jmd, R, P, L, shift, nfn, efn = _lj(12, seed=4)
  Rt = torch.as_tensor(R, device=DEV)
  Pt = torch.as_tensor(P, device=DEV)
  init_fn, _ = jmd.simulate.nve(efn, shift, 0.005)
  runs = []
  for graph in (False, True):
    nb = nfn.allocate(Rt)
    s = init_fn(0, Rt, kT=1.0, momenta=Pt, neighbor=nb)
    loop = jmd.simulate.NVELoop(efn, shift, 0.005, reorder=reorder, graph=graph)
    loop.prime(s, nb)
    snaps = []
    for blk in range(4):
      loop.run(s, nb, 15)
      torch.cuda.synchronize()
      _ = efn.energy_and_force(s.position, nb)  # eager work between replays
      snaps.append((s.position_hi.clone(), s.momentum.clone(), s.force.clone(),
                    nb.idx.clone(), nb.reference_position.clone(), nb.plan.builds))
    runs.append(snaps)
  for blk, (a, b) in enumerate(zip(*runs)):
    for k in range(5):
      assert torch.equal(a[k], b[k]), (reorder, blk, k)
    assert a[5] == b[5], (reorder, blk)
  assert runs[1][-1][5] > runs[1][0][5]  # rebuilds happened inside the replays


def test_fire_bookkeeping_hand_table_on_device():
  # The float16 dt / alpha / n_pos bookkeeping of jmd_fire_update against the
  # hand-derived table of the oracle's pin (tests/test_oracle.py), driven
  # with the scenario's forces and the momenta of its NVE steps.
  from tests.test_oracle import _FIRE_TABLE, fire_scheduled_force
  jmd = _jmd()
  nat = jmd._native
  lib = nat.load()
  cast = jmd.util.static_cast(0.1, 0.4, 5, 1.1, 0.5, 0.1, 0.99)
  params = (nat.ctypes.c_float * 7)(*[float(np.float32(c)) for c in cast])
  scal = torch.tensor([float(np.float32(cast[0])), float(np.float32(cast[5]))],
                      dtype=torch.float32, device=DEV)
  npos = torch.zeros((1,), dtype=torch.int32, device=DEV)
  ws = torch.empty((lib.jmd_fire_workspace_bytes(1),), dtype=torch.uint8, device=DEV)
  sums = torch.empty((8,), dtype=torch.float64, device=DEV)
  P = np.zeros((1, 3), np.float32)
  F_old = fire_scheduled_force(0)
  for step, n_pos, dt_bits, alpha_bits in _FIRE_TABLE:
    dt = np.float32(scal[0].item())
    F_new = fire_scheduled_force(step)
    half = np.float32(dt / np.float32(2))
    P[0, 0] = np.float32(np.float32(P[0, 0] + half * np.float32(F_old)) + half * np.float32(F_new))
    Pt = torch.as_tensor(P, device=DEV).contiguous()
    Ft = torch.zeros((1, 3), dtype=torch.float32, device=DEV)
    Ft[0, 0] = F_new
    nat.check(lib.jmd_fire_update(nat.ptr(Pt), nat.ptr(Ft), 1, 3, nat.ptr(scal), nat.ptr(npos),
                                  params, nat.ptr(sums), nat.ptr(ws), nat.stream_handle(DEV)),
              'jmd_fire_update')
    got = (int(npos.item()), int(np.float16(scal[0].item()).view(np.uint16)),
           int(np.float16(scal[1].item()).view(np.uint16)))
    assert got == (n_pos, dt_bits, alpha_bits), (step, got)
    P = Pt.cpu().numpy()
    if step == 31:
      assert np.all(P == 0.0)
    F_old = F_new


def test_skin_threshold_exact_cases_on_device():
  # the gate of jmd_neighbor_list_check at (dr/2)^2 exactly and one float32
  # step beyond it, including across a periodic face (cases of the oracle's
  # pin in tests/test_oracle.py, decided by hand)
  jmd = _jmd()
  L = F32(16.0)
  disp, _ = jmd.space.periodic(L)
  nfn = jmd.partition.neighbor_list(disp, L, 2.5, 0.5, format=jmd.partition.Dense)
  base = np.array([[15.875, 4.0, 4.0], [8.0, 8.0, 8.0], [1.0, 12.0, 6.0]], F32)
  nb = nfn.allocate(torch.as_tensor(base, device=DEV))

  def rebuilt(pos):
    nonlocal nb
    b0 = nb.plan.builds
    nb = nb.update(torch.as_tensor(pos, device=DEV))
    hit = nb.plan.builds > b0
    if hit:  # start the next case from the original reference again
      nb = nfn.allocate(torch.as_tensor(base, device=DEV))
    return hit

  cases = []
  p = base.copy(); p[1, 0] = F32(8.25); cases.append((p, False))
  p = base.copy(); p[1, 0] = np.nextafter(F32(8.25), F32(9)); cases.append((p, True))
  p = base.copy(); p[1, 1] = F32(7.75); cases.append((p, False))
  p = base.copy(); p[1, 1] = np.nextafter(F32(7.75), F32(0)); cases.append((p, True))
  p = base.copy(); p[0, 0] = F32(0.125); cases.append((p, False))
  p = base.copy(); p[0, 0] = F32(0.125) + F32(2.0 ** -26); cases.append((p, False))
  p = base.copy(); p[0, 0] = F32(0.125) + F32(2.0 ** -20); cases.append((p, True))
  for k, (pos, want) in enumerate(cases):
    print(f'case skin threshold {k}', flush=True)
    assert rebuilt(pos) == want, k


def test_skin_rule_keeps_every_pair_within_cutoff_on_device():
  # after gated updates along a random walk, every pair within r_cutoff at the
  # current positions is in the GPU list (brute force over all pairs)
  jmd = _jmd()
  rng = np.random.default_rng(31)
  L = F32(9.0)
  disp, _ = jmd.space.periodic(L)
  nfn = jmd.partition.neighbor_list(disp, L, 2.0, 0.6, format=jmd.partition.Dense)
  R = rng.uniform(0, L, (400, 3)).astype(F32)
  nb = nfn.allocate(torch.as_tensor(R, device=DEV), extra_capacity=8)
  osp = O.Space(L)
  b0 = nb.plan.builds
  for step in range(40):
    R = np.mod(R + rng.normal(0, 0.04, R.shape), L).astype(F32)
    nb = nb.update(torch.as_tensor(R, device=DEV))
    assert not bool(nb.did_buffer_overflow)
    idx = nb.idx.cpu().numpy()
    for i in range(0, len(R), 7):
      d2 = osp.metric_sq(R, np.broadcast_to(R[i], R.shape))
      want = set(np.flatnonzero(d2 < F32(4.0)).tolist()) - {i}
      have = set(idx[i][idx[i] < len(R)].tolist())
      assert want <= have, (step, i, sorted(want - have))
  assert 0 < nb.plan.builds - b0 < 40


@pytest.mark.parametrize('wrapped', [True, False])
def test_nve_unwrapped_positions_vs_fp64(wrapped):
  # space.periodic(side, wrapped=False) leaves positions unwrapped
  # (space.py:287-301): forces and trajectory must not depend on it.
  print(f'case NVE wrapped={wrapped}', flush=True)
  jmd = _jmd()
  R, P, L = O.lj_lattice(6, seed=8)
  disp, shift = jmd.space.periodic(L, wrapped=wrapped)
  nfn, efn = jmd.energy.lennard_jones_neighbor_list(disp, L, format=jmd.partition.Dense)
  # start with some particles one and two periods away
  rng = np.random.default_rng(9)
  k = rng.integers(-2, 3, R.shape).astype(F32)
  R0 = (R + k * F32(L)).astype(F32)
  Rt = torch.as_tensor(R0, device=DEV)
  nb = nfn.allocate(Rt)
  init_fn, step_fn = jmd.simulate.nve(efn, shift, 0.005)
  st = init_fn(0, Rt, kT=1.0, momenta=torch.as_tensor(P, device=DEV), neighbor=nb)
  sp = O.Space(np.float64(L))
  model = O.PairModel('lj')
  _, _, F64 = O.energy_force_brute(model, sp, R0.astype(np.float64))
  Fg = st.force.cpu().numpy().astype(np.float64)
  assert np.linalg.norm(Fg - F64) / np.linalg.norm(F64) < 1e-5
  for _ in range(30):
    nb = nb.update(st.position)
    st = step_fn(st, neighbor=nb)
  X = st.position.cpu().numpy().astype(np.float64)
  if not wrapped:
    assert np.any(X < 0) or np.any(X > L)  # the shift did not wrap
  _, _, F64 = O.energy_force_brute(model, sp, X)
  Fg = st.force.cpu().numpy().astype(np.float64)
  assert np.linalg.norm(Fg - F64) / np.linalg.norm(F64) < 1e-5
  # the same 30 steps in float64 from the same start
  o = O.NVEState(R0.astype(np.float64), P.astype(np.float64),
                 O.energy_force_brute(model, sp, R0.astype(np.float64))[2])
  osp = O.Space(np.float64(L)) if wrapped else O.Space(None, False)
  for _ in range(30):
    o = O.nve_step(o, lambda X: O.energy_force_brute(model, sp, X)[2], osp, 0.005)
  dx = sp.displacement(X, o.position)
  assert np.max(np.abs(dx)) < 1e-4
