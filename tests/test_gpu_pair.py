"""Pair energies and forces on the GPU against the float64 oracle.

Tolerances (float32 pair arithmetic, float64 accumulation):
  force:  ||F - F64|| / ||F64|| <= 1e-5 (norm-wise), and per particle
          max_i |F_i - F64_i| <= 5e-5 * rms(F64): float32 evaluation of
          r^-14 carries ~1e-6 relative error on the strongest pair of an atom,
          which sets the per-particle worst case
  forces that nearly cancel (a jammed packing at equilibrium) are judged
          against the per-particle sum of pair-force magnitudes instead
  energy: |E - E64| / |E64| <= 1e-6 (fixtures: the reference's own tolerance)
The oracle evaluates the same float32 positions in float64.
"""

import numpy as np
import pytest
import torch

from oracle import jaxmd_oracle as O
from tests.golden import fixtures

pytestmark = pytest.mark.gpu

F32 = np.float32
DEV = 'cuda:0'
FORCE_RTOL = 1e-5
FORCE_WORST = 5e-5


def _jmd():
  import jax_md_amd
  return jax_md_amd


def _force_err(F, F64):
  F = np.asarray(F, np.float64)
  rel = np.linalg.norm(F - F64) / np.linalg.norm(F64)
  rms = np.sqrt(np.mean(F64 ** 2))
  worst = np.max(np.abs(F - F64)) / rms
  return rel, worst


def _lj_setup(n_side, fmt='Dense', seed=0, **kw):
  jmd = _jmd()
  R, _, L = O.lj_lattice(n_side, seed=seed)
  disp, shift = jmd.space.periodic(L)
  nfn, efn = jmd.energy.lennard_jones_neighbor_list(
      disp, L, format=getattr(jmd.partition, fmt), **kw)
  Rt = torch.as_tensor(R, device=DEV)
  nbrs = nfn.allocate(Rt)
  return jmd, R, L, Rt, nbrs, efn


@pytest.mark.parametrize('n_side', [8, 16])
def test_lj_energy_force_vs_fp64(n_side):
  print(f'case LJ lattice n={n_side}', flush=True)
  jmd, R, L, Rt, nbrs, efn = _lj_setup(n_side)
  E, F = efn.energy_and_force(Rt, nbrs)
  m = O.PairModel('lj')
  E64, _, F64 = O.energy_force_brute(m, O.Space(np.float64(L)), R.astype(np.float64))
  rel, worst = _force_err(F.cpu().numpy(), F64)
  print(f'  force rel={rel:.3e} worst/rms={worst:.3e} E={float(E):.9f} E64={E64:.9f}')
  assert rel <= FORCE_RTOL and worst <= FORCE_WORST
  assert abs(float(E) - E64) <= 1e-6 * abs(E64)
  # the scalar energy_fn path and the force path agree with energy_and_force
  assert abs(float(efn(Rt, nbrs)) - float(E)) <= 1e-6 * abs(float(E))
  assert torch.equal(jmd.quantity.force(efn)(Rt, nbrs), F)


@pytest.mark.parametrize('fmt', ['Sparse', 'OrderedSparse'])
def test_lj_formats_agree(fmt):
  print(f'case LJ formats {fmt}', flush=True)
  _, R, L, Rt, nbrs, efn = _lj_setup(10)
  E_d, F_d = efn.energy_and_force(Rt, nbrs)
  _, _, _, _, nbrs_s, efn_s = _lj_setup(10, fmt)
  E_s, F_s = efn_s.energy_and_force(Rt, nbrs_s)
  assert abs(float(E_s) - float(E_d)) <= 1e-6 * abs(float(E_d))
  rel, worst = _force_err(F_s.cpu().numpy(), F_d.cpu().numpy().astype(np.float64))
  assert rel <= FORCE_RTOL and worst <= FORCE_WORST


@pytest.mark.parametrize('fmt', ['Sparse', 'OrderedSparse'])
def test_sparse_any_pair_order_and_per_particle(fmt):
  """The sparse kernel sums sender shares over runs of equal sender ids;
  a shuffled pair order (runs of one, the same id in several runs of one
  wavefront) must give the same forces and per-particle energies."""
  print(f'case sparse shuffled {fmt}', flush=True)
  import dataclasses
  jmd, R, L, Rt, nbrs, efn = _lj_setup(8, fmt, seed=3)
  disp, _ = jmd.space.periodic(L)
  _, efn_pp = jmd.energy.lennard_jones_neighbor_list(
      disp, L, format=getattr(jmd.partition, fmt), per_particle=True)
  g = torch.Generator().manual_seed(5)
  perm = torch.randperm(nbrs.idx.shape[1], generator=g).to(DEV)
  # half shuffled, half in receiver order with every other pair swapped
  # with its neighbour, so runs break up inside a wavefront
  idx = nbrs.idx.clone()
  half = idx.shape[1] // 2
  idx[:, :half] = nbrs.idx[:, perm[perm < half]]
  sw = torch.arange(half, idx.shape[1] - 1, 2, device=DEV)
  a, b = idx[:, sw].clone(), idx[:, sw + 1].clone()
  idx[:, sw], idx[:, sw + 1] = b, a
  shuffled = dataclasses.replace(nbrs, idx=idx)
  m = O.PairModel('lj')
  _, e64, F64 = O.energy_force_brute(m, O.Space(np.float64(L)), R.astype(np.float64))
  for nb in (nbrs, shuffled):
    E, F = efn.energy_and_force(Rt, nb)
    rel, worst = _force_err(F.cpu().numpy(), F64)
    print(f'  force rel={rel:.3e} worst={worst:.3e}')
    assert rel <= FORCE_RTOL and worst <= FORCE_WORST
    assert abs(float(E) - e64.sum()) <= 1e-6 * abs(e64.sum())
    if fmt == 'OrderedSparse':  # per-particle values are refused (smap.py:967-972)
      with pytest.raises(ValueError):
        efn_pp(Rt, nb)
      continue
    e = efn_pp(Rt, nb).cpu().numpy()
    print(f'  per-particle {np.max(np.abs(e - e64)) / np.max(np.abs(e64)):.2e}')
    assert np.max(np.abs(e - e64)) <= 1e-5 * np.max(np.abs(e64))


def test_lj_per_particle_and_autograd():
  print('case LJ per-particle + autograd', flush=True)
  jmd, R, L, Rt, nbrs, efn = _lj_setup(8)
  disp, _ = jmd.space.periodic(L)
  _, efn_pp = jmd.energy.lennard_jones_neighbor_list(disp, L, format=jmd.partition.Dense,
                                                      per_particle=True)
  e = efn_pp(Rt, nbrs)
  E, F = efn.energy_and_force(Rt, nbrs)
  assert e.shape == (R.shape[0],)
  assert abs(float(e.double().sum()) - float(E)) <= 1e-5 * abs(float(E))
  m = O.PairModel('lj')
  _, e64, _ = O.energy_force_brute(m, O.Space(np.float64(L)), R.astype(np.float64))
  assert np.max(np.abs(e.cpu().numpy() - e64)) <= 1e-5 * np.max(np.abs(e64))
  Rg = Rt.clone().requires_grad_(True)
  Eg = efn(Rg, nbrs)
  Eg.backward()
  assert torch.allclose(Rg.grad, -F, rtol=0, atol=0)


def test_lammps_lj_fixture():
  # quantity_test.py:436-461: E/N = -4.3523016 (LJ truncated at 2.5, 108 atoms)
  print('case LAMMPS LJ fixture', flush=True)
  jmd = _jmd()
  d = fixtures.lammps_lj()
  L = F32(d['box'])
  R = (d['fractional_position'] * np.float64(L)).astype(F32)
  disp, _ = jmd.space.periodic(L)
  nfn, efn = jmd.energy.lennard_jones_neighbor_list(disp, L, r_onset=2.5, r_cutoff=2.5,
                                                    dr_threshold=0.0,
                                                    format=jmd.partition.Dense)
  Rt = torch.as_tensor(R, device=DEV)
  nbrs = nfn.allocate(Rt)
  E = float(efn(Rt, nbrs))
  assert abs(E / len(R) - d['energy_per_atom']) < 5e-5
  m = O.PairModel('lj', r_onset=2.5, r_cutoff=2.5)
  _, _, F64 = O.energy_force_brute(m, O.Space(np.float64(L)), R.astype(np.float64))
  rel, worst = _force_err(jmd.quantity.force(efn)(Rt, nbrs).cpu().numpy(), F64)
  assert rel <= FORCE_RTOL and worst <= FORCE_WORST


@pytest.mark.parametrize('fmt', ['Dense', 'Sparse', 'OrderedSparse'])
def test_jammed_soft_sphere_fixture(fmt):
  # quantity_test.py:134-158 fixture: E = 0.45247561922261154 (species sigma)
  print(f'case jammed fixture {fmt}', flush=True)
  jmd = _jmd()
  st = fixtures.jammed_state()
  box = F32(st['box'][0, 0])
  R = st['real_position'].astype(F32)
  disp, _ = jmd.space.periodic(box)
  sigma = torch.as_tensor(st['sigma'], dtype=torch.float32, device=DEV)
  species = torch.as_tensor(st['species'], device=DEV)
  nfn, efn = jmd.energy.soft_sphere_neighbor_list(disp, box, species=species, sigma=sigma,
                                                  format=getattr(jmd.partition, fmt))
  Rt = torch.as_tensor(R, device=DEV)
  nbrs = nfn.allocate(Rt)
  E, F = efn.energy_and_force(Rt, nbrs)
  assert abs(float(E) - float(st['energy'])) <= 2e-5 * float(st['energy'])
  m = O.PairModel('soft_sphere', sigma=st['sigma'], species=st['species'])
  sp = O.Space(np.float64(box))
  E64, _, F64 = O.energy_force_brute(m, sp, R.astype(np.float64))
  assert abs(float(E) - E64) <= 1e-6 * E64
  # at mechanical equilibrium the net forces cancel to ~0: judge the error
  # against the pair-force scale
  scale = O.pair_force_scale(m, sp, R.astype(np.float64))
  err = np.abs(F.cpu().numpy() - F64).max()
  print(f'  max |dF| = {err:.3e}, rms pair-force scale = {np.sqrt(np.mean(scale**2)):.3e}')
  assert err <= FORCE_RTOL * np.sqrt(np.mean(scale ** 2))


def test_soft_sphere_per_particle_diameters_2d():
  # C5-style: per-particle diameters combined as 0.5 (d_i + d_j)
  print('case SS per-particle diameters', flush=True)
  jmd = _jmd()
  R, species, L = O.bidisperse_2d(2048, d_large=1.4, seed=3)
  diam = np.where(species == 0, 1.0, 1.4).astype(F32)
  disp, _ = jmd.space.periodic(L)
  dt = torch.as_tensor(diam, device=DEV)
  nfn, efn = jmd.energy.soft_sphere_neighbor_list(disp, L, sigma=dt,
                                                  format=jmd.partition.Dense)
  Rt = torch.as_tensor(R, device=DEV)
  nbrs = nfn.allocate(Rt)
  E, F = efn.energy_and_force(Rt, nbrs)
  m = O.PairModel('soft_sphere', sigma=diam.astype(np.float64))
  E64, _, F64 = O.energy_force_brute(m, O.Space(np.float64(L)), R.astype(np.float64))
  assert abs(float(E) - E64) <= 1e-6 * E64
  rel, worst = _force_err(F.cpu().numpy(), F64)
  assert rel <= FORCE_RTOL and worst <= FORCE_WORST


@pytest.mark.parametrize('fmt', ['Dense', 'Sparse', 'OrderedSparse'])
def test_lammps_lj_stress(fmt):
  """quantity.stress through the virial kernel against LAMMPS's stress
  (quantity_test.py:436-461, tol 5e-5) and the FP64 oracle (1e-5)."""
  print(f'case LAMMPS LJ stress {fmt}', flush=True)
  jmd = _jmd()
  d = fixtures.lammps_lj()
  L = F32(d['box'])
  R = (d['fractional_position'] * np.float64(L)).astype(F32)
  disp, _ = jmd.space.periodic(L)
  nfn, efn = jmd.energy.lennard_jones_neighbor_list(disp, L, r_onset=2.5, r_cutoff=2.5,
                                                    dr_threshold=0.0,
                                                    format=getattr(jmd.partition, fmt))
  Rt = torch.as_tensor(R, device=DEV)
  nbrs = nfn.allocate(Rt)
  V = torch.as_tensor(d['velocity'], dtype=torch.float32, device=DEV)
  C = jmd.quantity.stress(efn, Rt, L, velocity=V, neighbor=nbrs).cpu().numpy()
  v = d['stress_voigt']
  Cl = np.array([[v[0], v[3], v[4]], [v[3], v[1], v[5]], [v[4], v[5], v[2]]])
  assert np.allclose(C, Cl, atol=5e-5, rtol=5e-5), (C, Cl)
  m = O.PairModel('lj', r_onset=2.5, r_cutoff=2.5)
  sp = O.Space(np.float64(L))
  R64 = R.astype(np.float64)
  i, j = O.all_pairs(sp, R64, 2.5)
  Co = O.stress(O.pair_virial(m, sp, R64, i, j), float(L) ** 3, d['velocity'])
  print(f'  max |C - oracle| {np.abs(C - Co).max():.2e}, |C - LAMMPS| {np.abs(C - Cl).max():.2e}')
  assert np.allclose(C, Co, atol=1e-5 * np.abs(Co).max(), rtol=0)


@pytest.mark.parametrize('fmt', ['Dense', 'Sparse'])
def test_jammed_pressure(fmt):
  """quantity.pressure of the jammed packing (quantity_test.py:134-146, float32
  tolerance 2e-5)."""
  print(f'case jammed pressure {fmt}', flush=True)
  jmd = _jmd()
  st = fixtures.jammed_state()
  box = F32(st['box'][0, 0])
  R = st['real_position'].astype(F32)
  disp, _ = jmd.space.periodic(box)
  sigma = torch.as_tensor(st['sigma'], dtype=torch.float32, device=DEV)
  species = torch.as_tensor(st['species'], device=DEV)
  nfn, efn = jmd.energy.soft_sphere_neighbor_list(disp, box, species=species, sigma=sigma,
                                                  format=getattr(jmd.partition, fmt))
  Rt = torch.as_tensor(R, device=DEV)
  nbrs = nfn.allocate(Rt)
  P = float(jmd.quantity.pressure(efn, Rt, box, neighbor=nbrs))
  print(f'  P {P:.10f} vs {float(st["pressure"]):.10f}')
  assert abs(P - float(st['pressure'])) <= 2e-5 * abs(float(st['pressure'])) + 2e-5


@pytest.mark.parametrize('fmt', ['Dense', 'Sparse', 'OrderedSparse'])
def test_edge_energies_reduce_axis_empty(fmt):
  """smap.pair_neighbor_list with reduce_axis=() (smap.py:940-953): one
  energy per list entry over the reference's normalization (2, or 1 for
  OrderedSparse), shaped like the list's entries, 0 at padding; each entry
  within 1e-5 relative of the FP64 pair energy plus 1e-6 of the magnitude of
  the two Lennard-Jones terms that cancel near r = sigma (float32 rounding
  of 4 eps (s^12 - s^6)), their sum equal to the energy (1e-5).  Soft
  spheres with species in 2-D as well."""
  print(f'case edge energies {fmt}', flush=True)
  jmd, R, L, Rt, nbrs, efn = _lj_setup(8, fmt=fmt, seed=3)
  N = R.shape[0]
  disp, _ = jmd.space.periodic(L)
  fn = jmd.energy.multiplicative_isotropic_cutoff(jmd.energy.lennard_jones, 2.0, 2.5)
  edges_fn = jmd.smap.pair_neighbor_list(fn, jmd.space.canonicalize_displacement_or_metric(disp),
                                         reduce_axis=())
  out = edges_fn(Rt, nbrs).cpu().numpy().astype(np.float64)
  idx = nbrs.idx.cpu().numpy()
  m = O.PairModel('lj')
  sp = O.Space(np.float64(L))
  if fmt == 'Dense':
    assert out.shape == idx.shape
    rows = np.repeat(np.arange(N), idx.shape[1]).reshape(idx.shape)
    valid = idx < N
    i, j = rows[valid], idx[valid]
    want = O.pair_energies(m, sp, R.astype(np.float64), i, j, 0.5)
    got = out[valid]
    assert np.all(out[~valid] == 0)
  else:
    assert out.shape == (idx.shape[1],)
    valid = idx[0] < N
    w = 1.0 if fmt == 'OrderedSparse' else 0.5
    want = O.pair_energies(m, sp, R.astype(np.float64), idx[0][valid], idx[1][valid], w)
    got = out[valid]
    assert np.all(out[~valid] == 0)
  if fmt == 'Dense':
    ii, jj = i, j
  else:
    ii, jj = idx[0][valid], idx[1][valid]
  d = sp.displacement(R[jj].astype(np.float64), R[ii].astype(np.float64))
  s6 = 1.0 / np.sum(d * d, -1) ** 3
  terms = 0.5 * 4.0 * (s6 * s6 + s6)
  err = np.abs(got - want)
  bar = 1e-5 * np.abs(want) + 1e-6 * terms
  assert np.all(err <= bar), float(np.max(err / bar))
  E = float(efn(Rt, neighbor=nbrs))
  assert abs(out.sum() - E) <= 1e-5 * abs(E)
  # soft spheres with species, 2-D
  jmd = _jmd()
  R2, species, L2 = O.bidisperse_2d(400, seed=11)
  sig = np.array([[1.0, 1.2], [1.2, 1.4]], F32)
  d2, _ = jmd.space.periodic(L2)
  nfn2, efn2 = jmd.energy.soft_sphere_neighbor_list(
      d2, L2, species=torch.as_tensor(species, device=DEV),
      sigma=torch.as_tensor(sig, device=DEV), format=getattr(jmd.partition, fmt))
  R2t = torch.as_tensor(R2, device=DEV)
  nb2 = nfn2.allocate(R2t)
  ss = jmd.smap.pair_neighbor_list(jmd.energy.soft_sphere,
                                   jmd.space.canonicalize_displacement_or_metric(d2),
                                   species=torch.as_tensor(species, device=DEV),
                                   sigma=torch.as_tensor(sig, device=DEV), reduce_axis=())
  out2 = ss(R2t, nb2).cpu().numpy().astype(np.float64)
  E2 = float(efn2(R2t, neighbor=nb2))
  assert abs(out2.sum() - E2) <= 1e-5 * abs(E2)


@pytest.mark.parametrize('fmt', ['Sparse', 'OrderedSparse'])
def test_x64_over_sparse_lists(fmt):
  """The x64 form over Sparse-format lists (run as the Dense rows of the
  same pairs): energy and force against the FP64 oracle to 1e-12, and a
  20-step float64 NVE run over the Sparse list against the same run over a
  Dense list to 1e-12 in positions."""
  print(f'case x64 {fmt}', flush=True)
  jmd = _jmd()
  R, P, L = O.lj_lattice(6, seed=21)
  disp, shift = jmd.space.periodic(L)
  nfs, efs = jmd.energy.lennard_jones_neighbor_list(disp, L, format=getattr(jmd.partition, fmt))
  nfd, efd = jmd.energy.lennard_jones_neighbor_list(disp, L, format=jmd.partition.Dense)
  R64 = torch.as_tensor(R.astype(np.float64), device=DEV)
  nbs = nfs.allocate(R64.to(torch.float32))
  nbd = nfd.allocate(R64.to(torch.float32))
  E, F = efs.energy_and_force(R64, nbs)
  E64, _, F64 = O.energy_force_brute(O.PairModel('lj'), O.Space(np.float64(L)),
                                     R.astype(np.float64))
  assert F.dtype == torch.float64
  rel = np.linalg.norm(F.cpu().numpy() - F64) / np.linalg.norm(F64)
  print(f'  force rel {rel:.2e}, E {float(E)!r} vs {E64!r}')
  assert rel <= 1e-12 and abs(float(E) - E64) <= 1e-12 * abs(E64)
  out = []
  for nfn, efn, nb in ((nfs, efs, nbs), (nfd, efd, nbd)):
    init_fn, apply_fn = jmd.simulate.nve(efn, shift, 0.005)
    st = init_fn(0, R64.clone(), kT=1.0, momenta=torch.as_tensor(P.astype(np.float64),
                                                                 device=DEV), neighbor=nb)
    for _ in range(20):
      nb = nb.update(st.position.to(torch.float32))
      st = apply_fn(st, neighbor=nb)
    out.append(st.position.cpu().numpy())
  assert np.max(np.abs(out[0] - out[1])) <= 1e-12 * L


@pytest.mark.parametrize('fmt', ['Sparse', 'OrderedSparse'])
def test_sparse_lists_force_through_dense_rows(fmt):
  """A Sparse-format list from allocate / update carries the Dense rows of
  its pairs: energy and force run the row kernel, so two evaluations are
  bit-identical and equal the Dense list's force to 1e-6 (the same pairs,
  summed in another order); after updates that rebuild, the rows follow the
  list; a list written by anything else (here by hand) falls back to the
  pair kernel with the same result to 1e-5."""
  print(f'case sparse through dense rows {fmt}', flush=True)
  jmd, R, L, Rt, nbrs, efn = _lj_setup(12, fmt=fmt, seed=4)
  _, _, _, _, nbd, efd = _lj_setup(12, fmt='Dense', seed=4)
  assert jmd.partition.current_shadow(nbrs) is not None
  E1, F1 = efn.energy_and_force(Rt, nbrs)
  E2, F2 = efn.energy_and_force(Rt, nbrs)
  assert torch.equal(F1, F2) and float(E1) == float(E2)
  Ed, Fd = efd.energy_and_force(Rt, nbd)
  rel = float((F1 - Fd).norm() / Fd.norm())
  assert rel <= 1e-6 and abs(float(E1) - float(Ed)) <= 1e-6 * abs(float(Ed)), rel
  # moved beyond half the skin: both lists rebuild
  rng = np.random.default_rng(5)
  R2 = torch.as_tensor((R + rng.normal(0, 0.2, R.shape)).astype(F32) % F32(L), device=DEV)
  nbrs = nbrs.update(R2)
  nbd = nbd.update(R2)
  assert jmd.partition.current_shadow(nbrs) is not None
  _, F3 = efn.energy_and_force(R2, nbrs)
  _, Fd3 = efd.energy_and_force(R2, nbd)
  assert float((F3 - Fd3).norm() / Fd3.norm()) <= 1e-6
  E64, _, F64 = O.energy_force_brute(O.PairModel('lj'), O.Space(np.float64(L)),
                                     R2.cpu().numpy().astype(np.float64))
  assert float(np.linalg.norm(F3.cpu().numpy() - F64) / np.linalg.norm(F64)) <= 1e-5
  # a list written by hand: the pair kernel
  nbrs.idx.copy_(nbrs.idx.clone())
  assert jmd.partition.current_shadow(nbrs) is None
  _, F4 = efn.energy_and_force(R2, nbrs)
  assert float((F4 - F3).norm() / F3.norm()) <= 1e-5
