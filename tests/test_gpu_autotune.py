"""The autotuner and the knobs it turns, on the GPU.

Bars: lists built on a coarser cell grid (cell_size_factor) hold exactly the
oracle's neighbour sets (rows compared as sorted sets; the row order follows
that grid); forces with every threads-per-row setting within the pair tests'
bar of float64 (norm-wise 1e-5); the tuned configuration's lists equal the
oracle's sets, its forces meet the same bar, and a 60-step NVELoop run with it
follows the default configuration's trajectory to float32 round-off (1e-4 in
positions, the block-loop tests' bar).
"""

import numpy as np
import pytest
import torch

from oracle import jaxmd_oracle as O

pytestmark = pytest.mark.gpu

DEV = 'cuda:0'


def _setup(n_side=16, seed=3):
  import jax_md_amd as jmd
  R, P, L = O.lj_lattice(n_side, seed=seed)
  disp, shift = jmd.space.periodic(L)

  def make_fns(**kw):
    return jmd.energy.lennard_jones_neighbor_list(disp, L, format=jmd.partition.Dense, **kw)

  return jmd, R, P, L, shift, make_fns


def _sorted_rows(idx, N):
  a = np.sort(np.asarray(idx), axis=1)
  return a


def _oracle_rows(R, L):
  spec = O.NeighborListSpec(O.Space(L), L, 2.5, 0.5)
  return O.neighbor_list_build(spec, R).idx


def _same_sets(gpu_idx, oracle_idx, N):
  a = _sorted_rows(gpu_idx, N)
  b = _sorted_rows(oracle_idx, N)
  w = max(a.shape[1], b.shape[1])
  pad = lambda x: np.pad(x, ((0, 0), (0, w - x.shape[1])), constant_values=N)
  return np.array_equal(pad(a), pad(b))


def _force_rel(F, R, L):
  E64, _, F64 = O.energy_force_brute(O.PairModel('lj'), O.Space(np.float64(L)),
                                     R.astype(np.float64))
  F = F.cpu().numpy().astype(np.float64)
  return np.linalg.norm(F - F64) / np.linalg.norm(F64)


@pytest.mark.parametrize('factor', [1.2, 1.5])
def test_coarser_cells_keep_the_neighbour_sets(factor):
  print(f'case cell_size_factor={factor}', flush=True)
  jmd, R, P, L, shift, make_fns = _setup()
  N = R.shape[0]
  nfn, _ = make_fns(cell_size_factor=factor)
  nb = nfn.allocate(torch.as_tensor(R, device=DEV))
  assert not bool(nb.did_buffer_overflow)
  assert nb.plan.use_cell_list and nb.plan.cell_size3[0] >= factor * 3.0 * (1 - 1e-6)
  assert _same_sets(nb.idx.cpu().numpy(), _oracle_rows(R, L), N)
  # and after moving the particles (an update that rebuilds)
  R2 = np.mod(R + np.random.default_rng(1).normal(0, 0.3, R.shape), np.float32(L)).astype(np.float32)
  R2 = np.where(R2 >= L, np.float32(0), R2).astype(np.float32)
  nb = nb.update(torch.as_tensor(R2, device=DEV))
  assert not bool(nb.did_buffer_overflow)
  assert _same_sets(nb.idx.cpu().numpy(), _oracle_rows(R2, L), N)


@pytest.mark.parametrize('tpa', [4, 8, 16])
def test_threads_per_row_forces(tpa):
  print(f'case threads per row {tpa}', flush=True)
  jmd, R, P, L, shift, make_fns = _setup()
  nfn, efn = make_fns()
  Rt = torch.as_tensor(R, device=DEV)
  nb = nfn.allocate(Rt)
  try:
    jmd.autotune.set_threads_per_atom(tpa)
    _, F = efn.energy_and_force(Rt, nb)
  finally:
    jmd.autotune.set_threads_per_atom(0)
  rel = _force_rel(F, R, L)
  print(f'  force rel err {rel:.2e}')
  assert rel <= 1e-5


def test_tuned_configuration():
  print('case autotune', flush=True)
  jmd, R, P, L, shift, make_fns = _setup()
  N = R.shape[0]
  Rt = torch.as_tensor(R, device=DEV)
  Pt = torch.as_tensor(P, device=DEV)
  key = jmd.autotune.workload_key(Rt, L, 2.5, 0.5, jmd.partition.Dense, 'lj-test')
  cfg = jmd.autotune.tune_nve_loop(make_fns, Rt, Pt, shift, 0.005, key=key, trial_steps=10,
                                   repeats=2, tpas=(4, 8), cell_size_factors=(1.0, 1.2),
                                   capacity_multipliers=(1.25, 1.5))
  print(f'  tuned: {cfg.as_dict()}')
  assert cfg.threads_per_atom in (4, 8)
  assert cfg.cell_size_factor in (1.0, 1.2)
  assert cfg.capacity_multiplier in (1.25, 1.5)
  assert len(cfg.trials) >= 3 and cfg.us_per_step > 0
  assert torch.equal(Rt, torch.as_tensor(R, device=DEV))  # the caller's state is untouched
  assert jmd.autotune.tune_nve_loop(make_fns, Rt, Pt, shift, 0.005, key=key) is cfg
  try:
    cfg.apply()
    nfn, efn = make_fns(**cfg.neighbor_kwargs)
    nb = nfn.allocate(Rt)
    assert _same_sets(nb.idx.cpu().numpy(), _oracle_rows(R, L), N)
    _, F = efn.energy_and_force(Rt, nb)
    assert _force_rel(F, R, L) <= 1e-5
    states = []
    for kw in (cfg.neighbor_kwargs, {}):
      if not kw:
        jmd.autotune.set_threads_per_atom(0)
      nfn, efn = make_fns(**kw)
      nb = nfn.allocate(Rt)
      init_fn, _ = jmd.simulate.nve(efn, shift, 0.005)
      s = init_fn(0, Rt.clone(), kT=1.0, momenta=Pt, neighbor=nb)
      loop = jmd.simulate.NVELoop(efn, shift, 0.005)
      loop.prime(s, nb)
      loop.run(s, nb, 30)
      loop.run(s, nb, 30)
      assert not bool(nb.did_buffer_overflow)
      states.append(s)
    dR = O.Space(np.float64(L)).displacement(states[0].position_hi.cpu().numpy(),
                                             states[1].position_hi.cpu().numpy())
    print(f'  tuned vs default: max |dR| {np.abs(dR).max():.2e}')
    assert np.abs(dR).max() < 1e-4
  finally:
    jmd.autotune.set_threads_per_atom(0)
