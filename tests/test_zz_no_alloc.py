"""jmd_pair_energy_force_noalloc (the allocation-free energy/force entry
point) against jmd_pair_energy_force, both through the C ABI.

Two calls in a row, the second with a smaller system, so the shared scratch
array holds the first call's partials when the second runs.  Dense lists are
deterministic, so force, per-particle energy and total energy must be
bit-identical to the workspace entry point's; OrderedSparse forces sum with
float32 atomics (order varies run to run), so they are held to the force bar
of the pair tests (norm-wise 1e-5) and the float64 total to 1e-12 relative.
"""

import ctypes

import numpy as np
import pytest
import torch

from oracle import jaxmd_oracle as O

pytestmark = pytest.mark.gpu

DEV = 'cuda:0'


def _setup(n_side, fmt):
  import jax_md_amd as jmd
  R, _, L = O.lj_lattice(n_side, seed=n_side)
  disp, _ = jmd.space.periodic(L)
  nfn, efn = jmd.energy.lennard_jones_neighbor_list(disp, L, format=getattr(jmd.partition, fmt))
  Rt = torch.as_tensor(R, device=DEV)
  nbrs = nfn.allocate(Rt)
  return jmd, Rt, nbrs, efn


def _call(jmd, efn, Rt, nbrs, noalloc, want_atom):
  nat = jmd._native
  lib = nat.load()
  N, dim = Rt.shape
  pot, keep = efn._potential(Rt, {})
  F = torch.full((N, dim), float('nan'), dtype=torch.float32, device=DEV)
  e_tot = torch.full((1,), float('nan'), dtype=torch.float64, device=DEV)
  e_atom = (torch.full((N,), float('nan'), dtype=torch.float32, device=DEV)
            if want_atom else None)
  fmt = {jmd.partition.Dense: nat.DENSE, jmd.partition.Sparse: nat.SPARSE,
         jmd.partition.OrderedSparse: nat.ORDERED_SPARSE}[nbrs.format]
  args = [nat.ptr(Rt), N, dim, dim, efn.geometry_code(), nat.f3(efn.geometry.side_values(dim)),
          ctypes.byref(pot), nat.ptr(nbrs.idx), nbrs.idx.shape[1], fmt, nat.ptr(F),
          nat.ptr(e_atom), nat.ptr(e_tot), None, 0.0, None]
  if noalloc:
    rc = lib.jmd_pair_energy_force_noalloc(*args, nat.stream_handle(DEV))
  else:
    ws = torch.empty((lib.jmd_pair_workspace_bytes(N),), dtype=torch.uint8, device=DEV)
    rc = lib.jmd_pair_energy_force(*args, nat.ptr(ws), nat.stream_handle(DEV))
  nat.check(rc, 'noalloc' if noalloc else 'workspace')
  torch.cuda.synchronize()
  del keep
  return (F.cpu().numpy(), float(e_tot.cpu()[0]),
          None if e_atom is None else e_atom.cpu().numpy())


@pytest.mark.parametrize('fmt', ['Dense', 'OrderedSparse'])
def test_noalloc_matches_workspace_entry_point(fmt):
  for n_side in (16, 8):  # second call smaller: the scratch holds stale partials
    print(f'case noalloc {fmt} n_side={n_side}', flush=True)
    jmd, Rt, nbrs, efn = _setup(n_side, fmt)
    want_atom = fmt == 'Dense'
    F0, E0, A0 = _call(jmd, efn, Rt, nbrs, False, want_atom)
    F1, E1, A1 = _call(jmd, efn, Rt, nbrs, True, want_atom)
    assert np.isfinite(F1).all() and np.isfinite(E1)
    if fmt == 'Dense':
      assert np.array_equal(F0, F1)
      assert E0 == E1
      assert np.array_equal(A0, A1)
    else:
      assert np.linalg.norm(F1 - F0) / np.linalg.norm(F0) <= 1e-5
      assert abs(E1 - E0) <= 1e-12 * abs(E0)
    # and both against the float64 oracle (the pair tests' bar)
    R = Rt.cpu().numpy()
    L = float(efn.geometry.side_values(3)[0])
    E64, _, F64 = O.energy_force_brute(O.PairModel('lj'), O.Space(np.float64(L)),
                                       R.astype(np.float64))
    assert np.linalg.norm(F1 - F64) / np.linalg.norm(F64) <= 1e-5
    assert abs(E1 - E64) <= 1e-6 * abs(E64)


def test_noalloc_rejects_oversized_n():
  import jax_md_amd as jmd
  nat = jmd._native
  lib = nat.load()
  pot = nat.JmdPairPotential()
  # rejected before any pointer is read or any launch is enqueued
  rc = lib.jmd_pair_energy_force_noalloc(
      None, nat.PAIR_NOALLOC_MAX_N + 1, 3, 3, 1, nat.f3([10.0, 10.0, 10.0]), ctypes.byref(pot),
      None, 0, nat.DENSE, None, None, None, None, 0.0, None, nat.stream_handle(DEV))
  assert rc == nat.JMD_ERR_INVALID
