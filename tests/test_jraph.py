"""partition.to_jraph (partition.py:1185-1250), neighbor_list_mask
(:1167-1182) and to_dense (:1253-1271) on host tensors of oracle-built lists
(CPU only; tests/test_gpu_formats.py repeats them on GPU-built lists).

Pins are properties of the conversion: without a mask the edges are the
list's entries unchanged; with a mask the kept edges are exactly the valid,
masked-in pairs in their original order, with their edge features, and every
other edge is a padding edge (N, N) with zero features; the padding graph adds
one zero node and one zero global; n_edge counts both parts.
"""

import dataclasses

import numpy as np
import pytest
import torch

from oracle import jaxmd_oracle as O


def _list(fmt_code, seed=1):
  import jax_md_amd as jmd
  R, _, L = O.lj_lattice(5, seed=seed)
  spec = O.NeighborListSpec(O.Space(np.asarray(L, np.float32)), np.asarray(L, np.float32),
                            2.0, 0.3, format=fmt_code)
  o = O.neighbor_list_build(spec, R)
  fmt = {O.SPARSE: jmd.partition.Sparse, O.ORDERED_SPARSE: jmd.partition.OrderedSparse,
         O.DENSE: jmd.partition.Dense}[fmt_code]
  nb = jmd.partition.NeighborList(
      idx=torch.as_tensor(o.idx), reference_position=torch.as_tensor(R),
      error=jmd.partition.PartitionError(torch.zeros((), dtype=torch.uint8)),
      cell_list_capacity=None, max_occupancy=o.max_occupancy, format=fmt, cell_size=None,
      cell_list_fn=None, update_fn=None)
  return jmd, nb, R.shape[0]


@pytest.mark.parametrize('fmt', [O.SPARSE, O.ORDERED_SPARSE])
def test_to_jraph_without_mask(fmt):
  jmd, nb, N = _list(fmt)
  nodes = torch.arange(N * 2, dtype=torch.float32).reshape(N, 2)
  glob = torch.ones((1, 3))
  g = jmd.partition.to_jraph(nb, nodes=nodes, globals=glob)
  assert torch.equal(g.receivers, nb.idx[0]) and torch.equal(g.senders, nb.idx[1])
  valid = int((nb.idx[0] < N).sum())
  assert g.n_node.tolist() == [N, 1]
  assert g.n_edge.tolist() == [valid, nb.idx.shape[1] - valid]
  assert g.nodes.shape == (N + 1, 2) and torch.equal(g.nodes[:N], nodes)
  assert torch.all(g.nodes[N] == 0)
  assert g.globals.shape == (2, 3) and torch.all(g.globals[1] == 0)


@pytest.mark.parametrize('fmt', [O.SPARSE, O.ORDERED_SPARSE])
def test_to_jraph_with_mask_compacts_in_order(fmt):
  jmd, nb, N = _list(fmt, seed=3)
  E = nb.idx.shape[1]
  rng = np.random.default_rng(0)
  mask = torch.as_tensor(rng.uniform(size=E) < 0.6)
  feats = {'w': torch.as_tensor(rng.normal(size=(E, 4)).astype(np.float32))}
  g = jmd.partition.to_jraph(nb, mask=mask, edges=feats)
  keep = ((nb.idx[0] < N) & mask).numpy()
  k = int(keep.sum())
  r, s = nb.idx[0].numpy(), nb.idx[1].numpy()
  assert np.array_equal(g.receivers[:k].numpy(), r[keep])
  assert np.array_equal(g.senders[:k].numpy(), s[keep])
  assert np.all(g.receivers[k:].numpy() == N) and np.all(g.senders[k:].numpy() == N)
  assert np.array_equal(g.edges['w'][:k].numpy(), feats['w'].numpy()[keep])
  assert np.all(g.edges['w'][k:].numpy() == 0)
  assert g.n_edge.tolist() == [k, E - k]


def test_to_jraph_refuses_dense():
  jmd, nb, _ = _list(O.DENSE)
  with pytest.raises(ValueError):
    jmd.partition.to_jraph(nb)


def test_neighbor_list_mask_properties():
  """neighbor_list_mask (partition.py:1167-1182): the valid entries of each
  format, and mask_self removing exactly the self pairs of a list built
  without self-masking."""
  import jax_md_amd as jmd
  for fmt in (O.DENSE, O.SPARSE, O.ORDERED_SPARSE):
    _, nb, N = _list(fmt, seed=4)
    m = jmd.partition.neighbor_list_mask(nb)
    ids = nb.idx[0] if fmt != O.DENSE else nb.idx
    assert torch.equal(m, ids < N)
  # a Dense list with self entries (mask_self=False in the build)
  R, _, L = O.lj_lattice(5, seed=5)
  L32 = np.asarray(L, np.float32)
  spec = O.NeighborListSpec(O.Space(L32), L32, 2.0, 0.3, format=O.DENSE, mask_self=False)
  o = O.neighbor_list_build(spec, R)
  N = R.shape[0]
  nb = jmd.partition.NeighborList(
      idx=torch.as_tensor(o.idx), reference_position=torch.as_tensor(R),
      error=jmd.partition.PartitionError(torch.zeros((), dtype=torch.uint8)),
      cell_list_capacity=None, max_occupancy=o.max_occupancy, format=jmd.partition.Dense,
      cell_size=None, cell_list_fn=None, update_fn=None)
  all_valid = jmd.partition.neighbor_list_mask(nb)
  no_self = jmd.partition.neighbor_list_mask(nb, mask_self=True)
  assert int(all_valid.sum()) - int(no_self.sum()) == N
  rows = torch.arange(N)[:, None].expand_as(nb.idx)
  assert not torch.any(no_self & (nb.idx == rows))


def test_to_dense_holds_each_senders_receivers():
  """to_dense (partition.py:1253-1271) of a Sparse list: shape [N, max
  receiver count], and row i holds exactly the receivers of i's pairs (the
  Dense list's neighbour set of i); OrderedSparse lists are refused."""
  import jax_md_amd as jmd
  _, nb, N = _list(O.SPARSE, seed=6)
  dense = jmd.partition.to_dense(nb)
  r, s = nb.idx[0].numpy(), nb.idx[1].numpy()
  valid = r < N
  counts = np.bincount(r[valid], minlength=N)
  assert dense.shape == (N, counts.max())
  d = dense.numpy()
  for i in range(N):
    got = np.sort(d[i][d[i] < N])
    want = np.sort(r[valid & (s == i)])
    assert np.array_equal(got, want), i
  _, nbd, _ = _list(O.DENSE, seed=6)
  for i in range(N):
    assert np.array_equal(np.sort(d[i][d[i] < N]), np.sort(nbd.idx[i].numpy()[nbd.idx[i].numpy() < N]))
  _, nbo, _ = _list(O.ORDERED_SPARSE, seed=6)
  with pytest.raises(ValueError):
    jmd.partition.to_dense(nbo)
