"""partition.to_jraph (partition.py:1185-1250) on host tensors (CPU only).

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
