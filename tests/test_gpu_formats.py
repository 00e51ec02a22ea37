"""List-format utilities on GPU-built lists: neighbor_list_mask
(partition.py:1167-1182), to_dense (:1253-1271) and to_jraph (:1185-1250) on
device tensors.  Bars are exact: masks select exactly the valid entries,
to_dense rows hold exactly each particle's neighbour set of the Dense list
built from the same positions, to_jraph keeps the valid masked-in pairs in
order with their features and pads the rest."""

import numpy as np
import pytest
import torch

from oracle import jaxmd_oracle as O

pytestmark = pytest.mark.gpu

DEV = 'cuda:0'


def _lists(seed=0):
  import jax_md_amd as jmd
  R, _, L = O.lj_lattice(8, seed=seed)
  disp, _ = jmd.space.periodic(L)
  Rt = torch.as_tensor(R, device=DEV)
  out = {}
  for fmt in ('Dense', 'Sparse', 'OrderedSparse'):
    nfn = jmd.partition.neighbor_list(disp, L, 2.5, 0.5, format=getattr(jmd.partition, fmt))
    out[fmt] = nfn.allocate(Rt)
  nfn = jmd.partition.neighbor_list(disp, L, 2.5, 0.5, mask_self=False)
  out['DenseSelf'] = nfn.allocate(Rt)
  return jmd, out, R.shape[0]


def test_masks_and_to_dense_on_device():
  print('case formats on device', flush=True)
  jmd, nbs, N = _lists(1)
  for name, nb in nbs.items():
    m = jmd.partition.neighbor_list_mask(nb)
    assert m.device.type == 'cuda'
    ids = nb.idx[0] if name in ('Sparse', 'OrderedSparse') else nb.idx
    assert torch.equal(m, ids < N)
  ds = nbs['DenseSelf']
  k = int(jmd.partition.neighbor_list_mask(ds).sum() -
          jmd.partition.neighbor_list_mask(ds, mask_self=True).sum())
  assert k == N
  dense = jmd.partition.to_dense(nbs['Sparse'])
  assert dense.device.type == 'cuda'
  d = dense.cpu().numpy()
  ref = nbs['Dense'].idx.cpu().numpy()
  for i in range(N):
    assert np.array_equal(np.sort(d[i][d[i] < N]), np.sort(ref[i][ref[i] < N])), i
  assert dense.shape[1] == int((ref < N).sum(1).max())


@pytest.mark.parametrize('fmt', ['Sparse', 'OrderedSparse'])
def test_to_jraph_on_device(fmt):
  print(f'case to_jraph {fmt}', flush=True)
  jmd, nbs, N = _lists(2)
  nb = nbs[fmt]
  E = nb.idx.shape[1]
  rng = np.random.default_rng(3)
  mask = torch.as_tensor(rng.uniform(size=E) < 0.5, device=DEV)
  feats = {'w': torch.as_tensor(rng.normal(size=(E, 3)).astype(np.float32), device=DEV)}
  nodes = torch.ones((N, 2), device=DEV)
  g = jmd.partition.to_jraph(nb, mask=mask, nodes=nodes, edges=feats)
  r, s = nb.idx[0].cpu().numpy(), nb.idx[1].cpu().numpy()
  keep = (r < N) & mask.cpu().numpy()
  k = int(keep.sum())
  assert np.array_equal(g.receivers[:k].cpu().numpy(), r[keep])
  assert np.array_equal(g.senders[:k].cpu().numpy(), s[keep])
  assert torch.all(g.receivers[k:] == N) and torch.all(g.senders[k:] == N)
  assert np.array_equal(g.edges['w'][:k].cpu().numpy(), feats['w'].cpu().numpy()[keep])
  assert torch.all(g.edges['w'][k:] == 0)
  assert g.n_edge.tolist() == [k, E - k] and g.n_node.tolist() == [N, 1]
  assert g.nodes.shape == (N + 1, 2) and torch.all(g.nodes[N] == 0)
