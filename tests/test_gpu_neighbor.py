"""Neighbour-list parity on the GPU: libjaxmd_native.so against the oracle.

Bar: bit-exact.  Dense rows are compared element by element in the
reference's candidate order (not just as sorted sets), Sparse lists entry by
entry, and the error word, max_occupancy and cell capacity exactly.
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


def _fmt(jmd, code):
  return {O.DENSE: jmd.partition.Dense, O.SPARSE: jmd.partition.Sparse,
          O.ORDERED_SPARSE: jmd.partition.OrderedSparse}[code]


def _both(R, box, r_cut, dr, fmt=O.DENSE, cm=1.25, dc=False, ms=True, periodic=True):
  """Allocate on GPU and in the oracle with the same inputs."""
  jmd = _jmd()
  if periodic:
    disp, _ = jmd.space.periodic(box)
    osp = O.Space(np.asarray(box, F32))
  else:
    disp, _ = jmd.space.free()
    osp = O.Space(None, periodic=False)
  nfn = jmd.partition.neighbor_list(disp, box, r_cut, dr, capacity_multiplier=cm,
                                    disable_cell_list=dc, mask_self=ms, format=_fmt(jmd, fmt))
  spec = O.NeighborListSpec(osp, np.asarray(box, F32), r_cut, dr, cm, dc, ms, fmt)
  Rt = torch.as_tensor(R, device=DEV)
  g = nfn.allocate(Rt)
  o = O.neighbor_list_build(spec, R)
  return nfn, spec, g, o


def _assert_same(g, o, case):
  gi = g.idx.cpu().numpy()
  assert gi.shape == o.idx.shape, (case, gi.shape, o.idx.shape)
  assert g.max_occupancy == o.max_occupancy, case
  assert g.cell_list_capacity == o.cell_list_capacity, case
  assert int(g.error.code.item()) == o.error, (case, int(g.error.code.item()), o.error)
  if not (o.error & 3):
    bad = np.argwhere(gi != o.idx)
    assert bad.size == 0, (case, bad[:5], gi[tuple(bad[0])] if bad.size else None)


def _uniform(N, box, seed):
  rng = np.random.default_rng(seed)
  box = np.asarray(box, F32)
  return (box * rng.uniform(size=(N, box.size if box.ndim else 3))).astype(F32)


@pytest.mark.parametrize('fmt', [O.DENSE, O.SPARSE, O.ORDERED_SPARSE])
def test_lattice_c1_bit_exact(fmt):
  R, _, L = O.lj_lattice(16, seed=0)
  print(f'case lattice N=4096 fmt={fmt}', flush=True)
  _, _, g, o = _both(R, L, 2.5, 0.5, fmt)
  _assert_same(g, o, f'lattice fmt={fmt}')


@pytest.mark.parametrize('dim', [2, 3])
@pytest.mark.parametrize('fmt', [O.DENSE, O.SPARSE])
def test_random_vector_box_bit_exact(dim, fmt):
  box = np.array([9.0, 4.0, 7.25], F32) if dim == 3 else np.array([9.0, 4.25], F32)
  R = _uniform(1000, box, 11 + dim)
  print(f'case vector box dim={dim} fmt={fmt}', flush=True)
  _, _, g, o = _both(R, box, F32(1.23), 0.0, fmt, cm=1.1)
  _assert_same(g, o, f'vector box dim={dim} fmt={fmt}')


@pytest.mark.parametrize('fmt', [O.DENSE, O.SPARSE, O.ORDERED_SPARSE])
@pytest.mark.parametrize('ms', [True, False])
def test_all_pairs_path_bit_exact(fmt, ms):
  # cutoff >= L/3: the reference builds from all pairs (partition.py:904-909)
  R = _uniform(300, F32(5.0), 21)
  print(f'case all-pairs fmt={fmt} mask_self={ms}', flush=True)
  _, _, g, o = _both(R, F32(5.0), 2.0, 0.3, fmt, ms=ms)
  assert not o.use_cell_list
  _assert_same(g, o, f'all-pairs fmt={fmt} ms={ms}')


def test_free_space_bit_exact():
  R = _uniform(800, F32(20.0), 31)
  print('case free space', flush=True)
  _, _, g, o = _both(R, F32(20.0), 1.5, 0.2, O.DENSE, periodic=False)
  _assert_same(g, o, 'free space')


@pytest.mark.parametrize('r_cut,dc,cm', [(0.12, True, 1.5), (0.25, False, 1.5),
                                         (0.31, False, 1.5), (0.31, False, 1.0)])
@pytest.mark.parametrize('ms', [False, True])
@pytest.mark.parametrize('fmt', [O.DENSE, O.SPARSE, O.ORDERED_SPARSE])
def test_capacity_shapes(r_cut, dc, cm, ms, fmt):
  # partition_test.py:476-546: 20 coincident particles
  box = np.ones(3, F32)
  R = np.full((20, 3), 0.5, F32)
  print(f'case issue191 r={r_cut} dc={dc} cm={cm} ms={ms} fmt={fmt}', flush=True)
  nfn, spec, g, o = _both(R, box, F32(r_cut), F32(0.1) * F32(r_cut), fmt, cm=cm, dc=dc, ms=ms)
  _assert_same(g, o, 'issue191 allocate')
  R2 = (R + F32(0.1)).astype(F32)
  g2 = g.update(torch.as_tensor(R2, device=DEV))
  o2 = O.neighbor_list_update(spec, R2, o)
  _assert_same(g2, o2, 'issue191 update')


def test_update_skin_and_rebuild():
  R, _, L = O.lj_lattice(10, seed=3)
  nfn, spec, g, o = _both(R, L, 2.5, 0.5, O.DENSE)
  rng = np.random.default_rng(4)
  Rk = R.copy()
  for step in range(6):
    # small moves: below the skin most of the time, occasionally one large move
    Rk = (Rk + rng.normal(0, 0.03, Rk.shape)).astype(F32)
    if step == 3:
      Rk[5] += F32(0.4)
    Rk = np.mod(Rk, L).astype(F32)
    print(f'case update step={step}', flush=True)
    g = g.update(torch.as_tensor(Rk, device=DEV))
    o = O.neighbor_list_update(spec, Rk, o)
    _assert_same(g, o, f'update step {step}')
    assert np.array_equal(g.reference_position.cpu().numpy(), o.reference_position)


def test_forced_overflow_free_space():
  # partition_test.py:363-401
  jmd = _jmd()
  disp, _ = jmd.space.free()
  nfn = jmd.partition.neighbor_list(disp, 100.0, 3.0, 0.0)
  R = torch.tensor([[20.0, 20.0], [30.0, 30.0], [40.0, 40.0], [50.0, 50.0]], device=DEV)
  nbrs = nfn.allocate(R)
  assert nbrs.idx.dtype == torch.int32 and tuple(nbrs.idx.shape) == (4, 0)
  assert not bool(nbrs.did_buffer_overflow)
  R2 = torch.tensor([[20.0, 20.0], [20.0, 20.0], [40.0, 40.0], [50.0, 50.0]], device=DEV)
  nbrs = nbrs.update(R2)
  assert bool(nbrs.did_buffer_overflow)


def test_neighbor_overflow_sticky_then_reallocate():
  R, _, L = O.lj_lattice(8, seed=5)
  nfn, spec, g, o = _both(R, L, 2.5, 0.5, O.DENSE, cm=1.0)
  # compress a region: more neighbours than max_occupancy
  R2 = R.copy()
  R2[:40] = (R2[:40] * F32(0.9)).astype(F32)
  g2 = g.update(torch.as_tensor(R2, device=DEV))
  o2 = O.neighbor_list_update(spec, R2, o)
  assert int(g2.error.code.item()) == o2.error and (o2.error & 1)
  g3 = g2.update(torch.as_tensor(R, device=DEV))
  assert int(g3.error.code.item()) & 1  # sticky (partition.py:1066,1110)
  g4 = nfn.allocate(torch.as_tensor(R2, device=DEV))
  _assert_same(g4, O.neighbor_list_build(spec, R2), 'reallocate')


def test_cell_list_known_slots():
  # partition_test.py:61-89 and the scatter round trip
  jmd = _jmd()
  box = np.array([8.65, 8.0], F32)
  R = np.array([[0.25, 0.25], [8.5, 1.95], [8.1, 1.5], [3.7, 7.9]], F32)
  cl = jmd.partition.cell_list(box, F32(1.0)).allocate(torch.as_tensor(R, device=DEV))
  o = O.cell_list(R, box, F32(1.0))
  ids = cl.id_buffer.cpu().numpy()
  assert ids[0, 0, 0, 0] == 0 and ids[1, 7, 1, 0] == 1 and ids[1, 7, 0, 0] == 2
  assert ids[7, 3, 1, 0] == 3
  assert np.array_equal(ids.reshape(-1), o.id_buffer)
  assert np.array_equal(cl.position_buffer.cpu().numpy().reshape(-1, 2), o.position_buffer)
  assert np.array_equal(cl.particle_cell_id.cpu().numpy(), o.particle_cell_id)
  assert not bool(cl.did_buffer_overflow)


@pytest.mark.parametrize('dim', [2, 3])
def test_cell_list_random_roundtrip(dim):
  jmd = _jmd()
  R = (F32(9.0) * np.random.default_rng(dim).uniform(size=(1000, dim))).astype(F32)
  cl = jmd.partition.cell_list(F32(9.0), F32(1.0)).allocate(torch.as_tensor(R, device=DEV))
  o = O.cell_list(R, F32(9.0), F32(1.0))
  assert cl.cell_capacity == o.cell_capacity
  assert np.array_equal(cl.id_buffer.cpu().numpy().reshape(-1), o.id_buffer)
  out = np.zeros((1001, dim), F32)
  out[cl.id_buffer.cpu().numpy().reshape(-1)] = cl.position_buffer.cpu().numpy().reshape(-1, dim)
  assert np.array_equal(out[:-1], R)


def test_c2_scale_properties_and_sampled_rows():
  """N = 262,144 (64^3 lattice): symmetry, row sortedness-by-candidate-order
  checked on 3000 sampled rows against the oracle, and global counts."""
  R, _, L = O.lj_lattice(64, seed=0)
  jmd = _jmd()
  disp, _ = jmd.space.periodic(L)
  nfn = jmd.partition.neighbor_list(disp, L, 2.5, 0.5, format=jmd.partition.Dense)
  print('case C2 allocate', flush=True)
  g = nfn.allocate(torch.as_tensor(R, device=DEV))
  assert int(g.error.code.item()) == 0
  idx = g.idx.cpu().numpy()
  N = R.shape[0]
  spec = O.NeighborListSpec(O.Space(L), L, 2.5, 0.5)
  cl = O.build_cell_list_for(spec, R)
  assert g.cell_list_capacity == cl.cell_capacity
  rows = np.random.default_rng(0).choice(N, 3000, replace=False)
  want = O.neighbor_rows(spec, R, cl, rows)
  for r, w in zip(rows, want):
    got = idx[r][idx[r] < N]
    assert np.array_equal(got, w), r
  # symmetry on the full list: (i, j) present iff (j, i) present
  valid = idx < N
  ii = np.broadcast_to(np.arange(N)[:, None], idx.shape)[valid].astype(np.int64)
  jj = idx[valid].astype(np.int64)
  a = np.sort(ii * N + jj)
  b = np.sort(jj * N + ii)
  assert np.array_equal(a, b)
  counts = valid.sum(1)
  assert g.max_occupancy == int(F32(counts.max()) * F32(1.25))


@pytest.mark.parametrize('fmt', [O.DENSE, O.SPARSE])
def test_padded_layout_same_list(fmt):
  """A [N, 3] view of a 16-byte-padded [N, 4] buffer builds the same list."""
  jmd = _jmd()
  R, _, L = O.lj_lattice(12, seed=7)
  _, spec, g, o = _both(R, L, 2.5, 0.5, fmt)
  nfn = jmd.partition.neighbor_list(jmd.space.periodic(L)[0], L, 2.5, 0.5,
                                    format=_fmt(jmd, fmt))
  buf = torch.zeros((R.shape[0], 4), dtype=torch.float32, device=DEV)
  buf[:, :3] = torch.as_tensor(R, device=DEV)
  gp = nfn.allocate(buf[:, :3])
  _assert_same(gp, o, f'padded fmt={fmt}')
  assert torch.equal(gp.idx, g.idx)



@pytest.mark.parametrize('case', ['cps3', 'cps4', 'dim2', 'far', 'padded'])
def test_dense_cells_geometry_cases_with_updates(case):
  """Dense builds and gated updates against the oracle where the build's
  fast test needs care: 3 or 4 cells per side (the staged neighbour image is
  not always the nearest: minimum-image fold), 2-D, coordinates up to three
  periods outside the box (canonicalisation rounds at |x|: exact test for
  every candidate), and the padded [N, 4] layout."""
  jmd = _jmd()
  rng = np.random.default_rng(41)
  r_cut, dr = 2.5, 0.5
  if case in ('padded', 'far'):
    R, _, box = O.lj_lattice(12, seed=9)
  elif case == 'cps3':
    box = F32(9.5)  # list cutoff 3.0 -> 3 cells per side
    R = _uniform(600, box, 42)
  elif case == 'cps4':
    box = np.array([12.5, 13.0, 16.0], F32)  # 4, 4, 5 cells
    R = _uniform(1500, box, 43)
  else:
    box = np.array([20.0, 17.0], F32)
    R = _uniform(900, box, 44)
  if case == 'far':
    R = (R + F32(box) * rng.integers(-3, 4, size=R.shape).astype(F32)).astype(F32)
  disp, _ = jmd.space.periodic(box)
  nfn = jmd.partition.neighbor_list(disp, box, r_cut, dr, format=jmd.partition.Dense)
  osp = O.Space(np.asarray(box, F32))
  spec = O.NeighborListSpec(osp, np.asarray(box, F32), r_cut, dr, 1.25, False, True, O.DENSE)
  buf = None
  if case == 'padded':
    buf = torch.zeros((R.shape[0], 4), dtype=torch.float32, device=DEV)

  def dev_pos(X):
    if buf is None:
      return torch.as_tensor(X, device=DEV)
    buf[:, :3] = torch.as_tensor(X, device=DEV)
    return buf[:, :3]

  print(f'case dense geometry {case} N={R.shape[0]}', flush=True)
  g = nfn.allocate(dev_pos(R))
  o = O.neighbor_list_build(spec, R)
  assert o.use_cell_list
  _assert_same(g, o, f'{case} allocate')
  Rk = R.copy()
  for step in range(3):
    Rk = (Rk + rng.normal(0, 0.05, Rk.shape)).astype(F32)
    if step == 1:
      Rk[3] += F32(0.6)  # beyond the skin: forces a rebuild
    g = g.update(dev_pos(Rk))
    o = O.neighbor_list_update(spec, Rk, o)
    _assert_same(g, o, f'{case} update {step}')
    assert np.array_equal(g.reference_position.cpu().numpy(), o.reference_position)

def test_unwrapped_positions_bit_exact_sparse():
  """Coordinates several periods outside the box (Sparse phase kernels)."""
  R, _, L = O.lj_lattice(10, seed=12)
  rng = np.random.default_rng(13)
  R = (R + F32(L) * rng.integers(-3, 4, size=R.shape).astype(F32)).astype(F32)
  print('case unwrapped sparse', flush=True)
  _, _, g, o = _both(R, L, 2.5, 0.5, O.SPARSE)
  _assert_same(g, o, 'unwrapped sparse')


def test_sparse_formats_follow_dense_rows_at_64000():
  """N = 64,000 (many chunks of the row-offset scan): the Sparse list is the
  Dense rows flattened in row-major order (receiver j, sender i), and the
  OrderedSparse list the same with j < i (partition.py:982-1032); padding
  fills every entry after the total.  The Dense list's rows are checked
  against the oracle at this scale by the sampled-row test."""
  R, _, L = O.lj_lattice(40, seed=21)
  jmd = _jmd()
  disp, _ = jmd.space.periodic(L)
  Rt = torch.as_tensor(R, device=DEV)
  N = R.shape[0]
  print(f'case sparse vs dense rows N={N}', flush=True)
  d = jmd.partition.neighbor_list(disp, L, 2.5, 0.5, format=jmd.partition.Dense).allocate(Rt)
  idx = d.idx.cpu().numpy()
  valid = idx < N
  send = np.broadcast_to(np.arange(N, dtype=np.int32)[:, None], idx.shape)
  for fmt in (jmd.partition.Sparse, jmd.partition.OrderedSparse):
    keep = valid & (idx < send) if fmt is jmd.partition.OrderedSparse else valid
    want_r, want_s = idx[keep], send[keep]
    s = jmd.partition.neighbor_list(disp, L, 2.5, 0.5, format=fmt).allocate(Rt)
    assert int(s.error.code.item()) == 0
    got = s.idx.cpu().numpy()
    k = want_r.size
    print(f'  {fmt}: {k} pairs, capacity {got.shape[1]}')
    assert got.shape[1] >= k
    assert np.array_equal(got[0, :k], want_r) and np.array_equal(got[1, :k], want_s)
    assert np.all(got[:, k:] == N)


def test_custom_mask_function_known_answer_and_oracle():
  """custom_mask_function (partition.py:1079-1080): the reference's known
  answer (10 coincident particles, |id_i - id_j| > 3 kept: 42 entries,
  partition_test.py:403-459), and on a periodic lattice the masked Dense rows
  equal the oracle's entry by entry (order, width, error word) after
  allocate and after an update that rebuilds."""
  import jax_md_amd as jmd
  print('case custom mask', flush=True)
  disp, _ = jmd.space.free()
  mask = lambda idx: torch.where((idx - torch.arange(idx.shape[0], device=idx.device)[:, None]).abs()
                                 > 3, idx, torch.full_like(idx, idx.shape[0]))
  nfn = jmd.partition.neighbor_list(disp, 1.0, 3.0, 0.0, custom_mask_function=mask)
  R = torch.zeros((10, 3), dtype=torch.float32, device=DEV)
  nb = nfn.allocate(R)
  nb = nb.update(R)
  assert int((nb.idx != 10).sum()) == 42
  # periodic lattice, parity mask: keep pairs of different id parity
  Rl, _, L = O.lj_lattice(8, seed=23)
  N = Rl.shape[0]
  dispp, _ = jmd.space.periodic(L)
  pmask = lambda idx: torch.where((idx % 2) != (torch.arange(N, device=idx.device)[:, None] % 2),
                                  idx, torch.full_like(idx, N))
  omask = lambda cand, rows: np.where((cand % 2) != (rows[:, None] % 2), cand, N)
  nfn = jmd.partition.neighbor_list(dispp, L, 2.5, 0.5, custom_mask_function=pmask)
  spec = O.NeighborListSpec(O.Space(L), L, 2.5, 0.5, custom_mask_function=omask)
  nb = nfn.allocate(torch.as_tensor(Rl, device=DEV))
  onl = O.neighbor_list_build(spec, Rl)
  assert np.array_equal(nb.idx.cpu().numpy(), onl.idx)
  assert int(nb.error.code) == int(onl.error)
  R2 = np.mod(Rl + np.random.default_rng(24).normal(0, 0.4, Rl.shape), np.float32(L)).astype(np.float32)
  R2 = np.where(R2 >= L, np.float32(0), R2).astype(np.float32)
  nb = nb.update(torch.as_tensor(R2, device=DEV))
  onl2 = O.neighbor_list_update(spec, R2, onl)
  assert np.array_equal(nb.idx.cpu().numpy(), onl2.idx)
  assert int(nb.error.code) == int(onl2.error)
