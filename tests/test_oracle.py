"""Pins for the CPU oracle (oracle/jaxmd_oracle.py), CPU only.

Which check pins which oracle function:

  jnp_mod, periodic_displacement   numpy.mod cross-check; range and
                                   congruence invariants (test_wrap_*)
  square_distance                  exact (fsum) value within 2 ulp
  cell_dimensions, hash_constants,
  neighboring_cells, cell_list     reference known-answer slots of
                                   partition_test.py:61-89 (test_cell_emplace_2d)
                                   and the scatter round trip (:102-131)
  neighbor_list_build (Dense)      brute-force O(N^2) sets
                                   (partition_test.py:203-246), symmetry,
                                   cell path == all-pairs path, vector box
  neighbor_list_build (Sparse)     brute force (partition_test.py:259-301)
  capacities / error bits          partition_test.py:363-401 (overflow),
                                   :403-459 (custom mask = 42), :476-546
                                   (capacity shapes)
  pair_energy_force (SS)           jammed fixture energy 0.45247561922261154
  pair_energy_force (LJ)           LAMMPS E/N -4.3523016, closed forms
                                   (energy_test.py:325-347), list == brute force
  forces                           central finite differences of the energy
  smooth                           S(r_on)=1, S(r_c)=0, C1 continuity
  nve_step                         energy conservation, time reversibility
  fire_step                        static_cast float16 values (F4), monotone
                                   descent on a quadratic (minimize_test.py:111-150),
                                   and step by step a hand-derived float16
                                   dt / alpha / n_pos table for a scheduled force
                                   (growth after n_pos > n_min, the dt_max cap, a
                                   reset on F.P < 0, F.P == 0 twice)
  needs_rebuild                    (dr/2)^2 exactly and one float32 step beyond,
                                   across a periodic face (an offset the float32
                                   wrap rounds away does not rebuild); every pair
                                   within r_cutoff stays listed along a random
                                   walk; an adversarial pair that a dr threshold
                                   would miss
  force_vjp                        central differences of the force in R,
                                   sigma, epsilon; a dropped term fails
  force_vjp_per_particle           central differences in each particle's
                                   sigma and epsilon (force and energy); the
                                   row-only form fails
  fire_energy_gradient             central differences of fire_energy_after
                                   in r and x_0, also across momentum resets
  box_vjp, fire_energy_gradient    central differences in the box side (fixed
    (box=True)                     positions, and through 40 FIRE steps that
                                   fold particles across the boundary) and in
                                   the diameter ratio with the box following
                                   it; the fold term is material (24%)
  pair_energies                    closed-form pair energies per entry (switched
                                   LJ, species soft spheres); the entries sum
                                   to the brute-force energy
  nvt_step (Nose-Hoover chain)     invariant conserved for sy = 1, 3, 5, 7 on
                                   the jammed fixture (simulate_test.py:
                                   317-350), second order in dt
  pair_virial, stress, pressure    LAMMPS stress fixture (quantity_test.py:
                                   436-461), jammed pressure fixture
                                   (:134-146), central differences of an
                                   affine strain
  GeneralSpace, transform          diagonal box == periodic on x = T u; an
                                   FCC crystal has one E/N as a cubic cell and
                                   as a rotated triclinic primitive supercell
                                   (the transposed matrix fails); forces are
                                   real-space central differences; the shift's
                                   displacement returns dR
  fractional_cell_size             perpendicular widths from cross products
                                   (V / |a_j x a_k|) for triclinic boxes,
                                   scalar and vector closed forms
  neighbor_list_build (fractional) brute-force sets in the triclinic box
  multi_image_list                 a stencil two shells wider finds no more
                                   pairs; in a cell larger than twice the
                                   cutoff the pairs are the general-space
                                   neighbour sets; OrderedSparse holds one of
                                   each (i, j, s) / (j, i, -s)
"""

import math

import numpy as np
import pytest

from oracle import jaxmd_oracle as O
from tests.golden import fixtures

F32 = np.float32


# ---------------------------------------------------------------------------
# geometry
# ---------------------------------------------------------------------------


def test_wrap_matches_numpy_mod_f32():
  rng = np.random.default_rng(0)
  L = F32(16.929268)
  x = rng.uniform(-3 * L, 3 * L, 20000).astype(F32)
  assert np.array_equal(O.jnp_mod(x, L), np.mod(x, L).astype(F32))


def test_wrap_range_and_congruence():
  rng = np.random.default_rng(1)
  for L in (5.0, 67.71707, 1.0):
    d = rng.uniform(-2.5 * L, 2.5 * L, 10000)
    w = O.periodic_displacement(L, d)
    assert np.all(w >= -L / 2 - 1e-12) and np.all(w < L / 2 + 1e-12)
    k = (d - w) / L
    assert np.allclose(k, np.round(k), atol=1e-9)
    # float32 path: same congruence up to float32 rounding of L/2 terms
    w32 = O.periodic_displacement(F32(L), d.astype(F32))
    assert np.allclose(w32, w, atol=4 * np.spacing(F32(1.5 * L)))


def test_square_distance_rounding():
  rng = np.random.default_rng(2)
  d = rng.normal(size=(2000, 3)).astype(F32)
  got = O.square_distance(d)
  exact = np.array([math.fsum(float(v) ** 2 for v in row) for row in d])
  assert np.all(np.abs(got - exact) <= 2 * np.spacing(got))


# ---------------------------------------------------------------------------
# cell list
# ---------------------------------------------------------------------------


def test_cell_emplace_2d_known_slots():
  # partition_test.py:61-89; index 8 of an 8-cell axis is clamped to 7 by
  # JAX's gather, so [1, 8, k] reads cell (y=1, x=7).
  box = np.array([8.65, 8.0], F32)
  R = np.array([[0.25, 0.25], [8.5, 1.95], [8.1, 1.5], [3.7, 7.9]], F32)
  cl = O.cell_list(R, box, F32(1.0))
  cps = cl.cells_per_side
  shape = tuple(int(c) for c in cps[::-1]) + (cl.cell_capacity,)
  ids = cl.id_buffer.reshape(shape)
  pos = cl.position_buffer.reshape(shape + (2,))
  assert ids[0, 0, 0] == 0 and ids[1, 7, 1] == 1 and ids[1, 7, 0] == 2 and ids[7, 3, 1] == 3
  assert np.array_equal(pos[1, 7, 1], R[1]) and np.array_equal(pos[7, 3, 1], R[3])
  # scatter round trip (partition_test.py:82-88)
  out = np.zeros((5, 2), F32)
  out[cl.id_buffer] = cl.position_buffer
  assert np.array_equal(out[:-1], R)


@pytest.mark.parametrize('dim', [2, 3])
def test_cell_list_random_roundtrip(dim):
  rng = np.random.default_rng(3)
  R = (F32(9.0) * rng.uniform(size=(1000, dim))).astype(F32)
  cl = O.cell_list(R, F32(9.0), F32(1.0))
  out = np.zeros((1001, dim), F32)
  out[cl.id_buffer] = cl.position_buffer
  assert np.array_equal(out[:-1], R)
  assert not cl.did_buffer_overflow


# ---------------------------------------------------------------------------
# neighbour lists
# ---------------------------------------------------------------------------


def _brute_sets(R, side, cutoff, periodic=True):
  sp = O.Space(np.asarray(side, np.float64), periodic)
  i, j = O.all_pairs(sp, R.astype(np.float64), float(cutoff))
  N = R.shape[0]
  rows = [[] for _ in range(N)]
  for a, b in zip(i, j):
    rows[a].append(b)
  dR = sp.displacement(R[None, :, :].astype(np.float64), R[:, None, :].astype(np.float64))
  dist = np.sqrt(np.sum(dR ** 2, -1))
  return [np.sort(r) for r in rows], dist


def _assert_sets_match(rows, brute, dist, cutoff, tol=1e-5):
  for i, (a, b) in enumerate(zip(rows, brute)):
    sa, sb = set(a.tolist()), set(b.tolist())
    for j in sa ^ sb:
      assert abs(dist[i, j] - cutoff) < tol, (i, j, dist[i, j])


@pytest.mark.parametrize('dim', [2, 3])
def test_dense_vs_brute_vector_box(dim):
  # partition_test.py:203-246
  rng = np.random.default_rng(4)
  box = np.array([9.0, 4.0, 7.25], F32)[:dim] if dim == 3 else np.array([9.0, 4.25], F32)
  R = (box * rng.uniform(size=(1000, dim))).astype(F32)
  spec = O.NeighborListSpec(O.Space(box), box, F32(1.23), 0.0, 1.1)
  nl = O.neighbor_list_build(spec, R)
  assert nl.use_cell_list and nl.error == 0
  rows = O.sorted_rows(nl)
  brute, dist = _brute_sets(R, box, F32(1.23))
  _assert_sets_match(rows, brute, dist, 1.23)


def test_dense_lattice_symmetry_and_paths():
  R, _, L = O.lj_lattice(8, seed=5)
  spec = O.NeighborListSpec(O.Space(L), L, 2.5, 0.5)
  nl = O.neighbor_list_build(spec, R)
  rows = O.sorted_rows(nl)
  pairs = {(i, int(j)) for i, r in enumerate(rows) for j in r}
  assert all((j, i) in pairs for i, j in pairs)
  spec_nc = O.NeighborListSpec(O.Space(L), L, 2.5, 0.5, disable_cell_list=True)
  rows_nc = O.sorted_rows(O.neighbor_list_build(spec_nc, R))
  assert all(np.array_equal(a, b) for a, b in zip(rows, rows_nc))
  brute, dist = _brute_sets(R, L, 3.0)
  _assert_sets_match(rows, brute, dist, 3.0)


@pytest.mark.parametrize('fmt', [O.SPARSE, O.ORDERED_SPARSE])
def test_sparse_vs_brute(fmt):
  rng = np.random.default_rng(6)
  box = np.array([9.0, 4.0, 7.25], F32)
  R = (box * rng.uniform(size=(600, 3))).astype(F32)
  spec = O.NeighborListSpec(O.Space(box), box, F32(1.23), 0.0, 1.1, format=fmt)
  nl = O.neighbor_list_build(spec, R)
  N = R.shape[0]
  recv, send = nl.idx
  m = recv < N
  brute, dist = _brute_sets(R, box, F32(1.23))
  # sender-major order (partition.py:1000-1005)
  assert np.all(np.diff(send[m]) >= 0)
  for i in range(N):
    got = np.sort(recv[m][send[m] == i])
    want = brute[i] if fmt == O.SPARSE else brute[i][brute[i] < i]
    _assert_sets_match([got], [want], dist[i:i + 1], 1.23)


def test_overflow_forced_free_space():
  # partition_test.py:363-401
  sp = O.Space(None, periodic=False)
  spec = O.NeighborListSpec(sp, F32(100.0), 3.0, 0.0)
  R = np.array([[20, 20], [30, 30], [40, 40], [50, 50]], F32)
  nl = O.neighbor_list_build(spec, R)
  assert nl.error == 0 and nl.idx.shape == (4, 0)
  R2 = np.array([[20, 20], [20, 20], [40, 40], [50, 50]], F32)
  nl2 = O.neighbor_list_update(spec, R2, nl)
  assert nl2.error & (O.NEIGHBOR_LIST_OVERFLOW | O.CELL_LIST_OVERFLOW)


def test_custom_mask_known_answer():
  # partition_test.py:403-459: 10 coincident particles, |id_i - id_j| > 3 -> 42
  sp = O.Space(None, periodic=False)
  mask = lambda cand, rows: np.where(np.abs(cand - rows[:, None]) > 3, cand, 10)
  spec = O.NeighborListSpec(sp, F32(1.0), 3.0, 0.0, custom_mask_function=mask)
  R = np.zeros((10, 3), F32)
  nl = O.neighbor_list_build(spec, R)
  nl = O.neighbor_list_update(spec, R, nl)
  assert int(np.sum(nl.idx != 10)) == 42


@pytest.mark.parametrize('r_cut,dc,cm', [(0.12, True, 1.5), (0.25, False, 1.5),
                                         (0.31, False, 1.5), (0.31, False, 1.0)])
@pytest.mark.parametrize('mask_self', [False, True])
@pytest.mark.parametrize('fmt', [O.DENSE, O.SPARSE, O.ORDERED_SPARSE])
def test_capacity_shapes(r_cut, dc, cm, mask_self, fmt):
  # partition_test.py:476-546 (issue 191)
  box = np.ones(3, F32)
  R = np.full((20, 3), 0.5, F32)
  spec = O.NeighborListSpec(O.Space(box), box, F32(r_cut), F32(0.1) * F32(r_cut), cm, dc,
                            mask_self, fmt)
  nl = O.neighbor_list_build(spec, R)
  want = {O.DENSE: (20, 19) if mask_self else (20, 20),
          O.SPARSE: (2, 20 * 19) if mask_self else (2, 400),
          O.ORDERED_SPARSE: (2, 190)}[fmt]
  assert nl.idx.shape == want
  assert not (nl.error & 3)
  nl2 = O.neighbor_list_update(spec, (R + F32(0.1)).astype(F32), nl)
  assert not (nl2.error & 3) and nl2.idx.shape == want


# ---------------------------------------------------------------------------
# energies and forces
# ---------------------------------------------------------------------------


def test_jammed_fixture_energy():
  st = fixtures.jammed_state()
  box = np.diag(st['box'])
  sp = O.Space(box)
  m = O.PairModel('soft_sphere', sigma=st['sigma'], species=st['species'])
  E, _, F = O.energy_force_brute(m, sp, st['real_position'])
  assert abs(E - float(st['energy'])) < 1e-12
  # same through a Dense list built in float64 arithmetic
  spec = O.NeighborListSpec(O.Space(box), box.astype(F32), float(np.max(st['sigma'])), 0.2)
  nl = O.neighbor_list_build(spec, st['real_position'].astype(np.float64))
  i, j = O.pairs_from_dense(nl.idx)
  E2, _, F2 = O.pair_energy_force(m, sp, st['real_position'], i, j)
  assert abs(E2 - float(st['energy'])) < 1e-12
  assert np.allclose(F, F2, atol=1e-12)


def test_lammps_lj_energy():
  d = fixtures.lammps_lj()
  L = d['box']
  R = d['fractional_position'] * L
  m = O.PairModel('lj', r_onset=2.5, r_cutoff=2.5)
  E, _, _ = O.energy_force_brute(m, O.Space(np.float64(L)), R)
  assert abs(E / len(R) - d['energy_per_atom']) < 5e-5


def _lammps_stress(d):
  v = d['stress_voigt']
  return np.array([[v[0], v[3], v[4]], [v[3], v[1], v[5]], [v[4], v[5], v[2]]])


def test_lammps_lj_stress():
  """The virial stress against LAMMPS (quantity_test.py:436-461, tol 5e-5)."""
  d = fixtures.lammps_lj()
  L = d['box']
  R = d['fractional_position'] * L
  m = O.PairModel('lj', r_onset=2.5, r_cutoff=2.5)
  sp = O.Space(np.float64(L))
  i, j = O.all_pairs(sp, R, 2.5)
  C = O.stress(O.pair_virial(m, sp, R, i, j), L ** 3, d['velocity'])
  assert np.allclose(C, _lammps_stress(d), atol=5e-5, rtol=5e-5)


def test_jammed_pressure():
  """The virial pressure of the jammed packing (quantity_test.py:134-146)."""
  st = fixtures.jammed_state()
  L = np.diag(st['box']).astype(np.float64)
  sp = O.Space(L)
  m = O.PairModel('soft_sphere', sigma=st['sigma'], species=st['species'])
  R = st['real_position']
  i, j = O.all_pairs(sp, R, float(np.max(st['sigma'])))
  P = O.pressure(O.pair_virial(m, sp, R, i, j), float(np.prod(L)))
  assert abs(P - float(st['pressure'])) <= 1e-12 * abs(float(st['pressure']))


def test_virial_is_the_box_derivative():
  """W = dU/d eps with every displacement scaled by (I + eps): central
  differences of the energy under an affine box strain (a sign or transpose
  slip in W fails)."""
  R, L, model, cutoff = _d2_system('lj', 5)
  sp = O.Space(np.float64(L))
  i, j = O.all_pairs(sp, R, cutoff + 0.3)
  W = O.pair_virial(model, sp, R, i, j)
  h = 1e-6
  for a, b in ((0, 0), (0, 1), (2, 1)):
    E = []
    for sgn in (1, -1):
      eps = np.zeros((3, 3))
      eps[a, b] = sgn * h
      dR = sp.displacement(R[j], R[i]) @ (np.eye(3) + eps).T
      r2 = np.sum(dR * dR, -1)
      _, e = O.lj_dE_over_r(r2, 1.0, 1.0, *model.smoothing())
      E.append(0.5 * np.sum(e))
    fd = (E[0] - E[1]) / (2 * h)
    assert abs(fd - W[a, b]) <= 1e-6 * (abs(W[a, b]) + 1.0), (a, b, fd, W[a, b])


def test_lj_closed_forms():
  r_min = 2 ** (1 / 6)
  assert abs(O.lj_energy_pair(r_min) + 1.0) < 1e-12
  dEr, E = O.lj_dE_over_r(r_min ** 2, 1.0, 1.0, 2.0, 2.5)
  assert abs(E + 1.0) < 1e-12 and abs(dEr) < 1e-12
  # sigma, epsilon scaling
  assert abs(O.lj_energy_pair(2.0 * r_min, 2.0, 3.0) + 3.0) < 1e-12


def test_soft_sphere_closed_forms():
  for alpha in (2.0, 2.5, 3.0):
    assert abs(O.soft_sphere_energy(0.0, 1.0, 2.0, alpha) - 2.0 / alpha) < 1e-12
    assert O.soft_sphere_energy(1.0, 1.0, 2.0, alpha) == 0.0
    dEr, E = O.soft_sphere_dE_over_r(np.array([1.5 ** 2]), 1.0, 1.0, alpha)
    assert dEr[0] == 0.0 and E[0] == 0.0


def test_smoothing_c1():
  r_on, r_c = 2.0, 2.5
  assert O.smooth(r_on, r_on, r_c) == 1.0
  assert abs(O.smooth(r_c - 1e-12, r_on, r_c)) < 1e-10
  h = 1e-6
  for r in (r_on, r_c):
    lo = O.smooth(r - h, r_on, r_c)
    hi = O.smooth(r + h, r_on, r_c)
    assert abs(hi - lo) < 1e-9
    d_lo = (O.smooth(r - h, r_on, r_c) - O.smooth(r - 2 * h, r_on, r_c)) / h
    d_hi = (O.smooth(r + 2 * h, r_on, r_c) - O.smooth(r + h, r_on, r_c)) / h
    assert abs(d_hi - d_lo) < 1e-4


@pytest.mark.parametrize('kind', ['lj', 'soft_sphere'])
def test_force_is_minus_gradient(kind):
  rng = np.random.default_rng(7)
  if kind == 'lj':
    R, _, L = O.lj_lattice(6, jitter=0.1, seed=7)
    m = O.PairModel('lj')
  else:
    R, sp_, L = O.bidisperse_2d(40, seed=7)
    m = O.PairModel('soft_sphere', sigma=np.array([[1.0, 1.2], [1.2, 1.4]]), species=sp_)
  R = R.astype(np.float64)
  sp = O.Space(np.float64(L))
  E0, _, F = O.energy_force_brute(m, sp, R)
  h = 1e-5
  for _ in range(12):
    a = rng.integers(R.shape[0])
    c = rng.integers(R.shape[1])
    Rp, Rm = R.copy(), R.copy()
    Rp[a, c] += h
    Rm[a, c] -= h
    fd = -(O.energy_force_brute(m, sp, Rp)[0] - O.energy_force_brute(m, sp, Rm)[0]) / (2 * h)
    assert abs(fd - F[a, c]) < 1e-5 * max(1.0, abs(F[a, c]))
  assert np.allclose(F.sum(0), 0.0, atol=1e-9)


def test_list_energy_equals_brute_force():
  # energy_test.py:577-645: neighbour-list energy == pair energy
  R, _, L = O.lj_lattice(6, seed=8)
  sp = O.Space(np.float64(L))
  m = O.PairModel('lj')
  spec = O.NeighborListSpec(O.Space(L), L, 2.5, 0.5)
  nl = O.neighbor_list_build(spec, R)
  i, j = O.pairs_from_dense(nl.idx)
  E, e_atom, F = O.pair_energy_force(m, sp, R, i, j)
  Eb, _, Fb = O.energy_force_brute(m, sp, R)
  assert abs(E - Eb) < 1e-10 * abs(Eb)
  assert np.allclose(F, Fb, atol=1e-10)
  assert abs(e_atom.sum() - E) < 1e-9


# ---------------------------------------------------------------------------
# dynamics
# ---------------------------------------------------------------------------


def _lj_force_fn(L):
  sp = O.Space(np.float64(L))
  m = O.PairModel('lj')
  return sp, m, (lambda R: O.energy_force_brute(m, sp, R)[2])


def test_nve_second_order_and_reversibility():
  R, P, L = O.lj_lattice(5, seed=9)
  sp, m, force_fn = _lj_force_fn(L)
  R = R.astype(np.float64)
  P = P.astype(np.float64)

  def drift(dt, n):
    st = O.NVEState(R, P, force_fn(R))
    E0 = O.energy_force_brute(m, sp, R)[0] + O.kinetic_energy(P)
    for _ in range(n):
      st = O.nve_step(st, force_fn, sp, dt)
    return O.energy_force_brute(m, sp, st.position)[0] + O.kinetic_energy(st.momentum) - E0, st

  # velocity Verlet is second order: halving dt over the same time quarters
  # the energy error (a wrong kick or drift factor breaks this).
  e1, st = drift(0.005, 20)
  e2, _ = drift(0.0025, 40)
  assert 3.3 < e1 / e2 < 4.7, (e1, e2)
  back = O.NVEState(st.position, -st.momentum, st.force)
  for _ in range(20):
    back = O.nve_step(back, force_fn, sp, 0.005)
  dR = sp.displacement(back.position, R)
  assert np.max(np.abs(dR)) < 1e-8


def test_nve_dt_cast_to_float32():
  # simulate.py:235-236 (F5): dt is float32 even in the x64 run
  assert np.float64(np.float32(0.005)) != 0.005
  R = np.zeros((1, 3))
  P = np.ones((1, 3))
  st = O.NVEState(R, P, np.zeros((1, 3)))
  out = O.nve_step(st, lambda R: np.zeros_like(R), O.Space(None, False), 0.005)
  assert out.position[0, 0] == np.float64(np.float32(0.005))


def test_fire_static_cast_float16():
  # minimize.py:164-168, util.py:58-64 (F4)
  dt_start, dt_max, n_min, f_inc, f_dec, a0, f_a = O.FireParams(0.01, 0.1).cast()
  assert float(dt_start) == 0.01000213623046875
  assert float(dt_max) == 0.0999755859375
  assert float(f_inc) == 1.099609375 and float(f_a) == 0.990234375
  assert n_min.dtype == np.uint8 and dt_start.dtype == np.float16


def test_fire_descends_quadratic():
  # minimize_test.py:111-150
  rng = np.random.default_rng(10)
  R = rng.uniform(size=(64, 2))
  R0 = rng.uniform(size=(64, 2))
  force_fn = lambda R: -2.0 * (R - R0)
  energy = lambda R: float(np.sum((R - R0) ** 2))
  sp = O.Space(None, False)
  params = O.FireParams()
  st = O.fire_init(R, force_fn, params)
  E = energy(R)
  for _ in range(10):
    for _ in range(3):
      st = O.fire_step(st, force_fn, sp, params)
    E_new = energy(st.position)
    assert E_new < E
    E = E_new


# ---------------------------------------------------------------------------
# second derivatives (the adjoint's oracle), pinned by central differences
# ---------------------------------------------------------------------------


def _d2_system(kind, seed):
  rng = np.random.default_rng(seed)
  if kind == 'lj':
    R, _, L = O.lj_lattice(4, seed=seed)
    R = (R + rng.normal(0, 0.08, R.shape)).astype(np.float64) % L
    model = O.PairModel('lj', sigma=1.0, epsilon=1.0, fixed_smoothing=(2.0, 2.5))
    cutoff = 2.5
  else:
    L = 6.0
    R = rng.uniform(0, L, (40, 2))
    model = O.PairModel('soft_sphere', sigma=1.3, epsilon=1.0, alpha=2.0)
    cutoff = 1.3
  return R, L, model, cutoff


@pytest.mark.parametrize('kind', ['lj', 'soft_sphere'])
def test_force_vjp_matches_central_differences(kind):
  R, L, model, cutoff = _d2_system(kind, 3)
  sp = O.Space(np.float64(L))
  rng = np.random.default_rng(5)
  v = rng.normal(size=R.shape)
  i, j = O.all_pairs(sp, R, cutoff + 0.3)
  Jv, (ts, te), (es, ee) = O.force_vjp(model, sp, R, v, i, j)
  h = 1e-5

  def F(X, m=model):
    ii, jj = O.all_pairs(sp, X, cutoff + 0.3)
    return O.pair_energy_force(m, sp, X, ii, jj, 0.5)

  fd = (F(R + h * v)[2] - F(R - h * v)[2]) / (2 * h)
  assert np.linalg.norm(Jv - fd) <= 1e-6 * np.linalg.norm(fd)
  import dataclasses
  for name, t, e in (('sigma', ts, es), ('epsilon', te, ee)):
    base = float(getattr(model, name))
    mp = dataclasses.replace(model, **{name: base + h})
    mm = dataclasses.replace(model, **{name: base - h})
    Ep, _, Fp = F(R, mp)
    Em, _, Fm = F(R, mm)
    dF = (Fp - Fm) / (2 * h)
    assert abs(np.sum(v * dF) - t) <= 1e-6 * (abs(t) + np.linalg.norm(dF) * np.linalg.norm(v))
    assert abs((Ep - Em) / (2 * h) - e) <= 1e-6 * (abs(e) + 1.0)


@pytest.mark.parametrize('kind', ['lj', 'soft_sphere'])
def test_force_vjp_per_particle_matches_central_differences(kind):
  """Per-particle sigma and epsilon (diameters, p_ij = (p_i + p_j) / 2):
  v^T dF/dp_k and dE/dp_k for every k against central differences of the
  brute-force force and energy in p_k; swapping the pair's two halves for
  its first particle only (a dropped (delta_jk) term) fails."""
  import dataclasses
  R, L, model, cutoff = _d2_system(kind, 7)
  N = R.shape[0]
  rng = np.random.default_rng(9)
  sig0 = float(model.sigma) * (1.0 + 0.1 * rng.uniform(-1, 1, N))
  eps0 = float(model.epsilon) * (1.0 + 0.2 * rng.uniform(-1, 1, N))
  fixed = model.smoothing() if kind == 'lj' else None
  model = dataclasses.replace(model, sigma=sig0, epsilon=eps0, fixed_smoothing=fixed)
  sp = O.Space(np.float64(L))
  v = rng.normal(size=R.shape)
  reach = cutoff * 1.15 + 0.3
  i, j = O.all_pairs(sp, R, reach)
  cs, ce, es, ee = O.force_vjp_per_particle(model, sp, R, v, i, j)
  h = 1e-6

  def EF(m):
    ii, jj = O.all_pairs(sp, R, reach)
    E, _, F = O.pair_energy_force(m, sp, R, ii, jj, 0.5)
    return E, F

  for name, cot, eg in (('sigma', cs, es), ('epsilon', ce, ee)):
    base = np.asarray(getattr(model, name), np.float64)
    for k in rng.choice(N, 6, replace=False):
      dp = np.zeros(N)
      dp[k] = h
      Ep, Fp = EF(dataclasses.replace(model, **{name: base + dp}))
      Em, Fm = EF(dataclasses.replace(model, **{name: base - dp}))
      fd_v = np.sum(v * (Fp - Fm)) / (2 * h)
      fd_e = (Ep - Em) / (2 * h)
      assert abs(cot[k] - fd_v) <= 1e-5 * (abs(fd_v) + 1e-3 * np.abs(cot).max())
      assert abs(eg[k] - fd_e) <= 1e-5 * (abs(fd_e) + 1e-3 * np.abs(eg).max())
  # the scalar cotangent is the sum of the per-particle ones
  _, (ts, _), _ = O.force_vjp(dataclasses.replace(model, sigma=float(np.mean(sig0))), sp,
                              R, v, i, j)
  assert np.isfinite(ts)
  d = sp.displacement(R[j], R[i])
  _, _, gs, _, _, _ = O.pair_d2(model, np.sum(d * d, -1), model.pair_params(i, j, 'sigma'),
                                model.pair_params(i, j, 'epsilon'),
                                model.pair_params(i, j, 'alpha') if kind != 'lj' else 2.0)
  wrong = np.zeros(N)
  np.add.at(wrong, i, gs * np.sum(v[i] * d, -1))
  assert np.linalg.norm(wrong - cs) > 1e-2 * np.linalg.norm(cs)


def test_force_vjp_dropped_term_fails():
  """The pin is sharp: dropping the 2 g' (d.w) d term breaks the match."""
  R, L, model, cutoff = _d2_system('lj', 4)
  sp = O.Space(np.float64(L))
  v = np.random.default_rng(6).normal(size=R.shape)
  i, j = O.all_pairs(sp, R, cutoff + 0.3)
  Jv, _, _ = O.force_vjp(model, sp, R, v, i, j)
  d = sp.displacement(R[j], R[i])
  g = O.pair_d2(model, np.sum(d * d, -1), 1.0, 1.0)[0]
  partial = np.zeros_like(R)
  np.add.at(partial, i, g[:, None] * (v[j] - v[i]))
  assert np.linalg.norm(partial - Jv) > 1e-2 * np.linalg.norm(Jv)


def _fire_system(r, n=40, seed=3):
  R, species, L = O.bidisperse_2d(n, d_large=r, seed=seed)
  sig = np.array([[1.0, (1 + r) / 2], [(1 + r) / 2, r]], np.float64)
  return R.astype(np.float64), L, O.PairModel('soft_sphere', sigma=sig, species=species)


def test_fire_adjoint_matches_central_differences():
  """dE(x_T)/d r and dE(x_T)/dx_0 through FP64 FIRE, with sigma_ab(r) =
  [[1, (1+r)/2], [(1+r)/2, r]], against central differences of
  fire_energy_after (the dt / alpha decisions are fixed for |h| this small)."""
  r, T, h = 1.4, 24, 1e-6
  R, L, model = _fire_system(r)
  sp = O.Space(np.float64(L))
  par = O.FireParams(dt_start=0.01, dt_max=0.1)
  ds = np.array([[0.0, 0.5], [0.5, 1.0]])
  E, th, X = O.fire_energy_gradient(model, sp, R, T, par, ds)
  assert abs(E - O.fire_energy_after(model, sp, R, T, par)) <= 1e-12 * abs(E)

  def E_r(rr):
    return O.fire_energy_after(_fire_system(rr)[2], sp, R, T, par)

  fd = (E_r(r + h) - E_r(r - h)) / (2 * h)
  assert abs(th - fd) <= 1e-7 * abs(fd), (th, fd)
  v = np.random.default_rng(7).normal(size=R.shape)
  fdx = (O.fire_energy_after(model, sp, R + h * v, T, par)
         - O.fire_energy_after(model, sp, R - h * v, T, par)) / (2 * h)
  assert abs(np.sum(X * v) - fdx) <= 1e-7 * abs(fdx), (np.sum(X * v), fdx)


def test_fire_adjoint_across_momentum_resets():
  """With dt_start 0.1 the run passes F.P < 0 twice (momentum reset to 0);
  the gradient still matches central differences.  The bar of 1e-7 is sharp:
  dropping any one term of the recursion (the q term of Q, the F term or the
  |q|/|F| term of Fb, h Q, or the reset mask) moves dE/dr by 2e-3 to 1e-2
  relative on these systems."""
  r, T, h = 1.4, 40, 1e-6
  R, L, model = _fire_system(r)
  sp = O.Space(np.float64(L))
  par = O.FireParams(dt_start=0.1, dt_max=0.4)
  force = lambda X: O.energy_force_brute(model, sp, X)[2]
  st, tape = O.fire_init(R, force, par), []
  for _ in range(T):
    st = O.fire_step(st, force, sp, par, tape)
  assert sum(rec['keep'] == 0.0 for rec in tape) >= 2
  ds = np.array([[0.0, 0.5], [0.5, 1.0]])
  _, th, _ = O.fire_energy_gradient(model, sp, R, T, par, ds)
  fd = (O.fire_energy_after(_fire_system(r + h)[2], sp, R, T, par)
        - O.fire_energy_after(_fire_system(r - h)[2], sp, R, T, par)) / (2 * h)
  assert abs(th - fd) <= 1e-7 * abs(fd), (th, fd)


def test_pair_energies_per_entry():
  """pair_energies (the reduce_axis=() output, smap.py:940-953): each
  entry is the closed-form pair energy (switched LJ; soft sphere with
  species) times the normalization weight, and over a list holding both
  orders the entries sum to the brute-force total energy."""
  R, _, L = O.lj_lattice(5, seed=2)
  sp = O.Space(np.float64(L))
  R = R.astype(np.float64)
  m = O.PairModel('lj')
  i, j = O.all_pairs(sp, R, 2.5)
  e = O.pair_energies(m, sp, R, i, j, 0.5)
  r = np.sqrt(np.sum(sp.displacement(R[j], R[i]) ** 2, -1))
  # (the two forms round differently where the switch nearly cancels, r ~ r_c)
  assert np.allclose(e, 0.5 * O.lj_energy_smoothed(r, 1.0, 1.0, 2.0, 2.5), rtol=1e-9, atol=1e-15)
  assert abs(e.sum() - O.energy_force_brute(m, sp, R)[0]) <= 1e-12 * abs(e.sum())
  Rs, species, Ls = O.bidisperse_2d(60, seed=3)
  sig = np.array([[1.0, 1.2], [1.2, 1.4]])
  ms = O.PairModel('soft_sphere', sigma=sig, species=species)
  sps = O.Space(np.float64(Ls))
  Rs = Rs.astype(np.float64)
  i, j = O.all_pairs(sps, Rs, 1.4)
  es = O.pair_energies(ms, sps, Rs, i, j, 1.0)
  rs = np.sqrt(np.sum(sps.displacement(Rs[j], Rs[i]) ** 2, -1))
  want = O.soft_sphere_energy(rs, sig[species[i], species[j]], 1.0, 2.0)
  assert np.allclose(es, want, rtol=1e-12, atol=0) and np.any(es > 0)


def test_fire_box_adjoint_matches_central_differences():
  """dE(x_T)/dL for the side L of the periodic square, x_0 held in real
  coordinates (the meta-optimisation notebook's periodic_general(L(d) I,
  fractional_coordinates=False)), against central differences of
  fire_energy_after over L; then the derivative along the diameter ratio with
  the box following it, L(r) = sqrt(N pi (1 + r^2) / (8 phi)), against
  differences of E(r) with both sigma and L moving.  The run folds particles
  across the boundary 13 times; leaving out the fold term -X.m_k moves dE/dL
  by 24%, leaving out the pair image term (n = 0) changes dE/dL at fixed x."""
  r, T, h, n, phi = 1.4, 40, 1e-6, 40, 0.84
  R, L, model = _fire_system(r, seed=5)
  L = np.float64(L)
  par = O.FireParams(dt_start=0.1, dt_max=0.4)
  ds = np.array([[0.0, 0.5], [0.5, 1.0]])
  force = lambda X: O.energy_force_brute(model, O.Space(L), X)[2]
  st, tape = O.fire_init(R, force, par), []
  for _ in range(T):
    st = O.fire_step(st, force, O.Space(L), par, tape)
  assert sum(int(np.abs(np.rint((t['x_prev'] - t['x']) / L)).sum()) for t in tape) >= 10
  E, th, X, lb = O.fire_energy_gradient(model, O.Space(L), R, T, par, ds, box=True)
  E3, th3, X3 = O.fire_energy_gradient(model, O.Space(L), R, T, par, ds)
  assert E == E3 and th == th3 and np.array_equal(X, X3)
  fd = (O.fire_energy_after(model, O.Space(L + h), R, T, par)
        - O.fire_energy_after(model, O.Space(L - h), R, T, par)) / (2 * h)
  assert abs(lb - fd) <= 1e-7 * abs(fd), (lb, fd)
  # static part: dE/dL at fixed positions
  i, j = O.all_pairs(O.Space(L), R, r)
  dEdL = O.box_vjp(model, O.Space(L), R, np.zeros_like(R), i, j)[1]
  fds = (O.energy_force_brute(model, O.Space(L + h), R)[0]
         - O.energy_force_brute(model, O.Space(L - h), R)[0]) / (2 * h)
  assert abs(dEdL - fds) <= 1e-7 * abs(fds) and abs(dEdL) > 0.1
  # the box following the ratio
  Lr = lambda x: np.sqrt(n * np.pi * (1 + x * x) / (8 * phi))
  _, th2, _, lb2 = O.fire_energy_gradient(model, O.Space(Lr(r)), R, T, par, ds, box=True)
  total = th2 + lb2 * Lr(r) * r / (1 + r * r)
  Er = lambda x: O.fire_energy_after(_fire_system(x, seed=5)[2], O.Space(Lr(x)), R, T, par)
  fdt = (Er(r + h) - Er(r - h)) / (2 * h)
  assert abs(total - fdt) <= 1e-7 * abs(fdt), (total, fdt)
  assert abs(total - th2) > 0.1 * abs(fdt)  # a different derivative from the fixed box


def _jammed_nvt_setup(kT=1e-3, seed=0):
  st = fixtures.jammed_state()
  L = np.diag(st['box']).astype(np.float64)
  sp = O.Space(L)
  m = O.PairModel('soft_sphere', sigma=st['sigma'], species=st['species'])
  R = st['real_position'].astype(np.float64)
  rng = np.random.default_rng(seed)
  P = rng.normal(size=R.shape) * np.sqrt(2 * kT)
  P -= P.mean(0)
  return sp, m, R, P


@pytest.mark.parametrize('sy_steps', [1, 3, 5, 7])
def test_nvt_invariant_conserved(sy_steps):
  """The extended-system energy of the Nose-Hoover chain is conserved
  (simulate_test.py:317-350 on the jammed fixture, float64 bar 1e-7 there):
  here within 2e-6 over 200 steps started at twice the bath temperature;
  a wrong factor or index in the chain substep moves it by 1.5e-4 or more."""
  kT, dt = 1e-3, 1e-3
  sp, m, R, P = _jammed_nvt_setup(kT)
  force = lambda X: O.energy_force_brute(m, sp, X)[2]
  energy = lambda X: O.energy_force_brute(m, sp, X)[0]
  s, tau = O.nvt_init(R, P, force, kT, dt)
  H0 = O.nvt_invariant(s, energy(s.position), kT)
  for _ in range(200):
    s = O.nvt_step(s, force, sp, dt, kT, tau, sy_steps=sy_steps)
  H = O.nvt_invariant(s, energy(s.position), kT)
  assert abs(H - H0) <= 2e-6 * abs(H0), (H, H0)


def test_nvt_second_order():
  """Halving dt divides the invariant's drift over a fixed time by ~4."""
  kT = 1e-3
  sp, m, R, P = _jammed_nvt_setup(kT, seed=1)
  force = lambda X: O.energy_force_brute(m, sp, X)[2]
  energy = lambda X: O.energy_force_brute(m, sp, X)[0]
  drifts = []
  for dt, n in ((4e-3, 50), (2e-3, 100)):
    s, tau = O.nvt_init(R, P, force, kT, dt, tau=0.4)
    H0 = O.nvt_invariant(s, energy(s.position), kT)
    worst = 0.0
    for _ in range(n):
      s = O.nvt_step(s, force, sp, dt, kT, tau)
      worst = max(worst, abs(O.nvt_invariant(s, energy(s.position), kT) - H0))
    drifts.append(worst)
  ratio = drifts[0] / drifts[1]
  assert 3.0 < ratio < 5.5, (drifts, ratio)


# ---------------------------------------------------------------------------
# periodic_general with fractional coordinates
# ---------------------------------------------------------------------------


def test_general_space_diagonal_equals_periodic():
  rng = np.random.default_rng(11)
  Ls = np.array([7.5, 9.25, 11.0])
  u = rng.uniform(0, 1, (300, 3))
  g = O.GeneralSpace(np.diag(Ls).astype(F32))
  p = O.Space(Ls)
  d_g = g.displacement(u[:, None], u[None])
  d_p = p.displacement((u * Ls)[:, None], (u * Ls)[None])
  assert np.max(np.abs(d_g - d_p)) < 1e-12


def test_general_space_fcc_energy_invariant():
  # one crystal, two cells: the energy per atom cannot depend on the cell
  m = O.PairModel('lj')
  R, L = O.fcc_cubic(4)
  E1, _, _ = O.energy_force_brute(m, O.Space(np.float64(L)), R.astype(np.float64))
  u, T, _ = O.fcc_triclinic(6)
  E2, _, _ = O.energy_force_brute(m, O.GeneralSpace(T), u.astype(np.float64))
  assert abs(E2 / len(u) - E1 / len(R)) < 1e-8 * abs(E1 / len(R))
  # a plausible mistake (the transposed box) is caught
  E3, _, _ = O.energy_force_brute(m, O.GeneralSpace(T.T.copy()), u.astype(np.float64))
  assert abs(E3 / len(u) - E1 / len(R)) > 1e-2


def test_general_space_forces_are_real_space_gradients():
  m = O.PairModel('lj')
  u, T, _ = O.fcc_triclinic(6, jitter=0.08, seed=3)
  u = u.astype(np.float64)
  g = O.GeneralSpace(T)
  _, _, F = O.energy_force_brute(m, g, u)
  Ti = np.linalg.inv(T.astype(np.float64))
  h = 1e-6
  for i in (0, 17, 101):
    for k in range(3):
      e = np.zeros(3)
      e[k] = h
      up, um = u.copy(), u.copy()
      up[i] = up[i] + Ti @ e      # move particle i by +h along x_k in real space
      um[i] = um[i] - Ti @ e
      fd = -(O.energy_force_brute(m, g, up)[0] - O.energy_force_brute(m, g, um)[0]) / (2 * h)
      assert abs(fd - F[i, k]) < 1e-6 * max(1.0, np.abs(F).max())


def test_general_shift_moves_by_dR():
  rng = np.random.default_rng(5)
  _, T, _ = O.fcc_triclinic(6)
  g = O.GeneralSpace(T)
  u = rng.uniform(0, 1, (200, 3))
  dR = rng.normal(0, 0.3, (200, 3))
  v = g.shift(u, dR)
  assert np.all((v >= 0) & (v < 1))
  assert np.max(np.abs(g.displacement(v, u) - dR)) < 1e-6


def test_fractional_cell_size_from_perpendicular_widths():
  assert O.fractional_cell_size(F32(10.0), 2.5) == F32(0.25)
  assert O.fractional_cell_size(np.array([8.0, 10.0, 12.0], F32), 2.0) == F32(0.25)
  rng = np.random.default_rng(2)
  for _ in range(20):
    T = np.triu(rng.uniform(-2, 2, (3, 3))) + np.diag(rng.uniform(6, 14, 3))
    T = T.astype(F32)
    c = F32(rng.uniform(1.0, 2.5))
    a = [T[:, 0].astype(np.float64), T[:, 1].astype(np.float64), T[:, 2].astype(np.float64)]
    V = abs(np.linalg.det(T.astype(np.float64)))
    widths = [V / np.linalg.norm(np.cross(a[1], a[2])), V / np.linalg.norm(np.cross(a[2], a[0])),
              V / np.linalg.norm(np.cross(a[0], a[1]))]
    n = max(1.0, np.floor(min(widths) / c))
    got = O.fractional_cell_size(T, c)
    assert abs(got - 1.0 / n) < 1e-6, (got, 1.0 / n)


def test_general_neighbor_list_matches_brute_force():
  # 14^3 primitive cells: perpendicular width 12.5 > 4 cutoffs, 4 cells a side
  u, T, _ = O.fcc_triclinic(14, jitter=0.1, seed=9)
  g = O.GeneralSpace(T)
  spec = O.NeighborListSpec(g, T, 2.5, 0.5, fractional_coordinates=True)
  nl = O.neighbor_list_build(spec, u)
  assert nl.use_cell_list
  assert nl.error & O.MALFORMED_BOX   # the reference's predicate for a valid box
  rows = O.sorted_rows(nl)
  N = len(u)
  for i in range(0, N, 53):
    d2 = g.metric_sq(u, u[i][None, :])
    ref = np.nonzero((d2 < spec.cutoff_sq) & (np.arange(N) != i))[0]
    assert np.array_equal(rows[i], ref)


# ---------------------------------------------------------------------------
# multi-image lists
# ---------------------------------------------------------------------------


def _small_cell(seed=0, n=40):
  rng = np.random.default_rng(seed)
  T = np.array([[3.1, 0.7, -0.4], [0.0, 2.8, 0.5], [0.0, 0.0, 3.3]], F32)
  return rng.uniform(0, 1, (n, 3)).astype(F32), T


def test_multi_image_stencil_is_complete():
  u, T = _small_cell(1)
  c = 2.9
  n_max = O.multi_image_shift_n_max(T, c)
  ks, ii, jj, _, _ = O.multi_image_candidates(u, T, F32(c * c), O.multi_image_shifts(n_max))
  wide = O.multi_image_shifts(n_max + 2)
  kw, iw, jw, _, _ = O.multi_image_candidates(u, T, F32(c * c), wide)
  sh = O.multi_image_shifts(n_max)
  a = {(int(i), int(j), tuple(sh[k])) for k, i, j in zip(ks, ii, jj)}
  b = {(int(i), int(j), tuple(wide[k])) for k, i, j in zip(kw, iw, jw)}
  assert a == b and len(a) > 0


def test_multi_image_large_cell_equals_neighbor_sets():
  u, T, _ = O.fcc_triclinic(6, jitter=0.05, seed=2)   # widths 5.4 > 2 x 2.5
  c = 2.5
  n_max = O.multi_image_shift_n_max(T, c)
  idx, sh, occ = O.multi_image_list(u, T, F32(c * c), n_max, O.SPARSE, 10 ** 6)
  N = len(u)
  got = {(int(i), int(j)) for i, j in zip(idx[0][:occ], idx[1][:occ])}
  assert len(got) == occ     # one image per pair
  g = O.GeneralSpace(T)
  ref = set()
  for i in range(N):
    d2 = g.metric_sq(u.astype(np.float64), u[i].astype(np.float64)[None, :])
    for j in np.nonzero((d2 < c * c) & (np.arange(N) != i))[0]:
      ref.add((i, int(j)))
  assert got == ref


def test_multi_image_ordered_holds_each_pair_once():
  u, T = _small_cell(3)
  c = 2.9
  n_max = O.multi_image_shift_n_max(T, c)
  a, sa, na = O.multi_image_list(u, T, F32(c * c), n_max, O.SPARSE, 10 ** 6)
  b, sb, nb = O.multi_image_list(u, T, F32(c * c), n_max, O.ORDERED_SPARSE, 10 ** 6)
  assert na == 2 * nb
  full = {(int(i), int(j), tuple(s)) for i, j, s in zip(a[0][:na], a[1][:na], sa[:na])}
  for i, j, s in zip(b[0][:nb], b[1][:nb], sb[:nb]):
    assert (int(i), int(j), tuple(s)) in full
    assert (int(j), int(i), tuple(-s)) in full


# ---------------------------------------------------------------------------
# FIRE bookkeeping, step by step, against a hand-derived table
# ---------------------------------------------------------------------------

# One particle under a scheduled force, so that the sign of F.P at every step
# is known without integrating: F = +1 along x except F = -1 at step 31 and
# F = +1 again from step 32.  With P >= 0 before step 31 the NVE step gives
# F.P > 0 at steps 1-30; at step 31 P = P_mix + dt/2 - dt/2 > 0 meets F = -1,
# so F.P < 0 (reset); at step 32 P = 0 - dt/2 + dt/2 = 0 exactly, so F.P = 0
# (n_pos counts it, dt and alpha hold: minimize.py:209-220 tests > 0 and < 0
# for dt / alpha and >= 0 for n_pos and the momentum mask); from step 33
# F.P > 0 again, except at step 41 where F = 0 (F.P = 0 with n_pos > n_min
# and alpha below alpha_start: nothing but n_pos moves).  With the defaults (dt_start 0.1, dt_max 0.4, n_min 5,
# f_inc 1.1, f_dec 0.5, alpha_start 0.1, f_alpha 0.99, all float16 after
# static_cast, Bitzek et al. 2006): dt and alpha hold while n_pos <= 5, then
# each step dt = min(f16(dt * 1.1), f16(0.4)) and alpha = f16(alpha * 0.99)
# (products rounded to nearest even in float16: 0x2e66 * 0x3c66 = 0.109924...
# -> 0x2f09); dt reaches the cap f16(0.4) = 0x3666 at step 20; the reset
# halves it exactly (0x3266) and restores alpha = f16(0.1) = 0x2e66.
# Rows: (step, n_pos, dt bits, alpha bits), float16 bit patterns.
_FIRE_TABLE = [
    (1, 1, 0x2e66, 0x2e66), (2, 2, 0x2e66, 0x2e66), (3, 3, 0x2e66, 0x2e66),
    (4, 4, 0x2e66, 0x2e66), (5, 5, 0x2e66, 0x2e66), (6, 6, 0x2f09, 0x2e56),
    (7, 7, 0x2fbc, 0x2e46), (8, 8, 0x3041, 0x2e36), (9, 9, 0x30ad, 0x2e26),
    (10, 10, 0x3124, 0x2e17), (11, 11, 0x31a7, 0x2e08), (12, 12, 0x3237, 0x2df9),
    (13, 13, 0x32d5, 0x2dea), (14, 14, 0x3383, 0x2ddb), (15, 15, 0x3421, 0x2dcc),
    (16, 16, 0x348a, 0x2dbe), (17, 17, 0x34fe, 0x2db0), (18, 18, 0x357d, 0x2da2),
    (19, 19, 0x3609, 0x2d94), (20, 20, 0x3666, 0x2d86), (21, 21, 0x3666, 0x2d78),
    (22, 22, 0x3666, 0x2d6a), (23, 23, 0x3666, 0x2d5c), (24, 24, 0x3666, 0x2d4f),
    (25, 25, 0x3666, 0x2d42), (26, 26, 0x3666, 0x2d35), (27, 27, 0x3666, 0x2d28),
    (28, 28, 0x3666, 0x2d1b), (29, 29, 0x3666, 0x2d0e), (30, 30, 0x3666, 0x2d01),
    (31, 0, 0x3266, 0x2e66), (32, 1, 0x3266, 0x2e66), (33, 2, 0x3266, 0x2e66),
    (34, 3, 0x3266, 0x2e66), (35, 4, 0x3266, 0x2e66), (36, 5, 0x3266, 0x2e66),
    (37, 6, 0x3309, 0x2e56), (38, 7, 0x33bc, 0x2e46), (39, 8, 0x3441, 0x2e36),
    (40, 9, 0x34ad, 0x2e26), (41, 10, 0x34ad, 0x2e26), (42, 11, 0x3524, 0x2e17),
]


def fire_scheduled_force(step):
  """Force of the table's scenario at a given force evaluation (0 = init)."""
  return {31: -1.0, 41: 0.0}.get(step, 1.0)


def _f16_bits(x):
  return int(np.asarray(x, np.float16).view(np.uint16))


def test_fire_bookkeeping_hand_table():
  calls = [0]

  def force_fn(R):
    f = np.zeros_like(R)
    f[0, 0] = fire_scheduled_force(calls[0])
    calls[0] += 1
    return f

  sp = O.Space(None, False)
  params = O.FireParams()
  st = O.fire_init(np.zeros((1, 3)), force_fn, params)
  for step, n_pos, dt_bits, alpha_bits in _FIRE_TABLE:
    st = O.fire_step(st, force_fn, sp, params)
    got = (int(st.n_pos), _f16_bits(st.dt), _f16_bits(st.alpha))
    assert got == (n_pos, dt_bits, alpha_bits), (step, got)
    if step == 31:
      assert np.all(st.momentum == 0.0), 'momentum not reset when F.P < 0'
    if step == 32:
      # F.P == 0: the mask F.P >= 0 keeps P (which is exactly zero here)
      assert np.all(st.momentum == 0.0)
    if step in (30, 33):
      assert st.momentum[0, 0] > 0.0
  # structural facts of the table, independent of its bit patterns
  assert [r[1] for r in _FIRE_TABLE[:5]] == [1, 2, 3, 4, 5]
  assert _FIRE_TABLE[4][2] == _FIRE_TABLE[0][2] != _FIRE_TABLE[5][2]   # growth from n_pos 6
  assert _FIRE_TABLE[19][2] == _f16_bits(np.float16(0.4))               # the dt_max cap
  assert _FIRE_TABLE[30][2] == _f16_bits(np.float16(np.float16(0.4) * np.float16(0.5)))


# ---------------------------------------------------------------------------
# The skin rule any(metric_sq(R, R_ref) > (dr/2)^2) at its threshold
# ---------------------------------------------------------------------------


def test_skin_threshold_exact_cases():
  # partition.py:901,1141-1154: strict '>' against float32 (dr/2)^2
  F = np.float32
  spec = O.NeighborListSpec(O.Space(None, False), F(100.0), 2.5, 0.5)
  assert spec.threshold_sq == F(0.0625)
  ref = np.array([[1.0, 2.0, 3.0], [10.0, 10.0, 10.0]], F)
  at = ref.copy()
  at[0, 0] = F(1.25)                       # d = 0.25, d^2 = 0.0625 exactly: no rebuild
  assert not O.needs_rebuild(spec, at, ref)
  up = ref.copy()
  up[0, 0] = np.nextafter(F(1.25), F(2))   # d = 0.25 + 2^-23, d^2 = 0.0625 + 2^-24
  assert O.needs_rebuild(spec, up, ref)
  down = ref.copy()
  down[0, 0] = np.nextafter(F(1.25), F(0))
  assert not O.needs_rebuild(spec, down, ref)
  neg = ref.copy()
  neg[1, 2] = F(9.75)                      # d = -0.25 on another particle: no rebuild
  assert not O.needs_rebuild(spec, neg, ref)
  neg[1, 2] = np.nextafter(F(9.75), F(0))
  assert O.needs_rebuild(spec, neg, ref)
  # periodic, across the face (L = 16): the float32 difference 0.125 - 15.875
  # = -15.75 wraps to 0.25 exactly
  pspec = O.NeighborListSpec(O.Space(F(16.0)), F(16.0), 2.5, 0.5)
  pref = np.array([[15.875, 4.0, 4.0]], F)
  pos = np.array([[0.125, 4.0, 4.0]], F)
  assert not O.needs_rebuild(pspec, pos, pref)
  # one ulp of 0.125 is lost in the float32 difference (ulp(15.75) = 2^-20):
  # the reference's float32 wrap does not see it
  pos[0, 0] = F(0.125) + F(2.0 ** -26)
  assert not O.needs_rebuild(pspec, pos, pref)
  pos[0, 0] = F(0.125) + F(2.0 ** -20)     # d = 0.25 + 2^-20 after the wrap
  assert O.needs_rebuild(pspec, pos, pref)
  # dr = 0: every update rebuilds (partition.py:892,1156-1159)
  zspec = O.NeighborListSpec(O.Space(None, False), F(100.0), 2.5, 0.0)
  assert O.needs_rebuild(zspec, ref, ref)


def _pairs_within(space, R, rc):
  d2 = np.stack([space.metric_sq(R, np.broadcast_to(R[j], R.shape)) for j in range(len(R))], 1)
  i, j = np.nonzero(d2 < np.float32(rc) ** 2)
  keep = i != j
  return set(zip(i[keep].tolist(), j[keep].tolist()))


def test_skin_rule_keeps_every_pair_within_cutoff():
  # The guarantee the skin rule exists for: after any sequence of gated
  # updates, every pair within r_cutoff at the current positions is in the
  # list (no particle has moved dr/2 from where the list was built).
  rng = np.random.default_rng(31)
  L = np.float32(9.0)
  space = O.Space(L)
  spec = O.NeighborListSpec(space, L, 2.0, 0.6)
  R = rng.uniform(0, L, (160, 3)).astype(np.float32)
  nl = O.neighbor_list_build(spec, R)
  rebuilds = 0
  for step in range(40):
    R = np.mod(R + rng.normal(0, 0.04, R.shape), L).astype(np.float32)
    new = O.neighbor_list_update(spec, R, nl)
    rebuilds += new is not nl
    nl = new
    have = {(i, int(j)) for i, row in enumerate(nl.idx) for j in row if j < len(R)}
    missing = _pairs_within(space, R, 2.0) - have
    assert not missing, (step, sorted(missing)[:5])
  assert 0 < rebuilds < 40


def test_skin_rule_adversarial_pair():
  # Two particles just beyond the list cutoff c = r_c + dr move towards each
  # other by just under dr/2 each: no rebuild, and they end up just beyond
  # r_c.  A threshold of dr instead of dr/2 would let them move dr each
  # (also no rebuild) and end inside r_c, missing from the list.
  F = np.float32
  L = F(20.0)
  space = O.Space(L)
  spec = O.NeighborListSpec(space, L, 2.0, 0.5)
  R = np.array([[5.0, 5.0, 5.0], [F(5.0) + F(2.5) + F(0.001), 5.0, 5.0]], F)
  nl = O.neighbor_list_build(spec, R)
  assert np.all(nl.idx >= 2)                # not neighbours: beyond c = 2.5
  R2 = R.copy()
  R2[0, 0] += F(0.249)
  R2[1, 0] -= F(0.249)
  assert not O.needs_rebuild(spec, R2, nl.reference_position)
  assert space.metric_sq(R2[:1], R2[1:])[0] > F(4.0)   # still beyond r_c = 2
  R3 = R.copy()
  R3[0, 0] += F(0.26)                       # over dr/2: rebuilds
  assert O.needs_rebuild(spec, R3, nl.reference_position)
