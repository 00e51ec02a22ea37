"""The batched workloads at their full sizes on the GPU.

C3: 512 bidisperse 2-D soft-sphere systems (N log-uniform in [64, 4096],
553,041 particles), 500 FIRE steps in one launch.  Bars: no overflow; on 8
systems spread over the sizes, the plain per-system path (``nbrs.update`` +
``fire_descent``) reaches the same float16 dt / alpha / n_pos and positions
within 1e-5 after 50 steps, and after 500 steps positions within 1e-3 and
energies within 1e-4 (the two paths sum F.F, P.P and F.P in different
orders; the last-bit differences of the mixing grow over a long run); final energies of 4 of them within 1e-5 of the
FP64 oracle at the final positions (float32 pair energies of a minimised
packing, whose total is small against its terms).

C5: 64 replicas of N = 8192 (diameter ratios 1.2 .. 1.6, the box following
the ratio), 300 FIRE steps and the gradient of the mean final energy.
Bars: no overflow; every system's result is independent of the batch (two
replicas run alone give their batched forward state and gradient bit for
bit); the checkpoint spacing leaves E(x_T) and dE/dx_0 bit for bit and moves
dE/dr only by the grouping of its per-segment sums (1e-12); the mean-energy
scale is 1/64 of the sum's.  Parity of the gradient
itself rests on the FP64 adjoint at small sizes (tests/test_gpu_ensemble.py).
"""

import numpy as np
import pytest
import torch

from oracle import jaxmd_oracle as O

pytestmark = pytest.mark.gpu

DEV = 'cuda:0'
F32 = np.float32
FIRE = dict(dt_start=0.01, dt_max=0.1)


def _jmd():
  import jax_md_amd
  return jax_md_amd


def test_c3_full_size():
  jmd = _jmd()
  rng = np.random.default_rng(0)
  n_sys = 512
  sizes = np.floor(np.exp(rng.uniform(np.log(64), np.log(4096), n_sys))).astype(int)
  systems = [O.bidisperse_2d(int(n), seed=1000 + k) for k, n in enumerate(sizes)]
  sigma = np.array([[1.0, 1.2], [1.2, 1.4]], F32)
  print(f'case C3 full size: {n_sys} systems, {int(sizes.sum())} particles', flush=True)
  ens = jmd.ensemble.SoftSphereFireEnsemble([s[0] for s in systems], [s[2] for s in systems],
                                            [s[1] for s in systems], sigma=sigma, **FIRE)
  ens.allocate().run(50)
  at50 = (ens.position_hi.clone(), ens.fire_state.clone(), ens.fire_n_pos.clone())
  ens.run(450)
  assert int((ens.did_buffer_overflow != 0).sum()) == 0
  pick = np.argsort(sizes)[np.linspace(0, n_sys - 1, 8).astype(int)]
  E = ens.energy.cpu().numpy()
  for q, k in enumerate(pick):
    R, sp, L = systems[k]
    disp, shift = jmd.space.periodic(L)
    nfn, efn = jmd.energy.soft_sphere_neighbor_list(
        disp, L, species=torch.as_tensor(sp, device=DEV),
        sigma=torch.as_tensor(sigma, device=DEV), format=jmd.partition.Dense)
    init_fn, apply_fn = jmd.minimize.fire_descent(efn, shift, **FIRE)
    Rt = torch.as_tensor(R, device=DEV)
    nb = nfn.allocate(Rt)
    st = init_fn(Rt, neighbor=nb)
    sl = ens.system_slice(int(k))
    for step in range(500):
      nb = nb.update(st.position)
      st = apply_fn(st, neighbor=nb)
      if step == 49:
        dx = (at50[0][sl] - st.position_hi).abs().max().item()
        print(f'  system {k} (N={sizes[k]}), 50 steps: max |dx| {dx:.2e}', flush=True)
        assert dx <= 1e-5, (k, dx)
        assert float(at50[1][k, 0]) == float(st.fire_scalars[0]), k
        assert float(at50[1][k, 1]) == float(st.fire_scalars[1]), k
        assert int(at50[2][k]) == int(st.fire_n_pos[0]), k
    dx = (ens.position_hi[sl] - st.position_hi).abs().max().item()
    E_plain = float(efn(st.position, neighbor=nb))
    print(f'  system {k}, 500 steps: max |dx| {dx:.2e}, E {E[k]:.10f} plain {E_plain:.10f}',
          flush=True)
    assert dx <= 1e-3, (k, dx)
    assert abs(E[k] - E_plain) <= 1e-4 * abs(E_plain), (k, E[k], E_plain)
    if q % 2 == 0:
      model = O.PairModel('soft_sphere', sigma=sigma.astype(np.float64), species=sp)
      E64 = O.energy_force_brute(model, O.Space(np.float64(L)),
                                 ens.position[sl].cpu().numpy().astype(np.float64))[0]
      assert abs(E[k] - E64) <= 1e-5 * abs(E64) + 1e-9, (k, E[k], E64)


def _sigma_r(r):
  return np.array([[1.0, (1 + r) / 2], [(1 + r) / 2, r]], F32)


def test_c5_full_size():
  jmd = _jmd()
  n_rep, n, T, phi = 64, 8192, 300, 0.84
  r = [float(F32(1.2 + 0.4 * k / (n_rep - 1))) for k in range(n_rep)]
  systems = [O.bidisperse_2d(n, d_large=r[k], seed=2000 + k) for k in range(n_rep)]
  dsig = np.array([[0.0, 0.5], [0.5, 1.0]], F32)
  dLdr = np.array([np.sqrt(n * np.pi * (1 + x * x) / (8 * phi)) * x / (1 + x * x) for x in r])
  print('case C5 full size', flush=True)

  def make(ids):
    return jmd.ensemble.SoftSphereFireEnsemble(
        [systems[k][0] for k in ids], [systems[k][2] for k in ids],
        [systems[k][1] for k in ids], sigma=np.stack([_sigma_r(r[k]) for k in ids]),
        **FIRE).allocate()

  allk = list(range(n_rep))
  ens = make(allk)
  E, th, X = ens.energy_gradient(T, dsig, checkpoint_every=25, scale=1.0, dbox=dLdr)
  assert int((ens.did_buffer_overflow != 0).sum()) == 0
  lb = ens.box_adjoint.clone()
  # independent of the checkpoint spacing
  E2, th2, X2 = make(allk).energy_gradient(T, dsig, checkpoint_every=60, scale=1.0, dbox=dLdr)
  assert torch.equal(E, E2) and torch.equal(X, X2)
  assert torch.allclose(th, th2, rtol=1e-12, atol=0)
  # the mean-energy scale
  _, thm, _ = make(allk).energy_gradient(T, dsig, checkpoint_every=25, dbox=dLdr)
  assert torch.allclose(thm * n_rep, th, rtol=1e-12, atol=0)
  # each system's result is independent of the batch
  for k in (0, 41):
    one = make([k])
    E1, th1, X1 = one.energy_gradient(T, dsig, checkpoint_every=25, scale=1.0,
                                      dbox=dLdr[k:k + 1])
    sl = ens.system_slice(k)
    print(f'  replica {k}: dE/dr {float(th[k]):.9f} alone {float(th1[0]):.9f}', flush=True)
    assert float(E1[0]) == float(E[k])
    assert float(th1[0]) == float(th[k])
    assert float(one.box_adjoint[0]) == float(lb[k])
    assert torch.equal(X1, X[sl])
    assert torch.equal(one.position, ens.position[sl])
