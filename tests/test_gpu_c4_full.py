"""The adjoint workload at full size on the GPU: LJ N = 32,768 (32^3
lattice, rho 0.8442, jitter 0.05, kT 1), 200 NVE steps at dt 0.005,
checkpoints every 20, d E(x_T) / d(x_0, sigma, epsilon).

Bars, against central differences (h = 1e-7) of the C FP64 restatement
(oracle/md_oracle.c, the same velocity Verlet with the float32 dt): the x64
form within 1e-4 relative on sigma, epsilon and a random x_0 direction; the
float32 form within 1e-2 on sigma and 1e-3 on epsilon and the x_0 direction:
a chaotic liquid amplifies the float32 trajectory's rounding with the
horizon (on this system the float32 adjoint sits 1.3e-5 from the FP64
differences on sigma at 100 steps and 4e-3 at 200), while the x64 form
stays within 2e-5.
"""

import numpy as np
import pytest
import torch

from oracle import jaxmd_oracle as O

pytestmark = pytest.mark.gpu

DEV = 'cuda:0'


def test_c4_full_size_gradient_vs_fp64_differences():
  import jax_md_amd as jmd
  from oracle import c_oracle
  print('case C4 full size', flush=True)
  R, P, L = O.lj_lattice(32, seed=1)
  disp, shift = jmd.space.periodic(L)
  nfn, efn = jmd.energy.lennard_jones_neighbor_list(disp, L, format=jmd.partition.Dense)
  T, K, dt = 200, 20, 0.005
  dt32 = float(np.float32(dt))
  h = 1e-7
  rng = np.random.default_rng(5)
  u = rng.normal(size=R.shape)
  u /= np.linalg.norm(u)
  R64 = R.astype(np.float64)

  def e_after(R0, sigma=1.0, eps=1.0):
    sysc = c_oracle.LJSystem(R0, float(L), sigma=sigma, eps=eps)
    sysc.nve(P.astype(np.float64), dt32, T)
    return sysc.forces()[0]

  fd_x = (e_after(R64 + h * u) - e_after(R64 - h * u)) / (2 * h)
  fd_s = (e_after(R64, sigma=1 + h) - e_after(R64, sigma=1 - h)) / (2 * h)
  fd_e = (e_after(R64, eps=1 + h) - e_after(R64, eps=1 - h)) / (2 * h)
  for dtype, bars in ((torch.float64, (1e-4, 1e-4, 1e-4)), (torch.float32, (1e-2, 1e-3, 1e-3))):
    g = jmd.adjoint.nve_energy_gradient(
        efn, shift, nfn, torch.as_tensor(R, device=DEV).to(dtype),
        torch.as_tensor(P, device=DEV).to(dtype), dt, T, checkpoint_every=K)
    ad_x = float(np.sum(g.position.cpu().numpy() * u))
    errs = (abs(g.sigma - fd_s) / abs(fd_s), abs(g.epsilon - fd_e) / abs(fd_e),
            abs(ad_x - fd_x) / abs(fd_x))
    print(f'  {dtype}: sigma {g.sigma:.6e} (fd {fd_s:.6e}), epsilon {g.epsilon:.6e} '
          f'(fd {fd_e:.6e}), x0 {ad_x:.6e} (fd {fd_x:.6e}); rel errs {errs}', flush=True)
    assert errs[0] <= bars[0] and errs[1] <= bars[1] and errs[2] <= bars[2], (dtype, errs)
