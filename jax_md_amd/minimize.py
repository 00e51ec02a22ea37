"""FIRE descent (jax_md/minimize.py:124-226) on gfx950.

The NVE sub-step runs the native velocity-Verlet kernels with the step size
read from the device; the bookkeeping (float64 reductions of F^2, P^2, F.P,
the float16 dt/alpha update of util.static_cast, the mixing and the reset)
runs in jmd_fire_update without a host round trip.
"""

from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as onp
import torch

from . import _native as nat
from . import quantity, simulate, util


@dataclasses.dataclass
class FireDescentState:
  """FIRE state (minimize.py:97-121).  dt and alpha are held on the device as
  float32 values of float16 numbers (``fire_scalars``); n_pos likewise."""
  position: torch.Tensor
  momentum: torch.Tensor
  force: torch.Tensor
  mass: object
  fire_scalars: torch.Tensor          # float32 [2]: dt, alpha
  fire_n_pos: torch.Tensor            # int32 [1]
  position_hi: Optional[torch.Tensor] = None

  @property
  def dt(self):
    return self.fire_scalars[0].to(torch.float16)

  @property
  def alpha(self):
    return self.fire_scalars[1].to(torch.float16)

  @property
  def n_pos(self):
    return self.fire_n_pos[0]


def fire_descent(energy_or_force, shift_fn, dt_start=0.1, dt_max=0.4, n_min=5, f_inc=1.1,
                 f_dec=0.5, alpha_start=0.1, f_alpha=0.99):
  """Fast Inertial Relaxation Engine (minimize.py:124-226)."""
  cast = util.static_cast(dt_start, dt_max, n_min, f_inc, f_dec, alpha_start, f_alpha)
  params = (nat.ctypes.c_float * 7)(*[float(onp.float32(c)) for c in cast])
  dt_start_c, alpha_start_c = float(onp.float32(cast[0])), float(onp.float32(cast[5]))
  energy = quantity.native_energy_of(energy_or_force)
  shift = simulate.native_shift(shift_fn)
  ws_cache = {}

  def init_fn(R, mass=1.0, **kwargs) -> FireDescentState:
    nat.position_stride(R, 'R')
    R = simulate.new_positions(R.shape[0], R.shape[1], R.device, R)
    neighbor = kwargs.pop('neighbor', None)
    F = energy.force(R, neighbor, **kwargs)
    P = torch.zeros_like(R)
    scal = torch.tensor([dt_start_c, alpha_start_c], dtype=torch.float32, device=R.device)
    npos = torch.zeros((1,), dtype=torch.int32, device=R.device)
    return FireDescentState(R, P, F, mass, scal, npos, R.to(torch.float64))

  def apply_fn(state: FireDescentState, **kwargs) -> FireDescentState:
    lib = nat.load()
    neighbor = kwargs.pop('neighbor', None)
    nve_state = simulate.NVEState(state.position, state.momentum, state.force, state.mass,
                                  state.position_hi)
    s = simulate.velocity_verlet_native(energy, shift, nve_state, 0.0, neighbor,
                                        dt_device=state.fire_scalars[0:1], **kwargs)
    R = s.position
    N, dim = R.shape
    dev = R.device
    scal = state.fire_scalars.clone()
    npos = state.fire_n_pos.clone()
    key = str(dev)
    if key not in ws_cache:
      ws_cache[key] = (torch.empty((lib.jmd_fire_workspace_bytes(N),), dtype=torch.uint8,
                                   device=dev),
                       torch.empty((8,), dtype=torch.float64, device=dev))
    ws, scalars = ws_cache[key]
    nat.check(lib.jmd_fire_update(nat.ptr(s.momentum), nat.ptr(s.force), N, dim, nat.ptr(scal),
                                  nat.ptr(npos), params, nat.ptr(scalars), nat.ptr(ws),
                                  nat.stream_handle(dev)), 'jmd_fire_update')
    return FireDescentState(R, s.momentum, s.force, state.mass, scal, npos, s.position_hi)

  return init_fn, apply_fn
