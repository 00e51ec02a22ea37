"""Autotuner of the native NVE block loop: threads per Dense row, cell size
and row capacity, chosen by timing the loop on the device.

The reference fixes all three: cells are the list cutoff wide
(partition.py:146-188, ``cell_size = cutoff``), the cell capacity is 1.25x
the largest cell filling (partition.py:243-250) and the row capacity 1.25x
the largest row (partition.py:1092-1104, ``capacity_multiplier``); XLA picks
its own fusion layout for the pair reduction.  Here:

- ``threads_per_atom`` (lanes per Dense row in the force pass) is a knob of
  the pair kernel (``jmd_set_threads_per_atom``);
- ``cell_size_factor`` makes cells at least that multiple of the list cutoff
  (``partition.neighbor_list(..., cell_size_factor=)``): the neighbour sets
  are the same, the rows follow that grid's candidate order;
- ``capacity_multiplier`` is the reference's own argument.

``tune_nve_loop`` times ``trial_steps`` steps of ``simulate.NVELoop`` (HIP
graph replays, HIP events on the launching stream) for each candidate, on
copies of the given state, by coordinate descent: threads per row first, then
the cell size, then the capacity.  A candidate whose list overflows during
its trial is discarded.  Results are cached per workload (particle count,
dimension, density, cutoffs, list format, potential) for the process, and in
a JSON file when ``cache_path`` is given.
"""

from __future__ import annotations

import dataclasses
import json
import os
from typing import Callable, Dict, Optional, Sequence, Tuple

import numpy as onp
import torch

from . import _native as nat
from . import partition
from . import simulate

_current_tpa = 0  # threads per row the library was last told (0: its built-in choice)


def set_threads_per_atom(tpa: int) -> None:
  """Lanes per Dense row of the force pass for every later launch in this
  process (0 restores the built-in choice: 8 in 3-D, 2 in 2-D)."""
  global _current_tpa
  nat.check(nat.load().jmd_set_threads_per_atom(int(tpa)), 'jmd_set_threads_per_atom')
  _current_tpa = int(tpa)


@dataclasses.dataclass
class TunedConfig:
  threads_per_atom: int
  cell_size_factor: float
  capacity_multiplier: float
  us_per_step: float
  trials: list          # [(tpa, cell_size_factor, capacity_multiplier, us/step or None)]
  key: tuple

  @property
  def neighbor_kwargs(self) -> Dict[str, float]:
    """Keyword arguments for ``neighbor_list`` (or an energy wrapper's
    ``**neighbor_kwargs``)."""
    return {'cell_size_factor': self.cell_size_factor,
            'capacity_multiplier': self.capacity_multiplier}

  def apply(self) -> None:
    """Sets the chosen threads per row for the process."""
    set_threads_per_atom(self.threads_per_atom)

  def as_dict(self) -> dict:
    return {'threads_per_atom': self.threads_per_atom,
            'cell_size_factor': self.cell_size_factor,
            'capacity_multiplier': self.capacity_multiplier,
            'us_per_step': self.us_per_step,
            'trials': [list(t) for t in self.trials]}


_CACHE: Dict[tuple, TunedConfig] = {}


def workload_key(position: torch.Tensor, box, r_cutoff: float, dr_threshold: float,
                 fmt: partition.NeighborListFormat, kind: str) -> tuple:
  """What the tuned choice depends on: particle count and dimension, number
  density (3 significant digits), cutoff and skin, list format, potential."""
  N, dim = position.shape
  b = onp.asarray(box.detach().cpu().numpy() if isinstance(box, torch.Tensor) else box,
                  onp.float64)
  if b.ndim == 0:
    vol = float(b) ** dim
  elif b.ndim == 1:
    vol = float(onp.prod(b))
  else:
    vol = float(abs(onp.linalg.det(b)))
  rho = float(f'{N / vol:.3g}')
  return (int(N), int(dim), rho, round(float(r_cutoff), 6), round(float(dr_threshold), 6),
          fmt.name, str(kind))


def _load_file_cache(path):
  try:
    with open(path) as f:
      return json.load(f)
  except (OSError, ValueError):
    return {}


def _time_trial(make_fns, position, momentum, shift_fn, dt, kwargs, tpa, trial_steps,
                repeats) -> Optional[float]:
  """us per step of `trial_steps`-step NVELoop blocks with these settings, on
  copies of the state; None when the list overflowed."""
  set_threads_per_atom(tpa)
  neighbor_fn, energy_fn = make_fns(**kwargs)
  R = position.clone()
  nbrs = neighbor_fn.allocate(R)
  init_fn, _ = simulate.nve(energy_fn, shift_fn, dt)
  state = init_fn(0, R, kT=0.0, momenta=momentum.clone(), neighbor=nbrs)
  loop = simulate.NVELoop(energy_fn, shift_fn, dt)
  loop.prime(state, nbrs)
  loop.run(state, nbrs, trial_steps)      # eager: loads the kernels, records the graph
  loop.capture(state, nbrs, trial_steps)  # the continuing block's graph too
  stream = torch.cuda.current_stream(R.device)
  e0 = torch.cuda.Event(enable_timing=True)
  e1 = torch.cuda.Event(enable_timing=True)
  e0.record(stream)
  for _ in range(repeats):
    loop.run(state, nbrs, trial_steps)
  e1.record(stream)
  torch.cuda.synchronize(R.device)
  if bool(nbrs.did_buffer_overflow):
    return None
  return 1e3 * e0.elapsed_time(e1) / (repeats * trial_steps)


def tune_nve_loop(make_fns: Callable, position: torch.Tensor, momentum: torch.Tensor,
                  shift_fn, dt: float, key: Optional[tuple] = None, trial_steps: int = 40,
                  repeats: int = 3, tpas: Sequence[int] = (4, 8, 16),
                  cell_size_factors: Sequence[float] = (1.0, 1.15, 1.3),
                  capacity_multipliers: Sequence[float] = (1.25, 1.5),
                  cache_path: Optional[str] = None) -> TunedConfig:
  """Chooses threads per row, cell size factor and capacity multiplier for
  ``simulate.NVELoop`` by timing it.

  ``make_fns(**neighbor_kwargs) -> (neighbor_fn, energy_fn)`` builds the
  list and potential with the given ``cell_size_factor`` and
  ``capacity_multiplier`` (e.g. ``lambda **kw:
  energy.lennard_jones_neighbor_list(disp, box, format=Dense, **kw)``).
  ``position`` / ``momentum`` are the state to time from; they are not
  modified.  ``key`` (see ``workload_key``) caches the result; the process
  keeps its threads-per-row setting as it was.
  """
  if key is not None and key in _CACHE:
    return _CACHE[key]
  file_cache = _load_file_cache(cache_path) if (cache_path and key is not None) else {}
  skey = json.dumps(list(key)) if key is not None else None
  if skey is not None and skey in file_cache:
    d = file_cache[skey]
    cfg = TunedConfig(int(d['threads_per_atom']), float(d['cell_size_factor']),
                      float(d['capacity_multiplier']), float(d['us_per_step']),
                      [tuple(t) for t in d['trials']], key)
    _CACHE[key] = cfg
    return cfg
  saved_tpa = _current_tpa
  default_tpa = 8 if position.shape[1] == 3 else 2
  best = {'tpa': default_tpa if default_tpa in tpas else tpas[0],
          'cell_size_factor': cell_size_factors[0],
          'capacity_multiplier': (1.25 if 1.25 in capacity_multipliers
                                  else capacity_multipliers[0])}
  trials = []
  measured: Dict[Tuple, Optional[float]] = {}

  def measure(cand):
    t = (cand['tpa'], cand['cell_size_factor'], cand['capacity_multiplier'])
    if t not in measured:
      us = _time_trial(make_fns, position, momentum, shift_fn, dt,
                       {'cell_size_factor': t[1], 'capacity_multiplier': t[2]}, t[0],
                       trial_steps, repeats)
      measured[t] = us
      trials.append(t + (us,))
    return measured[t]

  try:
    for name, values in (('tpa', tpas), ('cell_size_factor', cell_size_factors),
                         ('capacity_multiplier', capacity_multipliers)):
      scores = []
      for v in values:
        cand = dict(best, **{name: v})
        us = measure(cand)
        scores.append((float('inf') if us is None else us, v))
      best_us, best_v = min(scores)
      if best_us < float('inf'):
        best[name] = best_v
  finally:
    set_threads_per_atom(saved_tpa)
  us = measured[(best['tpa'], best['cell_size_factor'], best['capacity_multiplier'])]
  cfg = TunedConfig(best['tpa'], best['cell_size_factor'], best['capacity_multiplier'],
                    float('nan') if us is None else us, trials, key)
  if key is not None:
    _CACHE[key] = cfg
    if cache_path:
      file_cache[skey] = cfg.as_dict()
      tmp = cache_path + '.tmp'
      with open(tmp, 'w') as f:
        json.dump(file_cache, f, indent=1)
      os.replace(tmp, cache_path)
  return cfg
