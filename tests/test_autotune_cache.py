"""The autotuner's host side, without a GPU: its cache key (density from
scalar, vector and matrix boxes) and its JSON file cache."""

import numpy as np


def test_autotune_workload_key_and_file_cache(tmp_path):
  """The tuner's cache key (density from scalar, vector and matrix boxes) and
  its JSON file cache: a cached workload returns without timing anything."""
  import json
  import torch
  from jax_md_amd import autotune, partition
  R = torch.zeros((1000, 3))
  k1 = autotune.workload_key(R, 10.0, 2.5, 0.5, partition.Dense, 'lj')
  k2 = autotune.workload_key(R, [10.0, 10.0, 10.0], 2.5, 0.5, partition.Dense, 'lj')
  k3 = autotune.workload_key(R, np.diag([10.0, 10.0, 10.0]), 2.5, 0.5, partition.Dense, 'lj')
  assert k1 == k2 == k3 and k1[2] == 1.0
  path = str(tmp_path / 'tune.json')
  entry = {'threads_per_atom': 16, 'cell_size_factor': 1.15, 'capacity_multiplier': 1.5,
           'us_per_step': 12.5, 'trials': [[16, 1.15, 1.5, 12.5]]}
  with open(path, 'w') as f:
    json.dump({json.dumps(list(k1)): entry}, f)
  def never(**kw):
    raise AssertionError('a cached workload must not be timed')
  cfg = autotune.tune_nve_loop(never, R, R, None, 0.005, key=k1, cache_path=path)
  assert (cfg.threads_per_atom, cfg.cell_size_factor, cfg.capacity_multiplier) == (16, 1.15, 1.5)
  assert cfg.neighbor_kwargs == {'cell_size_factor': 1.15, 'capacity_multiplier': 1.5}
  autotune._CACHE.clear()
