"""Loaders for the reference's own test fixtures, copied here as data:

* simulation_test_state.npy -- jammed 2D bidisperse soft-sphere packing,
  N=128, with its float64 energy and pressure
  (reference tests/data, loaded as in jax_md/test_util.py:347-358);
* lammps_lj_stress_test{,_states} -- 108 LJ atoms from LAMMPS with E/N and
  stress (jax_md/test_util.py:370-420).
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def jammed_state():
  with open(os.path.join(HERE, 'simulation_test_state.npy'), 'rb') as f:
    names = ['fractional_position', 'real_position', 'species', 'sigma', 'box', 'energy',
             'pressure']
    return {k: np.load(f, allow_pickle=False) for k in names}


def lammps_lj():
  with open(os.path.join(HERE, 'lammps_lj_stress_test_states')) as f:
    lines = f.read().split('\n')
  n = int(lines[3])
  L = float(lines[5].split(' ')[-1])
  R, V = [], []
  for line in lines[9:-1]:
    vals = [float(x) for x in line.split(' ')]
    R.append(vals[:3])
    V.append(vals[3:])
  with open(os.path.join(HERE, 'lammps_lj_stress_test')) as f:
    res = [float(x) for x in f.read().split()]
  assert len(R) == n
  return {'box': L, 'fractional_position': np.array(R), 'velocity': np.array(V),
          'energy_per_atom': res[1], 'stress_voigt': np.array(res[2:8])}
