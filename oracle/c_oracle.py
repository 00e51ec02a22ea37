"""ctypes wrapper of the C FP64 oracle (oracle/md_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/ and by bench.py's cpu_baseline
leg, never by the product package.
"""
import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, 'libmd_oracle.so')
_lib = None


def load():
  global _lib
  if _lib is None:
    if not os.path.exists(LIB):
      from . import build_oracle
      build_oracle.build()
    lib = ctypes.CDLL(LIB)
    lib.md_list_create.restype = c_void_p
    lib.md_list_create.argtypes = [c_int, c_double, c_double, c_double, c_int]
    lib.md_list_free.argtypes = [c_void_p]
    lib.md_list_build.argtypes = [c_void_p, POINTER(c_double)]
    lib.md_list_overflow.argtypes = [c_void_p]
    lib.md_list_max_count.argtypes = [c_void_p]
    lib.md_lj_forces.restype = c_double
    lib.md_lj_forces.argtypes = [c_void_p, POINTER(c_double), c_double, c_double, c_double,
                                 c_double, POINTER(c_double)]
    lib.md_nve_run.argtypes = [c_void_p, POINTER(c_double), POINTER(c_double),
                               POINTER(c_double), c_double, c_double, c_double, c_double,
                               c_double, c_int]
    lib.md_num_threads.restype = c_int
    _lib = lib
  return _lib


def _p(a):
  return a.ctypes.data_as(POINTER(c_double))


class LJSystem:
  """FP64 LJ system with a Dense Verlet list (sigma = epsilon = 1 by default)."""

  def __init__(self, R, L, r_on=2.0, r_c=2.5, skin=0.5, width=160, sigma=1.0, eps=1.0):
    self.lib = load()
    self.R = np.ascontiguousarray(R, np.float64).copy()
    self.n = self.R.shape[0]
    self.L = float(L)
    self.params = (sigma, eps, r_on, r_c)
    self.nl = self.lib.md_list_create(self.n, self.L, r_c + skin, skin, width)
    self.lib.md_list_build(self.nl, _p(self.R))
    self.F = np.zeros_like(self.R)

  def forces(self, R=None):
    if R is not None:
      self.R = np.ascontiguousarray(R, np.float64).copy()
      self.lib.md_list_build(self.nl, _p(self.R))
    E = self.lib.md_lj_forces(self.nl, _p(self.R), *self.params, _p(self.F))
    return E, self.F.copy()

  def nve(self, P, dt, n_steps):
    P = np.ascontiguousarray(P, np.float64).copy()
    self.lib.md_lj_forces(self.nl, _p(self.R), *self.params, _p(self.F))
    rebuilds = self.lib.md_nve_run(self.nl, _p(self.R), _p(P), _p(self.F), *self.params,
                                   float(dt), int(n_steps))
    return P, rebuilds

  @property
  def overflow(self):
    return bool(self.lib.md_list_overflow(self.nl))

  def __del__(self):
    try:
      self.lib.md_list_free(self.nl)
    except Exception:
      pass


def num_threads():
  return load().md_num_threads()
