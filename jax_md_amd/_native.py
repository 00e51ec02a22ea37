"""ctypes binding of libjaxmd_native.so (the C ABI in include/jaxmd_native.h).

The library is the only compute path of this package: when it is missing,
every entry point raises ``NativeLibraryError`` instead of falling back to
any other implementation.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, c_double, c_float, c_int32, c_int64, c_size_t, c_uint8,
                    c_void_p, c_char_p)

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get('JAX_MD_AMD_LIB', os.path.join(_HERE, 'libjaxmd_native.so'))

JMD_OK, JMD_ERR_INVALID, JMD_ERR_LAUNCH, JMD_ERR_UNSUPPORTED = 0, 1, 2, 3
# largest n jmd_pair_energy_force_noalloc accepts (JMD_PAIR_NOALLOC_MAX_N)
PAIR_NOALLOC_MAX_N = 2000000
DENSE, SPARSE, ORDERED_SPARSE = 0, 1, 2
DENSE_SPARSE_TEST = 0x10  # Dense rows with the Sparse formats' pair test
# ORed into a geometry code: every position lies in the box (positions
# written by a wrapping shift), which lets the force kernels skip folds
POSITIONS_IN_BOX = 0x100
LENNARD_JONES, SOFT_SPHERE = 0, 1
PARAM_SCALAR, PARAM_PER_PARTICLE, PARAM_SPECIES = 0, 1, 2


class NativeLibraryError(RuntimeError):
  pass


class JmdParam(ctypes.Structure):
  _fields_ = [('mode', c_int32), ('value', c_float), ('values', c_void_p)]


class JmdPairPotential(ctypes.Structure):
  _fields_ = [('kind', c_int32), ('sigma', JmdParam), ('epsilon', JmdParam),
              ('alpha', JmdParam), ('r_onset', c_float), ('r_cutoff', c_float),
              ('n_species', c_int32), ('species', c_void_p)]


_F3 = c_float * 3
_I3 = c_int32 * 3

# name -> (restype, argtypes)
_SIGNATURES = {
    'jmd_version': (c_char_p, []),
    'jmd_neighbor_workspace_bytes': (c_size_t, [c_int64, c_int32, c_int64, c_int32, c_int32]),
    'jmd_cell_list': (c_int32, [c_void_p, c_int64, c_int32, c_int32, POINTER(c_int32),
                                POINTER(c_float), c_int32, c_void_p, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_void_p, c_void_p]),
    'jmd_cell_filling_max': (c_int32, [c_void_p, c_int64, c_int32, c_int32, POINTER(c_int32),
                                       POINTER(c_float), c_void_p, c_void_p, c_void_p]),
    'jmd_neighbor_list_build': (c_int32, [c_void_p, c_int64, c_int32, c_int32, c_int32,
                                          POINTER(c_float), c_float, c_int32, POINTER(c_int32),
                                          POINTER(c_float), c_int32, c_int32, c_int32, c_void_p,
                                          c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                                          c_void_p, c_void_p]),
    'jmd_neighbor_list_check': (c_int32, [c_void_p, c_int32, c_void_p, c_int64, c_int32,
                                          c_int32, POINTER(c_float), c_float, c_void_p,
                                          c_void_p]),
    'jmd_pair_workspace_bytes': (c_size_t, [c_int64]),
    'jmd_set_threads_per_atom': (c_int32, [c_int32]),
    'jmd_pair_energy_force': (c_int32, [c_void_p, c_int64, c_int32, c_int32, c_int32,
                                        POINTER(c_float), POINTER(JmdPairPotential), c_void_p,
                                        c_int64, c_int32, c_void_p, c_void_p, c_void_p,
                                        c_void_p, c_float, c_void_p, c_void_p, c_void_p]),
    'jmd_pair_energy_force_noalloc': (c_int32, [c_void_p, c_int64, c_int32, c_int32, c_int32,
                                                POINTER(c_float), POINTER(JmdPairPotential),
                                                c_void_p, c_int64, c_int32, c_void_p, c_void_p,
                                                c_void_p, c_void_p, c_float, c_void_p,
                                                c_void_p]),
    'jmd_nve_force_kick_drift': (c_int32, [c_void_p, c_int64, c_int32, c_int32, c_int32,
                                           POINTER(c_float), POINTER(JmdPairPotential),
                                           c_void_p, c_int64, c_void_p, c_void_p, c_float,
                                           c_void_p, c_void_p, c_void_p, c_float, c_int32,
                                           POINTER(c_float), c_float, c_void_p, c_float,
                                           c_void_p, c_int32, c_void_p, c_void_p]),
    'jmd_nve_kick_drift': (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_int64,
                                     c_int32, c_int32, POINTER(c_float), c_float, c_float,
                                     c_void_p, c_void_p, c_void_p, c_int32, c_void_p,
                                     c_void_p, c_float, c_void_p, c_void_p]),
    'jmd_nve_list_update_kick_drift': (
        c_int32, [c_void_p, c_int64, c_int32, c_int32, c_int32, POINTER(c_float), c_float,
                  c_int32, POINTER(c_int32), POINTER(c_float), c_int32, c_int32, c_int32,
                  c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                  c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_int32, POINTER(c_float),
                  c_float, c_float, c_void_p, c_void_p, c_float, c_void_p, c_void_p]),
    'jmd_cell_order': (c_int32, [c_void_p, c_int64, c_int32, c_int32, POINTER(c_int32),
                                 POINTER(c_float), c_void_p, c_void_p, c_void_p, c_void_p]),
    'jmd_permute_rows': (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_int32, c_int32,
                                   c_int32, c_int32, c_void_p]),
    'jmd_remap_neighbor_ids': (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_int32,
                                         c_int64, c_void_p]),
    'jmd_adjoint_workspace_bytes': (c_size_t, [c_int64, c_int32]),
    'jmd_pair_force_vjp': (c_int32, [c_void_p, c_int64, c_int32, c_int32, c_int32,
                                     POINTER(c_float), POINTER(JmdPairPotential), c_void_p,
                                     c_int64, c_int32, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p]),
    'jmd_adjoint_param_acc_doubles': (c_size_t, [c_int64, c_int32]),
    'jmd_nve_adjoint_step': (c_int32, [c_void_p, c_int64, c_int32, c_int32, c_int32,
                                       POINTER(c_float), POINTER(JmdPairPotential), c_void_p,
                                       c_int64, c_void_p, c_void_p, c_void_p, c_float, c_float,
                                       c_void_p, c_float, c_void_p]),
    'jmd_adjoint_param_sum': (c_int32, [c_void_p, c_int64, c_int32, c_void_p, c_void_p]),
    'jmd_adjoint_param_count': (c_int64, [POINTER(JmdPairPotential), c_int64]),
    'jmd_adjoint_param_gather': (c_int32, [c_void_p, c_int64, c_int32,
                                           POINTER(JmdPairPotential), c_void_p, c_void_p]),
    'jmd_nve_adjoint_step_x64': (c_int32, [c_void_p, c_int64, c_int32, c_int32,
                                           POINTER(c_float), POINTER(JmdPairPotential), c_void_p,
                                           c_int64, c_void_p, c_void_p, c_void_p, c_float,
                                           c_float, c_void_p, c_float, c_void_p]),
    'jmd_pair_workspace_bytes_f64': (c_size_t, [c_int64]),
    'jmd_nvt_kick_drift': (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_int64,
                                     c_int32, c_int32, POINTER(c_float), c_float, c_float,
                                     c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_float,
                                     c_void_p, c_void_p, c_void_p]),
    'jmd_kick_kinetic_partials': (c_int32, [c_int64]),
    'jmd_kick_kinetic': (c_int32, [c_void_p, c_void_p, c_void_p, c_float, c_int64, c_int32,
                                   c_float, c_void_p, c_void_p]),
    'jmd_nhc_step': (c_int32, [c_int32, c_void_p, c_int32, c_int32, c_int32, c_float, c_float,
                               c_float, c_double, c_void_p, c_void_p, c_int32, c_void_p,
                               c_void_p]),
    'jmd_kick_kinetic_nhc': (c_int32, [c_void_p, c_void_p, c_void_p, c_float, c_int64, c_int32,
                                       c_float, c_void_p, c_void_p, c_int32, c_void_p, c_int32,
                                       c_int32, c_int32, c_float, c_float, c_float, c_double,
                                       c_void_p, c_void_p, c_void_p]),
    'jmd_scale_momentum': (c_int32, [c_void_p, c_int64, c_int32, c_void_p, c_void_p]),
    'jmd_pair_virial_workspace_bytes': (c_size_t, []),
    'jmd_multi_image_workspace_bytes': (c_size_t, [c_int64, c_int32]),
    'jmd_multi_image_max_displacement_sq': (c_int32, [c_void_p, c_void_p, c_int64, c_int32,
                                                      c_int32, POINTER(c_float),
                                                      POINTER(c_float), c_void_p, c_void_p]),
    'jmd_multi_image_build': (c_int32, [c_void_p, c_int64, c_int32, c_int32, POINTER(c_float),
                                        POINTER(c_int32), c_float, c_int32, c_void_p, c_void_p,
                                        c_int64, c_void_p, c_void_p, c_void_p, c_void_p]),
    'jmd_pair_virial': (c_int32, [c_void_p, c_int64, c_int32, c_int32, c_int32,
                                  POINTER(c_float), POINTER(JmdPairPotential), c_void_p, c_int64,
                                  c_int32, c_void_p, c_void_p, c_void_p]),
    'jmd_pair_edge_energies': (c_int32, [c_void_p, c_int64, c_int32, c_int32, c_int32,
                                         POINTER(c_float), POINTER(JmdPairPotential), c_void_p,
                                         c_int64, c_int32, c_void_p, c_void_p]),
    'jmd_pair_energy_force_f64': (c_int32, [c_void_p, c_int64, c_int32, c_int32,
                                            POINTER(c_double), POINTER(JmdPairPotential),
                                            c_void_p, c_int64, c_int32, c_void_p, c_void_p,
                                            c_void_p, c_double, c_void_p, c_void_p]),
    'jmd_nve_kick_drift_f64': (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_double,
                                         c_int64, c_int32, c_int32, POINTER(c_float), c_double,
                                         c_double, c_void_p, c_int32, c_void_p, c_float,
                                         c_void_p, c_void_p]),
    'jmd_fire_ensemble_max_cells': (c_int32, []),
    'jmd_fire_ensemble_2d': (c_int32, [c_void_p, c_void_p, c_int32, c_int64, c_int32, c_int32,
                                       c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_int32, c_void_p, POINTER(JmdPairPotential),
                                       POINTER(c_float), c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_void_p, c_void_p, c_void_p]),
    'jmd_fire_ensemble_adjoint_2d': (c_int32, [c_void_p, c_int32, c_int64, c_int32, c_int32,
                                               c_void_p, c_void_p, c_void_p, c_int32,
                                               POINTER(JmdPairPotential), c_void_p, c_void_p,
                                               c_void_p, c_void_p, c_void_p, c_void_p,
                                               c_void_p, c_void_p, c_void_p, c_void_p,
                                               c_void_p, c_void_p, c_float, c_void_p]),
    'jmd_fire_ensemble_set_threads': (c_int32, [c_int32]),
    'jmd_fire_workspace_bytes': (c_size_t, [c_int64]),
    'jmd_fire_update': (c_int32, [c_void_p, c_void_p, c_int64, c_int32, c_void_p, c_void_p,
                                  POINTER(c_float), c_void_p, c_void_p, c_void_p]),
}

EXPORTED = tuple(_SIGNATURES)

_lib = None


def load(path: str = LIB_PATH):
  """Loads the library once; raises NativeLibraryError when it is absent."""
  global _lib
  if _lib is not None:
    return _lib
  if not os.path.exists(path):
    raise NativeLibraryError(
        f'{path} is missing: build it with `python -m jax_md_amd.build_native` '
        '(the package has no other compute path).')
  lib = ctypes.CDLL(path)
  for name, (res, args) in _SIGNATURES.items():
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = args
  _lib = lib
  return lib


def check(status: int, what: str):
  if status != JMD_OK:
    kind = {JMD_ERR_INVALID: 'invalid argument', JMD_ERR_LAUNCH: 'HIP launch failure',
            JMD_ERR_UNSUPPORTED: 'unsupported request'}.get(status, f'status {status}')
    raise NativeLibraryError(f'{what} failed: {kind}')


def ptr(t):
  """Device pointer of a tensor (None -> NULL)."""
  if t is None:
    return None
  return c_void_p(t.data_ptr())


def stream_handle(device=None):
  return c_void_p(torch.cuda.current_stream(device).cuda_stream)


_REPLAY_STREAMS = {}


def replay_stream(device):
  """The device's second stream (default priority, one per device for the
  process), on which the gradients replay checkpointed segments beside the
  backward work of the current stream."""
  key = torch.device(device).index
  if key is None:
    key = torch.cuda.current_device()
  if key not in _REPLAY_STREAMS:
    _REPLAY_STREAMS[key] = torch.cuda.Stream(device=key)
  return _REPLAY_STREAMS[key]


def f3(vals):
  """float32 array of at least 3 entries (a box matrix passes 4 or 9)."""
  vals = list(vals)
  a = _F3() if len(vals) <= 3 else (c_float * len(vals))()
  for k, v in enumerate(vals):
    a[k] = float(v)
  return a


def d3(vals):
  a = (c_double * 3)()
  for k, v in enumerate(vals):
    a[k] = float(v)
  return a


def i3(vals):
  a = _I3()
  for k, v in enumerate(vals):
    a[k] = int(v)
  return a


def position_stride(t, name='position'):
  """Row stride of a position array: [N, dim] contiguous, or the [:, :3]
  view of a 16-byte-aligned [N, 4] buffer."""
  require_device_tensor(t, name, torch.float32, contiguous=False)
  if t.ndim != 2 or t.shape[1] not in (2, 3):
    raise ValueError(f'{name} must be [N, 2] or [N, 3] float32, got {tuple(t.shape)}')
  dim = t.shape[1]
  if t.shape[0] <= 1:
    return dim
  rs, cs = t.stride()
  if cs == 1 and (rs == dim or (dim == 3 and rs == 4 and t.data_ptr() % 16 == 0)):
    return rs
  raise ValueError(f'{name} must be contiguous or a [:, :3] view of an [N, 4] buffer')


def padded_positions(N, device, like=None):
  """[N, 3] float32 view of a zeroed [N, 4] buffer (16-byte rows)."""
  buf = torch.zeros((N, 4), dtype=torch.float32, device=device)
  v = buf[:, :3]
  if like is not None:
    v.copy_(like)
  return v


def require_device_tensor(t, name, dtype=None, contiguous=True):
  if not isinstance(t, torch.Tensor):
    raise TypeError(f'{name} must be a torch.Tensor on a ROCm device, got {type(t)}')
  if t.device.type != 'cuda':
    raise NativeLibraryError(
        f'{name} is on {t.device}; the native path runs on the GPU only')
  if dtype is not None and t.dtype != dtype:
    raise TypeError(f'{name} must be {dtype}, got {t.dtype}')
  if contiguous and not t.is_contiguous():
    raise ValueError(f'{name} must be contiguous')
  return t
