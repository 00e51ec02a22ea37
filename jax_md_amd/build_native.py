"""Builds libjaxmd_native.so (gfx950) in-tree with hipcc.

Each source is compiled to its own object in parallel (one hipcc process per
file, objects under ``jax_md_amd/_build``), then the objects are linked into
the shared library.  Every call recompiles.

Usage: python -m jax_md_amd.build_native  (or call build()).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OBJ = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'libjaxmd_native.so')

SOURCES = ['jmd_neighbor.hip', 'jmd_pair.hip', 'jmd_pair_dense_lj3.hip', 'jmd_pair_dense_ss3.hip',
           'jmd_pair_dense_u3.hip', 'jmd_pair_dense_2d.hip', 'jmd_adjoint.hip',
           'jmd_ensemble.hip', 'jmd_integrate.hip', 'jmd_layout.hip', 'jmd_x64.hip',
           'jmd_multi_image.hip']

FLAGS = [
    '--offload-arch=gfx950', '-O3', '-fPIC', '-std=c++17',
    # no packed-FP32 instructions anywhere in this library
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-command-line-argument',
]


def _run(cmd, verbose):
  if verbose:
    print(' '.join(cmd), flush=True)
  r = subprocess.run(cmd, capture_output=True, text=True)
  return r.returncode, r.stdout + r.stderr


def build(verbose: bool = False) -> str:
  hipcc = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
  os.makedirs(OBJ, exist_ok=True)
  srcs = list(SOURCES)
  objs = [os.path.join(OBJ, s.replace('.hip', '.o')) for s in srcs]
  cmds = [[hipcc] + FLAGS + ['-c', '-o', o, os.path.join(CSRC, s)] for s, o in zip(srcs, objs)]
  jobs = max(1, min(len(cmds), int(os.environ.get('MAX_JOBS', os.cpu_count() or 4))))
  with ThreadPoolExecutor(jobs) as ex:
    results = list(ex.map(lambda c: _run(c, verbose), cmds))
  failed = [(c, out) for c, (rc, out) in zip(cmds, results) if rc != 0]
  if failed:
    for c, out in failed:
      sys.stderr.write(' '.join(c) + '\n' + out)
    raise RuntimeError('hipcc failed building libjaxmd_native.so')
  tmp = LIB + '.tmp'
  rc, out = _run([hipcc, '--offload-arch=gfx950', '-shared', '-fPIC', '-o', tmp] + objs, verbose)
  if rc != 0:
    sys.stderr.write(out)
    raise RuntimeError('hipcc failed linking libjaxmd_native.so')
  os.replace(tmp, LIB)
  return LIB


if __name__ == '__main__':
  print(build(verbose=True))
