"""Builds libjaxmd_native.so (gfx950) in-tree with hipcc.

Usage: python -m jax_md_amd.build_native  (or call build()).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libjaxmd_native.so')

SOURCES = ['jmd_neighbor.hip', 'jmd_pair.hip', 'jmd_adjoint.hip', 'jmd_ensemble.hip',
           'jmd_integrate.hip', 'jmd_layout.hip', 'jmd_x64.hip', 'jmd_multi_image.hip']

FLAGS = [
    '--offload-arch=gfx950', '-O3', '-fPIC', '-shared', '-std=c++17',
    # no packed-FP32 instructions anywhere in this library
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-command-line-argument',
]


def build(verbose: bool = False) -> str:
  hipcc = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
  srcs = [os.path.join(CSRC, s) for s in SOURCES]
  tmp = LIB + '.tmp'
  cmd = [hipcc] + FLAGS + ['-o', tmp] + srcs
  if verbose:
    print(' '.join(cmd), flush=True)
  r = subprocess.run(cmd, capture_output=True, text=True)
  if r.returncode != 0:
    sys.stderr.write(r.stdout + r.stderr)
    raise RuntimeError('hipcc failed building libjaxmd_native.so')
  os.replace(tmp, LIB)
  return LIB


if __name__ == '__main__':
  print(build(verbose=True))
