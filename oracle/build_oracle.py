"""Builds the C oracle (oracle/libmd_oracle.so) with gcc -O3 -fopenmp.

TEST INFRASTRUCTURE ONLY.  The reference (jax-md) is pure Python and ships no
C/C++ sources, so there is no oracle/_ref build; this compiles only the
project's own FP64 restatement.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, 'libmd_oracle.so')


def build(verbose: bool = False) -> str:
  src = os.path.join(HERE, 'md_oracle.c')
  tmp = LIB + '.tmp'
  cmd = ['gcc', '-O3', '-fopenmp', '-fPIC', '-shared', '-std=c11', '-o', tmp, src, '-lm']
  if verbose:
    print(' '.join(cmd), flush=True)
  r = subprocess.run(cmd, capture_output=True, text=True)
  if r.returncode != 0:
    sys.stderr.write(r.stdout + r.stderr)
    raise RuntimeError('gcc failed building the C oracle')
  os.replace(tmp, LIB)
  return LIB


if __name__ == '__main__':
  print(build(verbose=True))
