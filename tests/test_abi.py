"""The C-ABI library loads and exports every symbol include/jaxmd_native.h
declares (no compute call: this runs without a GPU)."""

import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'jaxmd_native.h')


def declared_symbols():
  text = open(HEADER).read()
  text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
  return sorted(set(re.findall(r'\b(jmd_\w+)\s*\(', text)))


def test_header_declares_entry_points():
  names = declared_symbols()
  assert 'jmd_neighbor_list_build' in names and 'jmd_pair_energy_force' in names
  assert len(names) >= 12


def test_library_exports_every_declared_symbol():
  from jax_md_amd import _native
  if not os.path.exists(_native.LIB_PATH):
    pytest.skip('libjaxmd_native.so not built (run __graft_entry__.build())')
  lib = _native.load()
  for name in declared_symbols():
    assert hasattr(lib, name), name
  assert set(_native.EXPORTED) == set(declared_symbols())


def test_host_only_queries():
  from jax_md_amd import _native
  if not os.path.exists(_native.LIB_PATH):
    pytest.skip('libjaxmd_native.so not built')
  lib = _native.load()
  assert lib.jmd_version().startswith(b'jaxmd_native')
  assert lib.jmd_neighbor_workspace_bytes(1000, 3, 125, 0, 0) > 0
  # the slot array grows with the cell capacity
  assert (lib.jmd_neighbor_workspace_bytes(1000, 3, 125, 40, 0)
          >= lib.jmd_neighbor_workspace_bytes(1000, 3, 125, 0, 0) + 125 * 40 * 4 - 256)
  assert lib.jmd_pair_workspace_bytes(1000) >= 8 * 64
  assert lib.jmd_set_threads_per_atom(3) == _native.JMD_ERR_INVALID
  assert lib.jmd_set_threads_per_atom(0) == _native.JMD_OK


def test_product_path_fails_loudly_without_library(tmp_path):
  from jax_md_amd import _native
  with pytest.raises(_native.NativeLibraryError):
    saved = _native._lib
    try:
      _native._lib = None
      _native.load(str(tmp_path / 'missing.so'))
    finally:
      _native._lib = saved
