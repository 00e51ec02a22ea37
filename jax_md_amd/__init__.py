"""gfx950-native implementation of JAX MD's neighbour-list MD hot path.

Modules mirror the reference's: space, partition, smap, energy, quantity,
simulate, minimize, plus adjoint (force and trajectory gradients), ensemble (ragged batches
of systems in one launch) and autotune (threads per row, cell size, capacity).  Arrays are torch tensors on a ROCm device; every
per-particle computation runs in libjaxmd_native.so (HIP, gfx950).
"""

from . import util, space, partition, smap, energy, quantity, simulate, minimize  # noqa: F401
from . import adjoint, ensemble, custom_partition, autotune  # noqa: F401

__version__ = '0.1.0'
