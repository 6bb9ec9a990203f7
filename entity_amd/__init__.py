"""entity_amd -- MI355X-native (gfx950) implementation of Entity's per-timestep
PIC hot path: SR pusher, charge-conserving deposit, Minkowski Yee solver,
digital current filter, GR pusher.  Compute runs in hand-written HIP kernels
(libntt_hip.so, C ABI in include/ntt.h); this package holds the host-side
mirror of the reference's kernel interface.
"""
from . import _abi  # noqa: F401
from .containers import Fields, GridShape, Particles, N_GHOSTS  # noqa: F401

__all__ = ["Fields", "GridShape", "Particles", "N_GHOSTS"]
