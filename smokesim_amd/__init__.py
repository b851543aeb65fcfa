"""smokesim_amd -- MI355X-native smoke simulation step.

The product is ``libsmoke_hip.so`` (C-ABI in include/smoke_abi.h, gfx950 HIP
kernels).  This package is a thin ctypes binding used by tests, bench.py and
Python callers.  There is no CPU fallback: every call goes to the HIP
library, and loading fails loudly when the library is missing.

``Grid`` mirrors the reference's per-step operator surface
(common/physics/grid.h:31-84, physics.h:50-67): the same stage names, the
same settings structs and the same emitter/obstacle layouts.
"""
from ._abi import (  # noqa: F401
    Grid, Rollout, Batch, SmokeError, lib, lib_path, device_sync, calibrate_copy, PRECISION, FIELD, KERNEL,
    EMITTER_DTYPE, OBSTACLE_DTYPE, BOX_DTYPE,
)
