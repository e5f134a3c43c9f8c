"""Python host binding of libsph_hip.so (the C-ABI declared in include/sph_world.h).

The binding is ctypes over the C-ABI only: no torch types, no CPU fallback.  Loading fails
loudly when the HIP library is missing.
"""
from .abi import (  # noqa: F401
    SphError, SphParams, PassTimes, load_library, lib_path, DUMP_STAGES, KERNELS,
    params_for, fill_block, density_map_dims, fill_reference, reference_default_params, RNGS,
)
from .world import World  # noqa: F401
from .batch import Batch  # noqa: F401
from .server import Server  # noqa: F401
