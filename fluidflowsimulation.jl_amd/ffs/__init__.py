"""ffs: MI355X-native time-step engine for FluidFlowSimulation.jl's hot path.

The compute lives in libffs_hip.so (hand-written gfx950 HIP kernels behind
the C ABI of include/ffs/ffs.h); this package is the host-side mirror of the
reference's driver API.
"""
from . import _lib
from .config import Params, parameters, parse_expr, readglobal
from .fields import pad, padded_shape, read_field, unpad, write_field
from .ensemble import Ensemble
from .simulation import Simulation, footprint, run_simulation

__all__ = [
    "Ensemble",
    "Params",
    "Simulation",
    "footprint",
    "parameters",
    "parse_expr",
    "readglobal",
    "run_simulation",
    "pad",
    "padded_shape",
    "read_field",
    "unpad",
    "write_field",
]
