"""biofvmb -- MI355X (gfx950) HIP implementation of BioFVM-B's LOD diffusion-decay step
and agent source/sink update, driven through the C ABI in include/biofvmb.h."""
from ._lib import BioFVMBError, FP32, FP64, HEADER_PATH, LIB_PATH, load  # noqa: F401
from .solver import Batch, LODSolver  # noqa: F401

__all__ = ["LODSolver", "Batch", "BioFVMBError", "load", "LIB_PATH", "HEADER_PATH", "FP64", "FP32"]
