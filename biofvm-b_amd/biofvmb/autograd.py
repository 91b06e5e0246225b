"""Autograd over the time loop: torch.ops.biofvmb.lod_run / agent_tables, implemented in
C++ as libtorch ``torch::autograd::Function``s (host/bfvmb_autograd.cpp) over the C ABI's
adjoint (bfvmb_adjoint_forward / bfvmb_adjoint_backward).

    rho_T = lod_run(solver, rho0, voxel, temp1, temp2, nsteps, checkpoint_every=10)

is differentiable in ``rho0`` and the agent tables; ``agent_tables`` forms temp1/temp2 from
volumes and secretion/uptake/saturation rates with the reference's operation order
(BioFVM_basic_agent.cpp:155-176), so gradients reach the rates by the chain rule.
"""
from __future__ import annotations

import os

import torch

from ._lib import BioFVMBError, PKG_DIR, load as _load_c

TORCH_LIB_PATH = os.path.join(PKG_DIR, "lib", "libbiofvmb_torch.so")
_loaded = False


def load() -> None:
    """Load the operator library once (raises if it has not been built)."""
    global _loaded
    if _loaded:
        return
    if not os.path.exists(TORCH_LIB_PATH):
        raise BioFVMBError(f"{TORCH_LIB_PATH} is missing: build it with make -C biofvm-b_amd")
    _load_c()
    torch.ops.load_library(TORCH_LIB_PATH)
    _loaded = True


def lod_run(solver, rho0, voxel, temp1, temp2, nsteps: int, checkpoint_every: int = 10) -> torch.Tensor:
    """nsteps of (agent sources/sinks, LOD step) from ``rho0`` [nx,ny,nz,nsub] (FP64) on the
    context of ``solver`` (an LODSolver), with agents at ``voxel`` and tables temp1/temp2
    [n,nsub].  The context's field is rho_T afterwards; backward needs the context unchanged
    in between (another forward call or step makes it raise).  A CUDA ``rho0`` keeps the
    field and its gradient in HBM (the context then runs on torch's current stream)."""
    load()
    if rho0.is_cuda:  # the device path is stream-ordered on the context's stream
        solver.set_stream(torch.cuda.current_stream().cuda_stream)
    return torch.ops.biofvmb.lod_run(rho0, voxel, temp1, temp2, solver.handle, int(nsteps), int(checkpoint_every))


def agent_tables(volume, secretion, uptake, saturation, dt: float, voxel_volume: float):
    """(temp1, temp2) [n,nsub] of set_internal_uptake_constants, differentiable in the rates."""
    load()
    return torch.ops.biofvmb.agent_tables(volume, secretion, uptake, saturation, float(dt), float(voxel_volume))
