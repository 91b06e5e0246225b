"""ctypes binding of the C ABI in include/biofvmb.h (libbiofvmb.so, built in-tree).

The library is the only compute path: when it is missing or no gfx950 device is present
the calls fail with an error, they never fall back to a CPU implementation.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libbiofvmb.so")
HEADER_PATH = os.path.join(os.path.dirname(PKG_DIR), "include", "biofvmb.h")

OK, ERR_ARG, ERR_STATE, ERR_HIP, ERR_NOMEM, ERR_NODEV, ERR_IO = 0, 1, 2, 3, 4, 5, 6
FP64, FP32 = 0, 1

_c = ctypes
_dp = _c.POINTER(_c.c_double)
_ip = _c.POINTER(_c.c_int)
_bp = _c.POINTER(_c.c_uint8)
_vp = _c.c_void_p

# name -> (restype, argtypes); mirrors include/biofvmb.h one to one
SIGNATURES = {
    "bfvmb_version": (_c.c_char_p, []),
    "bfvmb_create": (_c.c_int, [_c.POINTER(_vp), _c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_double, _c.c_int, _c.c_int]),
    "bfvmb_destroy": (_c.c_int, [_vp]),
    "bfvmb_last_error": (_c.c_char_p, [_vp]),
    "bfvmb_set_stream": (_c.c_int, [_vp, _vp]),
    "bfvmb_set_coefficients": (_c.c_int, [_vp, _dp, _dp, _c.c_double]),
    "bfvmb_set_dirichlet": (_c.c_int, [_vp, _dp, _bp]),
    "bfvmb_set_dirichlet_nodes": (_c.c_int, [_vp, _c.c_int, _ip, _dp]),
    "bfvmb_upload_density": (_c.c_int, [_vp, _dp]),
    "bfvmb_download_density": (_c.c_int, [_vp, _dp]),
    "bfvmb_set_agents": (_c.c_int, [_vp, _c.c_int, _ip, _dp, _dp, _bp]),
    "bfvmb_set_agent_rates": (_c.c_int, [_vp, _c.c_int, _ip, _dp, _dp, _dp, _dp, _bp, _c.c_double]),
    "bfvmb_voxel_indices": (_c.c_int, [_c.c_int, _c.c_int, _c.c_int, _c.c_double, _dp, _c.c_int, _dp, _ip]),
    "bfvmb_simulate_cell_sources_and_sinks": (_c.c_int, [_vp]),
    "bfvmb_simulate_diffusion_decay": (_c.c_int, [_vp, _c.c_int]),
    "bfvmb_run": (_c.c_int, [_vp, _c.c_int]),
    "bfvmb_sync": (_c.c_int, [_vp]),
    "bfvmb_launch_phase": (_c.c_int, [_vp, _c.c_int, _c.c_int]),
    "bfvmb_set_launch": (_c.c_int, [_vp, _c.c_int, _c.c_int, _c.c_int, _c.c_int]),
    "bfvmb_run_graph": (_c.c_int, [_vp, _c.c_int]),
    "bfvmb_autotune": (_c.c_int, [_vp, _c.c_int, _ip]),
    "bfvmb_adjoint_forward": (_c.c_int, [_vp, _c.c_int, _c.c_int]),
    "bfvmb_write_matlab": (_c.c_int, [_vp, _c.c_char_p, _dp]),
    "bfvmb_set_bulk_rates": (_c.c_int, [_vp, _dp, _dp, _dp]),
    "bfvmb_compute_gradients": (_c.c_int, [_vp]),
    "bfvmb_download_gradients": (_c.c_int, [_vp, _dp]),
    "bfvmb_gradients_at": (_c.c_int, [_vp, _c.c_int, _ip, _dp]),
    "bfvmb_gradient_device_ptr": (_c.c_int, [_vp, _c.POINTER(_vp)]),
    "bfvmb_simulate_bulk_sources_and_sinks": (_c.c_int, [_vp, _c.c_double]),
    "bfvmb_batch_set_bulk_rates": (_c.c_int, [_vp, _c.c_int, _dp, _dp, _dp]),
    "bfvmb_read_matlab": (_c.c_int, [_vp, _c.c_char_p, _dp]),
    "bfvmb_adjoint_backward": (_c.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "bfvmb_device_info": (_c.c_int, [_c.c_int, _ip, _ip]),
    "bfvmb_density_device_ptr": (_c.c_int, [_vp, _c.POINTER(_vp)]),
    "bfvmb_batch_create": (_c.c_int, [_c.POINTER(_vp), _c.c_int, _ip, _ip, _ip, _ip, _dp, _c.c_int, _c.c_int]),
    "bfvmb_batch_destroy": (_c.c_int, [_vp]),
    "bfvmb_batch_last_error": (_c.c_char_p, [_vp]),
    "bfvmb_batch_set_stream": (_c.c_int, [_vp, _vp]),
    "bfvmb_batch_set_coefficients": (_c.c_int, [_vp, _c.c_int, _dp, _dp, _c.c_double]),
    "bfvmb_batch_set_dirichlet": (_c.c_int, [_vp, _c.c_int, _dp, _bp]),
    "bfvmb_batch_set_dirichlet_nodes": (_c.c_int, [_vp, _c.c_int, _c.c_int, _ip, _dp]),
    "bfvmb_batch_upload_density": (_c.c_int, [_vp, _c.c_int, _dp]),
    "bfvmb_batch_download_density": (_c.c_int, [_vp, _c.c_int, _dp]),
    "bfvmb_batch_set_agents": (_c.c_int, [_vp, _c.c_int, _c.c_int, _ip, _dp, _dp, _bp]),
    "bfvmb_batch_run": (_c.c_int, [_vp, _c.c_int, _c.c_int]),
    "bfvmb_batch_sync": (_c.c_int, [_vp]),
    "bfvmb_batch_set_launch": (_c.c_int, [_vp, _c.c_int, _c.c_int, _c.c_int, _c.c_int]),
    "bfvmb_batch_launch_phase": (_c.c_int, [_vp, _c.c_int, _c.c_int]),
    "bfvmb_batch_autotune": (_c.c_int, [_vp, _c.c_int, _ip]),
    "bfvmb_update_agent_positions": (_c.c_int, [_vp, _c.c_int, _vp, _dp, _c.c_int]),
    "bfvmb_densities_at": (_c.c_int, [_vp, _c.c_int, _vp, _vp, _c.c_int]),
    "bfvmb_densities_at_noalloc": (_c.c_int, [_vp, _c.c_int, _ip, _dp]),
    "bfvmb_upload_density_device": (_c.c_int, [_vp, _vp]),
    "bfvmb_download_density_device": (_c.c_int, [_vp, _vp]),
    "bfvmb_adjoint_backward_device": (_c.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "bfvmb_batch_update_agent_positions": (_c.c_int, [_vp, _c.c_int, _c.c_int, _vp, _dp, _c.c_int]),
    "bfvmb_assign_agent_positions": (_c.c_int, [_vp, _c.c_int, _vp, _dp, _c.c_int]),
    "bfvmb_batch_assign_agent_positions": (_c.c_int, [_vp, _c.c_int, _c.c_int, _vp, _dp, _c.c_int]),
}

_lib = None


class BioFVMBError(RuntimeError):
    pass


def load(path: str = LIB_PATH):
    """Load libbiofvmb.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise BioFVMBError(f"{path} is missing: build it with `make -C biofvm-b_amd` "
                               "or __graft_entry__.build()")
        # torch ships its own HIP runtime under the same soname (libamdhip64.so.7) as the one
        # the library links: whichever is loaded first serves the process, and torch finds no
        # GPU on the other one.  Load torch's first, so that torch tensors and streams keep
        # working next to the library.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = _c.CDLL(path)
        missing = []
        for name, (res, args) in SIGNATURES.items():
            try:
                fn = getattr(lib, name)
            except AttributeError:
                missing.append(name)
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = lib if not missing else _Partial(lib, path, missing)
    return _lib


class _Partial:
    """A library built from older sources than this binding: the entry points it lacks raise
    when called (instead of failing every use of the library at load time)."""

    def __init__(self, lib, path, missing):
        self._lib, self._path, self._missing = lib, path, set(missing)

    def __getattr__(self, name):
        if name in self._missing:
            def fail(*args, **kwargs):
                raise BioFVMBError(f"{self._path} has no {name}: rebuild it (make -C biofvm-b_amd)")
            return fail
        return getattr(self._lib, name)


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_ip)


def bptr(a: np.ndarray):
    return a.ctypes.data_as(_bp)
