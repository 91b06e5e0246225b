"""Python handle on one device-resident microenvironment (a ``bfvmb_ctx``).

Method names follow the reference's Microenvironment API for this path
(BioFVM_microenvironment.h:164-248): ``simulate_cell_sources_and_sinks``,
``simulate_diffusion_decay``, Dirichlet value/activation per substrate, and the field in
the reference's voxel-major layout ``[x][y][z][substrate]``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L

PRECISIONS = {"fp64": L.FP64, "fp32": L.FP32}


def voxel_indices(grid, dx, pos, bbox_min=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Voxel of each agent position (bfvmb_voxel_indices, host only): floor((p-min)/dx)
    clamped into the grid, -1 outside the bounding box."""
    lib = L.load()
    P = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    lo = np.ascontiguousarray(bbox_min, dtype=np.float64)
    out = np.empty(P.shape[0], dtype=np.int32)
    nx, ny, nz = (int(v) for v in grid)
    rc = lib.bfvmb_voxel_indices(nx, ny, nz, float(dx), L.dptr(lo), P.shape[0], L.dptr(P), L.iptr(out))
    if rc != L.OK:
        raise L.BioFVMBError(f"bfvmb_voxel_indices failed ({rc})")
    return out


class LODSolver:
    def __init__(self, nx: int, ny: int, nz: int, nsub: int, dx: float, precision: str = "fp64",
                 device: int = 0):
        self._lib = L.load()
        self.shape = (int(nx), int(ny), int(nz), int(nsub))
        self.dx = float(dx)
        self.precision = precision
        h = ctypes.c_void_p()
        rc = self._lib.bfvmb_create(ctypes.byref(h), nx, ny, nz, nsub, dx, PRECISIONS[precision], device)
        if rc != L.OK:
            raise L.BioFVMBError(f"bfvmb_create failed ({rc}): {self._lib.bfvmb_last_error(None).decode()}")
        self._h = h

    # -- plumbing -------------------------------------------------------------------------
    def _check(self, rc: int, what: str) -> None:
        if rc != L.OK:
            raise L.BioFVMBError(f"{what} failed ({rc}): {self._lib.bfvmb_last_error(self._h).decode()}")

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.bfvmb_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def nsub(self) -> int:
        return self.shape[3]

    # -- set-up ---------------------------------------------------------------------------
    def set_stream(self, stream_ptr: int | None) -> None:
        self._check(self._lib.bfvmb_set_stream(self._h, stream_ptr), "set_stream")

    def set_coefficients(self, diffusion, decay, dt: float) -> None:
        D = np.ascontiguousarray(diffusion, dtype=np.float64)
        lam = np.ascontiguousarray(decay, dtype=np.float64)
        assert D.size == self.nsub and lam.size == self.nsub
        self._check(self._lib.bfvmb_set_coefficients(self._h, L.dptr(D), L.dptr(lam), float(dt)), "set_coefficients")

    def set_dirichlet(self, value, active) -> None:
        v = np.ascontiguousarray(np.broadcast_to(value, (self.nsub,)), dtype=np.float64)
        a = np.ascontiguousarray(np.broadcast_to(active, (self.nsub,)), dtype=np.uint8)
        self._check(self._lib.bfvmb_set_dirichlet(self._h, L.dptr(v), L.bptr(a)), "set_dirichlet")

    def upload(self, rho) -> None:
        r = np.ascontiguousarray(rho, dtype=np.float64)
        assert r.size == np.prod(self.shape)
        self._check(self._lib.bfvmb_upload_density(self._h, L.dptr(r)), "upload_density")

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        self._check(self._lib.bfvmb_download_density(self._h, L.dptr(out)), "download_density")
        return out

    def set_agents(self, voxel, temp1, temp2, active=None) -> None:
        v = np.ascontiguousarray(voxel, dtype=np.int32)
        n = v.size
        t1 = np.ascontiguousarray(temp1, dtype=np.float64).reshape(n, self.nsub)
        t2 = np.ascontiguousarray(temp2, dtype=np.float64).reshape(n, self.nsub)
        a = None if active is None else np.ascontiguousarray(active, dtype=np.uint8)
        self._check(self._lib.bfvmb_set_agents(self._h, n, L.iptr(v), L.dptr(t1), L.dptr(t2),
                                               None if a is None else L.bptr(a)), "set_agents")

    def set_agent_rates(self, voxel, volume, secretion, uptake, saturation, dt: float, active=None) -> None:
        v = np.ascontiguousarray(voxel, dtype=np.int32)
        n = v.size
        vol = np.ascontiguousarray(np.broadcast_to(volume, (n,)), dtype=np.float64)
        S, U, T = (np.ascontiguousarray(x, dtype=np.float64).reshape(n, self.nsub) for x in (secretion, uptake, saturation))
        a = None if active is None else np.ascontiguousarray(active, dtype=np.uint8)
        self._check(self._lib.bfvmb_set_agent_rates(self._h, n, L.iptr(v), L.dptr(vol), L.dptr(S), L.dptr(U), L.dptr(T),
                                                    None if a is None else L.bptr(a), float(dt)), "set_agent_rates")

    def voxel_indices(self, pos, bbox_min=(0.0, 0.0, 0.0)) -> np.ndarray:
        """Agent voxel per position (-1 outside the bounding box), reference numbering."""
        return voxel_indices(self.shape[:3], self.dx, pos, bbox_min)

    def set_launch(self, yz_blocks: int = 0, x_blocks: int = 0, lds_blocks_per_cu: int = 0) -> None:
        self._check(self._lib.bfvmb_set_launch(self._h, yz_blocks, x_blocks, lds_blocks_per_cu), "set_launch")

    # -- stepping -------------------------------------------------------------------------
    def simulate_cell_sources_and_sinks(self) -> None:
        self._check(self._lib.bfvmb_simulate_cell_sources_and_sinks(self._h), "simulate_cell_sources_and_sinks")

    def simulate_diffusion_decay(self, nsteps: int = 1) -> None:
        self._check(self._lib.bfvmb_simulate_diffusion_decay(self._h, nsteps), "simulate_diffusion_decay")

    def run(self, nsteps: int) -> None:
        self._check(self._lib.bfvmb_run(self._h, nsteps), "run")

    def launch_phase(self, phase: int, count: int = 1) -> None:
        self._check(self._lib.bfvmb_launch_phase(self._h, phase, count), "launch_phase")

    def sync(self) -> None:
        self._check(self._lib.bfvmb_sync(self._h), "sync")

    def device_ptr(self) -> int:
        p = ctypes.c_void_p()
        self._check(self._lib.bfvmb_density_device_ptr(self._h, ctypes.byref(p)), "density_device_ptr")
        return int(p.value or 0)
