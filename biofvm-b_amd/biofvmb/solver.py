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


def _positions(pos, n):
    """(pointer, on_device, keep-alive) of an [n][3] FP64 position array: a torch CUDA tensor
    is passed as a device pointer, anything else is copied to a contiguous host array."""
    if hasattr(pos, "is_cuda") and pos.is_cuda:
        if pos.dtype != __import__("torch").float64 or not pos.is_contiguous() or pos.numel() != 3 * n:
            raise ValueError("device positions must be a contiguous float64 [n][3] tensor")
        return ctypes.c_void_p(pos.data_ptr()), 1, pos
    P = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    if P.shape[0] != n:
        raise ValueError(f"{P.shape[0]} positions for {n} agents")
    return ctypes.c_void_p(P.ctypes.data), 0, P


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

    @property
    def handle(self) -> int:
        """Address of the bfvmb_ctx (for the torch operators)."""
        return int(self._h.value or 0)

    # -- set-up ---------------------------------------------------------------------------
    def set_stream(self, stream_ptr: int | None) -> None:
        self._check(self._lib.bfvmb_set_stream(self._h, stream_ptr), "set_stream")

    def set_coefficients(self, diffusion, decay, dt: float) -> None:
        D = np.ascontiguousarray(diffusion, dtype=np.float64)
        lam = np.ascontiguousarray(decay, dtype=np.float64)
        assert D.size == self.nsub and lam.size == self.nsub
        self._check(self._lib.bfvmb_set_coefficients(self._h, L.dptr(D), L.dptr(lam), float(dt)), "set_coefficients")

    def set_dirichlet(self, value, active) -> None:
        """One value per substrate on every face voxel (value None: activations only)."""
        v = None if value is None else np.ascontiguousarray(np.broadcast_to(value, (self.nsub,)), dtype=np.float64)
        a = np.ascontiguousarray(np.broadcast_to(active, (self.nsub,)), dtype=np.uint8)
        self._check(self._lib.bfvmb_set_dirichlet(self._h, None if v is None else L.dptr(v), L.bptr(a)),
                    "set_dirichlet")

    def set_dirichlet_nodes(self, voxel, values) -> None:
        """Per-voxel Dirichlet values (add_dirichlet_node): voxel[n], values[n][nsub]."""
        v = np.ascontiguousarray(voxel, dtype=np.int32).reshape(-1)
        val = np.ascontiguousarray(values, dtype=np.float64).reshape(v.size, self.nsub)
        self._check(self._lib.bfvmb_set_dirichlet_nodes(self._h, v.size, L.iptr(v), L.dptr(val)),
                    "set_dirichlet_nodes")

    def _device_f64(self, t, what):
        import torch
        if t.dtype != torch.float64 or not t.is_contiguous() or t.numel() != int(np.prod(self.shape)):
            raise ValueError(f"{what}: a contiguous float64 CUDA tensor of {tuple(self.shape)} elements")
        torch.cuda.current_stream().synchronize()  # the tensor's producers are done
        return ctypes.c_void_p(t.data_ptr())

    def upload(self, rho) -> None:
        """Field [x][y][z][substrate]: a host array, or a CUDA float64 tensor (no host copy)."""
        if hasattr(rho, "is_cuda") and rho.is_cuda:
            self._check(self._lib.bfvmb_upload_density_device(self._h, self._device_f64(rho, "upload")),
                        "upload_density_device")
            self.sync()
            return
        r = np.ascontiguousarray(rho, dtype=np.float64)
        assert r.size == np.prod(self.shape)
        self._check(self._lib.bfvmb_upload_density(self._h, L.dptr(r)), "upload_density")

    def download(self, out=None):
        """The field as a host array, or into ``out`` (a CUDA float64 tensor, no host copy)."""
        if out is not None and hasattr(out, "is_cuda") and out.is_cuda:
            self._check(self._lib.bfvmb_download_density_device(self._h, self._device_f64(out, "download")),
                        "download_density_device")
            self.sync()
            return out
        res = np.empty(self.shape, dtype=np.float64)
        self._check(self._lib.bfvmb_download_density(self._h, L.dptr(res)), "download_density")
        return res

    def set_agents(self, voxel, temp1, temp2, active=None) -> None:
        v = np.ascontiguousarray(voxel, dtype=np.int32)
        n = v.size
        t1 = np.ascontiguousarray(temp1, dtype=np.float64).reshape(n, self.nsub)
        t2 = np.ascontiguousarray(temp2, dtype=np.float64).reshape(n, self.nsub)
        a = None if active is None else np.ascontiguousarray(active, dtype=np.uint8)
        self._nagents = n
        self._check(self._lib.bfvmb_set_agents(self._h, n, L.iptr(v), L.dptr(t1), L.dptr(t2),
                                               None if a is None else L.bptr(a)), "set_agents")

    def set_agent_rates(self, voxel, volume, secretion, uptake, saturation, dt: float, active=None) -> None:
        v = np.ascontiguousarray(voxel, dtype=np.int32)
        n = v.size
        vol = np.ascontiguousarray(np.broadcast_to(volume, (n,)), dtype=np.float64)
        S, U, T = (np.ascontiguousarray(x, dtype=np.float64).reshape(n, self.nsub) for x in (secretion, uptake, saturation))
        a = None if active is None else np.ascontiguousarray(active, dtype=np.uint8)
        self._nagents = n
        self._check(self._lib.bfvmb_set_agent_rates(self._h, n, L.iptr(v), L.dptr(vol), L.dptr(S), L.dptr(U), L.dptr(T),
                                                    None if a is None else L.bptr(a), float(dt)), "set_agent_rates")

    def densities_at(self, voxel, out=None):
        """Densities at the given voxels (reference numbering), [n][substrate] FP64, gathered
        from the resident field (bfvmb_densities_at).  ``voxel`` may be a CUDA int32 tensor;
        then ``out`` (a CUDA float64 tensor of n*nsub elements) is filled in stream order."""
        if hasattr(voxel, "is_cuda") and voxel.is_cuda:
            import torch
            n = voxel.numel()
            if out is None:
                out = torch.empty((n, self.nsub), dtype=torch.float64, device=voxel.device)
            if voxel.dtype != torch.int32 or not voxel.is_contiguous() or out.dtype != torch.float64 \
                    or not out.is_contiguous() or out.numel() != n * self.nsub:
                raise ValueError("device voxels: contiguous int32; out: contiguous float64 [n][nsub]")
            self._check(self._lib.bfvmb_densities_at(self._h, n, ctypes.c_void_p(voxel.data_ptr()),
                                                     ctypes.c_void_p(out.data_ptr()), 1), "densities_at")
            return out
        v = np.ascontiguousarray(voxel, dtype=np.int32).reshape(-1)
        res = np.empty((v.size, self.nsub), dtype=np.float64)
        self._check(self._lib.bfvmb_densities_at(self._h, v.size, ctypes.c_void_p(v.ctypes.data),
                                                 ctypes.c_void_p(res.ctypes.data), 0), "densities_at")
        return res

    def densities_at_noalloc(self, voxel):
        """densities_at from host arrays through bfvmb_densities_at_noalloc: staged in the
        library's static device scratch, no device allocation."""
        v = np.ascontiguousarray(voxel, dtype=np.int32).reshape(-1)
        res = np.empty((v.size, self.nsub), dtype=np.float64)
        self._check(self._lib.bfvmb_densities_at_noalloc(self._h, v.size, L.iptr(v), L.dptr(res)),
                    "densities_at_noalloc")
        return res

    def update_agent_positions(self, pos, bbox_min=(0.0, 0.0, 0.0)) -> None:
        """New positions ([n][3], host array or CUDA float64 tensor) of the agents last set:
        voxels and the grouping by voxel are recomputed on the device; agents outside the
        bounding box become inactive (bfvmb_update_agent_positions)."""
        n = getattr(self, "_nagents", 0)
        ptr, dev, keep = _positions(pos, n)
        lo = np.ascontiguousarray(bbox_min, dtype=np.float64)
        self._check(self._lib.bfvmb_update_agent_positions(self._h, n, ptr, L.dptr(lo), dev),
                    "update_agent_positions")
        del keep

    def assign_agent_positions(self, pos, bbox_min=(0.0, 0.0, 0.0)) -> None:
        """The same move by assign_position's rule: agents whose new position lies outside the
        bounding box keep their voxel and stay active (bfvmb_assign_agent_positions)."""
        n = getattr(self, "_nagents", 0)
        ptr, dev, keep = _positions(pos, n)
        lo = np.ascontiguousarray(bbox_min, dtype=np.float64)
        self._check(self._lib.bfvmb_assign_agent_positions(self._h, n, ptr, L.dptr(lo), dev),
                    "assign_agent_positions")
        del keep

    def compute_gradients(self) -> None:
        """Gradients of the current field into the device buffer (bfvmb_compute_gradients)."""
        self._check(self._lib.bfvmb_compute_gradients(self._h), "compute_gradients")

    def download_gradients(self) -> np.ndarray:
        """[x][y][z][substrate][axis] of the last compute_gradients."""
        out = np.empty(self.shape + (3,), dtype=np.float64)
        self._check(self._lib.bfvmb_download_gradients(self._h, L.dptr(out)), "download_gradients")
        return out

    def gradients_at(self, voxel) -> np.ndarray:
        """[k][substrate][axis] gradients of the current field at the listed voxels."""
        v = np.ascontiguousarray(voxel, dtype=np.int32)
        out = np.empty((v.size, self.nsub, 3), dtype=np.float64)
        self._check(self._lib.bfvmb_gradients_at(self._h, v.size, L.iptr(v), L.dptr(out)), "gradients_at")
        return out

    def set_bulk_rates(self, supply=None, target=None, uptake=None) -> None:
        """Per-voxel bulk supply S, target T, uptake U, each [x][y][z][substrate] (None = 0);
        all None removes the bulk phase (bfvmb_set_bulk_rates)."""
        keep = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape)
                for a in (supply, target, uptake)]
        self._check(self._lib.bfvmb_set_bulk_rates(self._h, *[None if k is None else L.dptr(k) for k in keep]),
                    "set_bulk_rates")

    def simulate_bulk_sources_and_sinks(self, dt: float) -> None:
        self._check(self._lib.bfvmb_simulate_bulk_sources_and_sinks(self._h, float(dt)),
                    "simulate_bulk_sources_and_sinks")

    def voxel_indices(self, pos, bbox_min=(0.0, 0.0, 0.0)) -> np.ndarray:
        """Agent voxel per position (-1 outside the bounding box), reference numbering."""
        return voxel_indices(self.shape[:3], self.dx, pos, bbox_min)

    def set_launch(self, sweep: int = -1, grid: int = 0, lds_blocks_per_cu: int = 0, ring: int = 0) -> None:
        """Launch shape of sweep 0 (x), 1 (y), 2 (z) or -1 (all); zeros = library default."""
        self._check(self._lib.bfvmb_set_launch(self._h, sweep, grid, lds_blocks_per_cu, ring), "set_launch")

    # -- stepping -------------------------------------------------------------------------
    def simulate_cell_sources_and_sinks(self) -> None:
        self._check(self._lib.bfvmb_simulate_cell_sources_and_sinks(self._h), "simulate_cell_sources_and_sinks")

    def simulate_diffusion_decay(self, nsteps: int = 1) -> None:
        self._check(self._lib.bfvmb_simulate_diffusion_decay(self._h, nsteps), "simulate_diffusion_decay")

    def run(self, nsteps: int, graph: bool = False) -> None:
        fn = self._lib.bfvmb_run_graph if graph else self._lib.bfvmb_run
        self._check(fn(self._h, nsteps), "run")

    def write_matlab(self, filename: str, bbox_min=(0.0, 0.0, 0.0)) -> None:
        """Microenvironment::write_to_matlab of the resident field (bfvmb_write_matlab)."""
        lo = np.ascontiguousarray(bbox_min, dtype=np.float64)
        self._check(self._lib.bfvmb_write_matlab(self._h, str(filename).encode(), L.dptr(lo)), "write_matlab")

    def read_matlab(self, filename: str, bbox_min=(0.0, 0.0, 0.0)) -> None:
        """Load a multiscale_microenvironment MAT v4 file into the field (bfvmb_read_matlab)."""
        lo = np.ascontiguousarray(bbox_min, dtype=np.float64)
        self._check(self._lib.bfvmb_read_matlab(self._h, str(filename).encode(), L.dptr(lo)), "read_matlab")

    def adjoint_forward(self, nsteps: int, checkpoint_every: int = 10) -> None:
        """``nsteps`` of ``run`` recording checkpoints for ``adjoint_backward``."""
        self._check(self._lib.bfvmb_adjoint_forward(self._h, nsteps, checkpoint_every), "adjoint_forward")

    def adjoint_backward(self, seed):
        """Gradients (d/drho_0 [x][y][z][substrate], d/dtemp1, d/dtemp2 [agent][substrate]) of
        sum(seed * rho_T) for the run recorded by ``adjoint_forward``.  A CUDA float64 seed
        gives CUDA gradients, with no host transfer (bfvmb_adjoint_backward_device)."""
        if hasattr(seed, "is_cuda") and seed.is_cuda:
            import torch
            na = getattr(self, "_nagents", 0)
            p = self._device_f64(seed, "adjoint_backward")
            g0 = torch.empty(tuple(self.shape), dtype=torch.float64, device=seed.device)
            g1 = torch.empty((na, self.nsub), dtype=torch.float64, device=seed.device)
            g2 = torch.empty((na, self.nsub), dtype=torch.float64, device=seed.device)
            torch.cuda.current_stream().synchronize()
            ptr = (lambda t: ctypes.c_void_p(t.data_ptr()) if t.numel() else None)
            self._check(self._lib.bfvmb_adjoint_backward_device(self._h, p, ptr(g0), ptr(g1), ptr(g2)),
                        "adjoint_backward_device")
            self.sync()
            return g0, g1, g2
        w = np.ascontiguousarray(seed, dtype=np.float64)
        assert w.size == np.prod(self.shape)
        na = getattr(self, "_nagents", 0)
        g0 = np.empty(self.shape, dtype=np.float64)
        g1 = np.empty((na, self.nsub), dtype=np.float64)
        g2 = np.empty((na, self.nsub), dtype=np.float64)
        self._check(self._lib.bfvmb_adjoint_backward(self._h, L.dptr(w), L.dptr(g0), L.dptr(g1) if na else None,
                                                     L.dptr(g2) if na else None), "adjoint_backward")
        return g0, g1, g2

    def autotune(self, reps: int = 3) -> np.ndarray:
        """Pick each sweep's launch shape by timing; returns [[wavefronts/CU, ring, us]] x 3
        (for the x sweep, "ring" is the kernel variant, see bfvmb_set_launch)."""
        out = np.zeros(9, dtype=np.int32)
        self._check(self._lib.bfvmb_autotune(self._h, reps, L.iptr(out)), "autotune")
        return out.reshape(3, 3)

    def launch_phase(self, phase: int, count: int = 1) -> None:
        self._check(self._lib.bfvmb_launch_phase(self._h, phase, count), "launch_phase")

    def sync(self) -> None:
        self._check(self._lib.bfvmb_sync(self._h), "sync")

    def device_ptr(self) -> int:
        p = ctypes.c_void_p()
        self._check(self._lib.bfvmb_density_device_ptr(self._h, ctypes.byref(p)), "density_device_ptr")
        return int(p.value or 0)


class Batch:
    """Many independent microenvironments of uneven sizes advanced together
    (``bfvmb_batch_*``): one kernel launch per phase covers every member, and
    ``run(..., graph=True)`` replays the whole run as one captured HIP graph."""

    def __init__(self, shapes, dx, precision: str = "fp64", device: int = 0):
        self._lib = L.load()
        shapes = [tuple(int(v) for v in s) for s in shapes]
        n = len(shapes)
        self.shapes = shapes
        self.precision = precision
        nx, ny, nz, ns = (np.ascontiguousarray([s[i] for s in shapes], dtype=np.int32) for i in range(4))
        dxa = np.ascontiguousarray(np.broadcast_to(dx, (n,)), dtype=np.float64)
        h = ctypes.c_void_p()
        rc = self._lib.bfvmb_batch_create(ctypes.byref(h), n, L.iptr(nx), L.iptr(ny), L.iptr(nz), L.iptr(ns),
                                          L.dptr(dxa), PRECISIONS[precision], device)
        if rc != L.OK:
            raise L.BioFVMBError(f"bfvmb_batch_create failed ({rc}): {self._lib.bfvmb_batch_last_error(None).decode()}")
        self._h = h
        self.dx = dxa

    def _check(self, rc, what):
        if rc != L.OK:
            raise L.BioFVMBError(f"{what} failed ({rc}): {self._lib.bfvmb_batch_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.bfvmb_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_ptr):
        self._check(self._lib.bfvmb_batch_set_stream(self._h, stream_ptr), "set_stream")

    def set_coefficients(self, i, diffusion, decay, dt):
        D = np.ascontiguousarray(diffusion, dtype=np.float64)
        lam = np.ascontiguousarray(decay, dtype=np.float64)
        self._check(self._lib.bfvmb_batch_set_coefficients(self._h, i, L.dptr(D), L.dptr(lam), float(dt)),
                    "set_coefficients")

    def update_agent_positions(self, i, pos, n, bbox_min=(0.0, 0.0, 0.0)):
        """bfvmb_batch_update_agent_positions for member i (n = length of its agent list)."""
        ptr, dev, keep = _positions(pos, n)
        lo = np.ascontiguousarray(bbox_min, dtype=np.float64)
        self._check(self._lib.bfvmb_batch_update_agent_positions(self._h, i, n, ptr, L.dptr(lo), dev),
                    "batch_update_agent_positions")
        del keep

    def assign_agent_positions(self, i, pos, n, bbox_min=(0.0, 0.0, 0.0)):
        """bfvmb_batch_assign_agent_positions for member i (n = length of its agent list)."""
        ptr, dev, keep = _positions(pos, n)
        lo = np.ascontiguousarray(bbox_min, dtype=np.float64)
        self._check(self._lib.bfvmb_batch_assign_agent_positions(self._h, i, n, ptr, L.dptr(lo), dev),
                    "batch_assign_agent_positions")
        del keep

    def set_bulk_rates(self, i, supply=None, target=None, uptake=None):
        keep = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(self.shapes[i])
                for a in (supply, target, uptake)]
        self._check(self._lib.bfvmb_batch_set_bulk_rates(self._h, i, *[None if k is None else L.dptr(k)
                                                                          for k in keep]), "set_bulk_rates")

    def set_dirichlet(self, i, value, active):
        ns = self.shapes[i][3]
        v = np.ascontiguousarray(np.broadcast_to(value, (ns,)), dtype=np.float64)
        a = np.ascontiguousarray(np.broadcast_to(active, (ns,)), dtype=np.uint8)
        self._check(self._lib.bfvmb_batch_set_dirichlet(self._h, i, L.dptr(v), L.bptr(a)), "set_dirichlet")

    def set_dirichlet_nodes(self, i, voxel, values):
        ns = self.shapes[i][3]
        v = np.ascontiguousarray(voxel, dtype=np.int32).reshape(-1)
        val = np.ascontiguousarray(values, dtype=np.float64).reshape(v.size, ns)
        self._check(self._lib.bfvmb_batch_set_dirichlet_nodes(self._h, i, v.size, L.iptr(v), L.dptr(val)),
                    "set_dirichlet_nodes")

    def upload(self, i, rho):
        r = np.ascontiguousarray(rho, dtype=np.float64)
        assert r.size == np.prod(self.shapes[i])
        self._check(self._lib.bfvmb_batch_upload_density(self._h, i, L.dptr(r)), "upload_density")

    def download(self, i):
        out = np.empty(self.shapes[i], dtype=np.float64)
        self._check(self._lib.bfvmb_batch_download_density(self._h, i, L.dptr(out)), "download_density")
        return out

    def set_agents(self, i, voxel, temp1, temp2, active=None):
        v = np.ascontiguousarray(voxel, dtype=np.int32)
        ns = self.shapes[i][3]
        t1 = np.ascontiguousarray(temp1, dtype=np.float64).reshape(v.size, ns)
        t2 = np.ascontiguousarray(temp2, dtype=np.float64).reshape(v.size, ns)
        a = None if active is None else np.ascontiguousarray(active, dtype=np.uint8)
        self._check(self._lib.bfvmb_batch_set_agents(self._h, i, v.size, L.iptr(v), L.dptr(t1), L.dptr(t2),
                                                     None if a is None else L.bptr(a)), "set_agents")

    def run(self, nsteps, graph=True):
        self._check(self._lib.bfvmb_batch_run(self._h, nsteps, 1 if graph else 0), "run")

    def sync(self):
        self._check(self._lib.bfvmb_batch_sync(self._h), "sync")

    def set_launch(self, sweep=-1, grid=0, lds_blocks_per_cu=0, ring=0):
        self._check(self._lib.bfvmb_batch_set_launch(self._h, sweep, grid, lds_blocks_per_cu, ring), "set_launch")

    def autotune(self, reps=3):
        """bfvmb_batch_autotune: [[wavefronts/CU, ring or x variant, us]] x 3."""
        out = np.zeros(9, dtype=np.int32)
        self._check(self._lib.bfvmb_batch_autotune(self._h, reps, L.iptr(out)), "autotune")
        return out.reshape(3, 3)

    def launch_phase(self, phase, count=1):
        self._check(self._lib.bfvmb_batch_launch_phase(self._h, phase, count), "launch_phase")
