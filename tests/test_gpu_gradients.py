"""Density gradients on the GPU (bfvmb_compute_gradients / download / gradients_at) against
the reference's own compute_all_gradient_vectors output (fixtures written by the compiled
reference) and the oracle: FP64 bit for bit; FP32 storage within 1e-5 of FP64 relative to
the gradient's scale."""
import os
import sys

import numpy as np
import pytest

import oracle as orc

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_golden as golden  # noqa: E402

pytestmark = pytest.mark.gpu


def _solver(shape, dx, rho, precision="fp64"):
    from biofvmb import LODSolver
    s = LODSolver(*shape, dx, precision)
    s.upload(rho)
    return s


@pytest.mark.parametrize("name", golden.GRAD_NAMES)
def test_gradients_match_reference(name):
    case, ref = golden.load(name)
    print(f"[gpu] gradients {name}", flush=True)
    rho = ref["fields"][case.nsteps]
    s = _solver(rho.shape, case.dx, rho)
    s.compute_gradients()
    assert np.array_equal(s.download_gradients(), ref["grad"])
    # at chosen voxels (corners, edges, interior), reference numbering
    nx, ny, nz, _ = rho.shape
    pts = [(0, 0, 0), (nx - 1, ny - 1, nz - 1), (nx // 2, ny // 2, nz // 2), (0, ny - 1, nz // 2)]
    vox = [(x * ny + y) * nz + z for x, y, z in pts]
    got = s.gradients_at(vox)
    for k, (x, y, z) in enumerate(pts):
        assert np.array_equal(got[k], ref["grad"][x, y, z])
    s.close()


@pytest.mark.parametrize("shape", [(40, 33, 27, 3), (1, 12, 9, 2), (5, 1, 1, 1), (64, 64, 1, 4)])
def test_gradients_match_oracle(shape):
    rng = np.random.default_rng(sum(shape))
    rho = rng.lognormal(0.0, 0.5, shape)
    s = _solver(shape, 12.5, rho)
    s.compute_gradients()
    want = orc.gradients(rho, 12.5)
    assert np.array_equal(s.download_gradients(), want)
    vox = rng.integers(0, np.prod(shape[:3]), 50)
    got = s.gradients_at(vox)
    assert np.array_equal(got, want.reshape(-1, shape[3], 3)[vox])
    s.close()


def test_gradients_fp32():
    case, ref = golden.load("grad_19x14x9x2")
    rho = ref["fields"][case.nsteps]
    s = _solver(rho.shape, case.dx, rho, "fp32")
    s.compute_gradients()
    got = s.download_gradients()
    scale = np.abs(ref["grad"]).max(axis=(0, 1, 2, 4), keepdims=True)
    err = np.max(np.abs(got - ref["grad"]) / scale)
    print(f"fp32 gradients: max error / scale {err:.2e}")
    assert err <= 1e-5
    s.close()


def test_gradients_state_errors():
    from biofvmb import BioFVMBError
    rho = np.ones((4, 4, 4, 1))
    s = _solver(rho.shape, 10.0, rho)
    with pytest.raises(BioFVMBError):
        s.download_gradients()  # nothing computed yet
    with pytest.raises(BioFVMBError):
        s.gradients_at([64])  # voxel out of range
    s.close()
