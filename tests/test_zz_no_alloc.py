"""bfvmb_densities_at_noalloc (host arrays staged in the library's static device scratch, no
device allocation) against bfvmb_densities_at on the same fields, bit for bit, called through
the C ABI: two calls in a row on one context (the second with fewer voxels, so stale scratch
contents would show), FP64 and FP32 contexts on the same GPU sharing the scratch, and the size
limit."""
import ctypes

import numpy as np
import pytest

import casefile as cf

pytestmark = pytest.mark.gpu


def _ctx(case, precision, steps):
    from biofvmb import LODSolver
    s = LODSolver(case.nx, case.ny, case.nz, case.nsub, case.dx, precision)
    s.set_coefficients(case.D, case.decay, case.dt)
    s.set_dirichlet(case.dvalue, case.dactive)
    s.upload(case.init)
    s.run(steps)
    return s


def _noalloc(s, voxel):
    from biofvmb import _lib as L
    v = np.ascontiguousarray(voxel, dtype=np.int32)
    out = np.full((v.size, s.nsub), np.nan)
    rc = L.load().bfvmb_densities_at_noalloc(ctypes.c_void_p(s.handle), v.size, L.iptr(v), L.dptr(out))
    return rc, out


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_gather_equals_allocating_gather(precision):
    case = cf.make_case(31, 22, 27, 3, seed=51, nagents=0, nsteps=3, with_agents=False)
    print(f"[gpu] densities_at_noalloc {precision}", flush=True)
    s = _ctx(case, precision, 3)
    other = _ctx(cf.make_case(9, 13, 7, 5, seed=52, nagents=0, nsteps=2, with_agents=False), precision, 2)
    rng = np.random.default_rng(3)
    nvox = case.nx * case.ny * case.nz
    for n in (5000, 37):  # the second call gathers fewer voxels
        voxel = rng.integers(0, nvox, n).astype(np.int32)
        rc, got = _noalloc(s, voxel)
        assert rc == 0
        assert np.array_equal(got, s.densities_at(voxel))
        # another context on the same GPU between the calls shares the scratch
        ov = rng.integers(0, 9 * 13 * 7, 211).astype(np.int32)
        rc, og = _noalloc(other, ov)
        assert rc == 0 and np.array_equal(og, other.densities_at(ov))
    s.close()
    other.close()


def test_noalloc_rejects_requests_beyond_the_scratch():
    from biofvmb import _lib as L
    case = cf.make_case(40, 40, 40, 4, seed=53, nagents=0, nsteps=1, with_agents=False)
    print("[gpu] densities_at_noalloc size limit", flush=True)
    s = _ctx(case, "fp64", 1)
    limit = 16 << 20
    n = limit // (4 + 8 * 4) + 1000  # more voxels than the scratch holds
    voxel = np.zeros(n, np.int32)
    rc, _ = _noalloc(s, voxel)
    assert rc == L.ERR_ARG
    assert b"scratch" in L.load().bfvmb_last_error(ctypes.c_void_p(s.handle))
    n_ok = (limit - 256) // (4 + 8 * 4) - 64  # fits
    voxel = np.arange(n_ok, dtype=np.int32) % (40 ** 3)
    rc, got = _noalloc(s, voxel)
    assert rc == 0 and np.array_equal(got, s.densities_at(voxel))
    rc, _ = _noalloc(s, np.array([40 ** 3], np.int32))  # out of the grid
    assert rc == L.ERR_ARG
    s.close()
