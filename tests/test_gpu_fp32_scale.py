"""FP32 storage path at a large grid (256^3 x 8 substrates, log-spaced D and decay, Dirichlet
on substrates 0-1, 200k agents in a centred sphere), against the FP64 path, which is bit for
bit the reference's arithmetic (every FP64 test checks that).  Reports max and RMS relative
error; the bound is 1e-5 max relative error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver(precision, n, ns, init, vox, rates, dt=0.01, dx=20.0):
    from biofvmb import LODSolver
    s = LODSolver(n, n, n, ns, dx, precision)
    s.set_coefficients(np.logspace(2, 5, ns), np.logspace(-3, 0, ns), dt)
    s.set_dirichlet(np.full(ns, 38.0), (np.arange(ns) < 2).astype(np.uint8))
    s.upload(init)
    s.set_agent_rates(vox, *rates, dt)
    return s


def test_fp32_vs_fp64_at_256cube_x8():
    n, ns, na, steps = 256, 8, 200_000, 10
    rng = np.random.default_rng(11)
    init = rng.lognormal(0.0, 0.5, (n, n, n, ns))
    d = rng.normal(size=(na, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos = n * 20.0 / 2 + d * (1600.0 * rng.uniform(0, 1, na) ** (1 / 3))[:, None]
    rates = (np.full(na, 4.0 / 3.0 * np.pi * 125.0), rng.uniform(0, 5, (na, ns)), rng.uniform(0, 20, (na, ns)),
             np.ones((na, ns)))
    out = {}
    for prec in ("fp64", "fp32"):
        print(f"[gpu] 256^3x8 {prec}", flush=True)
        from biofvmb.solver import voxel_indices
        vox = voxel_indices((n, n, n), 20.0, pos)
        vox[vox < 0] = 0
        s = _solver(prec, n, ns, init, vox, rates)
        s.run(steps, graph=True)
        out[prec] = s.download()
        s.close()
    rel = np.abs(out["fp32"] - out["fp64"]) / np.abs(out["fp64"])
    print(f"fp32 vs fp64 at 256^3x8, {steps} steps: max rel {rel.max():.3e}, rms rel {np.sqrt(np.mean(rel ** 2)):.3e}")
    assert rel.max() <= 1e-5
