"""The HIP library against the reference's own outputs (tests/golden/, written by the
compiled reference): FP64 must be bit-identical at every snapshot, agent mapping included."""
import os
import sys

import numpy as np
import pytest

import casefile as cf

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_golden as golden  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", golden.ALL_NAMES)
def test_fp64_matches_reference_fixture(name):
    from biofvmb import LODSolver
    case, ref = golden.load(name)
    print(f"[gpu] golden {name}", flush=True)
    value, active = cf.effective_dirichlet(case)
    s = LODSolver(case.nx, case.ny, case.nz, case.nsub, case.dx, "fp64")
    s.set_coefficients(case.D, case.decay, case.dt)
    if (np.asarray(case.dactive) >= 0).all():
        s.set_dirichlet(value, active)  # otherwise keep the library's defaults (100.0, active)
    s.upload(case.init)
    voxel = s.voxel_indices(case.pos, case.bbox[:3])
    voxel[voxel < 0] = 0
    assert np.array_equal(voxel, ref["voxel"])
    if case.with_agents:
        s.set_agent_rates(voxel, case.volume, case.S, case.U, case.T, case.dt)
    done = 0
    for step in sorted(ref["fields"]):
        s.run(step - done) if case.with_agents else s.simulate_diffusion_decay(step - done)
        done = step
        got = s.download()
        assert np.array_equal(got, ref["fields"][step]), f"{name} step {step}: {(got != ref['fields'][step]).sum()} differ"
    s.close()
