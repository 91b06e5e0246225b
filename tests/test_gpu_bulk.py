"""Bulk sources and sinks on the GPU (bfvmb_set_bulk_rates, the bulk phase of bfvmb_run and
bfvmb_simulate_bulk_sources_and_sinks) against the reference's own
Microenvironment::simulate_bulk_sources_and_sinks outputs (fixtures written by the compiled
reference) and the oracle: FP64 bit for bit, FP32 within 1e-5 max relative error."""
import os
import sys

import numpy as np
import pytest

import casefile as cf
import oracle as orc

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_golden as golden  # noqa: E402

pytestmark = pytest.mark.gpu


def _solver(case, precision="fp64"):
    from biofvmb import LODSolver
    s = LODSolver(case.nx, case.ny, case.nz, case.nsub, case.dx, precision)
    s.set_coefficients(case.D, case.decay, case.dt)
    s.set_dirichlet(*cf.effective_dirichlet(case))
    s.upload(case.init)
    if case.with_agents and case.nagents:
        voxel, t1, t2 = orc.agent_tables(case)
        s.set_agents(voxel, t1, t2)
    s.set_bulk_rates(*case.bulk)
    return s


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("name", golden.BULK_NAMES)
def test_bulk_run_matches_reference(name, graph):
    case, ref = golden.load(name)
    print(f"[gpu] bulk {name} graph={graph}", flush=True)
    s = _solver(case)
    done = 0
    for step in sorted(ref["fields"]):
        s.run(step - done, graph=graph)
        done = step
        assert np.array_equal(s.download(), ref["fields"][step]), f"{name} step {step}"
    s.close()


@pytest.mark.parametrize("name", golden.BULK_NAMES)
def test_bulk_fp32_error(name):
    case, ref = golden.load(name)
    s = _solver(case, "fp32")
    s.run(case.nsteps)
    want = ref["fields"][case.nsteps]
    err = np.max(np.abs(s.download() - want) / np.abs(want))
    print(f"fp32 bulk {name}: max rel {err:.2e}")
    assert err <= 1e-5
    s.close()


def test_standalone_bulk_call_and_removal():
    case, _ = golden.load("bulk_15x12x10x3")
    s = _solver(case)
    s.simulate_bulk_sources_and_sinks(0.037)
    want = np.array(case.init)
    orc.bulk_sources_sinks(want, 0.037, *case.bulk)
    assert np.array_equal(s.download(), want)
    # without rates the step is the plain time loop again
    s.set_bulk_rates()
    s.run(2)
    c2 = golden.load("bulk_15x12x10x3")[0]
    c2.bulk = None
    assert np.array_equal(s.download(), orc.run(c2, nsteps=2, rho=want))
    s.close()


def test_batch_members_with_and_without_bulk():
    from biofvmb import Batch
    cases = [golden.load(n)[0] for n in golden.BULK_NAMES]
    plain = cf.make_case(11, 13, 9, 2, seed=40, nagents=30, nsteps=3)
    cases.append(plain)
    b = Batch([(c.nx, c.ny, c.nz, c.nsub) for c in cases], [c.dx for c in cases], "fp64")
    for i, c in enumerate(cases):
        b.set_coefficients(i, c.D, c.decay, c.dt)
        b.set_dirichlet(i, *cf.effective_dirichlet(c))
        b.upload(i, c.init)
        if c.with_agents and c.nagents:
            voxel, t1, t2 = orc.agent_tables(c)
            b.set_agents(i, voxel, t1, t2)
        if c.bulk is not None:
            b.set_bulk_rates(i, *c.bulk)
    b.run(3, graph=True)
    for i, c in enumerate(cases):
        assert np.array_equal(b.download(i), orc.run(c, nsteps=3)), f"member {i}"
    b.close()
