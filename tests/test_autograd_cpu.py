"""The libtorch operator library (torch.ops.biofvmb.*) without a GPU: it loads, agent_tables
reproduces the reference's temp1/temp2 bit for bit (the oracle's restatement of
Basic_Agent::set_internal_uptake_constants), its backward passes torch's gradcheck, and
lod_run fails loudly when no device is present."""
import os

import numpy as np
import pytest
import torch

import casefile as cf
import oracle as orc


@pytest.fixture(scope="module")
def ag():
    from biofvmb import autograd
    autograd.load()
    return autograd


def test_operators_registered(ag):
    assert hasattr(torch.ops.biofvmb, "lod_run")
    assert hasattr(torch.ops.biofvmb, "agent_tables")


def test_agent_tables_match_reference_bitwise(ag):
    c = cf.make_case(10, 9, 8, 3, seed=4, nagents=200)
    dV = (c.dx * c.dx) * c.dx
    r1, r2 = orc.agent_constants(c.dt, c.volume, dV, c.S, c.U, c.T)
    t1, t2 = ag.agent_tables(torch.from_numpy(c.volume), torch.from_numpy(c.S), torch.from_numpy(c.U),
                             torch.from_numpy(c.T), c.dt, dV)
    assert np.array_equal(t1.numpy(), r1) and np.array_equal(t2.numpy(), r2)


def test_agent_tables_gradcheck(ag):
    g = torch.Generator().manual_seed(0)
    n, ns = 7, 3
    vol = torch.rand(n, generator=g, dtype=torch.float64) * 500 + 100
    S, U, T = (torch.rand(n, ns, generator=g, dtype=torch.float64, requires_grad=True) * s for s in (5.0, 20.0, 38.0))
    S, U, T = (x.detach().requires_grad_() for x in (S, U, T))
    f = lambda s, u, t: ag.agent_tables(vol, s, u, t, 0.01, 8000.0)  # noqa: E731
    assert torch.autograd.gradcheck(f, (S, U, T), eps=1e-6, atol=1e-10, rtol=1e-7)


def test_lod_run_without_gpu_raises(ag):
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from biofvmb import BioFVMBError, LODSolver
    with pytest.raises(BioFVMBError):
        LODSolver(4, 4, 4, 1, 20.0)
    with pytest.raises(RuntimeError, match="null context"):
        torch.ops.biofvmb.lod_run(torch.zeros(4, 4, 4, 1, dtype=torch.float64), torch.zeros(0, dtype=torch.int32),
                                  torch.zeros(0, 1, dtype=torch.float64), torch.zeros(0, 1, dtype=torch.float64), 0, 1, 1)


def test_library_file_in_tree(ag):
    assert os.path.dirname(ag.TORCH_LIB_PATH).endswith(os.path.join("biofvm-b_amd", "lib"))
