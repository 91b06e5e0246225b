"""bfvmb_densities_at: the densities nearest_density_vector reads (BioFVM_microenvironment.cpp:
984-992), gathered from the resident field, equal to the downloaded field at those voxels;
host and device index lists; the C++ host layer's nearest_density_vectors through the driver."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import casefile as cf
import oracle as orc

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_golden as golden  # noqa: E402

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "biofvm-b_amd", "lib", "bfvmb_driver")


def _run(precision):
    from biofvmb import LODSolver
    case = cf.make_case(31, 17, 23, 3, seed=12, nagents=500, nsteps=3, shared_voxels=True)
    s = LODSolver(case.nx, case.ny, case.nz, case.nsub, case.dx, precision)
    s.set_coefficients(case.D, case.decay, case.dt)
    s.set_dirichlet(case.dvalue, case.dactive)
    s.upload(case.init)
    s.set_agent_rates(orc.voxel_indices(case), case.volume, case.S, case.U, case.T, case.dt)
    s.run(case.nsteps)
    return case, s


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_gather_equals_downloaded_field(precision):
    case, s = _run(precision)
    print(f"[gpu] densities_at {precision}", flush=True)
    nvox = case.nx * case.ny * case.nz
    vox = np.concatenate([[0, nvox - 1, 5, 5], np.random.default_rng(1).integers(0, nvox, 1000)]).astype(np.int32)
    got = s.densities_at(vox)
    field = s.download().reshape(nvox, case.nsub)
    s.close()
    assert np.array_equal(got, field[vox])


def test_device_lists_and_out_of_range():
    import torch
    from biofvmb import BioFVMBError
    case, s = _run("fp64")
    print("[gpu] densities_at device lists", flush=True)
    nvox = case.nx * case.ny * case.nz
    vox = np.random.default_rng(2).integers(0, nvox, 700).astype(np.int32)
    vox[3] = nvox  # outside the grid: NaN on the device path
    dv = torch.tensor(vox, dtype=torch.int32, device="cuda:0")
    out = s.densities_at(dv)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    field = s.download().reshape(nvox, case.nsub)
    ok = np.arange(vox.size) != 3
    assert np.array_equal(got[ok], field[vox[ok]]) and np.all(np.isnan(got[3]))
    with pytest.raises(BioFVMBError):
        s.densities_at(vox)  # host lists are checked
    s.close()


@pytest.mark.parametrize("name", ["aniso_17x23x31x3", "multi_24x20x28x5", "plain_50cube"])
def test_host_layer_samples_match_snapshots(name):
    """Microenvironment::nearest_density_vectors of the C++ host layer (read from the device
    after each run segment) against the field the same driver writes at that snapshot."""
    case, ref = golden.load(name)
    print(f"[gpu] host layer samples {name}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        cpath, opath, spath = (os.path.join(d, f) for f in ("case.bin", "out.bin", "samples.bin"))
        case.write(cpath)
        subprocess.run([DRIVER, cpath, opath, "fp64", "--sample", spath], check=True, timeout=120)
        out = cf.read_output(opath, case)
        samples = np.fromfile(spath, dtype="<f8")
    steps = sorted(out["fields"])
    samples = samples.reshape(len(steps), case.nagents, case.nsub)
    for i, step in enumerate(steps):
        field = out["fields"][step].reshape(-1, case.nsub)
        assert np.array_equal(samples[i], field[out["voxel"]]), f"{name} step {step}"
        assert np.array_equal(out["fields"][step], ref["fields"][step])


@pytest.mark.parametrize("name", ["aniso_17x23x31x3", "multi_24x20x28x5"])
def test_host_layer_densities_right_after_the_cell_update(name):
    """Per-call host layer: simulate_cell_sources_and_sinks leaves its update waiting for the
    next simulate_diffusion_decay; nearest_density_vectors in between runs it first.  The
    densities at the agents' voxels after every step's cell update are the oracle's, bit for
    bit, and the run still ends on the reference's field."""
    case, ref = golden.load(name)
    print(f"[gpu] host layer densities after the cell update {name}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        cpath, opath, spath = (os.path.join(d, f) for f in ("case.bin", "out.bin", "cells.bin"))
        case.write(cpath)
        subprocess.run([DRIVER, cpath, opath, "fp64", "--cells-sample", spath], check=True, timeout=120)
        out = cf.read_output(opath, case)
        samples = np.fromfile(spath, dtype="<f8").reshape(case.nsteps, case.nagents, case.nsub)
    voxel, t1, t2 = orc.agent_tables(case)
    active = np.ones(case.nagents, np.uint8)
    rho = case.init.copy()
    for k in range(case.nsteps):
        orc.sources_sinks(rho, np.maximum(voxel, 0), active, t1, t2)
        assert np.array_equal(samples[k], rho.reshape(-1, case.nsub)[out["voxel"]]), f"{name} step {k + 1}"
        rho = orc.run(case, nsteps=1, rho=rho, with_agents=False)
    last = max(ref["fields"])
    assert np.array_equal(out["fields"][last], ref["fields"][last])
