"""The C++ host layer (biofvm-b_amd/host: Microenvironment / Basic_Agent over the C ABI),
run as a driver on the reference's own case files, against the reference's outputs:
agent mapping, source/sink constants and every snapshot must match bit for bit."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import casefile as cf

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_golden as golden  # noqa: E402

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "biofvm-b_amd", "lib", "bfvmb_driver")


@pytest.mark.parametrize("mode", ["per_call", "fused"])
@pytest.mark.parametrize("name", golden.ALL_NAMES)
def test_host_driver_matches_reference(name, mode):
    case, ref = golden.load(name)
    print(f"[gpu] host driver {name} {mode}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        cpath, opath = os.path.join(d, "case.bin"), os.path.join(d, "out.bin")
        case.write(cpath)
        args = [DRIVER, cpath, opath, "fp64"] + (["--fused"] if mode == "fused" else [])
        subprocess.run(args, check=True, timeout=120)
        out = cf.read_output(opath, case)
    assert np.array_equal(out["voxel"], ref["voxel"])
    assert np.array_equal(out["temp1"], ref["temp1"])
    assert np.array_equal(out["temp2"], ref["temp2"])
    for step, f in ref["fields"].items():
        assert np.array_equal(out["fields"][step], f), f"{name} step {step}: {(out['fields'][step] != f).sum()} differ"


@pytest.mark.parametrize("mode", ["per_call", "fused"])
@pytest.mark.parametrize("name", ["aniso_17x23x31x3", "multi_24x20x28x5", "long_x_300x17x19x1"])
def test_host_driver_autotuned(name, mode):
    """Microenvironment::autotune() before the first step (timed at the first solver call, on
    the live field, with a cell update waiting in per-call mode): bit for bit the reference."""
    case, ref = golden.load(name)
    print(f"[gpu] host driver autotuned {name} {mode}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        cpath, opath = os.path.join(d, "case.bin"), os.path.join(d, "out.bin")
        case.write(cpath)
        args = [DRIVER, cpath, opath, "fp64", "--autotune"] + (["--fused"] if mode == "fused" else [])
        subprocess.run(args, check=True, timeout=120)
        out = cf.read_output(opath, case)
    for step, f in ref["fields"].items():
        assert np.array_equal(out["fields"][step], f), f"{name} step {step}: {(out['fields'][step] != f).sum()} differ"


@pytest.mark.parametrize("mode", ["per_call", "fused"])
@pytest.mark.parametrize("name", golden.MAT_NAMES)
def test_host_driver_writes_reference_mat_file(name, mode):
    """Microenvironment::write_to_matlab of the host layer (written from the device) equals
    the file the reference wrote after the same run, byte for byte."""
    case, ref = golden.load(name)
    print(f"[gpu] host driver MAT {name} {mode}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        cpath, opath, mpath = (os.path.join(d, f) for f in ("case.bin", "out.bin", "out.mat"))
        case.write(cpath)
        args = [DRIVER, cpath, opath, "fp64", "--mat", mpath] + (["--fused"] if mode == "fused" else [])
        subprocess.run(args, check=True, timeout=120)
        with open(mpath, "rb") as f:
            assert f.read() == golden.load_mat(name)


@pytest.mark.parametrize("mode", ["per_call", "fused"])
@pytest.mark.parametrize("name", golden.BULK_NAMES)
def test_host_driver_bulk_matches_reference(name, mode):
    """The host layer's bulk_*_function slots + simulate_bulk_sources_and_sinks (per call)
    or update_rates + run (fused) against the reference's own bulk outputs."""
    case, ref = golden.load(name)
    print(f"[gpu] host driver bulk {name} {mode}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        cpath, opath, bpath = (os.path.join(d, f) for f in ("case.bin", "out.bin", "bulk.bin"))
        case.write(cpath)
        case.write_bulk(bpath)
        args = [DRIVER, cpath, opath, "fp64", "--bulk", bpath] + (["--fused"] if mode == "fused" else [])
        subprocess.run(args, check=True, timeout=120)
        out = cf.read_output(opath, case)
    for step, f in ref["fields"].items():
        assert np.array_equal(out["fields"][step], f), f"{name} step {step}"
