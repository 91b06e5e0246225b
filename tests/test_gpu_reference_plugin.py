"""Drop-in check: the reference's own library and driver (compiled from its sources by
oracle/Makefile, target ref-hip) with bfvmb_hip_LOD_3D in the diffusion_decay_solver slot
and bfvmb_hip_cell_sources_and_sinks in place of simulate_cell_sources_and_sinks
(biofvm-b_amd/host/bfvmb_reference_plugin.cpp) must reproduce the reference solver's
outputs bit for bit, both with the field synced every call and kept resident."""
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
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver_hip")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="reference build with the plugin (make -C oracle ref-hip) absent")
@pytest.mark.parametrize("mode", ["hip", "hipres", "hipall"])
@pytest.mark.parametrize("name", golden.ALL_NAMES)
def test_reference_driver_with_gpu_slot(name, mode):
    case, ref = golden.load(name)
    print(f"[gpu] reference driver + plugin {name} {mode}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        cpath, opath = os.path.join(d, "case.bin"), os.path.join(d, "out.bin")
        case.write(cpath)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([DRIVER, cpath, opath, mode, "1"], check=True, timeout=120, env=env)
        out = cf.read_output(opath, case)
    assert np.array_equal(out["voxel"], ref["voxel"])
    for step, f in ref["fields"].items():
        assert np.array_equal(out["fields"][step], f), f"{name} step {step}: {(out['fields'][step] != f).sum()} differ"


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="reference build with the plugin (make -C oracle ref-hip) absent")
@pytest.mark.parametrize("mode", ["hip", "hipres", "hipall"])
@pytest.mark.parametrize("name", ["aniso_17x23x31x3", "multi_24x20x28x5", "long_x_300x17x19x1", "dnodes_17x23x31x3"])
def test_reference_driver_autotuned(name, mode):
    """bfvmb_plugin_autotune before the first step: that step's solver call times the launch
    shapes on the live field (with a cell update waiting, in the resident modes) and restores
    it; every step then runs at the chosen shapes and stays bit for bit the reference's."""
    case, ref = golden.load(name)
    print(f"[gpu] reference driver + plugin, autotuned, {name} {mode}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        cpath, opath, npath = (os.path.join(d, f) for f in ("case.bin", "out.bin", "nodes.bin"))
        case.write(cpath)
        extra = ["--autotune"]
        if case.dnodes is not None:
            case.write_dnodes(npath)
            extra += ["--dnodes", npath]
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([DRIVER, cpath, opath, mode, "1"] + extra, check=True, timeout=120, env=env)
        out = cf.read_output(opath, case)
    for step, f in ref["fields"].items():
        assert np.array_equal(out["fields"][step], f), f"{name} step {step}: {(out['fields'][step] != f).sum()} differ"


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="reference build with the plugin (make -C oracle ref-hip) absent")
@pytest.mark.parametrize("mode", ["hip", "hipres", "hipall"])
@pytest.mark.parametrize("name", ["aniso_17x23x31x3", "multi_24x20x28x5", "dnodes_17x23x31x3", "plain_50cube"])
def test_reference_driver_fp32_storage(name, mode):
    """bfvmb_plugin_set_precision(M, BFVMB_FP32) in the reference's own driver: the field is
    stored in FP32 on the device and every snapshot stays within 1e-5 (max relative error) of
    the reference's FP64 solver."""
    case, ref = golden.load(name)
    print(f"[gpu] reference driver + plugin, FP32 storage, {name} {mode}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        cpath, opath, npath = (os.path.join(d, f) for f in ("case.bin", "out.bin", "nodes.bin"))
        case.write(cpath)
        extra = ["--fp32"]
        if case.dnodes is not None:
            case.write_dnodes(npath)
            extra += ["--dnodes", npath]
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([DRIVER, cpath, opath, mode, "1"] + extra, check=True, timeout=120, env=env)
        out = cf.read_output(opath, case)
    for step, f in ref["fields"].items():
        r = np.abs(out["fields"][step] - f) / np.maximum(np.abs(f), 1e-300)
        print(f"{name} {mode} step {step}: FP32 max rel {r.max():.3e}, RMS {np.sqrt(np.mean(r * r)):.3e}")
        assert r.max() <= 1e-5, f"{name} step {step}"
    assert not np.array_equal(out["fields"][max(ref["fields"])], ref["fields"][max(ref["fields"])])  # FP32 ran


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="reference build with the plugin (make -C oracle ref-hip) absent")
@pytest.mark.parametrize("mode", ["hip", "hipres", "hipall"])
@pytest.mark.parametrize("name", golden.BULK_NAMES)
def test_reference_driver_fp32_storage_bulk(name, mode):
    """FP32 storage with the bulk slot (bfvmb_hip_bulk_sources_and_sinks) in the reference's
    driver: within 1e-5 of the reference's FP64 run."""
    case, ref = golden.load(name)
    print(f"[gpu] reference driver + plugin, FP32 storage, bulk {name} {mode}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        cpath, opath, bpath = (os.path.join(d, f) for f in ("case.bin", "out.bin", "bulk.bin"))
        case.write(cpath)
        case.write_bulk(bpath)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([DRIVER, cpath, opath, mode, "1", "--bulk", bpath, "--fp32"], check=True, timeout=120, env=env)
        out = cf.read_output(opath, case)
    for step, f in ref["fields"].items():
        r = np.abs(out["fields"][step] - f) / np.maximum(np.abs(f), 1e-300)
        print(f"{name} {mode} step {step}: FP32 max rel {r.max():.3e}")
        assert r.max() <= 1e-5, f"{name} step {step}"


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="reference build with the plugin (make -C oracle ref-hip) absent")
@pytest.mark.parametrize("mode", ["hip", "hipres", "hipall"])
@pytest.mark.parametrize("name", golden.MAT_NAMES)
def test_reference_driver_mat_file(name, mode):
    """bfvmb_plugin_write_to_matlab in the reference's own driver (from the device when the
    field is resident) writes the reference's MAT file byte for byte."""
    case, ref = golden.load(name)
    print(f"[gpu] reference driver + plugin MAT {name} {mode}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        cpath, opath, mpath = (os.path.join(d, f) for f in ("case.bin", "out.bin", "out.mat"))
        case.write(cpath)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([DRIVER, cpath, opath, mode, "1", "--mat", mpath], check=True, timeout=120, env=env)
        with open(mpath, "rb") as f:
            assert f.read() == golden.load_mat(name)


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="reference build with the plugin (make -C oracle ref-hip) absent")
@pytest.mark.parametrize("mode", ["hip", "hipres", "hipall"])
@pytest.mark.parametrize("name", golden.BULK_NAMES)
def test_reference_driver_bulk(name, mode):
    """bfvmb_hip_bulk_sources_and_sinks in place of simulate_bulk_sources_and_sinks in the
    reference's own driver, with the rates coming from its bulk_*_function slots."""
    case, ref = golden.load(name)
    print(f"[gpu] reference driver + plugin bulk {name} {mode}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        cpath, opath, bpath = (os.path.join(d, f) for f in ("case.bin", "out.bin", "bulk.bin"))
        case.write(cpath)
        case.write_bulk(bpath)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([DRIVER, cpath, opath, mode, "1", "--bulk", bpath], check=True, timeout=120, env=env)
        out = cf.read_output(opath, case)
    for step, f in ref["fields"].items():
        assert np.array_equal(out["fields"][step], f), f"{name} step {step}"


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="reference build with the plugin (make -C oracle ref-hip) absent")
@pytest.mark.parametrize("mode", ["hipres", "hipall"])
@pytest.mark.parametrize("name", ["aniso_17x23x31x3", "multi_24x20x28x5"])
def test_reference_driver_field_right_after_the_cell_update(name, mode):
    """With the field resident the plugin holds each cell update back until the solver call;
    bfvmb_plugin_sync in between runs it first.  The reference's own driver reads
    p_density_vectors right after every step's cell update: the GPU slots give the
    reference's CPU path's fields there, bit for bit, and the same final fields."""
    case, ref = golden.load(name)
    print(f"[gpu] reference driver + plugin, field after the cell update {name} {mode}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        cpath = os.path.join(d, "case.bin")
        case.write(cpath)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        cells = {}
        for m in ("lod", mode):
            opath, kpath = os.path.join(d, f"out_{m}.bin"), os.path.join(d, f"cells_{m}.bin")
            subprocess.run([DRIVER, cpath, opath, m, "1", "--cells-out", kpath], check=True, timeout=120, env=env)
            cells[m] = np.fromfile(kpath, dtype="<f8")
            out = cf.read_output(opath, case)
            for step, f in ref["fields"].items():
                assert np.array_equal(out["fields"][step], f), f"{name} {m} step {step}"
    assert cells["lod"].size == case.nsteps * case.init.size
    assert np.array_equal(cells[mode], cells["lod"])
