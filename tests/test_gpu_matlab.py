"""bfvmb_write_matlab / bfvmb_read_matlab on the GPU: the file written from the resident
field is byte-identical to the reference's own Microenvironment::write_to_matlab output
(fixtures written by the compiled reference), reading it back restores the field exactly,
and files that do not fit the context's grid are refused without touching the field."""
import os
import struct
import sys

import numpy as np
import pytest

import matlab
import oracle as orc

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_golden as golden  # noqa: E402

pytestmark = pytest.mark.gpu


def _solver(case, precision="fp64"):
    from biofvmb import LODSolver
    import casefile as cf
    s = LODSolver(case.nx, case.ny, case.nz, case.nsub, case.dx, precision)
    s.set_coefficients(case.D, case.decay, case.dt)
    s.set_dirichlet(*cf.effective_dirichlet(case))
    s.upload(case.init)
    voxel, t1, t2 = orc.agent_tables(case)
    s.set_agents(voxel, t1, t2)
    return s


@pytest.mark.parametrize("name", golden.MAT_NAMES)
def test_written_file_equals_reference_file(name, tmp_path):
    case, ref = golden.load(name)
    print(f"[gpu] MAT write {name}", flush=True)
    s = _solver(case)
    s.run(case.nsteps)
    path = tmp_path / "out.mat"
    s.write_matlab(path, (case.x0,) * 3)
    assert path.read_bytes() == golden.load_mat(name)
    s.close()


@pytest.mark.parametrize("name", golden.MAT_NAMES)
def test_read_reference_file_restores_field(name, tmp_path):
    case, ref = golden.load(name)
    path = tmp_path / "ref.mat"
    path.write_bytes(golden.load_mat(name))
    s = _solver(case)
    s.read_matlab(path, (case.x0,) * 3)
    assert np.array_equal(s.download(), ref["fields"][case.nsteps])
    # stepping continues from the loaded field exactly as the reference would
    s.run(2)
    assert np.array_equal(s.download(), orc.run(case, nsteps=2, rho=ref["fields"][case.nsteps]))
    s.close()


def test_fp32_field_written_widened(tmp_path):
    case, ref = golden.load("mat_offset_13x9x11x3")
    s = _solver(case, "fp32")
    s.run(case.nsteps)
    got = s.download()
    path = tmp_path / "f.mat"
    s.write_matlab(path, (case.x0,) * 3)
    nm, m = matlab.read(path)
    want = matlab.matrix(ref["fields"][case.nsteps], case.dx, (case.x0,) * 3)
    assert np.array_equal(m[:, :4], want[:, :4])  # grid columns exact
    assert np.array_equal(m[:, 4:].reshape(got.shape), got)  # densities = the float field, widened
    rel = np.max(np.abs(m[:, 4:] - want[:, 4:]) / np.abs(want[:, 4:]))
    assert rel <= 1e-5
    # and read back: float(double(float)) is the identity
    s2 = _solver(case, "fp32")
    s2.read_matlab(path, (case.x0,) * 3)
    assert np.array_equal(s2.download(), got)
    s.close()
    s2.close()


def test_refuses_mismatched_files(tmp_path):
    from biofvmb import BioFVMBError
    case, ref = golden.load("mat_offset_13x9x11x3")
    good = golden.load_mat("mat_offset_13x9x11x3")
    s = _solver(case)
    before = s.download()
    cases = {
        "origin": (good, (0.0, 0.0, 0.0)),  # centers off the grid
        "type": (struct.pack("<I", 10) + good[4:], (case.x0,) * 3),  # float32 data
        "rows": (good[:4] + struct.pack("<I", 6) + good[8:], (case.x0,) * 3),
        "truncated": (good[:-16], (case.x0,) * 3),
    }
    for what, (data, origin) in cases.items():
        p = tmp_path / f"{what}.mat"
        p.write_bytes(data)
        with pytest.raises(BioFVMBError):
            s.read_matlab(p, origin)
        assert np.array_equal(s.download(), before), what
    with pytest.raises(BioFVMBError):
        s.read_matlab(tmp_path / "missing.mat")
    with pytest.raises(BioFVMBError):
        s.write_matlab(tmp_path / "no_such_dir" / "x.mat")
    s.close()


def test_multi_slab_roundtrip(tmp_path):
    """150x180x180x4: the columns stream through several 64 MB slabs each way."""
    from biofvmb import LODSolver
    nx, ny, nz, ns, dx = 150, 180, 180, 4, 20.0
    rng = np.random.default_rng(3)
    rho = rng.lognormal(0.0, 0.5, (nx, ny, nz, ns))
    s = LODSolver(nx, ny, nz, ns, dx, "fp64")
    s.upload(rho)
    path = tmp_path / "big.mat"
    s.write_matlab(path, (-100.0, 40.0, 0.0))
    head = path.read_bytes()[:47]
    assert struct.unpack_from("<5I", head, 0) == (0, 8, nx * ny * nz, 0, 27)
    m = np.fromfile(path, dtype="<f8", offset=47).reshape(nx * ny * nz, 8)
    assert np.array_equal(m[:, 4:].reshape(rho.shape), rho)
    cx = matlab.centers(nx, -100.0, dx)
    cz = matlab.centers(nz, 0.0, dx)
    assert np.array_equal(m[:, 0].reshape(nx, ny, nz)[:, 0, 0], cx)
    assert np.array_equal(m[:, 2].reshape(nx, ny, nz)[0, 0, :], cz)
    del m
    s2 = LODSolver(nx, ny, nz, ns, dx, "fp64")
    s2.read_matlab(path, (-100.0, 40.0, 0.0))
    assert np.array_equal(s2.download(), rho)
    s.close()
    s2.close()
