"""MAT v4 multiscale_microenvironment files: the oracle's writer reproduces the files the
reference itself wrote (Microenvironment::write_to_matlab, fixtures made by running it),
byte for byte, and its reader recovers the reference's field and grid."""
import os
import struct
import sys

import numpy as np
import pytest

import matlab

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_golden as golden  # noqa: E402


@pytest.mark.parametrize("name", golden.MAT_NAMES)
def test_writer_matches_reference_bytes(name):
    case, ref = golden.load(name)
    data = golden.load_mat(name)
    rho = ref["fields"][case.nsteps]
    assert matlab.to_bytes(rho, case.dx, (case.x0,) * 3) == data


@pytest.mark.parametrize("name", golden.MAT_NAMES)
def test_reader_recovers_field_and_grid(name):
    case, ref = golden.load(name)
    nm, m = matlab.parse(golden.load_mat(name))
    assert nm == "multiscale_microenvironment"
    assert m.shape == (case.nx * case.ny * case.nz, 4 + case.nsub)
    assert np.array_equal(m[:, 4:].reshape(ref["fields"][case.nsteps].shape), ref["fields"][case.nsteps])
    # centers: brute force over the mesh, reference voxel numbering x*ny*nz + y*nz + z
    v = 0
    for i in range(case.nx):
        for j in range(case.ny):
            for k in range(case.nz):
                c = m[v, :3]
                assert abs(c[0] - (case.x0 + (i + 0.5) * case.dx)) < 1e-9 * case.dx
                assert abs(c[1] - (case.x0 + (j + 0.5) * case.dx)) < 1e-9 * case.dx
                assert abs(c[2] - (case.x0 + (k + 0.5) * case.dx)) < 1e-9 * case.dx
                v += 1
    assert np.all(m[:, 3] == case.dx ** 3)


def test_header_layout():
    data = golden.load_mat("mat_1x1x2x1")
    assert struct.unpack_from("<5I", data, 0) == (0, 5, 2, 0, 27)
    assert data[20:47] == b"multiscale_microenvironment"
    assert len(data) == 47 + 2 * 5 * 8


def test_reader_rejects_other_formats():
    data = bytearray(golden.load_mat("mat_1x1x2x1"))
    for off, val in [(0, 10), (0, 1000), (12, 1)]:  # float data, big-endian, complex
        bad = bytearray(data)
        struct.pack_into("<I", bad, off, val)
        with pytest.raises(ValueError):
            matlab.parse(bytes(bad))
    with pytest.raises(ValueError):
        matlab.parse(bytes(data[:-8]))
