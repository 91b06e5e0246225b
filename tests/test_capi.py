"""CPU checks of the boundary: the built library loads, exports every entry point that
include/biofvmb.h declares, and fails loudly (no CPU fallback) where no gfx950 GPU exists."""
import ctypes
import re

import numpy as np
import pytest

import biofvmb
from biofvmb import _lib as L


def _declared():
    text = open(biofvmb.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bfvmb_\w+)\s*\(", text)))


def test_header_declares_the_bound_signatures():
    assert _declared() == sorted(L.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = biofvmb.load()
    for name in _declared():
        assert hasattr(lib, name), name
    assert b"gfx950" in lib.bfvmb_version()


def test_create_without_gpu_reports_no_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = biofvmb.load()
    h = ctypes.c_void_p()
    rc = lib.bfvmb_create(ctypes.byref(h), 8, 8, 8, 1, 20.0, L.FP64, 0)
    assert rc == L.ERR_NODEV
    assert h.value is None
    assert lib.bfvmb_last_error(None)
    with pytest.raises(biofvmb.BioFVMBError):
        biofvmb.LODSolver(8, 8, 8, 1, 20.0)


def test_create_rejects_bad_arguments():
    lib = biofvmb.load()
    h = ctypes.c_void_p()
    assert lib.bfvmb_create(ctypes.byref(h), 0, 8, 8, 1, 20.0, L.FP64, 0) == L.ERR_ARG
    assert lib.bfvmb_create(ctypes.byref(h), 8, 8, 8, 1, -1.0, L.FP64, 0) == L.ERR_ARG
    assert lib.bfvmb_create(ctypes.byref(h), 8, 8, 8, 1, 20.0, 7, 0) == L.ERR_ARG


def test_null_context_is_an_argument_error():
    lib = biofvmb.load()
    assert lib.bfvmb_run(None, 1) == L.ERR_ARG
    assert lib.bfvmb_sync(None) == L.ERR_ARG
    assert lib.bfvmb_destroy(None) == L.OK
    v = np.zeros(1, np.int32)
    p = np.zeros(3)
    assert lib.bfvmb_voxel_indices(0, 4, 4, 20.0, L.dptr(p), 1, L.dptr(p), L.iptr(v)) == L.ERR_ARG


def test_voxel_mapping_matches_reference_fixtures():
    """Host-side agent -> voxel mapping against the reference's own assign_position
    results (out-of-box agents: -1 here, voxel 0 in the reference)."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_golden as golden
    from biofvmb.solver import voxel_indices
    for name in golden.ALL_NAMES:
        case, ref = golden.load(name)
        got = voxel_indices((case.nx, case.ny, case.nz), case.dx, case.pos, case.bbox[:3])
        assert np.array_equal(np.where(got < 0, 0, got), ref["voxel"]), name
