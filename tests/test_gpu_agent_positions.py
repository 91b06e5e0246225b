"""Moving agents: bfvmb_update_agent_positions (voxel mapping and regrouping on the device)
against the oracle's Basic_Agent::update_voxel_index semantics (BioFVM_basic_agent.cpp:137-153):
FP64 bit for bit, FP32 within 1e-5, host and device position buffers, agents leaving the box
(deactivated for good), shared voxels (list order), single contexts and batch members."""
import numpy as np
import pytest

import casefile as cf
import oracle as orc

pytestmark = pytest.mark.gpu


def _case():
    return cf.make_case(37, 29, 23, 3, seed=21, nagents=3000, nsteps=2, shared_voxels=True)


def _moves(case, seed):
    """Two later position sets: random displacements (some agents leave the box), a group
    piled into one voxel (list order), and agents coming back into the box."""
    rng = np.random.default_rng(seed)
    ext = np.array([case.nx, case.ny, case.nz]) * case.dx
    p1 = case.pos + rng.normal(0.0, 3.0 * case.dx, case.pos.shape)
    p1[100:140] = ext * np.array([0.5, 0.25, 0.75])  # shared voxel
    p1[200:230] = -5.0  # outside: deactivated
    p2 = p1 + rng.normal(0.0, 2.0 * case.dx, p1.shape)
    p2[200:215] = ext * 0.5  # back inside: the reference keeps them inactive
    p2[300:320] = ext  # upper corner, inside
    return [p1, p2]


def _oracle(case, moves, steps):
    """steps[k] steps (sources/sinks then LOD) after the k-th position set."""
    voxel, t1, t2 = orc.agent_tables(case)
    active = np.ones(case.nagents, np.uint8)
    rho = case.init.copy()
    for k, n in enumerate(steps):
        if k > 0:
            voxel, active = orc.update_voxel_indices(case, moves[k - 1], active)
        for _ in range(n):
            orc.sources_sinks(rho, np.maximum(voxel, 0), active, t1, t2)
            rho = orc.run(case, nsteps=1, rho=rho, with_agents=False)
    return rho, active


def _gpu(case, moves, steps, precision, device_positions=False):
    from biofvmb import LODSolver
    s = LODSolver(case.nx, case.ny, case.nz, case.nsub, case.dx, precision)
    s.set_coefficients(case.D, case.decay, case.dt)
    s.set_dirichlet(case.dvalue, case.dactive)
    s.upload(case.init)
    s.set_agent_rates(orc.voxel_indices(case), case.volume, case.S, case.U, case.T, case.dt)
    for k, n in enumerate(steps):
        if k > 0:
            pos = moves[k - 1]
            if device_positions:
                import torch
                pos = torch.tensor(pos, dtype=torch.float64, device="cuda:0").contiguous()
                torch.cuda.synchronize()
            s.update_agent_positions(pos)
        s.run(n)
    out = s.download()
    s.close()
    return out


@pytest.mark.parametrize("device_positions", [False, True])
def test_moved_agents_fp64_bitwise(device_positions):
    case = _case()
    moves = _moves(case, 3)
    steps = [2, 2, 3]
    print(f"[gpu] moving agents fp64 device_positions={device_positions}", flush=True)
    ref, active = _oracle(case, moves, steps)
    assert 0 < int(active.sum()) < case.nagents
    got = _gpu(case, moves, steps, "fp64", device_positions)
    assert np.array_equal(got, ref)


def test_moved_agents_fp32_error():
    case = _case()
    moves = _moves(case, 4)
    steps = [1, 2, 2]
    print("[gpu] moving agents fp32", flush=True)
    ref, _ = _oracle(case, moves, steps)
    got = _gpu(case, moves, steps, "fp32")
    r = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
    print(f"fp32 moving agents: max rel {r.max():.3e}")
    assert r.max() <= 1e-5


def test_unmoved_positions_keep_results():
    """Re-sending the initial positions regroups the same agents the same way."""
    case = _case()
    print("[gpu] moving agents: same positions", flush=True)
    ref = _gpu(case, [case.pos], [2, 2], "fp64")
    want = orc.run(case, nsteps=4)
    assert np.array_equal(ref, want)


def test_batch_member_moves():
    from biofvmb import Batch
    case = _case()
    moves = _moves(case, 5)
    other = cf.make_case(11, 9, 13, 2, seed=2, nagents=50, nsteps=1)
    print("[gpu] moving agents in a batch member", flush=True)
    ref, _ = _oracle(case, moves, [1, 2])
    b = Batch([(other.nx, other.ny, other.nz, other.nsub), (case.nx, case.ny, case.nz, case.nsub)], case.dx, "fp64")
    for i, c in enumerate((other, case)):
        b.set_coefficients(i, c.D, c.decay, c.dt)
        b.set_dirichlet(i, c.dvalue, c.dactive)
        b.upload(i, c.init)
        voxel, t1, t2 = orc.agent_tables(c)
        b.set_agents(i, voxel, t1, t2)
    b.run(1)
    b.update_agent_positions(1, moves[0], case.nagents)
    b.run(2)
    assert np.array_equal(b.download(1), ref)
    assert np.array_equal(b.download(0), orc.run(other, nsteps=3))
    b.close()


def test_wrong_agent_count_is_rejected():
    from biofvmb import LODSolver, BioFVMBError
    case = _case()
    s = LODSolver(case.nx, case.ny, case.nz, case.nsub, case.dx, "fp64")
    s.set_coefficients(case.D, case.decay, case.dt)
    s.set_agent_rates(orc.voxel_indices(case), case.volume, case.S, case.U, case.T, case.dt)
    s._nagents = case.nagents - 1
    with pytest.raises(BioFVMBError):
        s.update_agent_positions(case.pos[:-1])
    s.close()


def test_adjoint_after_moves_matches_fresh_agent_list():
    """Gradients after a device regrouping equal those of a context given the moved agents'
    voxels directly (agents that left the box: inactive, zero rate gradients)."""
    from biofvmb import LODSolver
    case = cf.make_case(17, 13, 11, 2, seed=8, nagents=300, nsteps=4, shared_voxels=True)
    moves = _moves(case, 6)
    voxel, t1, t2 = orc.agent_tables(case)
    v1, a1 = orc.update_voxel_indices(case, moves[0], np.ones(case.nagents, np.uint8))
    seed = np.random.default_rng(3).normal(size=case.init.shape)
    print("[gpu] adjoint after agent moves", flush=True)
    out = []
    for moved in (True, False):
        s = LODSolver(case.nx, case.ny, case.nz, case.nsub, case.dx, "fp64")
        s.set_coefficients(case.D, case.decay, case.dt)
        s.set_dirichlet(case.dvalue, case.dactive)
        s.upload(case.init)
        if moved:
            s.set_agents(voxel, t1, t2)
            s.update_agent_positions(moves[0])
        else:
            s.set_agents(np.maximum(v1, 0), t1, t2, a1)
        s.adjoint_forward(4, 2)
        out.append(s.adjoint_backward(seed))
        s.close()
    for got, want in zip(out[0], out[1]):
        assert np.array_equal(got, want)
    dead = a1 == 0
    assert dead.any() and np.all(out[0][1][dead] == 0.0) and np.any(out[0][1][~dead] != 0.0)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_last_segment_shared_interior_voxel(precision):
    """Regression for the rebuilt voxel lists: the occupied voxel with the highest device
    offset (z*nx*ny + y*nx + x) lies inside the grid and holds several agents, and the field is
    compared right after the source/sink update, before any LOD step could overwrite a face, on
    the first regrouping and on a second one (with agents leaving the box)."""
    from biofvmb import LODSolver
    case = cf.make_case(19, 17, 15, 2, seed=31, nagents=400, nsteps=1)
    rng = np.random.default_rng(9)
    dx = case.dx
    moves = []
    for last, nlast, nout in (((16, 14, 12), 5, 0), ((15, 14, 12), 3, 40)):
        p = rng.uniform(0.0, 1.0, (case.nagents, 3)) * np.array([case.nx, case.ny, case.nz - 4]) * dx
        sel = rng.choice(case.nagents, nlast + nout, replace=False)
        p[sel[:nlast]] = (np.array(last) + 0.5) * dx
        p[sel[nlast:]] = -3.0 * dx
        moves.append(p)
    voxel, t1, t2 = orc.agent_tables(case)
    print(f"[gpu] last shared interior segment {precision}", flush=True)
    s = LODSolver(case.nx, case.ny, case.nz, case.nsub, case.dx, precision)
    s.set_coefficients(case.D, case.decay, case.dt)
    s.set_dirichlet(case.dvalue, np.zeros(case.nsub, np.uint8))
    s.set_agents(voxel, t1, t2)
    active = np.ones(case.nagents, np.uint8)
    rho = case.init.copy()
    s.upload(rho)
    for p in moves:
        v, active = orc.update_voxel_indices(case, p, active)
        s.update_agent_positions(p)
        orc.sources_sinks(rho, np.maximum(v, 0), active, t1, t2)
        s.simulate_cell_sources_and_sinks()
        got = s.download()
        if precision == "fp64":
            assert np.array_equal(got, rho)
        else:
            r = np.abs(got - rho) / np.maximum(np.abs(rho), 1e-300)
            assert r.max() <= 1e-6
        rho = got.astype(np.float64)  # continue from the library's field (FP32: its rounding)
    assert 0 < int(active.sum()) < case.nagents
    s.close()


# ---- both public routes of the reference, against its own run (tests/golden/moves_*.npz)
def _fixture():
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_golden as golden
    return golden.load_moves()


def _run_moves_capi(case, ref, precision, device_positions=False):
    """The fixture's moves through the C ABI: route 0 = bfvmb_assign_agent_positions, route 1 =
    bfvmb_update_agent_positions, before the step's cell update."""
    from biofvmb import LODSolver
    s = LODSolver(case.nx, case.ny, case.nz, case.nsub, case.dx, precision)
    s.set_coefficients(case.D, case.decay, case.dt)
    s.set_dirichlet(*cf.effective_dirichlet(case))
    s.upload(case.init)
    s.set_agent_rates(ref["voxel"], case.volume, case.S, case.U, case.T, case.dt)
    moves = {st: (r, p) for st, r, p in case.moves}
    out = {}
    for step in range(1, max(case.snaps) + 1):
        if step in moves:
            route, pos = moves[step]
            if device_positions:
                import torch
                pos = torch.tensor(pos, dtype=torch.float64, device="cuda:0").contiguous()
                torch.cuda.synchronize()
            (s.assign_agent_positions if route == 0 else s.update_agent_positions)(pos)
        s.run(1)
        if step in case.snaps:
            out[step] = s.download()
    s.close()
    return out


@pytest.mark.parametrize("device_positions", [False, True])
def test_both_move_routes_match_reference_fp64(device_positions):
    case, ref = _fixture()
    print(f"[gpu] reference moves fixture, device positions {device_positions}", flush=True)
    got = _run_moves_capi(case, ref, "fp64", device_positions)
    for step, f in ref["fields"].items():
        assert np.array_equal(got[step], f), f"step {step}: {(got[step] != f).sum()} differ"


def test_both_move_routes_fp32_error():
    case, ref = _fixture()
    print("[gpu] reference moves fixture fp32", flush=True)
    got = _run_moves_capi(case, ref, "fp32")
    for step, f in ref["fields"].items():
        r = np.abs(got[step] - f) / np.maximum(np.abs(f), 1e-300)
        print(f"step {step}: fp32 max rel {r.max():.3e}")
        assert r.max() <= 1e-5


@pytest.mark.parametrize("mode", ["per_call", "fused"])
def test_host_layer_moves_match_reference(mode):
    """Basic_Agent::assign_position / update_voxel_index of the C++ host layer, in the
    reference driver's order: voxels and activity after each move and the fields."""
    import os
    import subprocess
    import tempfile
    case, ref = _fixture()
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "biofvm-b_amd", "lib", "bfvmb_driver")
    print(f"[gpu] host layer moves {mode}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        cpath, opath, mpath, mo = (os.path.join(d, f) for f in ("case.bin", "out.bin", "moves.bin", "mo.bin"))
        case.write(cpath)
        case.write_moves(mpath)
        args = [exe, cpath, opath, "fp64", "--moves", mpath, mo] + (["--fused"] if mode == "fused" else [])
        subprocess.run(args, check=True, timeout=120)
        out = cf.read_output(opath, case)
        m = np.fromfile(mo, dtype=np.int32).reshape(len(case.moves), 2, case.nagents)
    assert np.array_equal(m[:, 0], ref["move_voxel"]) and np.array_equal(m[:, 1], ref["move_active"])
    for step, f in ref["fields"].items():
        assert np.array_equal(out["fields"][step], f), f"step {step}"


@pytest.mark.parametrize("mode", ["hip", "hipres", "hipall"])
def test_reference_driver_moves_with_gpu_slots(mode):
    """The reference's own agents moved by its own routes, with the GPU in its slots: the plugin
    re-uploads the agent list when voxels or activity changed."""
    import os
    import subprocess
    import tempfile
    case, ref = _fixture()
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "oracle", "_ref", "ref_driver_hip")
    if not os.path.exists(exe):
        pytest.skip("reference build with the plugin absent")
    print(f"[gpu] reference driver moves {mode}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        cpath, opath, mpath, mo = (os.path.join(d, f) for f in ("case.bin", "out.bin", "moves.bin", "mo.bin"))
        case.write(cpath)
        case.write_moves(mpath)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([exe, cpath, opath, mode, "1", "--moves", mpath, mo], check=True, timeout=120, env=env)
        out = cf.read_output(opath, case)
    for step, f in ref["fields"].items():
        assert np.array_equal(out["fields"][step], f), f"step {step}"


def test_batch_member_assign_moves():
    from biofvmb import Batch
    case, ref = _fixture()
    other = cf.make_case(11, 9, 13, 2, seed=2, nagents=50, nsteps=1)
    print("[gpu] batch member, both move routes", flush=True)
    b = Batch([(other.nx, other.ny, other.nz, other.nsub), (case.nx, case.ny, case.nz, case.nsub)], case.dx, "fp64")
    for i, c in enumerate((other, case)):
        b.set_coefficients(i, c.D, c.decay, c.dt)
        b.set_dirichlet(i, *cf.effective_dirichlet(c))
        b.upload(i, c.init)
        voxel, t1, t2 = orc.agent_tables(c)
        b.set_agents(i, voxel, t1, t2)
    moves = {st: (r, p) for st, r, p in case.moves}
    for step in range(1, max(case.snaps) + 1):
        if step in moves:
            route, pos = moves[step]
            (b.assign_agent_positions if route == 0 else b.update_agent_positions)(1, pos, case.nagents)
        b.run(1)
        if step in case.snaps:
            assert np.array_equal(b.download(1), ref["fields"][step]), step
    b.close()
