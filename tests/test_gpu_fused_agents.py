"""The step's cell sources and sinks inside the x sweep's launch (bfvmb_run, graphs, batches,
launch phase 4): bit-identical to the separate agent kernel followed by the x sweep
(bfvmb_simulate_cell_sources_and_sinks then bfvmb_simulate_diffusion_decay), in FP64 and
FP32, and in FP64 to the oracle.  Cases: dense agents (several batches of 256 segments per
tile), several agents per voxel (list order), agents on the x faces of active Dirichlet
substrates (their update is overwritten, as in the reference), row counts that leave a
partial tile, domains without agents beside domains with them, agents that move, and tiles
denser than the fusion limit (3072 occupied voxels per tile: the step then launches the agent
kernel and the x sweep separately, with the same results)."""
import numpy as np
import pytest

import casefile as cf
import oracle as orc

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "dense_33x21x13x3":  # 273 rows: four full tiles and one of 17 rows
        c = cf.make_case(33, 21, 13, 3, seed=41, nagents=6000, nsteps=3, shared_voxels=True,
                         dactive=[1, 0, 1])
    elif name == "xfaces_40x9x8x2":
        c = cf.make_case(40, 9, 8, 2, seed=42, nagents=900, nsteps=3, dactive=[1, 1])
        rng = np.random.default_rng(3)
        c.pos[:300, 0] = 0.5 * c.dx  # x = 0 column
        c.pos[300:600, 0] = (c.nx - 0.5) * c.dx  # x = nx - 1 column
        c.pos[600:650] = c.pos[650]  # one voxel with 51 agents
        c.pos[:, 1:] = np.clip(c.pos[:, 1:] + rng.normal(0, 0.1, (c.nagents, 2)), 0.0,
                               np.array([c.ny, c.nz]) * c.dx - 1e-9)
    elif name == "wide_150x70x3x2":  # 210 rows, long lines, sparse agents
        c = cf.make_case(150, 70, 3, 2, seed=43, nagents=400, nsteps=2, dactive=[0, 1])
    elif name == "dense_tiles_128x70x2x2":  # ~6000 occupied voxels per tile of 64 rows
        c = cf.make_case(128, 70, 2, 2, seed=45, nagents=16000, nsteps=2, shared_voxels=True, dactive=[1, 0])
    else:
        raise KeyError(name)
    return c


NAMES = ["dense_33x21x13x3", "xfaces_40x9x8x2", "wide_150x70x3x2", "dense_tiles_128x70x2x2"]


def _solver(c, precision):
    from biofvmb import LODSolver
    s = LODSolver(c.nx, c.ny, c.nz, c.nsub, c.dx, precision)
    s.set_coefficients(c.D, c.decay, c.dt)
    s.set_dirichlet(c.dvalue, c.dactive)
    s.upload(c.init)
    s.set_agent_rates(orc.voxel_indices(c), c.volume, c.S, c.U, c.T, c.dt)
    return s


def _separate(c, precision, nsteps):
    s = _solver(c, precision)
    for _ in range(nsteps):
        s.simulate_cell_sources_and_sinks()
        s.simulate_diffusion_decay(1)
    out = s.download()
    s.close()
    return out


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("name", NAMES)
def test_run_matches_separate_calls(name, precision):
    c = _case(name)
    print(f"[gpu] fused agents {name} {precision}", flush=True)
    want = _separate(c, precision, c.nsteps)
    for graph in (False, True):
        s = _solver(c, precision)
        s.run(c.nsteps, graph=graph)
        got = s.download()
        s.close()
        assert np.array_equal(got, want), (graph, int((got != want).sum()))
    if precision == "fp64":
        assert np.array_equal(want, orc.run(c))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_launch_phase_4_is_phases_0_and_1(precision):
    c = _case("dense_33x21x13x3")
    print(f"[gpu] launch phase 4 {precision}", flush=True)
    a = _solver(c, precision)
    a.launch_phase(0, 1)
    a.launch_phase(1, 1)
    b = _solver(c, precision)
    b.launch_phase(4, 1)
    assert np.array_equal(a.download(), b.download())
    a.close()
    b.close()


def test_batch_members_with_and_without_agents():
    from biofvmb import Batch
    cases = [_case("dense_33x21x13x3"), cf.make_case(17, 12, 9, 2, seed=44, nagents=0, nsteps=3),
             _case("wide_150x70x3x2")]
    print("[gpu] fused agents in a batch", flush=True)
    b = Batch([(c.nx, c.ny, c.nz, c.nsub) for c in cases], cases[0].dx, "fp64")
    for i, c in enumerate(cases):
        b.set_coefficients(i, c.D, c.decay, c.dt)
        b.set_dirichlet(i, c.dvalue, c.dactive)
        b.upload(i, c.init)
        if c.nagents:
            voxel, t1, t2 = orc.agent_tables(c)
            b.set_agents(i, voxel, t1, t2)
    b.run(3, graph=True)
    for i, c in enumerate(cases):
        assert np.array_equal(b.download(i), orc.run(c, nsteps=3)), i
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_moved_agents_then_run(precision):
    """The tile index follows the device regrouping of moved agents."""
    c = _case("dense_33x21x13x3")
    rng = np.random.default_rng(9)
    moved = c.pos + rng.normal(0.0, 2.0 * c.dx, c.pos.shape)
    print(f"[gpu] fused agents after moves {precision}", flush=True)
    outs = []
    for fused in (True, False):
        s = _solver(c, precision)
        s.run(1) if fused else (s.simulate_cell_sources_and_sinks(), s.simulate_diffusion_decay(1))
        s.update_agent_positions(moved)
        for _ in range(2):
            s.run(1) if fused else (s.simulate_cell_sources_and_sinks(), s.simulate_diffusion_decay(1))
        outs.append(s.download())
        s.close()
    assert np.array_equal(outs[0], outs[1])
