"""Batched microenvironments (uneven grids and substrate counts, one launch per phase for
all members), graph-captured steps, and the autotuner: every member must match its own
oracle run bit for bit in FP64 (FP32: 1e-5 max relative error), whatever the launch path."""
import numpy as np
import pytest

import casefile as cf
import oracle as orc

pytestmark = pytest.mark.gpu

SHAPES = [(17, 23, 31, 3), (40, 12, 9, 1), (5, 6, 70, 2), (33, 33, 33, 5), (1, 9, 4, 2), (64, 20, 48, 4), (130, 7, 11, 1)]


def _cases(nsteps):
    out = []
    for i, (nx, ny, nz, ns) in enumerate(SHAPES):
        c = cf.make_case(nx, ny, nz, ns, seed=100 + i, nagents=20 * (i + 1), nsteps=nsteps, shared_voxels=True,
                         outside=1)
        out.append(c)
    return out


def _batch(cases, precision):
    from biofvmb import Batch
    b = Batch([(c.nx, c.ny, c.nz, c.nsub) for c in cases], 20.0, precision)
    for i, c in enumerate(cases):
        b.set_coefficients(i, c.D, c.decay, c.dt)
        b.set_dirichlet(i, c.dvalue, c.dactive)
        b.upload(i, c.init)
        voxel, t1, t2 = orc.agent_tables(c)
        b.set_agents(i, voxel, t1, t2)
    return b


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


@pytest.mark.parametrize("graph", [False, True])
def test_batch_fp64_matches_oracle(graph):
    cases = _cases(5)
    print(f"[gpu] batch fp64 graph={graph}", flush=True)
    b = _batch(cases, "fp64")
    b.run(5, graph=graph)
    for i, c in enumerate(cases):
        assert np.array_equal(b.download(i), orc.run(c)), f"member {i} {SHAPES[i]}"
    b.close()


def test_batch_fp32_error():
    cases = _cases(5)
    b = _batch(cases, "fp32")
    b.run(5, graph=True)
    for i, c in enumerate(cases):
        err = _rel(b.download(i), orc.run(c))
        print(f"fp32 member {i} {SHAPES[i]}: max rel {err:.2e}")
        assert err <= 1e-5
    b.close()


def test_graph_replays_with_eager_work_between():
    """A captured run replayed several times, with eager steps and host syncs between the
    replays, equals the same number of eager steps (each replay checked)."""
    cases = _cases(0)
    ref = _batch(cases, "fp64")
    got = _batch(cases, "fp64")
    for rep in range(3):
        got.run(2, graph=True)   # replay of the cached 2-step graph (same parity each time)
        ref.run(2, graph=False)
        got.sync()
        for i in range(len(cases)):
            assert np.array_equal(got.download(i), ref.download(i)), f"replay {rep} member {i}"
        got.run(1, graph=False)  # eager step between replays flips the parity
        ref.run(1, graph=False)
    got.close()
    ref.close()


def test_single_context_graph_equals_eager():
    from biofvmb import LODSolver
    c = cf.make_case(24, 20, 28, 3, seed=7, nagents=300, nsteps=6, shared_voxels=True)
    outs = []
    for graph in (False, True):
        s = LODSolver(c.nx, c.ny, c.nz, c.nsub, c.dx, "fp64")
        s.set_coefficients(c.D, c.decay, c.dt)
        s.set_dirichlet(c.dvalue, c.dactive)
        s.upload(c.init)
        voxel, t1, t2 = orc.agent_tables(c)
        s.set_agents(voxel, t1, t2)
        s.run(3, graph=graph)
        s.run(3, graph=graph)
        outs.append(s.download())
        s.close()
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0], orc.run(c))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_autotune_keeps_state_and_results(precision):
    """Tuning leaves the fields untouched, and the tuned shapes (x-sweep variant, strided x
    grid, y/z rings) give the untuned bits: FP64 the oracle's, FP32 the default launch's."""
    cases = _cases(4)
    print(f"[gpu] autotune {precision}", flush=True)
    ref = _batch(cases, precision)
    ref.run(4)
    b = _batch(cases, precision)
    chosen = b.autotune(reps=2)
    print("autotune choice [bpc, ring, us] per sweep:", chosen.tolist())
    assert (chosen[:, 0] >= 1).all()
    b.run(4, graph=True)
    for i, c in enumerate(cases):
        got = b.download(i)
        assert np.array_equal(got, ref.download(i)), f"member {i}"
        if precision == "fp64":
            assert np.array_equal(got, orc.run(c)), f"member {i}"
        else:
            assert _rel(got, orc.run(c)) <= 1e-5
    b.close()
    ref.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_phases_compose_a_step(precision):
    """bfvmb_batch_launch_phase: sources/sinks, then the x, y and z sweeps, one phase at a
    time, give the bits of bfvmb_batch_run's steps (FP64: the oracle's)."""
    cases = _cases(3)
    print(f"[gpu] batch phases {precision}", flush=True)
    a = _batch(cases, precision)
    a.run(3)
    b = _batch(cases, precision)
    for _ in range(3):
        for phase in (0, 1, 2, 3):
            b.launch_phase(phase, 1)
    b.sync()
    for i, c in enumerate(cases):
        got = b.download(i)
        assert np.array_equal(got, a.download(i)), f"member {i} {SHAPES[i]}"
        if precision == "fp64":
            assert np.array_equal(got, orc.run(c))
    a.close()
    b.close()
