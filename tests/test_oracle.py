"""Pins the C restatement (oracle/bfvmb_oracle.c) before it is trusted as the oracle.

1. Bit for bit against the reference's own outputs (tests/golden/, written by the compiled
   reference: oracle/_ref/ref_driver via tests/golden/make_golden.py).
2. Against what the mathematics fixes, independently of the reference's code:
   dense numpy solves of each 1-D operator, D=0, Neumann cosine eigenmodes, mass
   conservation, a Dirichlet-valued field, and the agent update's closed form.
"""
import os
import sys

import numpy as np
import pytest

import casefile as cf
import oracle as orc

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_golden as golden  # noqa: E402


# ---------------------------------------------------------------- 1. reference fixtures
@pytest.mark.parametrize("name", golden.ALL_NAMES + golden.DNODES_NAMES)
def test_oracle_matches_reference_bitwise(name):
    case, ref = golden.load(name)
    voxel, t1, t2 = orc.agent_tables(case)
    assert np.array_equal(voxel, ref["voxel"])
    assert np.array_equal(t1, ref["temp1"])
    assert np.array_equal(t2, ref["temp2"])
    got = orc.run_snapshots(case)
    for s, f in ref["fields"].items():
        assert np.array_equal(got[s], f), f"{name} step {s}: {(got[s] != f).sum()} values differ"


@pytest.mark.parametrize("name", golden.ALL_NAMES + golden.DNODES_NAMES)
def test_reference_avx_solver_equals_scalar_solver(name):
    _, ref = golden.load(name)
    assert ref["avx_equal"]


@pytest.mark.parametrize("name", golden.DNODES_NAMES)
def test_dirichlet_node_fixtures_discriminate(name):
    """The per-voxel fixtures pin the node values: the same case with one value per substrate
    (nodes dropped) or with the nodes' values read at the wrong voxel (x and z swapped in the
    voxel index) differs from the reference's output."""
    case, ref = golden.load(name)
    last = max(ref["fields"])
    voxel, values = case.dnodes
    case.dnodes = None
    assert not np.array_equal(orc.run(case, nsteps=last), ref["fields"][last])
    x, rem = np.divmod(np.asarray(voxel), case.ny * case.nz)
    y, z = np.divmod(rem, case.nz)
    if case.nx == case.nz:
        pytest.skip("x/z swap is the identity map on this grid")
    ok = (z < case.nx) & (x < case.nz)
    case.dnodes = (((z * case.ny + y) * case.nz + x)[ok], np.asarray(values)[ok])
    assert not np.array_equal(orc.run(case, nsteps=last), ref["fields"][last])


@pytest.mark.parametrize("name", golden.DNODES_NAMES)
def test_dirichlet_nodes_fix_every_active_face_voxel(name):
    """Mathematics of apply_dirichlet_conditions: after a step, every face voxel of an active
    substrate holds exactly its node value (or the per-substrate value), interior nodes leave
    the field unchanged against the same case with them removed."""
    case, ref = golden.load(name)
    value, active = cf.effective_dirichlet(case)
    out = orc.run(case, nsteps=1)
    dv = orc.dirichlet_voxels(case)
    face = np.zeros((case.nx, case.ny, case.nz), bool)
    face[0] = face[-1] = True
    face[:, 0] = face[:, -1] = True
    face[:, :, 0] = face[:, :, -1] = True
    for d in range(case.nsub):
        if active[d]:
            assert np.array_equal(out[..., d][face], dv[..., d][face])
    voxel, values = case.dnodes
    on_face = face.reshape(-1)[np.asarray(voxel)]
    case.dnodes = (np.asarray(voxel)[on_face], np.asarray(values)[on_face])
    assert np.array_equal(orc.run(case, nsteps=1), out)


# ---------------------------------------------------------------- 2. mathematics
def _dense_operator(n, c1, c2):
    """The 1-D LOD operator of one substrate: tridiag(-c1, 1+2c1+c2, -c1) with Neumann ends
    (1+c1+c2), and 1+c2 for a single cell."""
    if n == 1:
        return np.array([[1.0 + c2]])
    A = np.diag(np.full(n, 1 + 2 * c1 + c2)) + np.diag(np.full(n - 1, -c1), 1) + np.diag(np.full(n - 1, -c1), -1)
    A[0, 0] = A[-1, -1] = 1 + c1 + c2
    return A


def _dense_lod_step(rho, D, decay, dt, dx, dvalue, dactive):
    nx, ny, nz, nsub = rho.shape
    out = rho.copy()

    def dirichlet(u):
        for d in range(nsub):
            if dactive[d]:
                u[0, :, :, d] = u[-1, :, :, d] = dvalue[d]
                u[:, 0, :, d] = u[:, -1, :, d] = dvalue[d]
                u[:, :, 0, d] = u[:, :, -1, d] = dvalue[d]

    dirichlet(out)
    for axis, n in enumerate((nx, ny, nz)):
        for d in range(nsub):
            c1 = D[d] * dt / dx / dx
            c2 = decay[d] * dt / 3
            A = _dense_operator(n, c1, c2)
            u = np.moveaxis(out[..., d], axis, 0)
            sol = np.linalg.solve(A, u.reshape(n, -1)).reshape(u.shape)
            out[..., d] = np.moveaxis(sol, 0, axis)
        dirichlet(out)
    return out


@pytest.mark.parametrize("shape", [(4, 5, 3, 2), (1, 6, 4, 1), (7, 1, 5, 3), (5, 4, 1, 2), (9, 8, 7, 1)])
def test_lod_step_equals_dense_solves(shape):
    nx, ny, nz, nsub = shape
    case = cf.make_case(nx, ny, nz, nsub, seed=sum(shape), nsteps=1, with_agents=False)
    value, active = cf.effective_dirichlet(case)
    want = _dense_lod_step(case.init, case.D, case.decay, case.dt, case.dx, value, active)
    got = orc.run(case)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=0)


def test_dense_check_catches_a_wrong_sign():
    """A planted mistake in the oracle's inputs (decay sign flipped) must fail the dense check."""
    case = cf.make_case(6, 5, 4, 2, seed=3, nsteps=1, with_agents=False, decay=[0.5, 0.2])
    value, active = cf.effective_dirichlet(case)
    want = _dense_lod_step(case.init, case.D, -case.decay, case.dt, case.dx, value, active)
    got = orc.run(case)
    assert not np.allclose(got, want, rtol=1e-12, atol=0)


def test_zero_diffusion_is_three_divisions():
    case = cf.make_case(6, 7, 8, 3, seed=21, nsteps=1, with_agents=False, D=[0.0, 0.0, 0.0], dactive=[0, 0, 0])
    c2 = case.decay * case.dt / 3.0
    want = ((case.init / (1.0 + c2)) / (1.0 + c2)) / (1.0 + c2)
    assert np.array_equal(orc.run(case), want)


@pytest.mark.parametrize("k", [(1, 0, 0), (0, 2, 0), (1, 1, 3)])
def test_neumann_cosine_mode_is_an_eigenvector(k):
    nx, ny, nz = 16, 12, 10
    D, decay, dt, dx = 1e3, 0.05, 0.01, 20.0
    c1, c2 = D * dt / dx / dx, decay * dt / 3
    axes = [np.cos(np.pi * kk * (np.arange(n) + 0.5) / n) for kk, n in zip(k, (nx, ny, nz))]
    mode = axes[0][:, None, None] * axes[1][None, :, None] * axes[2][None, None, :] + 2.0
    case = cf.make_case(nx, ny, nz, 1, seed=1, nsteps=1, with_agents=False, D=[D], decay=[decay], dactive=[0],
                        init=mode[..., None])
    got = orc.run(case)[..., 0]
    mu = [1 + c2 + 2 * c1 * (1 - np.cos(np.pi * kk / n)) for kk, n in zip(k, (nx, ny, nz))]
    want = (mode - 2.0) / np.prod(mu) + 2.0 / (1 + c2) ** 3
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-13)


def test_mass_is_conserved_without_decay_or_dirichlet():
    case, ref = golden.load("conserve_20x18x16x2")
    got = orc.run_snapshots(case)[10]
    np.testing.assert_allclose(got.sum(axis=(0, 1, 2)), case.init.sum(axis=(0, 1, 2)), rtol=1e-13)
    np.testing.assert_allclose(ref["fields"][10].sum(axis=(0, 1, 2)), case.init.sum(axis=(0, 1, 2)), rtol=1e-13)


def test_dirichlet_valued_field_stays_fixed():
    case = cf.make_case(10, 9, 8, 2, seed=4, nsteps=5, with_agents=False, decay=[0.0, 0.0], dvalue=[38.0, 7.5],
                        dactive=[1, 1], init=np.broadcast_to(np.array([38.0, 7.5]), (10, 9, 8, 2)))
    np.testing.assert_allclose(orc.run(case), case.init, rtol=1e-14, atol=0)


def test_single_agent_update_closed_form():
    rho = np.full((3, 3, 3, 2), 5.0)
    t1 = np.array([[0.3, 0.0]])
    t2 = np.array([[1.2, 2.0]])
    orc.sources_sinks(rho, np.array([13]), np.array([1]), t1, t2)
    flat = rho.reshape(-1, 2)
    assert flat[13, 0] == (5.0 + 0.3) / 1.2 and flat[13, 1] == 5.0 / 2.0
    assert np.all(np.delete(flat, 13, axis=0) == 5.0)


def test_agents_in_one_voxel_compose_in_list_order():
    rng = np.random.default_rng(0)
    t1 = rng.uniform(0, 1, (6, 1))
    t2 = rng.uniform(1, 2, (6, 1))
    rho = np.full((2, 2, 2, 1), 3.0)
    orc.sources_sinks(rho, np.full(6, 5), np.ones(6), t1, t2)
    r = 3.0
    for a in range(6):
        r = (r + t1[a, 0]) / t2[a, 0]
    assert rho.reshape(-1)[5] == r
    rho2 = np.full((2, 2, 2, 1), 3.0)
    orc.sources_sinks(rho2, np.full(6, 5), np.ones(6), t1[::-1].copy(), t2[::-1].copy())
    assert rho2.reshape(-1)[5] != r  # order matters, so the fixture checks order


def test_agent_constants_formula():
    S, U, T = np.array([[2.0, 0.5]]), np.array([[10.0, 3.0]]), np.array([[1.0, 4.0]])
    vol, dV, dt = 4.0 / 3.0 * np.pi * 125.0, 8000.0, 0.01
    t1, t2 = orc.agent_constants(dt, vol, dV, S, U, T)
    k = dt * vol / dV
    np.testing.assert_allclose(t1, S * T * k, rtol=1e-15)
    np.testing.assert_allclose(t2, 1 + k * (S + U), rtol=1e-15)


def test_voxel_mapping_brute_force():
    case = cf.make_case(13, 11, 9, 1, seed=5, nagents=400, outside=7)
    got = orc.voxel_indices(case)
    lo, hi = case.bbox[:3], case.bbox[3:]
    n = np.array([case.nx, case.ny, case.nz])
    inside = np.all((case.pos >= lo) & (case.pos <= hi), axis=1)
    c = np.clip(np.floor((case.pos - lo) / case.dx).astype(int), 0, n - 1)
    want = np.where(inside, (c[:, 0] * n[1] + c[:, 1]) * n[2] + c[:, 2], 0)
    assert np.array_equal(got, want)


def test_moved_agent_voxels_brute_force():
    """update_voxel_indices: inside the box the nearest voxel (as for the initial
    assignment), outside it voxel -1 and the agent inactive; inactive agents stay so."""
    case = cf.make_case(13, 11, 9, 1, seed=6, nagents=300)
    rng = np.random.default_rng(1)
    ext = np.array([case.nx, case.ny, case.nz]) * case.dx
    pos = rng.uniform(-0.2, 1.2, (300, 3)) * ext
    pos[:5] = ext  # upper corner: inside, clamped into the last voxel
    pos[5:8] = 0.0
    active = np.ones(300, np.uint8)
    active[10:20] = 0
    vox, act = orc.update_voxel_indices(case, pos, active)
    lo, hi = case.bbox[:3], case.bbox[3:]
    n = np.array([case.nx, case.ny, case.nz])
    inside = np.all((pos >= lo) & (pos <= hi), axis=1)
    c = np.clip(np.floor((pos - lo) / case.dx).astype(int), 0, n - 1)
    want = np.where(inside, (c[:, 0] * n[1] + c[:, 1]) * n[2] + c[:, 2], -1)
    assert np.array_equal(vox, want)
    assert np.array_equal(act, (inside & (active == 1)).astype(np.uint8))
    assert np.all(vox[:5] == case.nx * case.ny * case.nz - 1) and np.all(vox[5:8] == 0)
    assert 0 < inside.sum() < 300


# ---------------------------------------------------------------- bulk sources and sinks
@pytest.mark.parametrize("name", golden.BULK_NAMES)
def test_bulk_matches_reference_fixtures(name):
    """orc_bulk_sources_sinks inside the time loop, bit for bit against the reference's
    Microenvironment::simulate_bulk_sources_and_sinks run by its own driver."""
    case, ref = golden.load(name)
    for step, field in ref["fields"].items():
        assert np.array_equal(orc.run(case, nsteps=step), field), f"{name} step {step}"


def test_bulk_closed_forms():
    rng = np.random.default_rng(2)
    rho0 = rng.lognormal(0.0, 0.5, (4, 3, 5, 2))
    z = np.zeros_like(rho0)
    r = rho0.copy()
    orc.bulk_sources_sinks(r, 0.01, z, z, z)  # no rates: unchanged
    assert np.array_equal(r, rho0)
    U = rng.uniform(0.0, 20.0, rho0.shape)
    r = rho0.copy()
    orc.bulk_sources_sinks(r, 0.01, z, z, U)  # uptake only: one implicit decay step
    assert np.array_equal(r, rho0 / (U * 0.01 + 1.0))
    S = rng.uniform(0.1, 5.0, rho0.shape)
    T = rng.uniform(1.0, 40.0, rho0.shape)
    r = T.copy()
    orc.bulk_sources_sinks(r, 0.01, S, T, z)  # at the target with no uptake: a fixed point
    assert np.allclose(r, T, rtol=4e-16, atol=0)
    r = rho0.copy()  # long run with supply only converges to the target
    for _ in range(4000):
        orc.bulk_sources_sinks(r, 0.5, S, T, z)
    assert np.allclose(r, T, rtol=1e-12, atol=0)


# ---------------------------------------------------------------- gradients
@pytest.mark.parametrize("name", golden.GRAD_NAMES)
def test_gradients_match_reference_fixtures(name):
    """orc_gradients bit for bit against compute_all_gradient_vectors run by the reference."""
    case, ref = golden.load(name)
    rho = ref["fields"][case.nsteps]
    assert np.array_equal(orc.gradients(rho, case.dx), ref["grad"])


def test_gradients_exact_on_linear_and_quadratic_fields():
    """Centered differences are exact for quadratics in the interior, one-sided ones for
    linear fields at the ends (grid values chosen to be exact in binary)."""
    nx, ny, nz, dx = 6, 5, 4, 0.5
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    lin = (3.0 * x - 2.0 * y + 5.0 * z)[..., None] * 1.0
    g = orc.gradients(lin, dx)
    assert np.all(g[..., 0, 0] == 3.0 / dx) and np.all(g[..., 0, 1] == -2.0 / dx) and np.all(g[..., 0, 2] == 5.0 / dx)
    quad = (x * x * 1.0)[..., None]
    g = orc.gradients(quad, dx)
    assert np.array_equal(g[1:-1, ..., 0, 0], (2.0 * x[1:-1]) / dx)  # d(x^2)/dx = 2x exactly
    assert np.all(g[0, ..., 0, 0] == 1.0 / dx) and np.all(g[-1, ..., 0, 0] == (2 * nx - 3) / dx)
    assert np.all(g[..., 0, 1:] == 0.0)


def test_gradients_axis_of_length_one_is_zero():
    rho = np.random.default_rng(0).uniform(size=(1, 4, 1, 2))
    g = orc.gradients(rho, 20.0)
    assert np.all(g[..., 0] == 0.0) and np.all(g[..., 2] == 0.0) and np.any(g[..., 1] != 0.0)


def test_agent_moves_match_reference():
    """Both public routes that move an agent (assign_position, and position +
    update_voxel_index; BioFVM_basic_agent.cpp:112-153), as the reference applied them: the
    voxels and activity after every move, and the fields, bit for bit."""
    case, ref = golden.load_moves()
    snaps, after = orc.run_moves(case)
    for k, (voxel, active) in enumerate(after):
        assert np.array_equal(voxel, ref["move_voxel"][k]), f"move {k}: voxels"
        assert np.array_equal(active, ref["move_active"][k].astype(np.uint8)), f"move {k}: activity"
    for s, f in ref["fields"].items():
        assert np.array_equal(snaps[s], f), f"step {s}"
    # the routes differ on this case: swapping them changes the result
    case.moves = [(st, 1 - r, p) for st, r, p in case.moves]
    swapped, _ = orc.run_moves(case)
    last = max(ref["fields"])
    assert not np.array_equal(swapped[last], ref["fields"][last])
