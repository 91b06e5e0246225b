"""GPU parity of the LOD step and the agent source/sink update against the oracle.

The FP64 path must be bit-identical to the C restatement (itself bit-identical to the
compiled reference, see test_oracle.py); the FP32 path is held to 1e-5 max relative
error against the FP64 oracle (the tolerance of the FP32 storage path), with the error printed.
"""
import numpy as np
import pytest

import casefile as cf
import oracle as orc

pytestmark = pytest.mark.gpu


def _solver(case, precision):
    from biofvmb import LODSolver
    s = LODSolver(case.nx, case.ny, case.nz, case.nsub, case.dx, precision)
    s.set_coefficients(case.D, case.decay, case.dt)
    s.set_dirichlet(case.dvalue, case.dactive)
    s.upload(case.init)
    if case.with_agents and case.nagents:
        voxel = orc.voxel_indices(case)
        s.set_agent_rates(voxel, case.volume, case.S, case.U, case.T, case.dt)
    return s


def gpu_run(case, precision="fp64", nsteps=None):
    s = _solver(case, precision)
    s.run(case.nsteps if nsteps is None else nsteps)
    out = s.download()
    s.close()
    return out


def rel_err(a, b):
    den = np.maximum(np.abs(b), 1e-300)
    r = np.abs(a - b) / den
    return float(r.max()), float(np.sqrt(np.mean(r * r)))


CASES = {
    "aniso_17x23x31x3": dict(nx=17, ny=23, nz=31, nsub=3, seed=1, nagents=200, nsteps=5, shared_voxels=True, outside=3),
    "nx1": dict(nx=1, ny=5, nz=7, nsub=2, seed=2, nagents=10, nsteps=3),
    "ny1": dict(nx=2, ny=1, nz=9, nsub=4, seed=3, nagents=30, nsteps=4, shared_voxels=True),
    "nz1": dict(nx=6, ny=4, nz=1, nsub=2, seed=4, nagents=5, nsteps=3),
    "one_voxel": dict(nx=1, ny=1, nz=1, nsub=3, seed=5, nagents=4, nsteps=2),
    "chunks_130x70x33x2": dict(nx=130, ny=70, nz=33, nsub=2, seed=6, nagents=3000, nsteps=6, shared_voxels=True),
    "long_x_300x17x19x1": dict(nx=300, ny=17, nz=19, nsub=1, seed=7, nagents=100, nsteps=4),
    "long_x_700x5x6x2": dict(nx=700, ny=5, nz=6, nsub=2, seed=15, nagents=50, nsteps=3),
    "long_y_9x340x7x1": dict(nx=9, ny=340, nz=7, nsub=1, seed=16, nagents=50, nsteps=3),
    "long_z_5x6x700x2": dict(nx=5, ny=6, nz=700, nsub=2, seed=17, nagents=50, nsteps=3),
    "all_dirichlet_64x48x40x4": dict(nx=64, ny=48, nz=40, nsub=4, seed=8, nagents=2000, nsteps=8, dactive=[1, 1, 1, 1]),
    "no_dirichlet_40x41x42x2": dict(nx=40, ny=41, nz=42, nsub=2, seed=9, nagents=500, nsteps=8, dactive=[0, 0]),
    "plain_50cube_100steps": dict(nx=50, ny=50, nz=50, nsub=1, seed=10, nagents=1000, nsteps=100, D=[1e5], decay=[0.1],
                                    dvalue=[38.0], dactive=[1], init=38.0),
}


def _case(name):
    kw = dict(CASES[name])
    c = cf.make_case(**kw)
    if name == "plain_50cube_100steps":  # uptake 10/min, secretion 0
        c.U[:] = 10.0
        c.S[:] = 0.0
    return c


@pytest.mark.parametrize("name", list(CASES))
def test_fp64_bitwise_vs_oracle(name):
    case = _case(name)
    print(f"[gpu] fp64 {name} shape={case.nx, case.ny, case.nz, case.nsub} steps={case.nsteps}", flush=True)
    ref = orc.run(case)
    got = gpu_run(case, "fp64")
    assert got.shape == ref.shape
    if not np.array_equal(got, ref):
        mx, rms = rel_err(got, ref)
        bad = np.argwhere(got != ref)
        pytest.fail(f"{name}: {len(bad)} values differ, first {bad[:5].tolist()}, max rel {mx:.3e}, rms {rms:.3e}")


@pytest.mark.parametrize("name", list(CASES))
def test_fp32_error_vs_fp64_oracle(name):
    case = _case(name)
    print(f"[gpu] fp32 {name}", flush=True)
    ref = orc.run(case)
    got = gpu_run(case, "fp32")
    mx, rms = rel_err(got, ref)
    print(f"fp32 {name}: max rel {mx:.3e} rms rel {rms:.3e}")
    assert mx <= 1e-5, (mx, rms)


def test_phases_separately_match_oracle():
    """Sources/sinks alone, then diffusion alone, as two separate entry points."""
    case = _case("aniso_17x23x31x3")
    s = _solver(case, "fp64")
    voxel, t1, t2 = orc.agent_tables(case)
    rho = case.init.copy()
    orc.sources_sinks(rho, voxel, np.ones(case.nagents, np.uint8), t1, t2)
    s.simulate_cell_sources_and_sinks()
    assert np.array_equal(s.download(), rho)
    ref = orc.run(case, nsteps=3, rho=rho, with_agents=False)
    s.simulate_diffusion_decay(3)
    assert np.array_equal(s.download(), ref)
    s.close()


def test_set_agents_with_temps_and_inactive():
    """bfvmb_set_agents with precomputed temps; inactive agents are skipped."""
    case = _case("chunks_130x70x33x2")
    voxel, t1, t2 = orc.agent_tables(case)
    active = (np.arange(case.nagents) % 3 != 0).astype(np.uint8)
    rho = case.init.copy()
    orc.sources_sinks(rho, voxel, active, t1, t2)
    from biofvmb import LODSolver
    s = LODSolver(case.nx, case.ny, case.nz, case.nsub, case.dx, "fp64")
    s.set_coefficients(case.D, case.decay, case.dt)
    s.upload(case.init)
    s.set_agents(voxel, t1, t2, active)
    s.simulate_cell_sources_and_sinks()
    assert np.array_equal(s.download(), rho)
    s.close()


@pytest.mark.parametrize("name", ["chunks_130x70x33x2", "long_x_700x5x6x2", "long_z_5x6x700x2"])
@pytest.mark.parametrize("launch", [(7, 1, 16), (0, 2, 32), (3, 8, 64), (0, 4, 16), (64, 4, 32), (9, 6, 32)])
def test_launch_shapes_do_not_change_results(name, launch):
    """Strided grids, ring depths and smaller LDS line stores (more of each line recomputed
    from checkpoints) give the same bits as the default launch: FP64 the oracle's, FP32 the
    default FP32 launch's (within 1e-5 of the oracle)."""
    case = _case(name)
    print(f"[gpu] launch {launch} {name}", flush=True)
    ref = orc.run(case)
    s = _solver(case, "fp32")
    s.run(case.nsteps)
    base32 = s.download()
    s.close()
    for prec in ("fp64", "fp32"):
        s = _solver(case, prec)
        grid, bpc, ring = launch
        s.set_launch(-1, grid, bpc, ring)
        s.run(case.nsteps)
        got = s.download()
        s.close()
        if prec == "fp64":
            assert np.array_equal(got, ref)
        else:
            assert rel_err(got, ref)[0] <= 1e-5
            assert np.array_equal(got, base32)


@pytest.mark.parametrize("name", ["chunks_130x70x33x2", "long_x_700x5x6x2"])
def test_fp32_x_variants_agree_bitwise(name):
    """The FP32 x-sweep variants (1 or 3 column chunks in flight, with 17 or no register
    chunks; Thomas tables read with scalar loads or staged in LDS) at several LDS budgets and
    strided grids: each stores the same FP32-rounded forward results, so every launch gives
    the default launch's bits, within 1e-5 of the FP64 oracle."""
    case = _case(name)
    print(f"[gpu] fp32 x variants {name}", flush=True)
    ref = orc.run(case)
    s = _solver(case, "fp32")
    s.run(case.nsteps)
    base = s.download()
    s.close()
    assert rel_err(base, ref)[0] <= 1e-5
    for ring in (1, 2, 3, 4):
        for grid, bpc in ((0, 2), (5, 4), (0, 8)):
            s = _solver(case, "fp32")
            s.set_launch(0, grid, bpc, ring)
            s.run(case.nsteps)
            got = s.download()
            s.close()
            assert np.array_equal(got, base), (ring, grid, bpc)


def _tiny_case(nx, ny, nz, nsub, seed, mixed):
    """Subnormal fields around 1e-310 (numerators far below 2^-960, where the corrected
    reciprocal misrounds about one division in 400, some zero): the sweeps redo such tiles
    with true divisions.  ``mixed``: only the z < nz/2 half is tiny, so tiles hold tiny and
    ordinary lines together."""
    rng = np.random.default_rng(seed)
    init = rng.lognormal(0.0, 0.5, (nx, ny, nz, nsub)) * 1e-310
    init[rng.uniform(size=init.shape) < 0.1] *= 1e-5
    init[rng.uniform(size=init.shape) < 0.02] = 0.0
    if mixed:
        init[:, :, nz // 2:, :] = rng.lognormal(0.0, 0.5, (nx, ny, nz - nz // 2, nsub))
    dvalue = [1e-312] + [0.0] * (nsub - 1)
    return cf.make_case(nx, ny, nz, nsub, seed=seed, nagents=0, nsteps=3, init=init,
                        dvalue=dvalue, dactive=[1] + [0] * (nsub - 1), with_agents=False)


TINY = {
    "tiny_40x36x30x2": (40, 36, 30, 2, 31, False),
    "tiny_mixed_70x33x41x2": (70, 33, 41, 2, 32, True),
    "tiny_long_x_700x5x6x1": (700, 5, 6, 1, 33, False),
    "tiny_long_y_9x340x7x1": (9, 340, 7, 1, 34, True),
    "tiny_long_z_5x6x700x1": (5, 6, 700, 1, 35, False),
}


@pytest.mark.parametrize("name", list(TINY))
@pytest.mark.parametrize("launch", [None, (0, 8, 16)])
def test_fp64_tiny_numerators_bitwise(name, launch):
    """The exact-division redo (numerators below 2^-960, outside the corrected reciprocal's
    range) in every sweep, with head recomputation on the long lines and small line stores,
    bit for bit against the oracle (IEEE division, subnormals kept)."""
    case = _tiny_case(*TINY[name])
    print(f"[gpu] tiny numerators {name} launch={launch}", flush=True)
    ref = orc.run(case)
    assert np.count_nonzero(ref) and np.min(np.abs(ref[ref != 0])) < 2.0 ** -960
    s = _solver(case, "fp64")
    if launch is not None:
        s.set_launch(-1, *launch)
    s.run(case.nsteps)
    got = s.download()
    s.close()
    if not np.array_equal(got, ref):
        bad = np.argwhere(got != ref)
        pytest.fail(f"{name}: {len(bad)} values differ, first {bad[:5].tolist()}")


@pytest.mark.parametrize("shape", [(40, 36, 30, 2), (700, 5, 6, 1), (5, 6, 700, 1)])
def test_fp64_nonfinite_values_as_the_reference(shape):
    """Infinite values in the field: the reference's division gives +-inf (and inf - inf in
    the back substitution NaN), which the corrected reciprocal alone would turn into NaN
    earlier; the sweeps take the exact-division path for such lines.  Finite values bit for
    bit against the oracle, NaNs at the same places (NaN payloads are not compared)."""
    nx, ny, nz, nsub = shape
    rng = np.random.default_rng(41)
    init = rng.lognormal(0.0, 0.5, (nx, ny, nz, nsub))
    idx = tuple(rng.integers(1, max(2, n - 1), 6) for n in (nx, ny, nz)) + (rng.integers(0, nsub, 6),)
    init[idx] = np.where(np.arange(6) % 2 == 0, np.inf, -np.inf)
    case = cf.make_case(nx, ny, nz, nsub, seed=41, nagents=0, nsteps=2, init=init,
                        dvalue=[38.0] + [0.0] * (nsub - 1), dactive=[1] + [0] * (nsub - 1), with_agents=False)
    print(f"[gpu] non-finite values {shape}", flush=True)
    ref = orc.run(case)
    assert not np.isfinite(ref).all()  # infinities, and NaNs where opposite ones met
    got = gpu_run(case, "fp64")
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    fin = ~np.isnan(ref)
    assert np.array_equal(got[fin], ref[fin])


@pytest.mark.parametrize("shape", [(6, 351, 3, 2), (5, 360, 4, 1), (4, 3, 400, 2), (4, 520, 3, 1), (3, 4, 600, 1)])
def test_yz_table_modes_by_line_length(shape):
    """The y/z sweeps read their Thomas tables with scalar loads or from LDS depending on the
    launch's longest line (352-512 rows in FP64, 352 or more in the FP32 y sweep): lines on
    either side of both limits, with agents, FP64 bit for bit against the oracle and FP32
    within 1e-5 of it."""
    nx, ny, nz, nsub = shape
    case = cf.make_case(nx, ny, nz, nsub, seed=51, nagents=60, nsteps=3, shared_voxels=True)
    print(f"[gpu] y/z table modes {shape}", flush=True)
    ref = orc.run(case)
    assert np.array_equal(gpu_run(case, "fp64"), ref)
    assert rel_err(gpu_run(case, "fp32"), ref)[0] <= 1e-5
