"""The benchmark workloads at their full sizes (biofvmb.workloads), each against what pins it:

* drop-in FP64, 400^3 x 4 with 1M agents, 200 steps: the reference's own run of the same
  inputs (tests/golden/dropin_400cube_x4.npz: 4096 sampled voxels bit for bit, and per
  substrate the float sum, min, max and a position-weighted hash of every bit pattern at steps
  1, 10 and 200); the FP32-storage variant within 1e-5 (max relative error) of it;
* batch of 64 uneven microenvironments: every member of the autotuned, graph-replayed FP64
  batch bit for bit against a single context running the member alone, the smallest members
  against the oracle, and FP32 storage within 1e-5 of FP64 per member over 100 steps;
* adjoint at 128^3 x 2 with 10k agents, 50 steps, checkpoints every 10: the forward run bit for
  bit against the oracle, gradients in rho_0, secretion and uptake rates within 1e-6 of
  central differences of that forward map on 32 seeded coordinates; FP32 gradients reported;
* large FP32, 640^3 x 8 with 10M agents: within 1e-5 of the FP64 run of the same case on the
  same GPU.
"""
import os
import sys

import numpy as np
import pytest

import casefile as cf
import oracle as orc

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_golden as golden  # noqa: E402

pytestmark = pytest.mark.gpu


def _rel(a, b):
    r = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    return float(r.max()), float(np.sqrt(np.mean(r * r)))


# ------------------------------------------------------------------------- drop-in at scale
def _dropin_solver(precision):
    from biofvmb import LODSolver, workloads as wl
    w = wl.DROPIN
    init, pos, vol, S, U, T = wl.dropin_inputs()
    s = LODSolver(w["n"], w["n"], w["n"], w["nsub"], wl.DX, precision)
    s.set_coefficients(w["D"], w["decay"], wl.DT)
    s.set_dirichlet(w["dvalue"], w["dactive"])
    s.upload(init)
    del init
    voxel = s.voxel_indices(pos)
    voxel[voxel < 0] = 0  # none here: the sphere lies inside the domain
    s.set_agent_rates(voxel, vol, S, U, T, wl.DT)
    return s, voxel


@pytest.mark.skipif(not os.path.exists(os.path.join(golden.HERE, "dropin_400cube_x4.npz")),
                    reason="full-size drop-in fixture not generated (make_golden.py --dropin)")
def test_dropin_full_size_fp64_matches_reference_and_fp32_error():
    from biofvmb import workloads as wl
    ref = golden.load_dropin()
    idx = ref["sample_voxels"]
    print("[gpu] drop-in 400^3 x 4, 1M agents, 200 steps: fp64", flush=True)
    s, voxel = _dropin_solver("fp64")
    assert wl.field_checksums(voxel.astype(np.float64)[:, None])[0][3] == int(ref["agent_voxel_hash"])
    done = 0
    rho64 = None
    for step in [int(v) for v in ref["snaps"]]:
        s.run(step - done)
        done = step
        rho64 = s.download().reshape(-1, 4)
        assert np.array_equal(rho64[idx], ref[f"samples_{step}"]), f"step {step}: sampled voxels differ"
        cs = wl.field_checksums(rho64)
        assert np.array_equal(np.array([c[:3] for c in cs]), ref[f"sums_{step}"]), f"step {step}: sums"
        assert [c[3] for c in cs] == [int(h) for h in ref[f"hash_{step}"]], f"step {step}: bit hash"
        print(f"step {step}: bitwise (samples, sums, min, max, hash)", flush=True)
    s.close()
    print("[gpu] drop-in fp32 storage, 200 steps", flush=True)
    s, _ = _dropin_solver("fp32")
    s.run(done)
    rho32 = s.download().reshape(-1, 4)
    s.close()
    mx, rms = _rel(rho32, rho64)
    print(f"fp32 vs reference fp64 after {done} steps: max rel {mx:.3e}, rms {rms:.3e}")
    assert mx <= 1e-5


# ------------------------------------------------------------------------- batch of 64
def _batch(precision, members, tune=True):
    from biofvmb import Batch, workloads as wl
    b = Batch([m["shape"] for m in members], wl.DX, precision)
    for i, m in enumerate(members):
        dvalue, dactive, init, vox, t1, t2 = wl.batch_member_setup(i, m)
        b.set_coefficients(i, m["D"], m["decay"], wl.DT)
        b.set_dirichlet(i, dvalue, dactive)
        b.upload(i, init)
        b.set_agents(i, vox, t1, t2)
    if tune:
        b.autotune(reps=1)
    return b


def test_batch_members_equal_single_contexts_and_oracle():
    from biofvmb import LODSolver, workloads as wl
    members = wl.batch_members()
    steps = 20
    print(f"[gpu] batch of {len(members)} members, fp64, autotuned, graph, {steps} steps", flush=True)
    b = _batch("fp64", members)
    b.run(steps, graph=True)
    b.sync()
    for i, m in enumerate(members):
        dvalue, dactive, init, vox, t1, t2 = wl.batch_member_setup(i, m)
        s = LODSolver(*m["shape"], wl.DX, "fp64")
        s.set_coefficients(m["D"], m["decay"], wl.DT)
        s.set_dirichlet(dvalue, dactive)
        s.upload(init)
        s.set_agents(vox, t1, t2)
        s.run(steps)
        assert np.array_equal(b.download(i), s.download()), f"member {i} {m['shape']}"
        s.close()
    b.close()
    # the smallest members against the oracle, 2 steps
    order = sorted(range(len(members)), key=lambda i: np.prod(members[i]["shape"]))[:8]
    b = _batch("fp64", members, tune=False)
    b.run(2, graph=True)
    for i in order:
        m = members[i]
        nx, ny, nz, ns = m["shape"]
        dvalue, dactive, init, vox, t1, t2 = wl.batch_member_setup(i, m)
        na = len(m["pos"])
        case = cf.Case(nx, ny, nz, ns, wl.DX, wl.DT, m["D"], m["decay"], dvalue, dactive.astype(np.int32), init,
                       m["pos"], np.full(na, wl.AGENT_VOLUME), m["S"], m["U"], np.ones((na, ns)), 2, [2], True)
        want = orc.run(case, tables=(t1, t2))
        assert np.array_equal(b.download(i), want), f"member {i} {m['shape']} vs the oracle"
    b.close()


def test_batch_fp32_error_per_member():
    from biofvmb import workloads as wl
    members = wl.batch_members()
    steps = 100
    print(f"[gpu] batch of {len(members)} members, fp32 vs fp64, {steps} steps", flush=True)
    b64 = _batch("fp64", members)
    b32 = _batch("fp32", members)
    b64.run(steps, graph=True)
    b32.run(steps, graph=True)
    worst = 0.0
    for i, m in enumerate(members):
        mx, rms = _rel(b32.download(i), b64.download(i))
        worst = max(worst, mx)
        print(f"member {i:2d} {m['shape']}: fp32 max rel {mx:.3e}, rms {rms:.3e}")
        assert mx <= 1e-5, i
    print(f"worst member: {worst:.3e}")
    b64.close()
    b32.close()


# ------------------------------------------------------------------------- adjoint at 128^3 x 2
def _adjoint_case():
    from biofvmb import workloads as wl
    n, ns, na = 128, 2, 10_000
    rng = np.random.default_rng(5)
    init = rng.lognormal(0.0, 0.5, (n, n, n, ns))
    pos = rng.uniform(0, n * wl.DX, (na, 3))
    S, U = rng.uniform(0, 5, (na, ns)), rng.uniform(0, 20, (na, ns))
    return cf.Case(n, n, n, ns, wl.DX, wl.DT, np.array([1e5, 1e3]), np.array([0.1, 0.01]), np.array([0.0, 0.0]),
                   np.array([0, 0]), init, pos, np.full(na, wl.AGENT_VOLUME), S, U, np.ones((na, ns)), 50, [50], True)


def _forward(c, precision, rho0=None, S=None, U=None):
    """rho_T of the library's forward map for (rho0, S, U) (FP64: the reference's map)."""
    from biofvmb import LODSolver
    s = LODSolver(c.nx, c.ny, c.nz, c.nsub, c.dx, precision)
    s.set_coefficients(c.D, c.decay, c.dt)
    s.set_dirichlet(c.dvalue, np.asarray(c.dactive, np.uint8))
    s.upload(c.init if rho0 is None else rho0)
    s.set_agent_rates(orc.voxel_indices(c), c.volume, c.S if S is None else S, c.U if U is None else U, c.T, c.dt)
    s.run(c.nsteps)
    out = s.download()
    s.close()
    return out


def test_adjoint_full_size_matches_finite_differences():
    import torch
    from biofvmb import LODSolver, autograd as ag
    c = _adjoint_case()
    print("[gpu] adjoint 128^3 x 2, 10k agents, 50 steps, checkpoint 10", flush=True)
    base = _forward(c, "fp64")
    assert np.array_equal(base, orc.run(c)), "GPU FP64 forward differs from the oracle"
    rng = np.random.default_rng(31)
    pS, pU = c.S * rng.uniform(0.9, 1.1, c.S.shape), c.U * rng.uniform(0.9, 1.1, c.U.shape)
    tgt = _forward(c, "fp64", S=pS, U=pU)  # target: a run with perturbed rates
    dV = (c.dx * c.dx) * c.dx
    voxel = orc.voxel_indices(c)
    grads = {}
    for prec in ("fp64", "fp32"):
        s = LODSolver(c.nx, c.ny, c.nz, c.nsub, c.dx, prec)
        s.set_coefficients(c.D, c.decay, c.dt)
        s.set_dirichlet(c.dvalue, np.asarray(c.dactive, np.uint8))
        dev = "cuda"
        rho0 = torch.tensor(c.init, device=dev, requires_grad=True)
        S = torch.tensor(c.S, device=dev, requires_grad=True)
        U = torch.tensor(c.U, device=dev, requires_grad=True)
        t1, t2 = ag.agent_tables(torch.tensor(c.volume, device=dev), S, U, torch.tensor(c.T, device=dev), c.dt, dV)
        rT = ag.lod_run(s, rho0, torch.tensor(voxel, device=dev), t1, t2, c.nsteps, 10)
        if prec == "fp64":
            assert np.array_equal(rT.detach().cpu().numpy(), base)
        ((rT - torch.tensor(tgt, device=dev)) ** 2).sum().backward()
        grads[prec] = [x.grad.cpu().numpy() for x in (rho0, S, U)]
        s.close()
    gr, gS, gU = grads["fp64"]

    def fd_dir(f, x, e, h):
        a, b = f(x + h * e), f(x - h * e)
        return float(np.sum((a - b) * (a + b - 2.0 * tgt))) / (2 * h)

    coords = [("rho", tuple(int(rng.integers(0, n)) for n in c.init.shape)) for _ in range(16)]
    coords += [("S", (int(rng.integers(0, c.nagents)), int(rng.integers(0, c.nsub)))) for _ in range(8)]
    coords += [("U", (int(rng.integers(0, c.nagents)), int(rng.integers(0, c.nsub)))) for _ in range(8)]
    worst, worst32 = 0.0, 0.0
    for kind, idx in coords:
        if kind == "rho":
            e = np.zeros(c.init.shape)
            e[idx] = 1.0
            fd = fd_dir(lambda r: _forward(c, "fp64", rho0=r), c.init, e, 1e-3)  # L quadratic in rho0
            ad, ad32 = gr[idx], grads["fp32"][0][idx]
        else:
            x0 = c.S if kind == "S" else c.U
            e = np.zeros(x0.shape)
            e[idx] = max(abs(x0[idx]), 1e-3)
            f = (lambda x: _forward(c, "fp64", S=x)) if kind == "S" else (lambda x: _forward(c, "fp64", U=x))
            # Richardson-extrapolated central differences (truncation O(h^4)) at relative steps
            # of 2e-2 and 4e-2.  A rate enters through k*S, k*U (k = dt V/dV ~ 7e-4), compounded
            # over 50 steps: the truncation scales as (50 k h x)^4, ~1e-8 here, and these steps
            # lift the loss difference of the smallest single-rate gradients (~1e-9) well above
            # the rounding of rho_T, which steps of 4e-3 left at ~1e-6 relative
            fd = (4 * fd_dir(f, x0, e, 2e-2) - fd_dir(f, x0, e, 4e-2)) / 3
            k = 1 if kind == "S" else 2
            ad, ad32 = grads["fp64"][k][idx] * e[idx], grads["fp32"][k][idx] * e[idx]
        rel = abs(fd - ad) / max(abs(fd), 1e-300)
        rel32 = abs(fd - ad32) / max(abs(fd), 1e-300)
        if abs(fd) > 1e-12:
            worst, worst32 = max(worst, rel), max(worst32, rel32)
        print(f"{kind} {idx}: fd {fd:.10e} adjoint {ad:.10e} rel {rel:.2e} (fp32 {rel32:.2e})")
        assert abs(fd - ad) <= 1e-6 * abs(fd) or fd == ad == 0.0, (kind, idx, fd, ad)
    print(f"32 coordinates: fp64 worst relative error {worst:.2e}, fp32 {worst32:.2e}")


# ------------------------------------------------------------------------- large FP32
def test_large_fp32_against_fp64_on_the_gpu():
    import torch
    from biofvmb import LODSolver, workloads as wl
    n, ns = wl.LARGE["n"], wl.LARGE["nsub"]
    steps = 10
    D, decay, dvalue, dactive, pos, S, U = wl.large_inputs()
    na = len(pos)
    g = torch.Generator(device="cuda").manual_seed(7)
    init = torch.exp(0.5 * torch.randn((n, n, n, ns), generator=g, device="cuda", dtype=torch.float64))
    out = {}
    for prec in ("fp64", "fp32"):
        print(f"[gpu] 640^3 x 8, 10M agents, {steps} steps: {prec}", flush=True)
        s = LODSolver(n, n, n, ns, wl.DX, prec)
        s.set_coefficients(D, decay, wl.DT)
        s.set_dirichlet(dvalue, dactive)
        s.upload(init)
        vox = s.voxel_indices(pos)
        vox[vox < 0] = 0
        s.set_agent_rates(vox, np.full(na, wl.AGENT_VOLUME), S, U, np.ones((na, ns)), wl.DT)
        s.run(steps)
        out[prec] = s.download(torch.empty_like(init))
        s.close()
    del init
    b = out["fp64"]
    r = (out["fp32"] - b).abs_() / b.abs().clamp_min_(1e-300)
    mx = float(r.max())
    rms = float(torch.sqrt(torch.mean(r * r)))
    print(f"640^3 x 8 fp32 vs fp64 after {steps} steps: max rel {mx:.3e}, rms {rms:.3e}")
    assert mx <= 1e-5
