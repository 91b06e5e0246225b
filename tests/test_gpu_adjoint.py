"""Adjoint of the time loop (bfvmb_adjoint_forward / bfvmb_adjoint_backward) for the loss
L = sum((rho_T - target)^2) after nsteps of (sources/sinks, LOD step).  Gradients are checked against central finite differences
of the oracle's forward run (the reference's arithmetic), so a wrong transpose, a missed
Dirichlet face or a wrong agent order shows up as a mismatch with the reference's own map.

Tolerances: FP64 directional derivatives within 1e-6 relative (L is quadratic in rho_0,
so central differences are exact up to rounding; in temp1/temp2 the truncation error at a
relative step of 1e-5 is ~1e-10; rho_T is affine in temp1, so those differences are exact
up to rounding too).  FP32 storage: gradients within 1e-4 of FP64 (norm-wise).
"""
import numpy as np
import pytest

import casefile as cf
import oracle as orc

pytestmark = pytest.mark.gpu


def _case(nsteps=7, seed=3, **kw):
    args = dict(seed=seed, nagents=40, nsteps=nsteps, shared_voxels=True, outside=1, dactive=[1, 0, 1])
    args.update(kw)
    return cf.make_case(9, 8, 11, 3, **args)


def _solver(c, precision="fp64", tables=None):
    from biofvmb import LODSolver
    s = LODSolver(c.nx, c.ny, c.nz, c.nsub, c.dx, precision)
    s.set_coefficients(c.D, c.decay, c.dt)
    dv, da = cf.effective_dirichlet(c)
    s.set_dirichlet(dv, da)
    s.upload(c.init)
    voxel, t1, t2 = orc.agent_tables(c)
    if tables is not None:
        t1, t2 = tables
    s.set_agents(voxel, t1, t2)
    return s


def _loss(c, target, rho=None, tables=None):
    r = orc.run(c, rho=rho, tables=tables)
    return float(np.sum((r - target) ** 2))


def _target(c, seed=11):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 2.0, c.init.shape)


def _fd_dir(f, x, v, eps, tgt):
    """Central difference of L = sum((f(x) - tgt)^2) along v, f returning rho_T; the
    difference of the two losses is formed as sum((a - b)(a + b - 2 tgt)), which does not
    cancel."""
    a, b = f(x + eps * v), f(x - eps * v)
    return float(np.sum((a - b) * (a + b - 2.0 * tgt))) / (2 * eps)


def _adjoint(s, nsteps, tgt, every=10):
    """Squared loss and its gradients through the C ABI: forward, seed 2 (rho_T - target),
    backward."""
    s.adjoint_forward(nsteps, every)
    rT = s.download()
    g0, g1, g2 = s.adjoint_backward(2.0 * (rT - tgt))
    return float(np.sum((rT - tgt) ** 2)), g0, g1, g2, rT


def test_loss_and_field_restored():
    c = _case()
    tgt = _target(c)
    print("[gpu] adjoint loss 9x8x11x3", flush=True)
    s = _solver(c)
    L0, g0, g1, g2, rT = _adjoint(s, c.nsteps, tgt, every=3)
    assert np.array_equal(rT, orc.run(c))  # the recorded forward run is the reference's map
    assert L0 == _loss(c, tgt)
    # backward leaves the context at rho_T with the record valid: a second seed works
    assert np.array_equal(s.download(), rT)
    h0, h1, h2 = s.adjoint_backward(2.0 * (rT - tgt))
    assert np.array_equal(h0, g0) and np.array_equal(h1, g1) and np.array_equal(h2, g2)
    # linear in the seed
    k0, _, _ = s.adjoint_backward(-3.0 * (rT - tgt))
    assert np.linalg.norm(k0 + 1.5 * g0) <= 1e-13 * np.linalg.norm(g0)
    # more steps invalidate the record; the context keeps stepping exactly
    s.run(2)
    with pytest.raises(Exception, match="adjoint_backward"):
        s.adjoint_backward(rT)
    assert np.array_equal(s.download(), orc.run(c, nsteps=c.nsteps + 2))
    s.close()


def test_grad_rho0_matches_finite_differences():
    c = _case()
    tgt = _target(c)
    s = _solver(c)
    L, g0, _, _, _ = _adjoint(s, c.nsteps, tgt, every=3)
    s.close()
    rng = np.random.default_rng(5)
    f = lambda r: orc.run(c, rho=r)  # noqa: E731
    for k in range(4):
        v = rng.standard_normal(c.init.shape)
        fd = _fd_dir(f, c.init, v, 1e-3, tgt)
        ad = float(np.sum(g0 * v))
        print(f"dir {k}: fd {fd:.12e} adjoint {ad:.12e}")
        assert abs(fd - ad) <= 1e-6 * abs(fd)
    # single entries: an active Dirichlet face (gradient exactly 0 for an active substrate),
    # an interior voxel, an occupied voxel
    voxel, _, _ = orc.agent_tables(c)
    idx = np.unravel_index(int(voxel[0]), c.init.shape[:3])
    assert np.all(g0[0, :, :, 0] == 0.0) and np.all(g0[:, :, -1, 2] == 0.0)
    for pt in [(4, 3, 5, 1), (4, 3, 5, 0), tuple(idx) + (2,)]:
        e = np.zeros(c.init.shape)
        e[pt] = 1.0
        fd = _fd_dir(f, c.init, e, 1e-3, tgt)
        assert abs(fd - g0[pt]) <= 1e-6 * max(abs(fd), 1e-12), (pt, fd, g0[pt])


def test_grad_agent_tables_match_finite_differences():
    c = _case()
    tgt = _target(c)
    voxel, t1, t2 = orc.agent_tables(c)
    s = _solver(c)
    L, _, g1, g2, _ = _adjoint(s, c.nsteps, tgt, every=10)
    s.close()
    rng = np.random.default_rng(8)
    f1 = lambda x: orc.run(c, tables=(x, t2))  # noqa: E731
    f2 = lambda x: orc.run(c, tables=(t1, x))  # noqa: E731
    for k in range(3):
        v = rng.standard_normal(t1.shape) * np.abs(t1).max()
        fd = _fd_dir(f1, t1, v, 0.5, tgt)  # rho_T is affine in temp1: L is quadratic
        ad = float(np.sum(g1 * v))
        print(f"temp1 dir {k}: fd {fd:.10e} adjoint {ad:.10e}")
        assert abs(fd - ad) <= 1e-6 * abs(fd)
        v = rng.standard_normal(t2.shape) * np.abs(t2).max()
        fd = _fd_dir(f2, t2, v, 1e-5, tgt)
        ad = float(np.sum(g2 * v))
        print(f"temp2 dir {k}: fd {fd:.10e} adjoint {ad:.10e}")
        assert abs(fd - ad) <= 1e-6 * abs(fd)
    # agents sharing a voxel: each gets its own gradient, in the caller's order
    for a in (0, 1, c.nagents - 1):
        for sub in range(c.nsub):
            e = np.zeros(t1.shape)
            e[a, sub] = max(abs(t1[a, sub]), 1.0)
            fd = _fd_dir(f1, t1, e, 0.5, tgt)
            assert abs(fd - g1[a, sub] * e[a, sub]) <= 1e-6 * max(abs(fd), 1e-9), (a, sub)


def test_checkpoint_interval_does_not_change_results():
    c = _case(nsteps=12)
    tgt = _target(c)
    outs = []
    for k in (1, 5, 12, 40):
        s = _solver(c)
        outs.append(_adjoint(s, c.nsteps, tgt, every=k))
        s.close()
    for o in outs[1:]:
        for a, b in zip(o, outs[0]):
            assert np.array_equal(a, b)


def test_inactive_agents_get_zero_gradient():
    c = _case()
    tgt = _target(c)
    voxel, t1, t2 = orc.agent_tables(c)
    from biofvmb import LODSolver
    s = LODSolver(c.nx, c.ny, c.nz, c.nsub, c.dx, "fp64")
    s.set_coefficients(c.D, c.decay, c.dt)
    s.set_dirichlet(*cf.effective_dirichlet(c))
    s.upload(c.init)
    act = np.ones(c.nagents, dtype=np.uint8)
    act[::3] = 0
    s.set_agents(voxel, t1, t2, active=act)
    _, _, g1, g2, _ = _adjoint(s, c.nsteps, tgt)
    s.close()
    assert np.all(g1[::3] == 0.0) and np.all(g2[::3] == 0.0)
    assert np.any(g1[1::3] != 0.0)


def test_zero_steps_and_no_agents():
    c = _case(nagents=0)
    tgt = _target(c)
    s = _solver(c)
    L, g0, g1, g2, _ = _adjoint(s, 0, tgt)
    assert np.array_equal(g0, 2.0 * (c.init - tgt)) and g1.size == 0
    s.upload(c.init)
    L, g0, _, _, _ = _adjoint(s, 3, tgt, every=2)
    v = np.random.default_rng(2).standard_normal(c.init.shape)
    fd = _fd_dir(lambda r: orc.run(c, nsteps=3, rho=r), c.init, v, 1e-3, tgt)
    assert abs(fd - float(np.sum(g0 * v))) <= 1e-6 * abs(fd)
    s.close()


def test_fp32_gradients_close_to_fp64():
    c = _case()
    tgt = _target(c)
    s64 = _solver(c, "fp64")
    s32 = _solver(c, "fp32")
    L64, a0, a1, a2, _ = _adjoint(s64, c.nsteps, tgt)
    L32, b0, b1, b2, _ = _adjoint(s32, c.nsteps, tgt)
    s64.close()
    s32.close()
    assert abs(L32 - L64) <= 1e-5 * L64
    for a, b in ((a0, b0), (a1, b1), (a2, b2)):
        err = np.linalg.norm(a - b) / np.linalg.norm(a)
        print(f"fp32 gradient relative error {err:.2e}")
        assert err <= 1e-4


def _rates_case(nsteps=20):
    return cf.make_case(20, 18, 16, 2, seed=21, nagents=300, nsteps=nsteps, shared_voxels=True,
                        D=[1e5, 1e3], decay=[0.1, 0.01], dactive=[1, 0], dvalue=[38.0, 0.0])


def test_torch_autograd_field_and_rates():
    """torch.ops.biofvmb.lod_run + agent_tables: forward equals the reference run; the
    backward gradients in rho_0 and in the secretion/uptake rates match central differences
    of the oracle's map on 32 random coordinates (checkpoint every 10 steps)."""
    import torch
    from biofvmb import autograd as ag
    c = _rates_case()
    dV = (c.dx * c.dx) * c.dx
    voxel, _, _ = orc.agent_tables(c)
    # target: the reference run with rates perturbed by up to 10%
    rng = np.random.default_rng(31)
    pS, pU = c.S * rng.uniform(0.9, 1.1, c.S.shape), c.U * rng.uniform(0.9, 1.1, c.U.shape)
    tgt = orc.run(c, tables=orc.agent_constants(c.dt, c.volume, dV, pS, pU, c.T))
    s = _solver(c)
    dev = "cuda"
    rho0 = torch.tensor(c.init, device=dev, requires_grad=True)
    S = torch.tensor(c.S, device=dev, requires_grad=True)
    U = torch.tensor(c.U, device=dev, requires_grad=True)
    t1, t2 = ag.agent_tables(torch.tensor(c.volume, device=dev), S, U, torch.tensor(c.T, device=dev), c.dt, dV)
    rT = ag.lod_run(s, rho0, torch.tensor(voxel, device=dev), t1, t2, c.nsteps, 10)
    assert np.array_equal(rT.detach().cpu().numpy(), orc.run(c))
    loss = ((rT - torch.tensor(tgt, device=dev)) ** 2).sum()
    loss.backward()
    gr, gS, gU = (x.grad.cpu().numpy() for x in (rho0, S, U))
    s.close()

    def field_of_rates(Sx, Ux):
        return orc.run(c, tables=orc.agent_constants(c.dt, c.volume, dV, Sx, Ux, c.T))

    worst = 0.0
    coords = [("rho", tuple(rng.integers(0, n) for n in c.init.shape)) for _ in range(16)]
    coords += [("S", (int(rng.integers(0, c.nagents)), int(rng.integers(0, c.nsub)))) for _ in range(8)]
    coords += [("U", (int(rng.integers(0, c.nagents)), int(rng.integers(0, c.nsub)))) for _ in range(8)]
    for kind, idx in coords:
        if kind == "rho":
            e = np.zeros(c.init.shape)
            e[idx] = 1.0
            fd = _fd_dir(lambda r: orc.run(c, rho=r), c.init, e, 1e-3, tgt)
            ad = gr[idx]
        else:
            x0 = c.S if kind == "S" else c.U
            e = np.zeros(x0.shape)
            e[idx] = max(abs(x0[idx]), 1e-3)
            f = (lambda x: field_of_rates(x, c.U)) if kind == "S" else (lambda x: field_of_rates(c.S, x))
            # Richardson-extrapolated central differences (truncation O(h^4)) at relative
            # steps 4e-3 / 8e-3: the gradients of single rates span four decades, and
            # smaller steps leave the rounding of rho_T above 1e-6 for the smallest
            fd = (4 * _fd_dir(f, x0, e, 4e-3, tgt) - _fd_dir(f, x0, e, 8e-3, tgt)) / 3
            ad = (gS if kind == "S" else gU)[idx] * e[idx]
        rel = abs(fd - ad) / max(abs(fd), 1e-300)
        worst = max(worst, rel if abs(fd) > 1e-12 else 0.0)
        assert abs(fd - ad) <= 1e-6 * abs(fd) or fd == ad == 0.0, (kind, idx, fd, ad)
    print(f"32 coordinates: worst relative error {worst:.2e}")


def _bulk_case():
    c = _case()
    rng = np.random.default_rng(21)
    shape = c.init.shape
    c.bulk = (rng.uniform(0.0, 40.0, shape), rng.uniform(0.0, 3.0, shape), rng.uniform(0.0, 60.0, shape))
    return c


def test_gradients_with_bulk_rates_match_finite_differences():
    """With bulk sources/sinks in every step (rho <- (rho + dt T S)/(1 + dt (U + S)) before the
    agents), the reverse pass applies its adjoint and records the agents' inputs after it."""
    c = _bulk_case()
    tgt = _target(c)
    voxel, t1, t2 = orc.agent_tables(c)
    print("[gpu] adjoint with bulk rates", flush=True)
    s = _solver(c)
    s.set_bulk_rates(*c.bulk)
    L, g0, g1, g2, rT = _adjoint(s, c.nsteps, tgt, every=3)
    s.close()
    assert np.array_equal(rT, orc.run(c))
    rng = np.random.default_rng(9)
    f0 = lambda r: orc.run(c, rho=r)  # noqa: E731
    f1 = lambda x: orc.run(c, tables=(x, t2))  # noqa: E731
    f2 = lambda x: orc.run(c, tables=(t1, x))  # noqa: E731
    for k in range(3):
        v = rng.standard_normal(c.init.shape)
        fd, ad = _fd_dir(f0, c.init, v, 1e-3, tgt), float(np.sum(g0 * v))
        print(f"bulk rho0 dir {k}: fd {fd:.12e} adjoint {ad:.12e}")
        assert abs(fd - ad) <= 1e-6 * abs(fd)
        v = rng.standard_normal(t1.shape) * np.abs(t1).max()
        fd, ad = _fd_dir(f1, t1, v, 0.5, tgt), float(np.sum(g1 * v))
        assert abs(fd - ad) <= 1e-6 * abs(fd)
        v = rng.standard_normal(t2.shape) * np.abs(t2).max()
        fd, ad = _fd_dir(f2, t2, v, 1e-5, tgt), float(np.sum(g2 * v))
        assert abs(fd - ad) <= 1e-6 * abs(fd)


def test_repeated_backward_calls_replay_the_same_graphs():
    """The reverse-pass state and its segment graphs are kept with the record: several seeds
    in a row, with eager work on the context in between, give what fresh records give."""
    c = _case(nsteps=11)
    tgt = _target(c)
    print("[gpu] adjoint repeated backward", flush=True)
    s = _solver(c)
    s.adjoint_forward(c.nsteps, 4)
    rT = s.download()
    seeds = [2.0 * (rT - tgt), np.random.default_rng(4).standard_normal(rT.shape), 2.0 * (rT - tgt)]
    outs = [s.adjoint_backward(w) for w in seeds]
    s.close()
    for k, w in enumerate(seeds):
        f = _solver(c)
        f.adjoint_forward(c.nsteps, 4)
        want = f.adjoint_backward(w)
        f.close()
        for a, b in zip(outs[k], want):
            assert np.array_equal(a, b)


def test_new_records_reuse_state_only_when_valid():
    """Records taken one after another on one context (other start parity, other checkpoint
    interval, new agent tables, bulk rates added) give what a fresh context gives."""
    c = _case(nsteps=9)
    tgt = _target(c)
    voxel, t1, t2 = orc.agent_tables(c)
    print("[gpu] adjoint records in a row", flush=True)

    def fresh(nsteps, every, tables=None, bulk=None, steps_before=0):
        f = _solver(c, tables=tables)
        if bulk is not None:
            f.set_bulk_rates(*bulk)
        f.run(steps_before)
        out = _adjoint(f, nsteps, tgt, every=every)
        f.close()
        return out

    bulk = _bulk_case().bulk
    s = _solver(c)
    plan = [(9, 3, None, None, 0), (9, 3, None, None, 1), (5, 2, None, None, 0), (9, 3, (t1 * 1.5, t2), None, 0),
            (9, 3, (t1 * 1.5, t2), bulk, 0)]
    done = 0
    for nsteps, every, tables, bk, before in plan:
        s.upload(c.init)
        if tables is not None:
            s.set_agents(voxel, *tables)
        if bk is not None:
            s.set_bulk_rates(*bk)
        s.run(before)
        got = _adjoint(s, nsteps, tgt, every=every)
        want = fresh(nsteps, every, tables, bk, before)
        for a, b in zip(got, want):
            assert np.array_equal(a, b), (nsteps, every, before)
        done += 1
    s.close()
    assert done == len(plan)


def test_device_pointers_give_the_host_results():
    """bfvmb_upload/download_density_device and bfvmb_adjoint_backward_device (CUDA tensors,
    no host transfer) against the host-pointer calls, bit for bit, with inactive agents."""
    import torch
    c = _case(nsteps=8)
    tgt = _target(c)
    voxel, t1, t2 = orc.agent_tables(c)
    act = np.ones(c.nagents, np.uint8)
    act[::4] = 0
    print("[gpu] adjoint device pointers", flush=True)
    outs = []
    for dev in (False, True):
        from biofvmb import LODSolver
        s = LODSolver(c.nx, c.ny, c.nz, c.nsub, c.dx, "fp64")
        s.set_coefficients(c.D, c.decay, c.dt)
        s.set_dirichlet(*cf.effective_dirichlet(c))
        s.set_agents(voxel, t1, t2, act)
        if dev:
            s.upload(torch.tensor(c.init, device="cuda:0"))
        else:
            s.upload(c.init)
        s.adjoint_forward(c.nsteps, 3)
        if dev:
            rT = s.download(torch.empty(c.init.shape, dtype=torch.float64, device="cuda:0"))
            g = s.adjoint_backward(2.0 * (rT - torch.tensor(tgt, device="cuda:0")))
            outs.append([rT.cpu().numpy()] + [x.cpu().numpy() for x in g])
        else:
            rT = s.download()
            outs.append([rT] + list(s.adjoint_backward(2.0 * (rT - tgt))))
        s.close()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
    assert np.all(outs[1][2][::4] == 0.0) and np.any(outs[1][2][1::4] != 0.0)


def test_torch_op_cuda_and_cpu_tensors_agree():
    import torch
    from biofvmb import autograd as ag
    c = _case(nsteps=6)
    tgt = _target(c)
    voxel, t1, t2 = orc.agent_tables(c)
    print("[gpu] torch op cuda vs cpu tensors", flush=True)
    res = []
    for dev in ("cpu", "cuda"):
        s = _solver(c)
        r0 = torch.tensor(c.init, device=dev, requires_grad=True)
        a1 = torch.tensor(t1, device=dev, requires_grad=True)
        a2 = torch.tensor(t2, device=dev, requires_grad=True)
        rT = ag.lod_run(s, r0, torch.tensor(voxel, device=dev), a1, a2, c.nsteps, 4)
        ((rT - torch.tensor(tgt, device=dev)) ** 2).sum().backward()
        res.append([x.detach().cpu().numpy() for x in (rT, r0.grad, a1.grad, a2.grad)])
        s.close()
    for a, b in zip(*res):
        assert np.array_equal(a, b)
