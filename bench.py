"""Benchmark of the BioFVM-B hot path on one MI355X.

Workload (drop-in FP64 at scale): 400x400x400 voxels, dx = 20 um, 4 substrates with
D = {1e5, 1e3, 2.5e3, 5e4} um^2/min and decay = {0.1, 0.01, 0.4, 1e-3} /min, Dirichlet
38.0 on the faces of substrate 0 only, initial fields lognormal(0, 0.5), 1M agents uniform
in a centred sphere of radius 3 mm with uptake U~U[0,20], secretion S~U[0,5] toward
T = 1, dt = 0.01 min.  One step = cell sources/sinks + one LOD step (the demo time loop).
The FP32-storage variant of the same case is timed beside it.

Metric: voxel-substrate updates/s per LOD step = 400^3 * 4 * steps / wall time, with the
field resident in HBM when the timed region starts.  The roofline entry is for the
slowest sweep kernel: algorithmic bytes per launch = 400^3 * 4 * (read + write) * 8 B,
divided by its average launch time from HIP events on the stream the kernel runs on.
The CPU baseline times the reference itself (oracle/_ref, 1 rank x N OpenMP threads) on
a bounded sample of the same case, or the C restatement if the reference binary is absent.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "biofvm-b_amd"))

N, NSUB, DX, DT = 400, 4, 20.0, 0.01
D = [1e5, 1e3, 2.5e3, 5e4]
DECAY = [0.1, 0.01, 0.4, 1e-3]
N_AGENTS = 1_000_000
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_inputs(n=N, nagents=N_AGENTS, seed=2024):
    from biofvmb import workloads
    return workloads.dropin_inputs(n=n, nagents=nagents, seed=seed)


def setup_solver(precision, init, pos, vol, S, U, T, stream):
    from biofvmb import LODSolver
    s = LODSolver(N, N, N, NSUB, DX, precision)
    s.set_stream(stream)
    s.set_coefficients(D, DECAY, DT)
    s.set_dirichlet([38.0, 0.0, 0.0, 0.0], [1, 0, 0, 0])
    s.upload(init)
    voxel = s.voxel_indices(pos)
    voxel[voxel < 0] = 0
    s.set_agent_rates(voxel, vol, S, U, T, DT)
    return s


def time_steps(s, torch, steps, warmup):
    s.run(warmup)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    a.record()
    s.run(steps)
    b.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return a.elapsed_time(b) / 1e3, wall


def time_phases(s, torch, count=10):
    """Kernel times inside the step: `count` steps launched kernel by kernel as `run()` launches
    them (FP64: the x sweep with the cell sources and sinks in its launch, then the y and z
    sweeps; FP32: the agent kernel, then the three sweeps) with a HIP event on the library's
    stream between consecutive kernels, so each kernel runs after the kernel it follows in a
    step (a sweep launched right after another launch of itself runs 3-7% faster than in the
    step).  Returns ms per launch of each kernel and "step", their sum."""
    if s.precision == "fp64":
        phases = ((4, "sweep_x_with_agents"), (2, "sweep_y"), (3, "sweep_z"))
    else:
        phases = ((0, "sources_sinks"), (1, "sweep_x"), (2, "sweep_y"), (3, "sweep_z"))
    for p, _ in phases:
        s.launch_phase(p, 1)
    n = len(phases)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(n * count + 1)]
    ev[0].record()
    for k in range(count):
        for i, (p, _) in enumerate(phases):
            s.launch_phase(p, 1)
            ev[n * k + i + 1].record()
    torch.cuda.synchronize()
    out = {name: sum(ev[n * k + i].elapsed_time(ev[n * k + i + 1]) for k in range(count)) / count
           for i, (_, name) in enumerate(phases)}
    out["step"] = ev[0].elapsed_time(ev[n * count]) / count
    return out


def bench_batched(torch, stream, steps=20, seed=3):
    """64 independent microenvironments of uneven sizes: nx, ny, nz ~ U{32..256}, nsub ~
    U{1..8}, D ~ logU[1e2, 1e5], decay ~ logU[1e-3, 1], U{0..50000} agents each; the
    autotuner picks the launch shapes, then `steps` steps run as one captured graph, in FP64
    and in FP32 storage."""
    from biofvmb import Batch, workloads
    members = workloads.batch_members(seed=seed)
    shapes = [mb["shape"] for mb in members]
    m = len(members)
    total_vs = sum(nx * ny * nz * ns for nx, ny, nz, ns in shapes)
    out = {"members": m, "voxel_substrates": total_vs, "steps": steps}
    for prec in ("fp64", "fp32"):
        b = Batch(shapes, DX, prec)
        b.set_stream(stream)
        for i, mb in enumerate(members):
            dvalue, dactive, init, vox, t1, t2 = workloads.batch_member_setup(i, mb)
            b.set_coefficients(i, mb["D"], mb["decay"], DT)
            b.set_dirichlet(i, dvalue, dactive)
            b.upload(i, init)
            b.set_agents(i, vox, t1, t2)
        chosen = b.autotune(reps=2)
        b.run(steps, graph=True)  # capture + warm-up
        torch.cuda.synchronize()
        a, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        b.run(steps, graph=True)
        e.record()
        torch.cuda.synchronize()
        sec = a.elapsed_time(e) / 1e3
        phases = time_phases(b, torch, count=3)  # the tuned shapes, one phase at a time
        b.close()
        out[prec] = {"value": total_vs * steps / sec, "ms_per_step": sec * 1e3 / steps,
                     "autotune": np.asarray(chosen).tolist(), "phases_ms": phases}
        log(f"batched {prec}: {total_vs * steps / sec:.4e} vs/s, {sec * 1e3 / steps:.3f} ms/step")
    return out


def bench_c5(torch, stream, steps=10, nagents=10_000_000, seed=9, precisions=("fp32",), tune=True):
    """The large workload, 640^3 x 8 substrates (log-spaced D and decay, Dirichlet on
    substrates 0-1, 10M agents in a centred 5 mm sphere), initial field 1: FP32 storage (the
    bulk path) and, when asked, FP64, with per-phase kernel times.  With `tune` the
    autotuner picks the launch shapes first (its trial launches restore the field)."""
    from biofvmb import LODSolver, workloads
    n, ns = workloads.LARGE["n"], workloads.LARGE["nsub"]
    Dm, lam, dvalue, dactive, pos, S, U = workloads.large_inputs(nagents=nagents, seed=seed)
    out = {"workload": f"{n}^3x{ns}, {nagents} agents"}
    for prec in precisions:
        s = LODSolver(n, n, n, ns, DX, prec)
        s.set_stream(stream)
        s.set_coefficients(Dm, lam, DT)
        s.set_dirichlet(dvalue, dactive)
        ones = torch.ones((n, n, n, ns), dtype=torch.float64, device="cuda:0")  # no 16.8 GB host array
        s.upload(ones)
        del ones
        vox = s.voxel_indices(pos)
        vox[vox < 0] = 0
        s.set_agent_rates(vox, np.full(nagents, workloads.AGENT_VOLUME), S, U, np.ones((nagents, ns)), DT)
        chosen = s.autotune(reps=2) if tune else None
        gpu_s, wall = time_steps(s, torch, steps, 2)
        phases = time_phases(s, torch, count=3)
        s.close()
        vs = n ** 3 * ns * steps / gpu_s
        eb = 8 if prec == "fp64" else 4
        log(f"c5 {prec}: {vs:.4e} vs/s, {gpu_s * 1e3 / steps:.3f} ms/step, phases {phases}")
        out[prec] = {"value": vs, "ms_per_step": gpu_s * 1e3 / steps, "phases_ms": phases,
                     "autotune": None if chosen is None else np.asarray(chosen).tolist(),
                     "step_hbm_frac": round(n ** 3 * ns * 6 * eb / (gpu_s / steps) / 1e9 / HBM_PEAK_GBS, 4)}
    first = out[precisions[0]]
    out.update(value=first["value"], ms_per_step=first["ms_per_step"], step_hbm_frac=first["step_hbm_frac"])
    return out


def bench_adjoint(torch, stream, n=128, nsteps=50, every=10, nagents=10_000, seed=5):
    """Gradient of sum((rho_T - target)^2) after 50 steps at 128^3 x 2 with 10k agents:
    forward run with checkpoints every 10 steps, then the reverse pass."""
    from biofvmb import LODSolver
    rng = np.random.default_rng(seed)
    ns = 2
    s = LODSolver(n, n, n, ns, DX, "fp64")
    s.set_stream(stream)
    s.set_coefficients([1e5, 1e3], [0.1, 0.01], DT)
    s.set_dirichlet([0.0, 0.0], [0, 0])
    init = rng.lognormal(0.0, 0.5, (n, n, n, ns))
    s.upload(init)
    pos = rng.uniform(0, n * DX, (nagents, 3))
    vox = s.voxel_indices(pos)
    s.set_agent_rates(vox, np.full(nagents, 4.0 / 3.0 * np.pi * 125.0), rng.uniform(0, 5, (nagents, ns)),
                      rng.uniform(0, 20, (nagents, ns)), np.ones((nagents, ns)), DT)
    target = rng.lognormal(0.0, 0.5, (n, n, n, ns))
    vs = n ** 3 * ns * nsteps
    out = {"workload": f"{n}^3x{ns}, {nagents} agents, {nsteps} steps, checkpoint every {every}"}
    # launch shapes by the autotuner: a sweep here has 512 tiles, fewer than the default
    # shapes' wavefront slots, and fewer wavefronts per CU keep whole 128-column lines in LDS
    out["autotune"] = s.autotune().tolist()
    # fields resident in HBM (CUDA tensors: seed and gradients never cross PCIe), then the
    # same through host arrays
    init_d = torch.tensor(init, device="cuda:0")
    target_d = torch.tensor(target, device="cuda:0")
    rT_d = torch.empty_like(init_d)
    for device in (True, False):
        for rep in range(2):  # the first run captures the segment graphs
            s.upload(init_d if device else init)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            s.adjoint_forward(nsteps, every)
            s.sync()
            t1 = time.perf_counter()
            rT = s.download(rT_d) if device else s.download()
            t2 = time.perf_counter()
            s.adjoint_backward(2.0 * (rT - (target_d if device else target)))
            torch.cuda.synchronize()
            t3 = time.perf_counter()
        key = "device" if device else "host_arrays"
        out[key] = {"forward_s": t1 - t0, "backward_s": t3 - t2, "value": vs / ((t1 - t0) + (t3 - t2))}
        log(f"adjoint ({key}): forward {t1 - t0:.4f} s, backward {t3 - t2:.4f} s, {out[key]['value']:.3e} vs/s")
    s.close()
    out["value"] = out["device"]["value"]
    out["unit"] = ("voxel-substrate updates/s (forward + reverse pass, fields in HBM; host_arrays: seed and "
                   "gradients through host memory)")
    return out


def bench_agent_moves(torch, stream, n=400, nagents=1_000_000, reps=10, seed=11):
    """Moving agents in the drop-in workload's domain (400^3 x 4, 1M agents in a 3 mm
    sphere): every rep moves all agents (positions already in HBM) and regroups the
    source/sink list on the device (bfvmb_update_agent_positions)."""
    from biofvmb import LODSolver
    rng = np.random.default_rng(seed)
    ns = 4
    s = LODSolver(n, n, n, ns, DX, "fp64")
    s.set_stream(stream)
    s.set_coefficients([1e5, 1e3, 2.5e3, 5e4], [0.1, 0.01, 0.4, 1e-3], DT)
    d = rng.normal(size=(nagents, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos = n * DX / 2.0 + d * (3000.0 * rng.uniform(0, 1, nagents) ** (1.0 / 3.0))[:, None]
    vox = s.voxel_indices(pos)
    s.set_agent_rates(vox, np.full(nagents, 4.0 / 3.0 * np.pi * 125.0), rng.uniform(0, 5, (nagents, ns)),
                      rng.uniform(0, 20, (nagents, ns)), np.ones((nagents, ns)), DT)
    moves = [torch.tensor(pos + rng.normal(0.0, DX, pos.shape), dtype=torch.float64, device="cuda:0")
             for _ in range(reps)]
    s.update_agent_positions(moves[0])  # set-up of the device regrouping state
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for m in moves:
        s.update_agent_positions(m)
    t1 = time.perf_counter()
    s.close()
    sec = (t1 - t0) / reps
    out = {"workload": f"{n}^3x{ns}, {nagents} agents moved by N(0, dx) per axis, positions in HBM",
           "ms_per_update": sec * 1e3, "value": nagents / sec, "unit": "agents regrouped/s (incl. one read-back)"}
    log(f"agent moves: {sec * 1e3:.3f} ms per update of {nagents} agents")
    return out


def bench_per_call(steps=20, precision="fp64"):
    """The drop-in workload's shape (400^3 x 4, 1M agents in the 3 mm sphere) through the C++
    host layer (biofvmb::Microenvironment, no reference code), one step per call as the
    reference's drivers call the slots (simulate_cell_sources_and_sinks, then
    simulate_diffusion_decay), against one run(agents, dt, steps) call; a separate process
    (biofvm-b_amd/lib/bfvmb_driver --bench-dropin)."""
    exe = os.path.join(ROOT, "biofvm-b_amd", "lib", "bfvmb_driver")
    out = subprocess.run([exe, "--bench-dropin", str(N), str(N_AGENTS), str(steps), precision], capture_output=True,
                         text=True, check=True, timeout=600)
    rec = json.loads(out.stdout.strip().splitlines()[-1])
    rec["value"] = N ** 3 * NSUB / (rec["per_call_ms_per_step"] / 1e3)
    rec["unit"] = "voxel-substrate updates/s, one step per call (inputs drawn with std::mt19937_64)"
    log(f"per-call host layer: {rec['per_call_ms_per_step']:.3f} ms/step, run(): {rec['run_ms_per_step']:.3f} ms/step")
    return rec


def traffic_record(dom):
    """HBM bytes per launch of kernel `dom` from the FETCH_SIZE / WRITE_SIZE profile in
    profiles/, with the sha256 of the sweep source it was measured on; null when the source
    has changed since (the counters then describe other kernels)."""
    import hashlib
    tpath = os.path.join(ROOT, "profiles", "traffic_400cube_fp64.json")
    if not os.path.exists(tpath):
        return None, None
    rec = json.load(open(tpath))
    src = os.path.join(ROOT, "biofvm-b_amd", "csrc", "lod.hip")
    cur = hashlib.sha256(open(src, "rb").read()).hexdigest()
    meta = {"profile": "profiles/traffic_400cube_fp64.json", "lod_hip_sha256": rec.get("lod_hip_sha256"),
            "current_lod_hip_sha256": cur, "commit": rec.get("commit")}
    if rec.get("lod_hip_sha256") != cur:
        meta["stale"] = True
        return None, meta
    return rec.get("bytes_per_launch", {}).get(dom), meta


def copy_ceiling(torch, reps=20):
    """Streaming ceilings of this GPU for a field-sized buffer (400^3 x 4 doubles, 2.048 GB):
    device copy (read + write), read-only (sum) and write-only (fill), each timed with HIP
    events over `reps` launches; GB/s counted as bytes moved."""
    n = N ** 3 * NSUB
    a = torch.rand(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    out = {}
    for name, fn, bytes_moved in (("copy", lambda: b.copy_(a), 16 * n), ("read", lambda: a.sum(), 8 * n),
                                  ("write", lambda: b.fill_(1.0), 8 * n)):
        for _ in range(3):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        sec = e0.elapsed_time(e1) / 1e3 / reps
        out[name] = {"GB_per_s": bytes_moved / sec / 1e9, "ms": sec * 1e3, "frac_of_peak": bytes_moved / sec / 1e9 / HBM_PEAK_GBS}
        log(f"{name}: {bytes_moved / sec / 1e9:.1f} GB/s")
    del a, b
    return out


def cpu_baseline(repeats=3):
    """The reference (oracle/_ref) on a bounded sample: 200^3 x 4 with 125k agents (1/8 of
    the domain, same densities), 16 steps, median of three runs (≈20 s of CPU work on 16
    threads); falls back to the C restatement (1 thread)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import casefile as cf
    n, steps = 200, 16
    init, pos, vol, S, U, T = make_inputs(n=n, nagents=N_AGENTS // 8, seed=7)
    case = cf.Case(n, n, n, NSUB, DX, DT, np.array(D), np.array(DECAY), np.array([38.0, 0, 0, 0]),
                   np.array([1, 0, 0, 0]), init, pos, vol, S, U, T, steps, [], True)
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        threads = os.cpu_count() or 1
    threads = max(1, min(threads, 16))
    sample = f"{n}^3x{NSUB}, {N_AGENTS // 8} agents, {steps} steps (sources/sinks + LOD), median of {repeats} runs"
    if os.path.exists(cf.REF_DRIVER):
        try:
            rates = []
            for _ in range(repeats):
                with tempfile.TemporaryDirectory() as d:
                    out = cf.run_reference(case, d, solver="lod", threads=threads, timing=True)
                rec = json.loads(out["stdout"].strip().splitlines()[-1])
                rates.append(n ** 3 * NSUB * steps / (rec["ms_total"] / 1e3))
                log(f"cpu baseline run: {rates[-1]:.3e} vs/s")
            return {"value": float(np.median(rates)), "unit": "voxel-substrate updates/s", "cores": threads,
                    "kind": "reference", "sample": sample + f", 1 rank x {threads} OpenMP threads, LOD_3D solver"}
        except Exception as e:  # reference binary not runnable here: use the restatement
            log(f"reference CPU run failed ({e}); timing the C restatement instead")
    import oracle as orc
    steps = 2
    t0 = time.perf_counter()
    orc.run(case, nsteps=steps)
    dt = time.perf_counter() - t0
    return {"value": n ** 3 * NSUB * steps / dt, "unit": "voxel-substrate updates/s", "cores": 1, "kind": "port",
            "sample": f"{n}^3x{NSUB}, {N_AGENTS // 8} agents, {steps} steps, C restatement, 1 thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip the batched and adjoint workloads")
    ap.add_argument("--no-c5", action="store_true", help="skip the FP32 640^3 x 8 workload")
    ap.add_argument("--c5-only", action="store_true",
                    help="time only the 640^3 x 8 workload, FP32 and FP64, and print its JSON object")
    ap.add_argument("--c5-defaults", action="store_true",
                    help="with --c5-only: the library's default launch shapes instead of the autotuner's")
    ap.add_argument("--batched-only", action="store_true",
                    help="time only the batched workload and print its JSON object (for profiling)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="time only the adjoint workload and print its JSON object (for profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 1 GiB device copy (reference for HBM byte counters)")
    ap.add_argument("--copy-ceiling", action="store_true",
                    help="only time streaming copy / read / write of a field-sized buffer and print them")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the LOD path is built and measured on a single MI355X "
              f"(--gpus {args.gpus} is not supported; multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    import torch
    stream = torch.cuda.current_stream().cuda_stream
    if args.copy_calibration:
        a = torch.ones(128 * 1024 * 1024, dtype=torch.float64, device="cuda")  # 1 GiB
        b = torch.empty_like(a)
        for _ in range(2):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.copy_ceiling:
        print(json.dumps({"copy_ceiling": copy_ceiling(torch), "bytes": N ** 3 * NSUB * 8}), flush=True)
        return
    if args.c5_only:
        print(json.dumps(bench_c5(torch, stream, precisions=("fp32", "fp64"), tune=not args.c5_defaults)), flush=True)
        return
    if args.batched_only:
        print(json.dumps(bench_batched(torch, stream)), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps(bench_adjoint(torch, stream)), flush=True)
        return
    log("generating inputs ...")
    inputs = make_inputs()

    results = {}
    for prec in ("fp64", "fp32"):
        log(f"{prec}: set-up")
        s = setup_solver(prec, *inputs, stream)
        gpu_s, wall = time_steps(s, torch, args.steps, args.warmup)
        phases = time_phases(s, torch, count=20)
        s.close()
        vs = N ** 3 * NSUB * args.steps / gpu_s
        results[prec] = {"value": vs, "ms_per_step": gpu_s * 1e3 / args.steps, "wall_s": wall, "phases_ms": phases}
        log(f"{prec}: {vs:.4e} vs/s, {gpu_s * 1e3 / args.steps:.3f} ms/step, phases {phases}")

    f64 = results["fp64"]
    dom = max(("sweep_x_with_agents", "sweep_y", "sweep_z"), key=lambda k: f64["phases_ms"][k])
    # field read + write; the x sweep's launch also carries the agents' bytes (index, two
    # coefficients per substrate, voxel read + write per substrate)
    bytes_per_launch = N ** 3 * NSUB * 2 * 8
    if dom == "sweep_x_with_agents":
        bytes_per_launch += N_AGENTS * (4 + 2 * NSUB * 8 + 2 * NSUB * 8)
    achieved = bytes_per_launch / (f64["phases_ms"][dom] / 1e3) / 1e9
    traffic, traffic_meta = traffic_record(dom)  # FETCH_SIZE + WRITE_SIZE counters (profiles/)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "algorithmic_bytes_per_launch": bytes_per_launch, "traffic_measurement": traffic_meta}
    step_gbs = N ** 3 * NSUB * 48 / (f64["ms_per_step"] / 1e3) / 1e9
    extra = {}
    if not args.no_extra:
        extra["batched"] = bench_batched(torch, stream)
        b = extra["batched"]
        b["fp32_vs_fp64_batch"] = b["fp32"]["value"] / b["fp64"]["value"]  # the same members in FP64
        b["fp32_vs_fp64_dropin"] = b["fp32"]["value"] / f64["value"]
        extra["adjoint"] = bench_adjoint(torch, stream)
        extra["agent_moves"] = bench_agent_moves(torch, stream)
        extra["per_call_host_layer"] = bench_per_call()
    if not args.no_c5 and not args.no_extra:
        extra["fp32_640cube_x8"] = bench_c5(torch, stream)
    cpu = None if args.no_cpu else cpu_baseline()
    line = {
        "metric": "voxel-substrate updates/s per LOD step (FP64 vs FP32/batched) + % HBM peak",
        "value": f64["value"],
        "unit": "voxel-substrate updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lognormal fields, 1M agents in a 3 mm sphere)",
        "config": {"workload": "drop-in FP64 LOD step + agent sources/sinks, 400^3 voxels x 4 substrates, 1M agents",
                   "grid": [N, N, N], "substrates": NSUB, "agents": N_AGENTS, "dt": DT, "dx": DX},
        "step_hbm_frac": round(step_gbs / HBM_PEAK_GBS, 4),
        "phases_ms": f64["phases_ms"],
        "fp32": {"value": results["fp32"]["value"], "ms_per_step": results["fp32"]["ms_per_step"],
                 "vs_fp64": results["fp32"]["value"] / f64["value"], "phases_ms": results["fp32"]["phases_ms"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
