"""Seeded synthetic inputs of the benchmark workloads, shared by bench.py and the parity tests
so that both see the same bytes.  Each generator is a pure function of its seed (numpy's
PCG64 stream), so a fixture written on one machine matches a run on another.

* drop-in FP64 at scale: 400^3 voxels, dx = 20 um, 4 substrates (D = {1e5, 1e3, 2.5e3, 5e4}
  um^2/min, decay = {0.1, 0.01, 0.4, 1e-3} /min), Dirichlet 38.0 on substrate 0 only,
  lognormal(0, 0.5) initial fields, 1M agents uniform in a centred 3 mm sphere with uptake
  U[0, 20], secretion U[0, 5] toward 1.0;
* batch: 64 microenvironments, nx, ny, nz ~ U{32..256}, nsub ~ U{1..8}, D ~ logU[1e2, 1e5],
  decay ~ logU[1e-3, 1], U{0..50000} agents each;
* adjoint: 128^3 x 2 (D = {1e5, 1e3}, decay = {0.1, 0.01}), 10k agents, 50 steps;
* large FP32: 640^3 x 8 (D log-spaced 1e2..1e5, decay 1e-3..1), Dirichlet on substrates 0-1,
  10M agents in a centred 5 mm sphere.
"""
from __future__ import annotations

import numpy as np

DX, DT = 20.0, 0.01
AGENT_VOLUME = 4.0 / 3.0 * np.pi * 125.0  # radius 5 um, the reference examples' cells

DROPIN = dict(n=400, nsub=4, D=(1e5, 1e3, 2.5e3, 5e4), decay=(0.1, 0.01, 0.4, 1e-3),
              dvalue=(38.0, 0.0, 0.0, 0.0), dactive=(1, 0, 0, 0), nagents=1_000_000, radius=3000.0)
LARGE = dict(n=640, nsub=8, nagents=10_000_000, radius=5000.0)


def sphere_positions(rng, nagents, center, radius):
    """Uniform positions in a ball: isotropic direction, radius ~ R * U^(1/3)."""
    d = rng.normal(size=(nagents, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = radius * rng.uniform(0, 1, nagents) ** (1.0 / 3.0)
    return center + d * r[:, None]


def dropin_inputs(n=400, nagents=1_000_000, seed=2024):
    """(init [n][n][n][4], pos [a][3], volume [a], S, U, T [a][4]) of the drop-in workload; a
    smaller n scales the sphere with the domain."""
    nsub = DROPIN["nsub"]
    rng = np.random.default_rng(seed)
    init = rng.lognormal(0.0, 0.5, (n, n, n, nsub))
    pos = sphere_positions(rng, nagents, n * DX / 2.0, DROPIN["radius"] * (n / DROPIN["n"]))
    S = rng.uniform(0, 5, (nagents, nsub))
    U = rng.uniform(0, 20, (nagents, nsub))
    T = np.ones((nagents, nsub))
    vol = np.full(nagents, AGENT_VOLUME)
    return init, pos, vol, S, U, T


def batch_members(m=64, seed=3):
    """The batch workload: a list of dicts with shape (nx, ny, nz, nsub), D, decay, agent
    positions, secretion S and uptake U (saturation 1)."""
    rng = np.random.default_rng(seed)
    shapes = [tuple(int(v) for v in rng.integers(32, 257, 3)) + (int(rng.integers(1, 9)),) for _ in range(m)]
    out = []
    for (nx, ny, nz, ns) in shapes:
        D = 10.0 ** rng.uniform(2, 5, ns)
        decay = 10.0 ** rng.uniform(-3, 0, ns)
        na = int(rng.integers(0, 50001))
        pos = rng.uniform(0, 1, (na, 3)) * (np.array([nx, ny, nz]) * DX)
        out.append(dict(shape=(nx, ny, nz, ns), D=D, decay=decay, pos=pos, S=rng.uniform(0, 5, (na, ns)),
                        U=rng.uniform(0, 20, (na, ns))))
    return out


def batch_member_setup(i, member):
    """Dirichlet (value 10 on even substrates), initial field (constant 1 + 0.01 i) and agent
    tables of member i as the batch workload sets them: (dvalue, dactive, init, voxel, temp1,
    temp2).  temp1/temp2 follow set_internal_uptake_constants with T = 1."""
    from .solver import voxel_indices
    nx, ny, nz, ns = member["shape"]
    dvalue = np.full(ns, 10.0)
    dactive = (np.arange(ns) % 2 == 0).astype(np.uint8)
    init = np.full((nx, ny, nz, ns), 1.0 + 0.01 * i)
    vox = voxel_indices((nx, ny, nz), DX, member["pos"])
    vox[vox < 0] = 0
    k = (DT * AGENT_VOLUME) / (DX * DX * DX)
    t1 = (member["S"] * 1.0) * k
    t2 = 1.0 + k * member["U"] + k * member["S"]
    return dvalue, dactive, init, vox, t1, t2


def large_inputs(nagents=10_000_000, seed=9):
    """Coefficients and agents of the large FP32 workload: (D, decay, dvalue, dactive, pos, S, U);
    the initial field is 1 everywhere."""
    ns, n = LARGE["nsub"], LARGE["n"]
    rng = np.random.default_rng(seed)
    D = np.logspace(2, 5, ns)
    decay = np.logspace(-3, 0, ns)
    dvalue = np.full(ns, 38.0)
    dactive = (np.arange(ns) < 2).astype(np.uint8)
    pos = sphere_positions(rng, nagents, n * DX / 2.0, LARGE["radius"])
    return D, decay, dvalue, dactive, pos, rng.uniform(0, 5, (nagents, ns)), rng.uniform(0, 20, (nagents, ns))


def field_checksums(rho):
    """Per-substrate summary of a field [..][nsub] that pins every bit: float64 sum (numpy's
    pairwise order), min, max, and a position-weighted 64-bit hash of the bit patterns
    (sum of bits[i] * (2 i + 1) mod 2^64, i = voxel index)."""
    ns = rho.shape[-1]
    flat = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1, ns)
    w = (2 * np.arange(flat.shape[0], dtype=np.uint64) + np.uint64(1))
    out = []
    for d in range(ns):
        col = np.ascontiguousarray(flat[:, d])
        bits = col.view(np.uint64)
        h = int(np.sum(bits * w, dtype=np.uint64))
        out.append((float(np.sum(col)), float(np.min(col)), float(np.max(col)), h))
    return out
