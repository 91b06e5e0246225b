"""ctypes front end of the C restatement (``oracle/bfvmb_oracle.c``) -- test infrastructure.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use this module,
and only to compare against; the product path (biofvm-b_amd/) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbfvmb_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_bp = ctypes.POINTER(ctypes.c_ubyte)


def build() -> None:
    subprocess.run(["make", "-C", HERE, "-s", "libbfvmb_oracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        i, d = ctypes.c_int, ctypes.c_double
        L.orc_lod_constants.argtypes = [i, _dp, _dp, d, d, _dp, _dp, _dp, _dp]
        L.orc_thomas_axis.argtypes = [i, i, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_apply_dirichlet.argtypes = [i, i, i, i, _dp, _dp, _bp]
        L.orc_sweep.argtypes = [i, i, i, i, i, _dp, _dp, _dp, _dp]
        L.orc_lod_step.argtypes = [i, i, i, i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _bp]
        L.orc_agent_constants.argtypes = [i, d, d, d, _dp, _dp, _dp, _dp, _dp]
        L.orc_sources_sinks.argtypes = [i, i, _ip, _bp, _dp, _dp, _dp]
        L.orc_voxel_index.argtypes = [_dp, _dp, d, d, d, i, i, i]
        L.orc_voxel_index.restype = i
        L.orc_run.argtypes = [i, i, i, i, d, d, _dp, _dp, _dp, _bp, i, _ip, _bp, _dp, _dp, i, _dp, _dp]
        L.orc_run_dv.argtypes = [i, i, i, i, d, d, _dp, _dp, _dp, _dp, _bp, i, _ip, _bp, _dp, _dp, i, _dp, _dp]
        L.orc_apply_dirichlet_voxels.argtypes = [i, i, i, i, _dp, _dp, _dp, _bp]
        L.orc_bulk_sources_sinks.argtypes = [ctypes.c_size_t, i, d, _dp, _dp, _dp, _dp]
        L.orc_gradients.argtypes = [i, i, i, i, d, _dp, _dp]
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _b(a):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    return a, a.ctypes.data_as(_bp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def lod_constants(D, decay, dt, dx):
    D, pD = _d(D)
    dc, pdc = _d(decay)
    n = D.size
    out = [np.empty(n) for _ in range(4)]
    lib().orc_lod_constants(n, pD, pdc, dt, dx, *[o.ctypes.data_as(_dp) for o in out])
    return out  # c1, c2, c3, c3a


def thomas_axis(n, c1, c2, c3, c3a):
    nsub = len(c1)
    denom = np.empty((n, nsub))
    c = np.empty((n, nsub))
    args = [_d(a)[1] for a in (c1, c2, c3, c3a)]
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (c1, c2, c3, c3a)]
    args = [k.ctypes.data_as(_dp) for k in keep]
    lib().orc_thomas_axis(n, nsub, *args, denom.ctypes.data_as(_dp), c.ctypes.data_as(_dp))
    return denom, c


def apply_dirichlet(rho, dvalue, dactive):
    nx, ny, nz, nsub = rho.shape
    dv, pdv = _d(dvalue)
    da, pda = _b(dactive)
    lib().orc_apply_dirichlet(nx, ny, nz, nsub, rho.ctypes.data_as(_dp), pdv, pda)


def sweep(axis, rho, c1, denom, c):
    nx, ny, nz, nsub = rho.shape
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (c1, denom, c)]
    lib().orc_sweep(axis, nx, ny, nz, nsub, rho.ctypes.data_as(_dp), *[k.ctypes.data_as(_dp) for k in keep])


def agent_constants(dt, volume, voxel_volume, S, U, T):
    S, U, T = (np.ascontiguousarray(a, dtype=np.float64) for a in (S, U, T))
    n, nsub = S.shape
    t1 = np.empty((n, nsub))
    t2 = np.empty((n, nsub))
    vol = np.broadcast_to(np.asarray(volume, dtype=np.float64), (n,))
    for a in range(n):
        lib().orc_agent_constants(nsub, dt, float(vol[a]), voxel_volume, S[a].ctypes.data_as(_dp),
                                  U[a].ctypes.data_as(_dp), T[a].ctypes.data_as(_dp),
                                  t1[a].ctypes.data_as(_dp), t2[a].ctypes.data_as(_dp))
    return t1, t2


def voxel_indices(case):
    """Reference agent -> voxel mapping; out-of-box agents stay at voxel 0 (active)."""
    bbox, pbbox = _d(case.bbox)
    out = np.empty(case.nagents, dtype=np.int32)
    for a in range(case.nagents):
        p, pp = _d(case.pos[a])
        v = lib().orc_voxel_index(pp, pbbox, case.dx, case.dx, case.dx, case.nx, case.ny, case.nz)
        out[a] = 0 if v < 0 else v
    return out


def update_voxel_indices(case, pos, active):
    """Basic_Agent::update_voxel_index (BioFVM_basic_agent.cpp:137-153) for agents at new
    positions ``pos`` ([n][3], lower grid corner case.x0): a position outside the bounding
    box deactivates the agent (voxel -1); otherwise the nearest voxel, clamped into the grid
    (BioFVM_parallel.cpp:418-522).  An inactive agent stays inactive.  Returns (voxel, active)."""
    bbox, pbbox = _d(case.bbox)
    P = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    voxel = np.empty(P.shape[0], dtype=np.int32)
    act = np.array(active, dtype=np.uint8, copy=True).reshape(-1)
    for a in range(P.shape[0]):
        p, pp = _d(P[a])
        v = lib().orc_voxel_index(pp, pbbox, case.dx, case.dx, case.dx, case.nx, case.ny, case.nz)
        if v < 0:
            act[a] = 0
        voxel[a] = v
    return voxel, act


def assign_positions(case, pos, voxel, active):
    """Basic_Agent::assign_position (BioFVM_basic_agent.cpp:112-128) for every agent at new
    positions ``pos``: a position outside the bounding box is rejected (the agent keeps its
    voxel and its activity); otherwise update_voxel_index's nearest voxel.  The activity flag
    is never changed.  Returns (voxel, active)."""
    bbox, pbbox = _d(case.bbox)
    P = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    vox = np.array(voxel, dtype=np.int32, copy=True).reshape(-1)
    for a in range(P.shape[0]):
        p, pp = _d(P[a])
        v = lib().orc_voxel_index(pp, pbbox, case.dx, case.dx, case.dx, case.nx, case.ny, case.nz)
        if v >= 0:
            vox[a] = v
    return vox, np.array(active, dtype=np.uint8, copy=True).reshape(-1)


def run_moves(case):
    """The case's snapshots with its agent moves (case.moves) applied before the cell update
    of their steps, as the reference driver applies them; also the (voxel, active) after each
    move.  Voxels of agents deactivated by update_voxel_index read -1, as the reference's do."""
    voxel, t1, t2 = agent_tables(case)
    active = np.ones(case.nagents, np.uint8)
    rho = np.array(case.init, dtype=np.float64, copy=True)
    snaps, after = {}, []
    moves = {step: (route, pos) for step, route, pos in case.moves}
    for s in range(1, max(case.snaps) + 1):
        if s in moves:
            route, pos = moves[s]
            if route == 0:
                voxel, active = assign_positions(case, pos, voxel, active)
            else:
                voxel, active = update_voxel_indices(case, pos, active)
            after.append((voxel.copy(), active.copy()))
        sources_sinks(rho, np.maximum(voxel, 0), active, t1, t2)
        rho = run(case, nsteps=1, rho=rho, with_agents=False)
        if s in case.snaps:
            snaps[s] = rho.copy()
    return snaps, after


def agent_tables(case):
    voxel = voxel_indices(case)
    dV = (case.dx * case.dx) * case.dx
    t1, t2 = agent_constants(case.dt, case.volume, dV, case.S, case.U, case.T)
    return voxel, t1, t2


def sources_sinks(rho, voxel, active, t1, t2):
    nsub = rho.shape[-1]
    v, pv = _i(voxel)
    a, pa = _b(active)
    k1, p1 = _d(t1)
    k2, p2 = _d(t2)
    lib().orc_sources_sinks(len(v), nsub, pv, pa, p1, p2, rho.ctypes.data_as(_dp))


def run(case, nsteps=None, rho=None, with_agents=None, tables=None):
    """Run ``case`` with the C restatement; returns the AoS field after ``nsteps``.
    ``tables`` = (temp1, temp2) replaces the agent constants derived from the case."""
    nsteps = case.nsteps if nsteps is None else nsteps
    with_agents = case.with_agents if with_agents is None else with_agents
    rho = np.array(case.init if rho is None else rho, dtype=np.float64, order="C", copy=True)
    voxel, t1, t2 = agent_tables(case)
    if tables is not None:
        t1, t2 = (np.ascontiguousarray(t, dtype=np.float64).reshape(t1.shape) for t in tables)
    na = case.nagents if with_agents else 0
    active = np.ones(max(na, 1), dtype=np.uint8)
    D, pD = _d(case.D)
    dc, pdc = _d(case.decay)
    from casefile import effective_dirichlet
    dvalue, dactive = effective_dirichlet(case)
    dv, pdv = _d(dvalue)
    da, pda = _b(dactive)
    v, pv = _i(voxel if na else np.zeros(1))
    a, pa = _b(active)
    k1, p1 = _d(t1 if na else np.zeros(1))
    k2, p2 = _d(t2 if na else np.zeros(1))
    work = np.empty((3 * (case.nx + case.ny + case.nz) + 4) * case.nsub)
    dvox = dirichlet_voxels(case)
    pvox = None if dvox is None else dvox.ctypes.data_as(_dp)
    if case.bulk is None:
        lib().orc_run_dv(case.nx, case.ny, case.nz, case.nsub, case.dx, case.dt, pD, pdc, pdv, pvox, pda, na, pv,
                         pa, p1, p2, nsteps, rho.ctypes.data_as(_dp), work.ctypes.data_as(_dp))
        return rho
    # per step: bulk sources/sinks, then the cell update and the LOD step
    for _ in range(nsteps):
        bulk_sources_sinks(rho, case.dt, *case.bulk)
        lib().orc_run_dv(case.nx, case.ny, case.nz, case.nsub, case.dx, case.dt, pD, pdc, pdv, pvox, pda, na, pv,
                         pa, p1, p2, 1, rho.ctypes.data_as(_dp), work.ctypes.data_as(_dp))
    return rho


def dirichlet_voxels(case):
    """dirichlet_value_vectors ([nx][ny][nz][nsub]) after the case's per-voxel Dirichlet nodes
    (add_dirichlet_node / update_dirichlet_node, BioFVM_microenvironment.cpp:226-290, applied in
    list order over the per-substrate values), or None when the case has none."""
    if getattr(case, "dnodes", None) is None:
        return None
    from casefile import effective_dirichlet
    dvalue, _ = effective_dirichlet(case)
    out = np.empty((case.nx * case.ny * case.nz, case.nsub))
    out[:] = dvalue
    voxel, values = case.dnodes
    for v, val in zip(np.asarray(voxel), np.asarray(values, dtype=np.float64).reshape(-1, case.nsub)):
        out[int(v)] = val
    return np.ascontiguousarray(out.reshape(case.nx, case.ny, case.nz, case.nsub))


def gradients(rho, dx):
    """[nx][ny][nz][nsub][3] gradients of the AoS field (orc_gradients)."""
    r = np.ascontiguousarray(rho, dtype=np.float64)
    nx, ny, nz, ns = r.shape
    g = np.empty((nx, ny, nz, ns, 3))
    lib().orc_gradients(nx, ny, nz, ns, dx, r.ctypes.data_as(_dp), g.ctypes.data_as(_dp))
    return g


def bulk_sources_sinks(rho, dt, supply, target, uptake):
    """In place on the AoS field (orc_bulk_sources_sinks)."""
    keep = [np.ascontiguousarray(a, dtype=np.float64).reshape(rho.shape) for a in (supply, target, uptake)]
    nvox = rho.size // rho.shape[-1]
    lib().orc_bulk_sources_sinks(nvox, rho.shape[-1], dt,
                                 *[k.ctypes.data_as(_dp) for k in keep], rho.ctypes.data_as(_dp))


def run_snapshots(case):
    """Fields at each of ``case.snaps`` (ascending), by repeated stepping."""
    out, rho, done = {}, None, 0
    for s in sorted(case.snaps):
        rho = run(case, nsteps=s - done, rho=case.init if rho is None else rho)
        done = s
        out[s] = rho.copy()
    return out
