"""Case files for the reference driver (oracle/ref_driver.cpp) -- test infrastructure only.

A case is one microenvironment set-up plus a run: grid, substrates, Dirichlet, initial
field in the reference's voxel-major layout ``[x][y][z][d]``, agents and snapshot steps.
The same ``Case`` object feeds the reference driver, the C restatement
(``oracle/bfvmb_oracle.c``) and the HIP library under test, so all three see identical
inputs.
"""
from __future__ import annotations

import dataclasses
import os
import subprocess

import numpy as np

MAGIC_IN = b"BFVMCAS1"
MAGIC_OUT = b"BFVMOUT1"
HERE = os.path.dirname(os.path.abspath(__file__))
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")


@dataclasses.dataclass
class Case:
    nx: int
    ny: int
    nz: int
    nsub: int
    dx: float
    dt: float
    D: np.ndarray
    decay: np.ndarray
    dvalue: np.ndarray
    dactive: np.ndarray
    init: np.ndarray  # float64, shape (nx, ny, nz, nsub)
    pos: np.ndarray  # float64 (n, 3)
    volume: np.ndarray  # float64 (n,)
    S: np.ndarray  # float64 (n, nsub)
    U: np.ndarray
    T: np.ndarray
    nsteps: int
    snaps: list
    with_agents: bool = True
    x0: float = 0.0
    bulk: object = None  # (supply, target, uptake), each [nx, ny, nz, nsub]; bulk update per step
    dnodes: object = None  # (voxel[n], values[n][nsub]): per-voxel Dirichlet values, in list order
    moves: object = None  # [(step, route, positions [n][3])]: agents moved before that step's cell
                          # update by assign_position (route 0) or position + update_voxel_index (1)

    @property
    def nagents(self) -> int:
        return int(self.pos.shape[0])

    @property
    def bbox(self) -> np.ndarray:
        return np.array([self.x0, self.x0, self.x0, self.x0 + self.nx * self.dx,
                         self.x0 + self.ny * self.dx, self.x0 + self.nz * self.dx])

    def write_bulk(self, path: str) -> None:
        """Bulk supply, target, uptake rates as the drivers' --bulk file."""
        with open(path, "wb") as f:
            for a in self.bulk:
                np.ascontiguousarray(a, dtype=np.float64).tofile(f)

    def write_dnodes(self, path: str) -> None:
        """Per-voxel Dirichlet nodes as the drivers' --dnodes file: n, voxel[n], values[n][nsub]."""
        voxel, values = self.dnodes
        with open(path, "wb") as f:
            np.array([len(voxel)], dtype=np.int32).tofile(f)
            np.ascontiguousarray(voxel, dtype=np.int32).tofile(f)
            np.ascontiguousarray(values, dtype=np.float64).tofile(f)

    def write_moves(self, path: str) -> None:
        """Agent moves as the drivers' --moves file: nmoves, then per move step, route, positions."""
        with open(path, "wb") as f:
            np.array([len(self.moves)], dtype=np.int32).tofile(f)
            for step, route, pos in self.moves:
                np.array([step, route], dtype=np.int32).tofile(f)
                np.ascontiguousarray(pos, dtype=np.float64).reshape(self.nagents, 3).tofile(f)

    def write(self, path: str) -> None:
        n = self.nagents
        with open(path, "wb") as f:
            f.write(MAGIC_IN)
            np.array([self.nx, self.ny, self.nz, self.nsub, self.nsteps, int(self.with_agents), n,
                      len(self.snaps)], dtype=np.int32).tofile(f)
            np.array([self.x0, self.x0, self.x0, self.dx, self.dt], dtype=np.float64).tofile(f)
            for a in (self.D, self.decay, self.dvalue):
                np.asarray(a, dtype=np.float64).tofile(f)
            np.asarray(self.dactive, dtype=np.int32).tofile(f)
            np.ascontiguousarray(self.init, dtype=np.float64).tofile(f)
            for a in (self.pos, self.volume, self.S, self.U, self.T):
                np.ascontiguousarray(a, dtype=np.float64).tofile(f)
            np.asarray(self.snaps, dtype=np.int32).tofile(f)


def effective_dirichlet(case: Case):
    """(value, active) per substrate as the reference holds them: a negative activation in a
    case means "setters not called", leaving resize_space's defaults (100.0, active)."""
    dact = np.asarray(case.dactive)
    value = np.where(dact < 0, 100.0, np.asarray(case.dvalue, dtype=np.float64))
    active = np.where(dact < 0, 1, dact != 0).astype(np.uint8)
    return value, active


def read_output(path: str, case: Case) -> dict:
    n, nsub = case.nagents, case.nsub
    nv = case.nx * case.ny * case.nz
    with open(path, "rb") as f:
        if f.read(8) != MAGIC_OUT:
            raise ValueError("bad reference output magic")
        na = int(np.fromfile(f, np.int32, 1)[0])
        assert na == n
        voxel = np.fromfile(f, np.int32, n)
        t1 = np.fromfile(f, np.float64, n * nsub).reshape(n, nsub)
        t2 = np.fromfile(f, np.float64, n * nsub).reshape(n, nsub)
        ns = int(np.fromfile(f, np.int32, 1)[0])
        fields = {}
        for _ in range(ns):
            step = int(np.fromfile(f, np.int32, 1)[0])
            fields[step] = np.fromfile(f, np.float64, nv * nsub).reshape(case.nx, case.ny, case.nz, nsub)
    return {"voxel": voxel, "temp1": t1, "temp2": t2, "fields": fields}


def run_reference(case: Case, workdir: str, solver: str = "lod", granularity: int = 1,
                  threads: int = 1, timing: bool = False, mat: bool = False, grad: bool = False) -> dict:
    """Run the compiled reference on ``case``; returns its agent mapping and snapshots, and
    with ``mat`` the bytes of Microenvironment::write_to_matlab after the last step."""
    cpath = os.path.join(workdir, "case.bin")
    opath = os.path.join(workdir, "out.bin")
    mpath = os.path.join(workdir, "out.mat")
    case.write(cpath)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    args = [REF_DRIVER, cpath, opath, solver, str(granularity)] + (["--time"] if timing else [])
    if mat:
        args += ["--mat", mpath]
    gpath = os.path.join(workdir, "grad.bin")
    if grad:
        args += ["--grad", gpath]
    if case.bulk is not None:
        bpath = os.path.join(workdir, "bulk.bin")
        case.write_bulk(bpath)
        args += ["--bulk", bpath]
    if case.dnodes is not None:
        dpath = os.path.join(workdir, "dnodes.bin")
        case.write_dnodes(dpath)
        args += ["--dnodes", dpath]
    if case.moves is not None:
        mpath, mopath = os.path.join(workdir, "moves.bin"), os.path.join(workdir, "moves_out.bin")
        case.write_moves(mpath)
        args += ["--moves", mpath, mopath]
    proc = subprocess.run(args, env=env, capture_output=True, text=True, check=True)
    out = read_output(opath, case)
    out["stdout"] = proc.stdout
    if mat:
        with open(mpath, "rb") as f:
            out["mat"] = f.read()
    if grad:
        out["grad"] = np.fromfile(gpath, dtype=np.float64).reshape(case.nx, case.ny, case.nz, case.nsub, 3)
    if case.moves is not None:
        m = np.fromfile(mopath, dtype=np.int32).reshape(len(case.moves), 2, case.nagents)
        out["move_voxel"], out["move_active"] = m[:, 0], m[:, 1]
    return out


def make_case(nx, ny, nz, nsub, *, seed=0, dx=20.0, dt=0.01, nagents=0, nsteps=1, snaps=None,
              D=None, decay=None, dvalue=None, dactive=None, init="lognormal", volume=None,
              agent_box=None, shared_voxels=False, outside=0, with_agents=True) -> Case:
    """Seeded synthetic case.  ``init`` is "lognormal" (mu=0, sigma=0.5), a float, or an
    array; agents are uniform in the domain (or ``agent_box`` = (lo, hi) per axis)."""
    rng = np.random.default_rng(seed)
    D = np.asarray(D if D is not None else 10.0 ** rng.uniform(2, 5, nsub), dtype=np.float64)
    decay = np.asarray(decay if decay is not None else 10.0 ** rng.uniform(-3, 0, nsub), dtype=np.float64)
    dvalue = np.asarray(dvalue if dvalue is not None else rng.uniform(0.5, 40.0, nsub), dtype=np.float64)
    dactive = np.asarray(dactive if dactive is not None else rng.integers(0, 2, nsub), dtype=np.int32)
    if isinstance(init, str):
        init = rng.lognormal(0.0, 0.5, (nx, ny, nz, nsub))
    elif np.isscalar(init):
        init = np.full((nx, ny, nz, nsub), float(init))
    init = np.asarray(init, dtype=np.float64).reshape(nx, ny, nz, nsub)
    ext = np.array([nx, ny, nz], dtype=np.float64) * dx
    if agent_box is None:
        pos = rng.uniform(0.0, 1.0, (nagents, 3)) * ext
    else:
        lo, hi = agent_box
        pos = rng.uniform(lo, hi, (nagents, 3))
    if shared_voxels and nagents:
        # pile several agents into a few voxels so that list order matters
        k = max(1, nagents // 8)
        pos[:k] = pos[0]
        pos[k:2 * k] = pos[1] if nagents > 1 else pos[0]
    if outside and nagents:
        pos[-outside:] = ext * 1.5  # outside the bounding box: reference keeps voxel 0
    vol = np.full(nagents, 4.0 / 3.0 * np.pi * 125.0) if volume is None else np.broadcast_to(volume, (nagents,)).astype(np.float64)
    S = rng.uniform(0.0, 5.0, (nagents, nsub))
    U = rng.uniform(0.0, 20.0, (nagents, nsub))
    T = np.ones((nagents, nsub))
    return Case(nx, ny, nz, nsub, dx, dt, D, decay, dvalue, dactive, init, pos, vol, S, U, T,
                nsteps, list(snaps if snaps is not None else [nsteps]), with_agents)
