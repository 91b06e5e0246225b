"""MAT v4 "multiscale_microenvironment" files: numpy restatement of the reference's writer
and reader.  Test infrastructure (the oracle of the device writer/reader), not product.

Writer: Microenvironment::write_to_matlab (BioFVM_microenvironment.cpp:1079-1104) through
write_matlab4_header (BioFVM_matlab.cpp:535-579); the MPI-IO variant
(BioFVM_parallel.cpp:610-678) writes the same bytes at world size 1.
  header: five uint32 (little-endian): type = 1000*0 + 100*0 + 10*0 + 0 (doubles, full),
          rows = 4 + nsub, cols = number of voxels, imag = 0, name length (no NUL);
          then the name bytes.
  data:   one column of rows doubles per voxel, voxels in mesh order (z fastest,
          index x*ny*nz + y*nz + z; BioFVM_parallel.cpp:176-200): center x, y, z, volume,
          then the densities.
Centers: x0 + (i + 0.5)*dx where the local start is x0 + 0*(nodes*dx) at world size 1
(BioFVM_parallel.cpp:127-142); volume dV = (dx*dy)*dz (:153).

Reader: read_matlab (BioFVM_matlab.cpp:71-200): rejects anything but little-endian full
real double-class matrices.
"""
import struct

import numpy as np

NAME = "multiscale_microenvironment"


def centers(n, start, dx):
    """Voxel centers along one axis, with the reference's operation order."""
    local_start = start + (0 * n * dx)  # coords = 0 at world size 1
    return local_start + (np.arange(n, dtype=np.float64) + 0.5) * dx


def matrix(rho, dx, origin=(0.0, 0.0, 0.0)):
    """The (4 + nsub) x nvox matrix, as columns per voxel; rho is [nx][ny][nz][nsub]."""
    nx, ny, nz, nsub = rho.shape
    cx, cy, cz = centers(nx, origin[0], dx), centers(ny, origin[1], dx), centers(nz, origin[2], dx)
    cols = np.empty((nx, ny, nz, 4 + nsub))
    cols[..., 0] = cx[:, None, None]
    cols[..., 1] = cy[None, :, None]
    cols[..., 2] = cz[None, None, :]
    cols[..., 3] = (dx * dx) * dx
    cols[..., 4:] = rho
    return cols.reshape(nx * ny * nz, 4 + nsub)


def to_bytes(rho, dx, origin=(0.0, 0.0, 0.0), name=NAME):
    cols = matrix(np.asarray(rho, dtype=np.float64), dx, origin)
    nb = name.encode()
    head = struct.pack("<5I", 0, cols.shape[1], cols.shape[0], 0, len(nb)) + nb
    return head + np.ascontiguousarray(cols, dtype="<f8").tobytes()


def write(path, rho, dx, origin=(0.0, 0.0, 0.0), name=NAME):
    with open(path, "wb") as f:
        f.write(to_bytes(rho, dx, origin, name))


def parse(data):
    """-> (name, matrix[cols][rows]) of a MAT v4 double matrix; ValueError otherwise."""
    if len(data) < 20:
        raise ValueError("truncated MAT v4 header")
    typ, rows, cols, imag, nlen = struct.unpack_from("<5I", data, 0)
    if typ // 1000 != 0 or (typ // 100) % 10 != 0 or (typ // 10) % 10 != 0 or typ % 10 != 0:
        raise ValueError(f"unsupported MAT v4 type {typ}")
    if imag != 0:
        raise ValueError("complex MAT v4 matrix")
    name = data[20:20 + nlen].decode("latin-1")
    off = 20 + nlen
    need = off + 8 * rows * cols
    if len(data) < need:
        raise ValueError("truncated MAT v4 data")
    m = np.frombuffer(data, dtype="<f8", count=rows * cols, offset=off).reshape(cols, rows)
    return name, m


def read(path):
    with open(path, "rb") as f:
        return parse(f.read())
