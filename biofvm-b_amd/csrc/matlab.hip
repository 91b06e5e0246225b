// matlab.hip -- MAT v4 "multiscale_microenvironment" files written from / read into the
// device-resident field.
//
// Microenvironment::write_to_matlab (BioFVM_microenvironment.cpp:1079-1104, header by
// write_matlab4_header, BioFVM_matlab.cpp:535-579; the MPI-IO variant
// BioFVM_parallel.cpp:610-678 writes the same bytes at world size 1):
//   five little-endian uint32 -- type 0 (full real double matrix), rows = 4 + nsub,
//   cols = voxels, imag 0, name length -- then the name without a terminator, then one
//   column per voxel in mesh order (index x*ny*nz + y*nz + z): center x, y, z, volume dV,
//   densities.  Centers are start + (i + 0.5)*dx, the start being x_start + 0*(nodes*dx)
//   at world size 1 (BioFVM_parallel.cpp:127-142); dV = dx*dy*dz (:153).
//
// The columns are assembled on the GPU straight from the [substrate][z][y][x] field, one
// slab of x planes at a time, and stream to the file through two pinned host buffers so
// that the device->host copy of one slab overlaps the write of the previous one.  The
// reader runs the same pipeline backwards, checks every center and volume against the
// context's grid, and only then makes the read field current.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "bfvmb_internal.hpp"

namespace bfvmb {
namespace {

constexpr char kMatName[] = "multiscale_microenvironment";
constexpr size_t kSlabBytes = size_t(64) << 20;  // target size of one slab of columns

struct Grid {
    int nx, ny, nz, nsub;
    double sx, sy, sz;  // local starts (x_start + 0*(nodes*dx))
    double dx, dv;
};

__device__ __forceinline__ double center(double start, int i, double dx)
{
    return start + ((double)i + 0.5) * dx;
}

// slab [xa, xa+xc) of soa[d][z][y][x]  ->  cols[row][4 + nsub], row = ((x-xa)*ny + y)*nz + z
template <class T>
__global__ __launch_bounds__(256) void k_mat_pack(const T *__restrict__ soa, Grid g, int xa, int xc,
                                                  double *__restrict__ cols)
{
    __shared__ double tile[32][33];
    const int z0 = blockIdx.x * 32, x0 = blockIdx.y * 32, y = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
    const size_t nxy = (size_t)g.nx * g.ny, nvox = nxy * g.nz;
    const int R = 4 + g.nsub;
    for (int d = 0; d < g.nsub; ++d) {
        for (int q = 0; q < 32; q += 8) {
            const int x = x0 + tx, z = z0 + ty + q;
            if (x < xc && z < g.nz)
                tile[tx][ty + q] = (double)soa[(size_t)d * nvox + (size_t)z * nxy + (size_t)y * g.nx + xa + x];
        }
        __syncthreads();
        for (int q = 0; q < 32; q += 8) {
            const int x = x0 + ty + q, z = z0 + tx;
            if (x < xc && z < g.nz)
                cols[(((size_t)x * g.ny + y) * g.nz + z) * R + 4 + d] = tile[ty + q][tx];
        }
        __syncthreads();
    }
    for (int q = 0; q < 32; q += 8) {
        const int x = x0 + ty + q, z = z0 + tx;
        if (x < xc && z < g.nz) {
            double *c = cols + (((size_t)x * g.ny + y) * g.nz + z) * R;
            c[0] = center(g.sx, xa + x, g.dx);
            c[1] = center(g.sy, y, g.dx);
            c[2] = center(g.sz, z, g.dx);
            c[3] = g.dv;
        }
    }
}

// inverse of k_mat_pack into `soa`; counts columns whose center or volume is off the grid
template <class T>
__global__ __launch_bounds__(256) void k_mat_unpack(const double *__restrict__ cols, Grid g, int xa, int xc,
                                                    T *__restrict__ soa, unsigned int *bad)
{
    __shared__ double tile[32][33];
    const int z0 = blockIdx.x * 32, x0 = blockIdx.y * 32, y = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const size_t nxy = (size_t)g.nx * g.ny, nvox = nxy * g.nz;
    const int R = 4 + g.nsub;
    {
        const int x = x0 + ty, z = z0 + tx;
        unsigned int nbad = 0;
        for (int q = 0; q < 32; q += 8) {
            const int xq = x + q;
            if (xq < xc && z < g.nz) {
                const double *c = cols + (((size_t)xq * g.ny + y) * g.nz + z) * R;
                const double tol = 1e-6 * g.dx;
                nbad += !(fabs(c[0] - center(g.sx, xa + xq, g.dx)) <= tol) ||
                        !(fabs(c[1] - center(g.sy, y, g.dx)) <= tol) ||
                        !(fabs(c[2] - center(g.sz, z, g.dx)) <= tol) || !(fabs(c[3] - g.dv) <= 1e-9 * g.dv);
            }
        }
        if (nbad)
            atomicAdd(bad, nbad);
    }
    for (int d = 0; d < g.nsub; ++d) {
        for (int q = 0; q < 32; q += 8) {
            const int x = x0 + ty + q, z = z0 + tx;
            if (x < xc && z < g.nz)
                tile[ty + q][tx] = cols[(((size_t)x * g.ny + y) * g.nz + z) * R + 4 + d];
        }
        __syncthreads();
        for (int q = 0; q < 32; q += 8) {
            const int x = x0 + tx, z = z0 + ty + q;
            if (x < xc && z < g.nz)
                soa[(size_t)d * nvox + (size_t)z * nxy + (size_t)y * g.nx + xa + x] = (T)tile[tx][ty + q];
        }
        __syncthreads();
    }
}

Grid grid_of(const Domain &d, const double *origin)
{
    Grid g;
    g.nx = d.nx;
    g.ny = d.ny;
    g.nz = d.nz;
    g.nsub = d.nsub;
    const double o[3] = {origin ? origin[0] : 0.0, origin ? origin[1] : 0.0, origin ? origin[2] : 0.0};
    const int n[3] = {d.nx, d.ny, d.nz};
    double s[3];
    for (int a = 0; a < 3; ++a) {
        const int zero = 0;  // the MPI coordinate at world size 1
        s[a] = o[a] + (zero * n[a] * d.dx);
    }
    g.sx = s[0];
    g.sy = s[1];
    g.sz = s[2];
    g.dx = d.dx;
    g.dv = (d.dx * d.dx) * d.dx;
    return g;
}

// Slab pipeline resources: two device column buffers, two pinned host buffers, two events.
struct Pipe {
    Group *g;
    int planes = 1;        // x planes per slab
    size_t slab_bytes = 0;  // bytes of one full slab of columns
    double *dev[2] = {nullptr, nullptr};
    double *host[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    int init(const Domain &d)
    {
        const size_t plane = (size_t)d.ny * d.nz * (4 + d.nsub) * sizeof(double);
        planes = (int)std::max<size_t>(1, std::min<size_t>((size_t)d.nx, kSlabBytes / plane));
        slab_bytes = plane * planes;
        int rc;
        for (int b = 0; b < 2; ++b)
            if ((rc = hip_check(g, hipMalloc((void **)&dev[b], slab_bytes), "MAT slab buffer")) ||
                (rc = hip_check(g, hipHostMalloc((void **)&host[b], slab_bytes, hipHostMallocDefault),
                                "MAT pinned buffer")) ||
                (rc = hip_check(g, hipEventCreateWithFlags(&ev[b], hipEventDisableTiming), "MAT event")))
                return rc;
        return BFVMB_OK;
    }
    ~Pipe()
    {
        if (g->stream)
            (void)hipStreamSynchronize(g->stream);
        else
            (void)hipDeviceSynchronize();
        for (int b = 0; b < 2; ++b) {
            (void)hipFree(dev[b]);
            (void)hipHostFree(host[b]);
            if (ev[b])
                (void)hipEventDestroy(ev[b]);
        }
    }
};

dim3 slab_grid(const Domain &d, int xc)
{
    return dim3((unsigned)((d.nz + 31) / 32), (unsigned)((xc + 31) / 32), (unsigned)d.ny);
}

struct File {
    FILE *f = nullptr;
    ~File()
    {
        if (f)
            fclose(f);
    }
};

int io_error(Group *g, const std::string &what, const char *path)
{
    return set_error(g, BFVMB_ERR_IO, what + ": " + path + (errno ? std::string(" (") + strerror(errno) + ")" : ""));
}

template <class T>
int write_t(Group *g, const char *path, const double *origin)
{
    Domain &d = g->doms[0];
    const Grid gr = grid_of(d, origin);
    const int R = 4 + d.nsub;
    Pipe p{g};
    int rc = p.init(d);
    if (rc)
        return rc;
    File out;
    errno = 0;
    if (!(out.f = fopen(path, "wb")))
        return io_error(g, "cannot open", path);
    const uint32_t head[5] = {0u, (uint32_t)R, (uint32_t)d.nvox, 0u, (uint32_t)std::strlen(kMatName)};
    if (fwrite(head, sizeof(uint32_t), 5, out.f) != 5 || fwrite(kMatName, 1, head[4], out.f) != head[4])
        return io_error(g, "write failed", path);
    const T *soa = static_cast<const T *>(d.field[g->cur]);
    const int nslab = (d.nx + p.planes - 1) / p.planes;
    const size_t plane_doubles = (size_t)d.ny * d.nz * R;
    auto flush = [&](int c) -> int {  // wait for slab c's copy, write it
        const int b = c & 1;
        const int xc = std::min(p.planes, d.nx - c * p.planes);
        int r = hip_check(g, hipEventSynchronize(p.ev[b]), "MAT copy");
        if (r)
            return r;
        const size_t n = plane_doubles * xc;
        if (fwrite(p.host[b], sizeof(double), n, out.f) != n)
            return io_error(g, "write failed", path);
        return BFVMB_OK;
    };
    for (int c = 0; c < nslab; ++c) {
        const int b = c & 1;
        const int xa = c * p.planes, xc = std::min(p.planes, d.nx - xa);
        hipLaunchKernelGGL(k_mat_pack<T>, slab_grid(d, xc), dim3(32, 8), 0, g->stream, soa, gr, xa, xc, p.dev[b]);
        if ((rc = hip_check(g, hipGetLastError(), "MAT pack")) ||
            (rc = hip_check(g, hipMemcpyAsync(p.host[b], p.dev[b], plane_doubles * xc * sizeof(double),
                                              hipMemcpyDeviceToHost, g->stream), "MAT copy")) ||
            (rc = hip_check(g, hipEventRecord(p.ev[b], g->stream), "MAT event")))
            return rc;
        if (c > 0 && (rc = flush(c - 1)))
            return rc;
    }
    if ((rc = flush(nslab - 1)))
        return rc;
    errno = 0;
    if (fclose(out.f) != 0) {
        out.f = nullptr;
        return io_error(g, "close failed", path);
    }
    out.f = nullptr;
    return BFVMB_OK;
}

template <class T>
int read_t(Group *g, const char *path, const double *origin)
{
    Domain &d = g->doms[0];
    const Grid gr = grid_of(d, origin);
    const int R = 4 + d.nsub;
    File in;
    errno = 0;
    if (!(in.f = fopen(path, "rb")))
        return io_error(g, "cannot open", path);
    uint32_t head[5];
    if (fread(head, sizeof(uint32_t), 5, in.f) != 5)
        return set_error(g, BFVMB_ERR_IO, std::string("truncated MAT v4 header: ") + path);
    // read_matlab's format checks (BioFVM_matlab.cpp:90-120), restricted to doubles
    if (head[0] != 0u)
        return set_error(g, BFVMB_ERR_IO, "not a little-endian full double MAT v4 matrix (type " +
                                              std::to_string(head[0]) + "): " + path);
    if (head[3] != 0u)
        return set_error(g, BFVMB_ERR_IO, std::string("complex MAT v4 matrix: ") + path);
    if (head[1] != (uint32_t)R || head[2] != (uint32_t)d.nvox)
        return set_error(g, BFVMB_ERR_ARG, "MAT file holds " + std::to_string(head[1]) + " x " +
                                               std::to_string(head[2]) + ", the context needs " + std::to_string(R) +
                                               " x " + std::to_string(d.nvox) + ": " + path);
    if (fseek(in.f, (long)head[4], SEEK_CUR) != 0)
        return io_error(g, "truncated name", path);
    Pipe p{g};
    int rc = p.init(d);
    if (rc)
        return rc;
    unsigned int *bad = nullptr;
    if ((rc = hip_check(g, hipMalloc((void **)&bad, sizeof(unsigned int)), "MAT check counter")))
        return rc;
    struct Free {
        void *q;
        ~Free() { (void)hipFree(q); }
    } free_bad{bad};
    if ((rc = hip_check(g, hipMemsetAsync(bad, 0, sizeof(unsigned int), g->stream), "MAT check counter")))
        return rc;
    // the file is unpacked into the spare buffer and made current only when it checks out
    T *soa = static_cast<T *>(d.field[g->cur ^ 1]);
    const int nslab = (d.nx + p.planes - 1) / p.planes;
    const size_t plane_doubles = (size_t)d.ny * d.nz * R;
    for (int c = 0; c < nslab; ++c) {
        const int b = c & 1;
        const int xa = c * p.planes, xc = std::min(p.planes, d.nx - xa);
        if (c >= 2 && (rc = hip_check(g, hipEventSynchronize(p.ev[b]), "MAT copy")))  // slab c-2 done
            return rc;
        const size_t n = plane_doubles * xc;
        if (fread(p.host[b], sizeof(double), n, in.f) != n)
            return set_error(g, BFVMB_ERR_IO, std::string("truncated MAT v4 data: ") + path);
        if ((rc = hip_check(g, hipMemcpyAsync(p.dev[b], p.host[b], n * sizeof(double), hipMemcpyHostToDevice,
                                              g->stream), "MAT copy")))
            return rc;
        hipLaunchKernelGGL(k_mat_unpack<T>, slab_grid(d, xc), dim3(32, 8), 0, g->stream, p.dev[b], gr, xa, xc, soa,
                           bad);
        if ((rc = hip_check(g, hipGetLastError(), "MAT unpack")) ||
            (rc = hip_check(g, hipEventRecord(p.ev[b], g->stream), "MAT event")))
            return rc;
    }
    unsigned int nbad = 0;
    if ((rc = hip_check(g, hipMemcpyAsync(&nbad, bad, sizeof(nbad), hipMemcpyDeviceToHost, g->stream), "MAT check")) ||
        (rc = hip_check(g, hipStreamSynchronize(g->stream), "MAT sync")))
        return rc;
    if (nbad)
        return set_error(g, BFVMB_ERR_ARG, std::to_string(nbad) +
                                               " voxel columns have centers or volumes off the context's grid: " + path);
    g->cur ^= 1;
    g->epoch++;
    return BFVMB_OK;
}

}  // namespace

int domain_write_matlab(Group *g, const char *path, const double *origin)
{
    if (g->doms.size() != 1 || !path)
        return set_error(g, BFVMB_ERR_ARG, "write_matlab: bad argument");
    return g->precision == BFVMB_FP64 ? write_t<double>(g, path, origin) : write_t<float>(g, path, origin);
}

int domain_read_matlab(Group *g, const char *path, const double *origin)
{
    if (g->doms.size() != 1 || !path)
        return set_error(g, BFVMB_ERR_ARG, "read_matlab: bad argument");
    return g->precision == BFVMB_FP64 ? read_t<double>(g, path, origin) : read_t<float>(g, path, origin);
}

}  // namespace bfvmb

extern "C" int bfvmb_write_matlab(bfvmb_ctx *ctx, const char *filename, const double *bbox_min)
{
    if (!ctx)
        return BFVMB_ERR_ARG;
    return bfvmb::domain_write_matlab(ctx, filename, bbox_min);
}

extern "C" int bfvmb_read_matlab(bfvmb_ctx *ctx, const char *filename, const double *bbox_min)
{
    if (!ctx)
        return BFVMB_ERR_ARG;
    return bfvmb::domain_read_matlab(ctx, filename, bbox_min);
}
