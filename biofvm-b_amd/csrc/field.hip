// field.hip -- agent source/sink kernel and the layout transposes between the reference's
// voxel-major host field and the device's [substrate][z][y][x] field.
#include <mutex>
#include <string>

#include "bfvmb_internal.hpp"

namespace bfvmb {
namespace {

__device__ __forceinline__ int find_domain_of(const long long *prefix, int ndom, long long t)
{
    int lo = 0, hi = ndom - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Cell sources and sinks (Basic_Agent::simulate_secretion_and_uptake,
// BioFVM_basic_agent.cpp:279-300, looped by BioFVM_microenvironment.cpp:1155-1169).
// One thread per (domain, substrate, occupied voxel) applies that voxel's agents in list
// order:  r = (r + temp1[a]) / temp2[a].  The arithmetic is FP64 on both storage paths, so
// the FP64 path matches the reference bit for bit and the FP32 path rounds once per voxel.
template <class T>
__global__ __launch_bounds__(256) void k_agents(const DomainDesc<T> *__restrict__ descs,
                                                const long long *__restrict__ prefix, int ndom, int parity)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= prefix[ndom])
        return;
    const int di = find_domain_of(prefix, ndom, i);
    const DomainDesc<T> &D = descs[di];
    const long long local = i - prefix[di];
    const int s = (int)(local / D.nseg);
    const int k = (int)(local - (long long)s * D.nseg);
    const size_t nvox = (size_t)D.nx * D.ny * D.nz;
    T *p = D.field[parity] + (size_t)s * nvox + D.seg_off[k];
    double r = (double)*p;
    const double *a1 = D.t1 + (size_t)s * D.nact;
    const double *a2 = D.t2 + (size_t)s * D.nact;
    for (int a = D.seg_start[k]; a < D.seg_start[k + 1]; ++a) {
        r = r + a1[a];
        r = r / a2[a];
    }
    *p = (T)r;
}

// First segment of each x-sweep tile (64 consecutive rows r = z*ny + y of a substrate): the
// smallest k with seg_off[k] >= 64 * t * nx, t = 0 .. ntiles (t = ntiles: nseg).  The list is
// sorted by voxel, so a tile's occupied voxels are the segments [xseg[t], xseg[t + 1]).
__global__ __launch_bounds__(256) void k_xseg(const long long *__restrict__ seg_off, int nseg, int nx,
                                              long long nrows, int ntiles, int *__restrict__ xseg)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntiles)
        return;
    const long long v = min((long long)t * kWave, nrows) * nx;
    int lo = 0, hi = nseg;  // first k in [lo, hi] with seg_off[k] >= v
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (seg_off[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    xseg[t] = lo;
}

// Bulk sources and sinks (Microenvironment::simulate_bulk_sources_and_sinks,
// BioFVM_microenvironment.cpp:1110-1153), one thread per (domain, substrate, voxel):
//   rho = fma(dt, T*S, rho) / (((U + S)*dt) + 1)
// with T*S and U+S precomputed; the reference's compiled loop fuses only the numerator's
// multiply-add.  FP64 arithmetic on both storage paths.
template <class T>
__global__ __launch_bounds__(256) void k_bulk(const DomainDesc<T> *__restrict__ descs,
                                              const long long *__restrict__ prefix, int ndom, int parity, double dt)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= prefix[ndom])
        return;
    const int di = find_domain_of(prefix, ndom, i);
    const DomainDesc<T> &D = descs[di];
    const long long e = i - prefix[di];
    const double h = dt > 0.0 ? dt : D.dt;
    T *p = D.field[parity] + e;
    const double num = fma(h, D.bulk_ts[e], (double)*p);
    const double den = D.bulk_us[e] * h + 1.0;
    *p = (T)(num / den);
}

// rho_aos[((x*ny + y)*nz + z)*nsub + d]  ->  soa[d][z][y][x], via 32x32 (x,z) LDS tiles.
template <class T>
__global__ __launch_bounds__(256) void k_aos_to_soa(const double *__restrict__ aos, T *__restrict__ soa,
                                                    int nx, int ny, int nz, int nsub)
{
    __shared__ double tile[32][33];
    const int z0 = blockIdx.x * 32, x0 = blockIdx.y * 32, y = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
    const size_t nxy = (size_t)nx * ny, nvox = nxy * nz;
    for (int d = 0; d < nsub; ++d) {
        for (int q = 0; q < 32; q += 8) {
            const int x = x0 + ty + q, z = z0 + tx;
            if (x < nx && z < nz)
                tile[ty + q][tx] = aos[(((size_t)x * ny + y) * nz + z) * nsub + d];
        }
        __syncthreads();
        for (int q = 0; q < 32; q += 8) {
            const int x = x0 + tx, z = z0 + ty + q;
            if (x < nx && z < nz)
                soa[(size_t)d * nvox + (size_t)z * nxy + (size_t)y * nx + x] = (T)tile[tx][ty + q];
        }
        __syncthreads();
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_soa_to_aos(const T *__restrict__ soa, double *__restrict__ aos,
                                                    int nx, int ny, int nz, int nsub)
{
    __shared__ double tile[32][33];
    const int z0 = blockIdx.x * 32, x0 = blockIdx.y * 32, y = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const size_t nxy = (size_t)nx * ny, nvox = nxy * nz;
    for (int d = 0; d < nsub; ++d) {
        for (int q = 0; q < 32; q += 8) {
            const int x = x0 + tx, z = z0 + ty + q;
            if (x < nx && z < nz)
                tile[tx][ty + q] = (double)soa[(size_t)d * nvox + (size_t)z * nxy + (size_t)y * nx + x];
        }
        __syncthreads();
        for (int q = 0; q < 32; q += 8) {
            const int x = x0 + ty + q, z = z0 + tx;
            if (x < nx && z < nz)
                aos[(((size_t)x * ny + y) * nz + z) * nsub + d] = tile[ty + q][tx];
        }
        __syncthreads();
    }
}

// Densities at listed voxels (reference numbering x*ny*nz + y*nz + z), out[k][substrate] in
// FP64: the values Microenvironment::nearest_density_vector(voxel) points at
// (BioFVM_microenvironment.cpp:989-992), read from the resident field.  A voxel outside the
// grid reads as NaN (device-side lists are not checked on the host).
template <class T>
__global__ __launch_bounds__(256) void k_densities_at(const T *__restrict__ f, int nx, int ny, int nz, int nsub,
                                                      const int *__restrict__ voxel, int n, double *__restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)n * nsub)
        return;
    const int k = (int)(i / nsub), s = (int)(i - (long long)k * nsub);
    const long long v = voxel[k];
    const long long nyz = (long long)ny * nz, nvox = nyz * nx;
    if (v < 0 || v >= nvox) {
        out[i] = __builtin_nan("");
        return;
    }
    const long long x = v / nyz, rem = v - x * nyz, y = rem / nz, z = rem - y * nz;
    out[i] = (double)f[(size_t)s * nvox + (size_t)((z * ny + y) * nx + x)];
}

template <class T>
int launch_agents_t(Group *g)
{
    int rc = sync_descriptors(g);
    if (rc)
        return rc;
    const int ndom = (int)g->doms.size();
    const long long work = g->prefix[0][ndom];
    if (work == 0)
        return BFVMB_OK;
    const unsigned grid = (unsigned)((work + 255) / 256);
    hipLaunchKernelGGL(k_agents<T>, dim3(grid), dim3(256), 0, g->stream, static_cast<const DomainDesc<T> *>(g->d_desc),
                       g->d_prefix, ndom, g->cur);
    return hip_check(g, hipGetLastError(), "agents launch");
}

dim3 transpose_grid(const Domain &d)
{
    return dim3((unsigned)((d.nz + 31) / 32), (unsigned)((d.nx + 31) / 32), (unsigned)d.ny);
}

}  // namespace

int domain_densities_at(Group *g, int i, int n, const int *voxel, double *out, int on_device)
{
    if (i < 0 || (size_t)i >= g->doms.size() || n < 0 || (n > 0 && (!voxel || !out)))
        return set_error(g, BFVMB_ERR_ARG, "densities_at: bad argument");
    Domain &d = g->doms[i];
    if (!on_device)
        for (int k = 0; k < n; ++k)
            if (voxel[k] < 0 || (size_t)voxel[k] >= d.nvox)
                return set_error(g, BFVMB_ERR_ARG, "densities_at: voxel index out of range");
    if (n == 0)
        return BFVMB_OK;
    const int *dv = voxel;
    double *dout = out;
    int *tv = nullptr;
    double *tout = nullptr;
    const size_t obytes = (size_t)n * d.nsub * sizeof(double);
    int rc = BFVMB_OK;
    if (!on_device &&
        (!(rc = hip_check(g, hipMalloc((void **)&tv, n * sizeof(int)), "density voxels")) &&
         !(rc = hip_check(g, hipMalloc((void **)&tout, obytes), "density output")) &&
         !(rc = hip_check(g, hipMemcpyAsync(tv, voxel, n * sizeof(int), hipMemcpyHostToDevice, g->stream), "voxels")))) {
        dv = tv;
        dout = tout;
    }
    if (!rc) {
        const unsigned grid = (unsigned)(((long long)n * d.nsub + 255) / 256);
        if (g->precision == BFVMB_FP64)
            hipLaunchKernelGGL(k_densities_at<double>, dim3(grid), dim3(256), 0, g->stream,
                               static_cast<const double *>(d.field[g->cur]), d.nx, d.ny, d.nz, d.nsub, dv, n, dout);
        else
            hipLaunchKernelGGL(k_densities_at<float>, dim3(grid), dim3(256), 0, g->stream,
                               static_cast<const float *>(d.field[g->cur]), d.nx, d.ny, d.nz, d.nsub, dv, n, dout);
        rc = hip_check(g, hipGetLastError(), "densities_at launch");
    }
    if (!rc && !on_device &&
        !(rc = hip_check(g, hipMemcpyAsync(out, tout, obytes, hipMemcpyDeviceToHost, g->stream), "densities")))
        rc = hip_check(g, hipStreamSynchronize(g->stream), "densities sync");
    (void)hipFree(tv);
    (void)hipFree(tout);
    return rc;
}

// ---- densities at listed voxels without allocating (bfvmb_densities_at_noalloc).  The voxel
// list and the gathered values are staged in `scratch`, a file-scope device array that every
// GPU the library runs on holds until the process ends.  Host code never passes the array's
// name to a kernel or a copy: that is a host-side stand-in, not the device address.  group_init
// looks the address up once per device (hipGetSymbolAddress, with the device current) and
// every later use goes through that pointer.  All calls on one GPU share the array, so calls
// that overlap on one GPU (two host threads, two streams) race; one call at a time is safe.
constexpr size_t kScratchBytes = BFVMB_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int scratch_setup(Group *g)
{
    if (g->device < 0 || g->device >= kMaxDevices)
        return set_error(g, BFVMB_ERR_NODEV, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[g->device])
        return BFVMB_OK;
    void *p = nullptr;
    int rc = hip_check(g, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch address");
    if (rc)
        return rc;
    if (!p)
        return set_error(g, BFVMB_ERR_HIP, "scratch address: null");
    g_scratch_ptr[g->device] = p;
    return BFVMB_OK;
}

int domain_densities_at_noalloc(Group *g, int i, int n, const int *voxel, double *out)
{
    if (i < 0 || (size_t)i >= g->doms.size() || n < 0 || (n > 0 && (!voxel || !out)))
        return set_error(g, BFVMB_ERR_ARG, "densities_at_noalloc: bad argument");
    Domain &d = g->doms[i];
    for (int k = 0; k < n; ++k)
        if (voxel[k] < 0 || (size_t)voxel[k] >= d.nvox)
            return set_error(g, BFVMB_ERR_ARG, "densities_at_noalloc: voxel index out of range");
    const size_t vbytes = ((size_t)n * sizeof(int) + 255) / 256 * 256;
    const size_t obytes = (size_t)n * d.nsub * sizeof(double);
    if (vbytes + obytes > kScratchBytes)
        return set_error(g, BFVMB_ERR_ARG,
                         "densities_at_noalloc: " + std::to_string(n) + " voxels x " + std::to_string(d.nsub) +
                             " substrates need " + std::to_string(vbytes + obytes) + " bytes of scratch, above its " +
                             std::to_string(kScratchBytes) + " (use bfvmb_densities_at)");
    if (n == 0)
        return BFVMB_OK;
    unsigned char *base;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        base = static_cast<unsigned char *>(g->device >= 0 && g->device < kMaxDevices ? g_scratch_ptr[g->device]
                                                                                       : nullptr);
    }
    if (!base)
        return set_error(g, BFVMB_ERR_STATE, "densities_at_noalloc: scratch not set up on this device");
    int *tv = reinterpret_cast<int *>(base);
    double *tout = reinterpret_cast<double *>(base + vbytes);
    int rc = hip_check(g, hipMemcpyAsync(tv, voxel, (size_t)n * sizeof(int), hipMemcpyHostToDevice, g->stream),
                       "voxels");
    if (rc)
        return rc;
    const unsigned grid = (unsigned)(((long long)n * d.nsub + 255) / 256);
    if (g->precision == BFVMB_FP64)
        hipLaunchKernelGGL(k_densities_at<double>, dim3(grid), dim3(256), 0, g->stream,
                           static_cast<const double *>(d.field[g->cur]), d.nx, d.ny, d.nz, d.nsub, tv, n, tout);
    else
        hipLaunchKernelGGL(k_densities_at<float>, dim3(grid), dim3(256), 0, g->stream,
                           static_cast<const float *>(d.field[g->cur]), d.nx, d.ny, d.nz, d.nsub, tv, n, tout);
    if ((rc = hip_check(g, hipGetLastError(), "densities_at_noalloc launch")) ||
        (rc = hip_check(g, hipMemcpyAsync(out, tout, obytes, hipMemcpyDeviceToHost, g->stream), "densities")))
        return rc;
    return hip_check(g, hipStreamSynchronize(g->stream), "densities sync");
}

int launch_xseg(Group *g, const Domain &d)
{
    if (d.nseg == 0 || d.xseg == nullptr)
        return BFVMB_OK;
    const long long nrows = (long long)d.ny * d.nz;
    const int ntiles = (int)((nrows + kWave - 1) / kWave);
    hipLaunchKernelGGL(k_xseg, dim3((unsigned)((ntiles + 1 + 255) / 256)), dim3(256), 0, g->stream, d.seg_off, d.nseg,
                       d.nx, nrows, ntiles, d.xseg);
    return hip_check(g, hipGetLastError(), "agent tile index launch");
}

int launch_agents(Group *g)
{
    if (g->precision == BFVMB_FP64)
        return launch_agents_t<double>(g);
    return launch_agents_t<float>(g);
}

int launch_aos_to_soa_f64(Group *g, Domain &d, const double *d_aos, double *d_soa)
{
    hipLaunchKernelGGL(k_aos_to_soa<double>, transpose_grid(d), dim3(32, 8), 0, g->stream, d_aos, d_soa, d.nx, d.ny,
                       d.nz, d.nsub);
    return hip_check(g, hipGetLastError(), "aos->soa launch");
}

int launch_bulk(Group *g, double dt)
{
    int rc = sync_descriptors(g);
    if (rc)
        return rc;
    const int ndom = (int)g->doms.size();
    const long long work = g->prefix[4][ndom];
    if (work == 0)
        return BFVMB_OK;
    const unsigned grid = (unsigned)((work + 255) / 256);
    const long long *pre = g->d_prefix + (size_t)4 * (ndom + 1);
    if (g->precision == BFVMB_FP64)
        hipLaunchKernelGGL(k_bulk<double>, dim3(grid), dim3(256), 0, g->stream,
                           static_cast<const DomainDesc<double> *>(g->d_desc), pre, ndom, g->cur, dt);
    else
        hipLaunchKernelGGL(k_bulk<float>, dim3(grid), dim3(256), 0, g->stream,
                           static_cast<const DomainDesc<float> *>(g->d_desc), pre, ndom, g->cur, dt);
    return hip_check(g, hipGetLastError(), "bulk launch");
}

int launch_aos_to_soa(Group *g, Domain &d, const double *d_aos, void *d_soa)
{
    if (g->precision == BFVMB_FP64)
        hipLaunchKernelGGL(k_aos_to_soa<double>, transpose_grid(d), dim3(32, 8), 0, g->stream, d_aos,
                           static_cast<double *>(d_soa), d.nx, d.ny, d.nz, d.nsub);
    else
        hipLaunchKernelGGL(k_aos_to_soa<float>, transpose_grid(d), dim3(32, 8), 0, g->stream, d_aos,
                           static_cast<float *>(d_soa), d.nx, d.ny, d.nz, d.nsub);
    return hip_check(g, hipGetLastError(), "aos->soa launch");
}

int launch_soa_to_aos(Group *g, Domain &d, const void *d_soa, double *d_aos)
{
    if (g->precision == BFVMB_FP64)
        hipLaunchKernelGGL(k_soa_to_aos<double>, transpose_grid(d), dim3(32, 8), 0, g->stream,
                           static_cast<const double *>(d_soa), d_aos, d.nx, d.ny, d.nz, d.nsub);
    else
        hipLaunchKernelGGL(k_soa_to_aos<float>, transpose_grid(d), dim3(32, 8), 0, g->stream,
                           static_cast<const float *>(d_soa), d_aos, d.nx, d.ny, d.nz, d.nsub);
    return hip_check(g, hipGetLastError(), "soa->aos launch");
}

}  // namespace bfvmb
