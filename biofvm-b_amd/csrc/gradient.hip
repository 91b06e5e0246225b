// gradient.hip -- density gradients of the device-resident field.
//
// Microenvironment::compute_all_gradient_vectors (BioFVM_microenvironment.cpp:1243-1386) and
// compute_gradient_vector (:1388-1533), per voxel, substrate and axis:
//   interior  (rho[+1] - rho[-1]) / (2*dx),   first  (rho[1] - rho[0]) / dx,
//   last      (rho[n-1] - rho[n-2]) / dx,
// every subtraction and division rounded on its own (FP64 arithmetic on both storage
// paths).  An axis of length 1 gives 0: the reference leaves z at its initial 0 then and
// reads outside the field for x or y.
//
// The full field is written as [substrate*3 + axis][z][y][x] in the field's precision, so
// that the voxel-major download ([voxel][substrate][axis], gradient_vectors' layout) is the
// field transpose with 3*nsub components.  One thread per (substrate, voxel); the six
// neighbour loads of a wavefront are coalesced along x and hit cache for y and z.
#include "bfvmb_internal.hpp"

namespace bfvmb {
namespace {

__device__ __forceinline__ double axis_diff(const double lo, const double mid, const double hi, int i, int n,
                                            double dx, double two_dx)
{
    if (n == 1)
        return 0.0;
    if (i == 0)
        return (hi - mid) / dx;
    if (i == n - 1)
        return (mid - lo) / dx;
    return (hi - lo) / two_dx;
}

template <class T>
__global__ __launch_bounds__(256) void k_gradients(const T *__restrict__ f, int nx, int ny, int nz, int nsub,
                                                   double dx, T *__restrict__ grad)
{
    const size_t nxy = (size_t)nx * ny, nvox = nxy * nz;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nvox * nsub)
        return;
    const int s = (int)(e / nvox);
    const size_t idx = e - (size_t)s * nvox;
    const int x = (int)(idx % nx);
    const int y = (int)((idx / nx) % ny);
    const int z = (int)(idx / nxy);
    const T *p = f + e;
    const double two_dx = dx * 2.0;
    const double c = (double)p[0];
    const double xm = x > 0 ? (double)p[-1] : c, xp = x < nx - 1 ? (double)p[1] : c;
    const double ym = y > 0 ? (double)p[-(ptrdiff_t)nx] : c, yp = y < ny - 1 ? (double)p[nx] : c;
    const double zm = z > 0 ? (double)p[-(ptrdiff_t)nxy] : c, zp = z < nz - 1 ? (double)p[nxy] : c;
    T *g = grad + (size_t)s * 3 * nvox + idx;
    g[0] = (T)axis_diff(xm, c, xp, x, nx, dx, two_dx);
    g[nvox] = (T)axis_diff(ym, c, yp, y, ny, dx, two_dx);
    g[2 * nvox] = (T)axis_diff(zm, c, zp, z, nz, dx, two_dx);
}

// gradients at listed voxels (reference numbering), out[k][substrate][axis] in FP64
template <class T>
__global__ __launch_bounds__(256) void k_gradients_at(const T *__restrict__ f, int nx, int ny, int nz, int nsub,
                                                      double dx, const int *__restrict__ voxel, int n,
                                                      double *__restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)n * nsub)
        return;
    const int k = (int)(i / nsub), s = (int)(i - (long long)k * nsub);
    const int v = voxel[k];
    const int x = v / (ny * nz), rem = v - x * ny * nz, y = rem / nz, z = rem - y * nz;
    const size_t nxy = (size_t)nx * ny;
    const T *p = f + (size_t)s * nxy * nz + (size_t)z * nxy + (size_t)y * nx + x;
    const double two_dx = dx * 2.0;
    const double c = (double)p[0];
    const double xm = x > 0 ? (double)p[-1] : c, xp = x < nx - 1 ? (double)p[1] : c;
    const double ym = y > 0 ? (double)p[-(ptrdiff_t)nx] : c, yp = y < ny - 1 ? (double)p[nx] : c;
    const double zm = z > 0 ? (double)p[-(ptrdiff_t)nxy] : c, zp = z < nz - 1 ? (double)p[nxy] : c;
    double *o = out + (size_t)i * 3;
    o[0] = axis_diff(xm, c, xp, x, nx, dx, two_dx);
    o[1] = axis_diff(ym, c, yp, y, ny, dx, two_dx);
    o[2] = axis_diff(zm, c, zp, z, nz, dx, two_dx);
}

template <class T>
int compute_t(Group *g, Domain &d)
{
    const size_t n = d.nvox * d.nsub;
    if (!d.grad && hip_check(g, hipMalloc(&d.grad, 3 * n * sizeof(T)), "gradient buffer"))
        return BFVMB_ERR_NOMEM;
    hipLaunchKernelGGL(k_gradients<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, g->stream,
                       static_cast<const T *>(d.field[g->cur]), d.nx, d.ny, d.nz, d.nsub, d.dx,
                       static_cast<T *>(d.grad));
    return hip_check(g, hipGetLastError(), "gradient launch");
}

}  // namespace

int domain_compute_gradients(Group *g)
{
    if (g->doms.size() != 1)
        return set_error(g, BFVMB_ERR_ARG, "compute_gradients: single microenvironment only");
    Domain &d = g->doms[0];
    if (d.nvox * d.nsub >= (size_t)UINT32_MAX * 256)
        return set_error(g, BFVMB_ERR_ARG, "compute_gradients: field too large");
    const int rc = g->precision == BFVMB_FP64 ? compute_t<double>(g, d) : compute_t<float>(g, d);
    if (rc)
        return rc;
    d.grad_valid = true;
    return BFVMB_OK;
}

int domain_download_gradients(Group *g, double *out)
{
    if (g->doms.size() != 1 || !out)
        return set_error(g, BFVMB_ERR_ARG, "download_gradients: bad argument");
    Domain &d = g->doms[0];
    if (!d.grad || !d.grad_valid)
        return set_error(g, BFVMB_ERR_STATE, "download_gradients: compute_gradients has not been called");
    // the transpose with 3*nsub components maps [s*3+a][z][y][x] to [voxel][s][a]
    Domain view;
    view.nx = d.nx;
    view.ny = d.ny;
    view.nz = d.nz;
    view.nsub = 3 * d.nsub;
    const size_t bytes = d.nvox * view.nsub * sizeof(double);
    double *stage = nullptr;
    int rc;
    if ((rc = hip_check(g, hipMalloc((void **)&stage, bytes), "gradient staging")))
        return rc;
    if (!(rc = launch_soa_to_aos(g, view, d.grad, stage)))
        rc = hip_check(g, hipMemcpyAsync(out, stage, bytes, hipMemcpyDeviceToHost, g->stream), "gradient download");
    if (!rc)
        rc = hip_check(g, hipStreamSynchronize(g->stream), "gradient sync");
    (void)hipFree(stage);
    return rc;
}

int domain_gradients_at(Group *g, int n, const int *voxel, double *out)
{
    if (g->doms.size() != 1 || n < 0 || (n > 0 && (!voxel || !out)))
        return set_error(g, BFVMB_ERR_ARG, "gradients_at: bad argument");
    Domain &d = g->doms[0];
    for (int k = 0; k < n; ++k)
        if (voxel[k] < 0 || (size_t)voxel[k] >= d.nvox)
            return set_error(g, BFVMB_ERR_ARG, "gradients_at: voxel index out of range");
    if (n == 0)
        return BFVMB_OK;
    int *dv = nullptr;
    double *dout = nullptr;
    const size_t obytes = (size_t)n * d.nsub * 3 * sizeof(double);
    int rc;
    if (!(rc = hip_check(g, hipMalloc((void **)&dv, n * sizeof(int)), "gradient voxels")) &&
        !(rc = hip_check(g, hipMalloc((void **)&dout, obytes), "gradient output")) &&
        !(rc = hip_check(g, hipMemcpyAsync(dv, voxel, n * sizeof(int), hipMemcpyHostToDevice, g->stream), "voxels"))) {
        const unsigned grid = (unsigned)(((long long)n * d.nsub + 255) / 256);
        if (g->precision == BFVMB_FP64)
            hipLaunchKernelGGL(k_gradients_at<double>, dim3(grid), dim3(256), 0, g->stream,
                               static_cast<const double *>(d.field[g->cur]), d.nx, d.ny, d.nz, d.nsub, d.dx, dv, n,
                               dout);
        else
            hipLaunchKernelGGL(k_gradients_at<float>, dim3(grid), dim3(256), 0, g->stream,
                               static_cast<const float *>(d.field[g->cur]), d.nx, d.ny, d.nz, d.nsub, d.dx, dv, n,
                               dout);
        if (!(rc = hip_check(g, hipGetLastError(), "gradient launch")) &&
            !(rc = hip_check(g, hipMemcpyAsync(out, dout, obytes, hipMemcpyDeviceToHost, g->stream), "gradients")))
            rc = hip_check(g, hipStreamSynchronize(g->stream), "gradient sync");
    }
    (void)hipFree(dv);
    (void)hipFree(dout);
    return rc;
}

}  // namespace bfvmb

extern "C" int bfvmb_compute_gradients(bfvmb_ctx *ctx)
{
    return ctx ? bfvmb::domain_compute_gradients(ctx) : BFVMB_ERR_ARG;
}

extern "C" int bfvmb_download_gradients(bfvmb_ctx *ctx, double *grad_voxel_major)
{
    return ctx ? bfvmb::domain_download_gradients(ctx, grad_voxel_major) : BFVMB_ERR_ARG;
}

extern "C" int bfvmb_gradients_at(bfvmb_ctx *ctx, int n, const int *voxel, double *grad)
{
    return ctx ? bfvmb::domain_gradients_at(ctx, n, voxel, grad) : BFVMB_ERR_ARG;
}

extern "C" int bfvmb_gradient_device_ptr(bfvmb_ctx *ctx, void **ptr)
{
    if (!ctx || !ptr || ctx->doms.size() != 1)
        return BFVMB_ERR_ARG;
    *ptr = ctx->doms[0].grad;
    return BFVMB_OK;
}
