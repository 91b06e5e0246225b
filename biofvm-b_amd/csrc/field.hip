// field.hip -- agent source/sink kernel and the layout transposes between the reference's
// voxel-major host field and the device's [substrate][z][y][x] field.
#include "bfvmb_internal.hpp"

namespace bfvmb {
namespace {

// Cell sources and sinks (Basic_Agent::simulate_secretion_and_uptake,
// BioFVM_basic_agent.cpp:279-300, looped by BioFVM_microenvironment.cpp:1155-1169).
// One thread per (substrate, occupied voxel) applies that voxel's agents in list order:
//     r = (r + temp1[a]) / temp2[a]
// The arithmetic is FP64 on both storage paths, so the FP64 path matches the reference
// bit for bit and the FP32 path rounds once per voxel per step.
template <class T>
__global__ __launch_bounds__(256) void k_agents(T *rho, const int *__restrict__ seg_start,
                                                const long long *__restrict__ seg_off,
                                                const double *__restrict__ t1,
                                                const double *__restrict__ t2, int nseg, int nsub,
                                                int nact, size_t nvox)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nseg * nsub)
        return;
    const int s = (int)(i / nseg);
    const int k = (int)(i - (long long)s * nseg);
    T *p = rho + (size_t)s * nvox + seg_off[k];
    double r = (double)*p;
    const double *a1 = t1 + (size_t)s * nact;
    const double *a2 = t2 + (size_t)s * nact;
    for (int a = seg_start[k]; a < seg_start[k + 1]; ++a) {
        r = r + a1[a];
        r = r / a2[a];
    }
    *p = (T)r;
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

template <class T>
int launch_agents_t(bfvmb_ctx *ctx)
{
    const AgentCSR &A = ctx->agents;
    if (A.nseg == 0)
        return BFVMB_OK;
    long long work = (long long)A.nseg * ctx->nsub;
    unsigned grid = (unsigned)((work + 255) / 256);
    hipLaunchKernelGGL(k_agents<T>, dim3(grid), dim3(256), 0, ctx->stream,
                       static_cast<T *>(ctx->field[ctx->cur]), A.seg_start, A.seg_off, A.t1, A.t2,
                       A.nseg, ctx->nsub, A.nact, ctx->nvox);
    return hip_check(ctx, hipGetLastError(), "agents launch");
}

dim3 transpose_grid(const bfvmb_ctx *ctx)
{
    return dim3((unsigned)((ctx->nz + 31) / 32), (unsigned)((ctx->nx + 31) / 32), (unsigned)ctx->ny);
}

}  // namespace

int launch_agents(bfvmb_ctx *ctx)
{
    if (ctx->precision == BFVMB_FP64)
        return launch_agents_t<double>(ctx);
    return launch_agents_t<float>(ctx);
}

int launch_aos_to_soa(bfvmb_ctx *ctx, const double *d_aos, void *d_soa)
{
    if (ctx->precision == BFVMB_FP64)
        hipLaunchKernelGGL(k_aos_to_soa<double>, transpose_grid(ctx), dim3(32, 8), 0, ctx->stream, d_aos,
                           static_cast<double *>(d_soa), ctx->nx, ctx->ny, ctx->nz, ctx->nsub);
    else
        hipLaunchKernelGGL(k_aos_to_soa<float>, transpose_grid(ctx), dim3(32, 8), 0, ctx->stream, d_aos,
                           static_cast<float *>(d_soa), ctx->nx, ctx->ny, ctx->nz, ctx->nsub);
    return hip_check(ctx, hipGetLastError(), "aos->soa launch");
}

int launch_soa_to_aos(bfvmb_ctx *ctx, const void *d_soa, double *d_aos)
{
    if (ctx->precision == BFVMB_FP64)
        hipLaunchKernelGGL(k_soa_to_aos<double>, transpose_grid(ctx), dim3(32, 8), 0, ctx->stream,
                           static_cast<const double *>(d_soa), d_aos, ctx->nx, ctx->ny, ctx->nz, ctx->nsub);
    else
        hipLaunchKernelGGL(k_soa_to_aos<float>, transpose_grid(ctx), dim3(32, 8), 0, ctx->stream,
                           static_cast<const float *>(d_soa), d_aos, ctx->nx, ctx->ny, ctx->nz, ctx->nsub);
    return hip_check(ctx, hipGetLastError(), "soa->aos launch");
}

}  // namespace bfvmb
