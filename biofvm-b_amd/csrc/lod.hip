// lod.hip -- the three Thomas sweeps of one LOD diffusion-decay step on gfx950.
//
// The reference (BioFVM_parallel.cpp:693-1395) solves, for every grid line along x, then
// y, then z, the tridiagonal system
//     forward   u[0] /= den[0];   u[i] = fma(c1, u[i-1], u[i]) / den[i]
//     backward  u[i] = fma(-c[i], u[i+1], u[i]),   i = n-2 .. 0
// and overwrites all six domain faces with the Dirichlet value of each active substrate
// before the x sweep and after every sweep (BioFVM_microenvironment.cpp:381-484).
//
// Here the field is [substrate][z][y][x] (x contiguous) and every sweep reads one buffer
// and writes the other, so the loads of the forward pass can run ahead of its stores.
//  * y and z sweeps: one line per lane, consecutive lanes on consecutive x (y sweep) or
//    consecutive xy (z sweep), so each step of the recurrence is one coalesced row access;
//    the next U rows are prefetched into registers while the current U are solved.
//  * x sweep: lines lie along the contiguous axis; a wavefront stages 64 lines x W
//    columns in LDS with coalesced row loads, each lane then walks its own line in LDS
//    (transposed access, padded pitch), and the tile goes back with coalesced stores.
//  * Dirichlet is fused: the x sweep applies it on load (the reference's first call) and
//    every sweep applies it on store; a line lying entirely on an active face is written
//    with the value directly (the reference solves it and then overwrites all of it).
// FP64 uses true division so every value is bit-identical to the reference's; FP32 uses
// the FP64-computed reciprocal of each denominator.
#include "bfvmb_internal.hpp"

namespace bfvmb {
namespace {

template <class T>
__device__ __forceinline__ T fwd_step(T c1, T prev, T v, T den, T rden)
{
    if constexpr (sizeof(T) == 8) {
        (void)rden;
        return fma(c1, prev, v) / den;
    } else {
        (void)den;
        return fmaf(c1, prev, v) * rden;
    }
}

template <class T>
__device__ __forceinline__ T first_step(T v, T den, T rden)
{
    if constexpr (sizeof(T) == 8) {
        (void)rden;
        return v / den;
    } else {
        (void)den;
        return v * rden;
    }
}

template <class T>
__device__ __forceinline__ T back_step(T cc, T next, T v)
{
    if constexpr (sizeof(T) == 8)
        return fma(-cc, next, v);
    else
        return fmaf(-cc, next, v);
}

// y (AXIS=1) or z (AXIS=2) sweep.  Block = 256 lanes = 256 lines of one substrate; the
// grid strides over line tiles so that the number of resident tiles can be bounded.
template <class T, int AXIS, int U>
__global__ __launch_bounds__(256) void k_sweep_yz(SweepArgs<T> a)
{
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const size_t nxy = (size_t)nx * ny;
    const size_t nvox = nxy * nz;
    const int n = AXIS == 1 ? ny : nz;                      // line length
    const long long lines_per_s = AXIS == 1 ? (long long)nx * nz : (long long)nxy;
    const long long tiles_per_s = (lines_per_s + 255) / 256;
    const long long ntiles = tiles_per_s * a.nsub;
    const size_t stride = AXIS == 1 ? (size_t)nx : nxy;

    for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int s = (int)(tile / tiles_per_s);
        const long long line = (tile - (long long)s * tiles_per_s) * 256 + threadIdx.x;
        if (line >= lines_per_s)
            continue;
        int x, o;  // o = z (y sweep) or y (z sweep)
        size_t base;
        if (AXIS == 1) {
            o = (int)(line / nx);
            x = (int)(line - (long long)o * nx);
            base = (size_t)s * nvox + (size_t)o * nxy + x;
        } else {
            o = (int)(line / nx);
            x = (int)(line - (long long)o * nx);
            base = (size_t)s * nvox + (size_t)line;
        }
        const int on = AXIS == 1 ? nz : ny;
        const bool act = a.dact[s] != 0;
        const T dv = a.dval[s];
        const T *src = a.in + base;
        T *dst = a.out + base;
        if (act && (x == 0 || x == nx - 1 || o == 0 || o == on - 1)) {
            for (int k = 0; k < n; ++k)
                dst[(size_t)k * stride] = dv;
            continue;
        }
        const T c1 = a.c1[s];
        const T *den = a.tab.den + (size_t)s * n;
        const T *cc = a.tab.cc + (size_t)s * n;
        const T *rden = a.tab.rden + (size_t)s * n;

        // forward elimination, U rows in flight ahead of the ones being solved
        T cur[U];
#pragma unroll
        for (int j = 0; j < U; ++j)
            if (j < n)
                cur[j] = src[(size_t)j * stride];
        T y = first_step(cur[0], den[0], rden[0]);
        dst[0] = y;
        for (int k0 = 0; k0 < n; k0 += U) {
            T nxt[U];
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const int k = k0 + U + j;
                if (k < n)
                    nxt[j] = src[(size_t)k * stride];
            }
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const int k = k0 + j;
                if (k >= 1 && k < n) {
                    y = fwd_step(c1, y, cur[j], den[k], rden[k]);
                    dst[(size_t)k * stride] = y;
                }
            }
#pragma unroll
            for (int j = 0; j < U; ++j)
                cur[j] = nxt[j];
        }
        // back substitution; the last element is y[n-1] itself
        if (act)
            dst[(size_t)(n - 1) * stride] = dv;
        T xv = y;
        for (int k0 = n - 2; k0 >= 0; k0 -= U) {
            T v[U];
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const int k = k0 - j;
                if (k >= 0)
                    v[j] = dst[(size_t)k * stride];
            }
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const int k = k0 - j;
                if (k >= 0) {
                    xv = back_step(cc[k], xv, v[j]);
                    dst[(size_t)k * stride] = (act && k == 0) ? dv : xv;
                }
            }
        }
    }
}

// x sweep.  Block = one wavefront = 64 lines (consecutive rows (y,z) of one substrate),
// staged through LDS W columns at a time.  The forward results of all but the last chunk
// are parked in the output buffer and read back by the backward pass.
template <class T, int W>
__global__ __launch_bounds__(64) void k_sweep_x(SweepArgs<T> a)
{
    __shared__ T tile[64][W + 1];
    constexpr int RB = 16;  // rows loaded per batch (loads in flight per lane)
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const size_t nvox = (size_t)nx * ny * nz;
    const int nrows = ny * nz;
    const int tiles_per_s = (nrows + 63) / 64;
    const long long ntiles = (long long)tiles_per_s * a.nsub;
    const int lane = threadIdx.x;
    const int last_c0 = ((nx - 1) / W) * W;

    for (long long t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int s = (int)(t / tiles_per_s);
        const int r0 = (int)(t - (long long)s * tiles_per_s) * 64;
        const int my_row = r0 + lane;
        const bool row_ok = my_row < nrows;
        const int yy = my_row % ny, zz = my_row / ny;
        const bool act = a.dact[s] != 0;
        const T dv = a.dval[s];
        const bool whole = act && (yy == 0 || yy == ny - 1 || zz == 0 || zz == nz - 1);
        const T *src = a.in + (size_t)s * nvox;
        T *dst = a.out + (size_t)s * nvox;
        const T c1 = a.c1[s];
        const T *den = a.tab.den + (size_t)s * nx;
        const T *cc = a.tab.cc + (size_t)s * nx;
        const T *rden = a.tab.rden + (size_t)s * nx;
        const int rows_here = min(64, nrows - r0);

        T y = T(0);
        for (int c0 = 0; c0 < nx; c0 += W) {
            const int w = min(W, nx - c0);
            for (int rb = 0; rb < rows_here; rb += RB) {
                T r[RB];
#pragma unroll
                for (int q = 0; q < RB; ++q)
                    if (rb + q < rows_here && lane < w)
                        r[q] = src[(size_t)(r0 + rb + q) * nx + c0 + lane];
#pragma unroll
                for (int q = 0; q < RB; ++q)
                    if (rb + q < rows_here && lane < w)
                        tile[rb + q][lane] = r[q];
            }
            __syncthreads();
            if (row_ok && !whole) {
                for (int j = 0; j < w; ++j) {
                    const int gx = c0 + j;
                    T v = tile[lane][j];
                    if (act && (gx == 0 || gx == nx - 1))
                        v = dv;
                    y = gx == 0 ? first_step(v, den[0], rden[0]) : fwd_step(c1, y, v, den[gx], rden[gx]);
                    tile[lane][j] = y;
                }
            }
            __syncthreads();
            if (c0 != last_c0) {
                for (int rr = 0; rr < rows_here; ++rr)
                    if (lane < w)
                        dst[(size_t)(r0 + rr) * nx + c0 + lane] = tile[rr][lane];
                __syncthreads();
            }
        }
        T xv = y;
        for (int c0 = last_c0; c0 >= 0; c0 -= W) {
            const int w = min(W, nx - c0);
            if (c0 != last_c0) {
                for (int rb = 0; rb < rows_here; rb += RB) {
                    T r[RB];
#pragma unroll
                    for (int q = 0; q < RB; ++q)
                        if (rb + q < rows_here && lane < w)
                            r[q] = dst[(size_t)(r0 + rb + q) * nx + c0 + lane];
#pragma unroll
                    for (int q = 0; q < RB; ++q)
                        if (rb + q < rows_here && lane < w)
                            tile[rb + q][lane] = r[q];
                }
                __syncthreads();
            }
            if (row_ok) {
                for (int j = w - 1; j >= 0; --j) {
                    const int gx = c0 + j;
                    if (whole) {
                        tile[lane][j] = dv;
                        continue;
                    }
                    if (gx != nx - 1)
                        xv = back_step(cc[gx], xv, tile[lane][j]);
                    tile[lane][j] = (act && (gx == 0 || gx == nx - 1)) ? dv : xv;
                }
            }
            __syncthreads();
            for (int rr = 0; rr < rows_here; ++rr)
                if (lane < w)
                    dst[(size_t)(r0 + rr) * nx + c0 + lane] = tile[rr][lane];
            __syncthreads();
        }
    }
}

template <class T>
SweepArgs<T> make_args(bfvmb_ctx *ctx, int axis)
{
    CoefLayout L = coef_layout(ctx);
    const T *coef = static_cast<const T *>(ctx->coef);
    SweepArgs<T> a;
    a.in = static_cast<const T *>(ctx->field[ctx->cur]);
    a.out = static_cast<T *>(ctx->field[ctx->cur ^ 1]);
    a.c1 = coef + L.c1;
    a.tab.den = coef + L.den[axis];
    a.tab.cc = coef + L.cc[axis];
    a.tab.rden = coef + L.rden[axis];
    a.dval = static_cast<const T *>(ctx->d_dval);
    a.dact = ctx->d_dact;
    a.nx = ctx->nx;
    a.ny = ctx->ny;
    a.nz = ctx->nz;
    a.nsub = ctx->nsub;
    return a;
}

template <class T>
int launch_sweep_t(bfvmb_ctx *ctx, int axis)
{
    SweepArgs<T> a = make_args<T>(ctx, axis);
    if (axis == 0) {
        constexpr int W = 64;
        long long tiles = (long long)((ctx->ny * ctx->nz + 63) / 64) * ctx->nsub;
        long long grid = ctx->x_blocks > 0 ? std::min<long long>(ctx->x_blocks, tiles) : tiles;
        hipLaunchKernelGGL((k_sweep_x<T, W>), dim3((unsigned)grid), dim3(64), 0, ctx->stream, a);
    } else {
        long long lines = axis == 1 ? (long long)ctx->nx * ctx->nz : (long long)ctx->nx * ctx->ny;
        long long tiles = ((lines + 255) / 256) * ctx->nsub;
        long long grid = ctx->yz_blocks > 0 ? std::min<long long>(ctx->yz_blocks, tiles) : tiles;
        if (axis == 1)
            hipLaunchKernelGGL((k_sweep_yz<T, 1, 8>), dim3((unsigned)grid), dim3(256), 0, ctx->stream, a);
        else
            hipLaunchKernelGGL((k_sweep_yz<T, 2, 8>), dim3((unsigned)grid), dim3(256), 0, ctx->stream, a);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return hip_check(ctx, e, "sweep launch");
    ctx->cur ^= 1;
    return BFVMB_OK;
}

}  // namespace

int launch_sweep(bfvmb_ctx *ctx, int axis)
{
    if (ctx->precision == BFVMB_FP64)
        return launch_sweep_t<double>(ctx, axis);
    return launch_sweep_t<float>(ctx, axis);
}

}  // namespace bfvmb
