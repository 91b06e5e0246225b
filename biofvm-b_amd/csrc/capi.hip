// capi.hip -- the extern "C" boundary declared in include/biofvmb.h.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>

#include "bfvmb_internal.hpp"

namespace bfvmb {

int set_error(bfvmb_ctx *ctx, int code, const std::string &msg)
{
    if (ctx)
        ctx->err = msg;
    return code;
}

int hip_check(bfvmb_ctx *ctx, hipError_t e, const char *what)
{
    if (e == hipSuccess)
        return BFVMB_OK;
    return set_error(ctx, e == hipErrorOutOfMemory ? BFVMB_ERR_NOMEM : BFVMB_ERR_HIP,
                     std::string(what) + ": " + hipGetErrorString(e));
}

CoefLayout coef_layout(const bfvmb_ctx *ctx)
{
    CoefLayout L;
    const size_t s = (size_t)ctx->nsub;
    const int n[3] = {ctx->nx, ctx->ny, ctx->nz};
    size_t off = 0;
    L.c1 = off;
    off += s;
    for (int a = 0; a < 3; ++a) {
        L.den[a] = off;
        off += s * n[a];
        L.cc[a] = off;
        off += s * n[a];
        L.rden[a] = off;
        off += s * n[a];
    }
    L.total = off;
    return L;
}

namespace {

thread_local std::string g_create_error;

template <class T>
void pack(std::vector<unsigned char> &blob, size_t idx, double v)
{
    T t = (T)v;
    std::memcpy(blob.data() + idx * sizeof(T), &t, sizeof(T));
}

int upload_dirichlet(bfvmb_ctx *ctx)
{
    std::vector<unsigned char> vals(ctx->nsub * ctx->elem);
    for (int d = 0; d < ctx->nsub; ++d) {
        if (ctx->elem == 8)
            pack<double>(vals, d, ctx->dval[d]);
        else
            pack<float>(vals, d, ctx->dval[d]);
    }
    int rc = hip_check(ctx, hipMemcpyAsync(ctx->d_dval, vals.data(), vals.size(), hipMemcpyHostToDevice, ctx->stream),
                       "dirichlet values");
    if (rc)
        return rc;
    rc = hip_check(ctx, hipMemcpyAsync(ctx->d_dact, ctx->dact.data(), ctx->nsub, hipMemcpyHostToDevice, ctx->stream),
                   "dirichlet flags");
    if (rc)
        return rc;
    return hip_check(ctx, hipStreamSynchronize(ctx->stream), "dirichlet sync");
}

void free_agents(bfvmb_ctx *ctx)
{
    AgentCSR &A = ctx->agents;
    (void)hipFree(A.seg_start);
    (void)hipFree(A.seg_off);
    (void)hipFree(A.t1);
    (void)hipFree(A.t2);
    A = AgentCSR();
}

int ensure_stage(bfvmb_ctx *ctx)
{
    if (ctx->stage)
        return BFVMB_OK;
    return hip_check(ctx, hipMalloc(&ctx->stage, ctx->nvox * ctx->nsub * sizeof(double)), "staging buffer");
}

int lod_steps(bfvmb_ctx *ctx, int nsteps, bool with_agents)
{
    if (!ctx->have_coef)
        return set_error(ctx, BFVMB_ERR_STATE, "bfvmb_set_coefficients has not been called");
    if (nsteps < 0)
        return set_error(ctx, BFVMB_ERR_ARG, "nsteps < 0");
    for (int s = 0; s < nsteps; ++s) {
        int rc;
        if (with_agents && (rc = launch_agents(ctx)))
            return rc;
        for (int axis = 0; axis < 3; ++axis)
            if ((rc = launch_sweep(ctx, axis)))
                return rc;
    }
    return BFVMB_OK;
}

}  // namespace
}  // namespace bfvmb

using namespace bfvmb;

extern "C" {

const char *bfvmb_version(void) { return "biofvmb-hip 0.1.0 (gfx950)"; }

const char *bfvmb_last_error(const bfvmb_ctx *ctx)
{
    if (!ctx)
        return g_create_error.c_str();
    return ctx->err.c_str();
}

int bfvmb_create(bfvmb_ctx **out, int nx, int ny, int nz, int nsub, double dx, int precision, int device)
{
    if (!out)
        return BFVMB_ERR_ARG;
    *out = nullptr;
    if (nx < 1 || ny < 1 || nz < 1 || nsub < 1 || !(dx > 0.0) ||
        (precision != BFVMB_FP64 && precision != BFVMB_FP32)) {
        g_create_error = "bfvmb_create: bad grid, substrate count, dx or precision";
        return BFVMB_ERR_ARG;
    }
    if ((long long)ny * nz >= (1LL << 31) || (long long)nx * ny >= (1LL << 31) || (long long)nx * nz >= (1LL << 31)) {
        g_create_error = "bfvmb_create: plane too large";
        return BFVMB_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0 || device < 0 || device >= ndev) {
        g_create_error = "bfvmb_create: no HIP device";
        return BFVMB_ERR_NODEV;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess || std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        g_create_error = std::string("bfvmb_create: device is not gfx950: ") + prop.gcnArchName;
        return BFVMB_ERR_NODEV;
    }
    bfvmb_ctx *ctx = new (std::nothrow) bfvmb_ctx;
    if (!ctx)
        return BFVMB_ERR_NOMEM;
    ctx->nx = nx;
    ctx->ny = ny;
    ctx->nz = nz;
    ctx->nsub = nsub;
    ctx->dx = dx;
    ctx->precision = precision;
    ctx->device = device;
    ctx->nvox = (size_t)nx * ny * nz;
    ctx->elem = precision == BFVMB_FP64 ? 8 : 4;
    ctx->dval.assign(nsub, 100.0);
    ctx->dact.assign(nsub, 1);
    ctx->max_lds_per_block = prop.sharedMemPerBlock;
    int rc = BFVMB_OK;
    const size_t bytes = ctx->nvox * nsub * ctx->elem;
    if ((rc = hip_check(ctx, hipSetDevice(device), "hipSetDevice")) ||
        (rc = hip_check(ctx, hipStreamCreate(&ctx->own_stream), "hipStreamCreate")) ||
        (rc = hip_check(ctx, hipMalloc(&ctx->field[0], bytes), "field buffer 0")) ||
        (rc = hip_check(ctx, hipMalloc(&ctx->field[1], bytes), "field buffer 1")) ||
        (rc = hip_check(ctx, hipMalloc(&ctx->d_dval, nsub * ctx->elem), "dirichlet values")) ||
        (rc = hip_check(ctx, hipMalloc((void **)&ctx->d_dact, nsub), "dirichlet flags"))) {
        g_create_error = ctx->err;
        bfvmb_destroy(ctx);
        return rc;
    }
    ctx->stream = ctx->own_stream;
    if ((rc = hip_check(ctx, (hipError_t)sweep_lds_attributes(ctx->max_lds_per_block), "LDS attributes")) ||
        (rc = hip_check(ctx, hipMemsetAsync(ctx->field[0], 0, bytes, ctx->stream), "zero field")) ||
        (rc = upload_dirichlet(ctx))) {
        g_create_error = ctx->err;
        bfvmb_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return BFVMB_OK;
}

int bfvmb_destroy(bfvmb_ctx *ctx)
{
    if (!ctx)
        return BFVMB_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream)
        (void)hipStreamSynchronize(ctx->stream);
    free_agents(ctx);
    (void)hipFree(ctx->field[0]);
    (void)hipFree(ctx->field[1]);
    (void)hipFree(ctx->coef);
    (void)hipFree(ctx->d_dval);
    (void)hipFree(ctx->d_dact);
    (void)hipFree(ctx->stage);
    if (ctx->own_stream)
        (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return BFVMB_OK;
}

int bfvmb_set_stream(bfvmb_ctx *ctx, void *stream)
{
    if (!ctx)
        return BFVMB_ERR_ARG;
    int rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "stream switch");
    ctx->stream = stream ? static_cast<hipStream_t>(stream) : ctx->own_stream;
    return rc;
}

int bfvmb_set_coefficients(bfvmb_ctx *ctx, const double *diffusion, const double *decay, double dt)
{
    if (!ctx || !diffusion || !decay || !(dt > 0.0))
        return set_error(ctx, BFVMB_ERR_ARG, "bfvmb_set_coefficients: bad argument");
    const int ns = ctx->nsub;
    const int n[3] = {ctx->nx, ctx->ny, ctx->nz};
    std::vector<double> c1(ns), c2(ns), c3(ns), c3a(ns);
    lod_constants(ns, diffusion, decay, dt, ctx->dx, c1.data(), c2.data(), c3.data(), c3a.data());
    CoefLayout L = coef_layout(ctx);
    std::vector<double> blob(L.total);
    for (int d = 0; d < ns; ++d)
        blob[L.c1 + d] = c1[d];
    for (int a = 0; a < 3; ++a) {
        std::vector<double> den((size_t)ns * n[a]), cc((size_t)ns * n[a]);
        thomas_axis(n[a], ns, c1.data(), c2.data(), c3.data(), c3a.data(), den.data(), cc.data());
        for (size_t i = 0; i < den.size(); ++i) {
            blob[L.den[a] + i] = den[i];
            blob[L.cc[a] + i] = cc[i];
            blob[L.rden[a] + i] = 1.0 / den[i];
        }
    }
    const size_t bytes = blob.size() * sizeof(double);
    int rc;
    if (!ctx->coef && (rc = hip_check(ctx, hipMalloc((void **)&ctx->coef, bytes), "coefficient block")))
        return rc;
    if ((rc = hip_check(ctx, hipMemcpyAsync(ctx->coef, blob.data(), bytes, hipMemcpyHostToDevice, ctx->stream),
                        "coefficient upload")) ||
        (rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "coefficient sync")))
        return rc;
    ctx->D.assign(diffusion, diffusion + ns);
    ctx->decay.assign(decay, decay + ns);
    ctx->dt = dt;
    ctx->have_coef = true;
    return BFVMB_OK;
}

int bfvmb_set_dirichlet(bfvmb_ctx *ctx, const double *value, const uint8_t *active)
{
    if (!ctx || !value || !active)
        return set_error(ctx, BFVMB_ERR_ARG, "bfvmb_set_dirichlet: null argument");
    ctx->dval.assign(value, value + ctx->nsub);
    for (int d = 0; d < ctx->nsub; ++d)
        ctx->dact[d] = active[d] ? 1 : 0;
    return upload_dirichlet(ctx);
}

int bfvmb_upload_density(bfvmb_ctx *ctx, const double *rho)
{
    if (!ctx || !rho)
        return set_error(ctx, BFVMB_ERR_ARG, "bfvmb_upload_density: null argument");
    int rc;
    if ((rc = ensure_stage(ctx)) ||
        (rc = hip_check(ctx, hipMemcpyAsync(ctx->stage, rho, ctx->nvox * ctx->nsub * sizeof(double),
                                            hipMemcpyHostToDevice, ctx->stream), "density upload")) ||
        (rc = launch_aos_to_soa(ctx, static_cast<const double *>(ctx->stage), ctx->field[ctx->cur])))
        return rc;
    return hip_check(ctx, hipStreamSynchronize(ctx->stream), "upload sync");
}

int bfvmb_download_density(bfvmb_ctx *ctx, double *rho)
{
    if (!ctx || !rho)
        return set_error(ctx, BFVMB_ERR_ARG, "bfvmb_download_density: null argument");
    int rc;
    if ((rc = ensure_stage(ctx)) ||
        (rc = launch_soa_to_aos(ctx, ctx->field[ctx->cur], static_cast<double *>(ctx->stage))) ||
        (rc = hip_check(ctx, hipMemcpyAsync(rho, ctx->stage, ctx->nvox * ctx->nsub * sizeof(double),
                                            hipMemcpyDeviceToHost, ctx->stream), "density download")))
        return rc;
    return hip_check(ctx, hipStreamSynchronize(ctx->stream), "download sync");
}

int bfvmb_set_agents(bfvmb_ctx *ctx, int n, const int *voxel, const double *temp1, const double *temp2,
                     const uint8_t *active)
{
    if (!ctx || n < 0 || (n > 0 && (!voxel || !temp1 || !temp2)))
        return set_error(ctx, BFVMB_ERR_ARG, "bfvmb_set_agents: bad argument");
    const int ns = ctx->nsub;
    std::vector<int> order;
    order.reserve(n);
    for (int a = 0; a < n; ++a) {
        if (active && !active[a])
            continue;
        if (voxel[a] < 0 || (size_t)voxel[a] >= ctx->nvox)
            return set_error(ctx, BFVMB_ERR_ARG, "bfvmb_set_agents: voxel index out of range");
        order.push_back(a);
    }
    std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return voxel[p] < voxel[q]; });
    const int nact = (int)order.size();
    std::vector<int> seg_start;
    std::vector<long long> seg_off;
    std::vector<double> t1((size_t)ns * nact), t2((size_t)ns * nact);
    const long long nyz = (long long)ctx->ny * ctx->nz, nxy = (long long)ctx->nx * ctx->ny;
    for (int i = 0; i < nact; ++i) {
        const int a = order[i];
        if (i == 0 || voxel[a] != voxel[order[i - 1]]) {
            const long long v = voxel[a];
            const long long x = v / nyz, rem = v - x * nyz, y = rem / ctx->nz, z = rem - y * ctx->nz;
            seg_start.push_back(i);
            seg_off.push_back(z * nxy + y * ctx->nx + x);
        }
        for (int d = 0; d < ns; ++d) {
            t1[(size_t)d * nact + i] = temp1[(size_t)a * ns + d];
            t2[(size_t)d * nact + i] = temp2[(size_t)a * ns + d];
        }
    }
    seg_start.push_back(nact);
    int rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "agents sync");
    if (rc)
        return rc;
    free_agents(ctx);
    AgentCSR &A = ctx->agents;
    if (nact == 0)
        return BFVMB_OK;
    A.nseg = (int)seg_off.size();
    A.nact = nact;
    if ((rc = hip_check(ctx, hipMalloc(&A.seg_start, seg_start.size() * sizeof(int)), "agent segments")) ||
        (rc = hip_check(ctx, hipMalloc(&A.seg_off, seg_off.size() * sizeof(long long)), "agent voxels")) ||
        (rc = hip_check(ctx, hipMalloc(&A.t1, t1.size() * sizeof(double)), "agent temp1")) ||
        (rc = hip_check(ctx, hipMalloc(&A.t2, t2.size() * sizeof(double)), "agent temp2")) ||
        (rc = hip_check(ctx, hipMemcpy(A.seg_start, seg_start.data(), seg_start.size() * sizeof(int), hipMemcpyHostToDevice), "agent upload")) ||
        (rc = hip_check(ctx, hipMemcpy(A.seg_off, seg_off.data(), seg_off.size() * sizeof(long long), hipMemcpyHostToDevice), "agent upload")) ||
        (rc = hip_check(ctx, hipMemcpy(A.t1, t1.data(), t1.size() * sizeof(double), hipMemcpyHostToDevice), "agent upload")) ||
        (rc = hip_check(ctx, hipMemcpy(A.t2, t2.data(), t2.size() * sizeof(double), hipMemcpyHostToDevice), "agent upload"))) {
        free_agents(ctx);
        return rc;
    }
    return BFVMB_OK;
}

int bfvmb_set_agent_rates(bfvmb_ctx *ctx, int n, const int *voxel, const double *volume, const double *secretion,
                          const double *uptake, const double *saturation, const uint8_t *active, double dt)
{
    if (!ctx || n < 0 || (n > 0 && (!voxel || !volume || !secretion || !uptake || !saturation)))
        return set_error(ctx, BFVMB_ERR_ARG, "bfvmb_set_agent_rates: bad argument");
    const int ns = ctx->nsub;
    const double dV = (ctx->dx * ctx->dx) * ctx->dx;
    std::vector<double> t1((size_t)n * ns), t2((size_t)n * ns);
    for (int a = 0; a < n; ++a) {
        // BioFVM_basic_agent.cpp:155-176
        const double k = dt * volume[a] / dV;
        for (int d = 0; d < ns; ++d) {
            const size_t i = (size_t)a * ns + d;
            t1[i] = ((0.0 + secretion[i]) * saturation[i]) * k;
            t2[i] = std::fma(k, uptake[i], std::fma(k, secretion[i], 1.0));
        }
    }
    return bfvmb_set_agents(ctx, n, voxel, t1.data(), t2.data(), active);
}

int bfvmb_voxel_indices(int nx, int ny, int nz, double dx, const double *bbox_min, int n, const double *pos,
                        int *voxel)
{
    if (nx < 1 || ny < 1 || nz < 1 || !(dx > 0.0) || n < 0 || !bbox_min || (n > 0 && (!pos || !voxel)))
        return BFVMB_ERR_ARG;
    const int nn[3] = {nx, ny, nz};
    double lo[3], hi[3];
    int g[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = bbox_min[a];
        hi[a] = bbox_min[a] + nn[a] * dx;
        g[a] = (int)((hi[a] - lo[a]) / dx);  // voxels per axis as the reference counts them
    }
    const long long nv = (long long)nx * ny * nz;
    for (int i = 0; i < n; ++i) {
        const double *p = pos + 3 * (size_t)i;
        bool inside = true;
        for (int a = 0; a < 3; ++a)
            if (p[a] < lo[a] || p[a] > hi[a])
                inside = false;
        if (!inside) {
            voxel[i] = -1;
            continue;
        }
        long long c[3];
        for (int a = 0; a < 3; ++a) {
            long long v = (long long)std::floor((p[a] - lo[a]) / dx);
            if (v >= g[a])
                v = g[a] - 1;
            if (v < 0)
                v = 0;
            c[a] = v;
        }
        long long v = c[0] * g[1] * g[2] + c[1] * g[2] + c[2];
        if (v < 0)
            v = 0;
        if (v >= nv)
            v = nv - 1;
        voxel[i] = (int)v;
    }
    return BFVMB_OK;
}

int bfvmb_simulate_cell_sources_and_sinks(bfvmb_ctx *ctx)
{
    if (!ctx)
        return BFVMB_ERR_ARG;
    return launch_agents(ctx);
}

int bfvmb_simulate_diffusion_decay(bfvmb_ctx *ctx, int nsteps)
{
    if (!ctx)
        return BFVMB_ERR_ARG;
    return lod_steps(ctx, nsteps, false);
}

int bfvmb_run(bfvmb_ctx *ctx, int nsteps)
{
    if (!ctx)
        return BFVMB_ERR_ARG;
    return lod_steps(ctx, nsteps, true);
}

int bfvmb_sync(bfvmb_ctx *ctx)
{
    if (!ctx)
        return BFVMB_ERR_ARG;
    return hip_check(ctx, hipStreamSynchronize(ctx->stream), "bfvmb_sync");
}

int bfvmb_launch_phase(bfvmb_ctx *ctx, int phase, int count)
{
    if (!ctx || phase < 0 || phase > 3 || count < 0)
        return set_error(ctx, BFVMB_ERR_ARG, "bfvmb_launch_phase: bad argument");
    if (phase > 0 && !ctx->have_coef)
        return set_error(ctx, BFVMB_ERR_STATE, "bfvmb_set_coefficients has not been called");
    for (int i = 0; i < count; ++i) {
        int rc = phase == 0 ? launch_agents(ctx) : launch_sweep(ctx, phase - 1);
        if (rc)
            return rc;
    }
    return BFVMB_OK;
}

int bfvmb_set_launch(bfvmb_ctx *ctx, int yz_blocks, int x_blocks, int lds_blocks_per_cu)
{
    if (!ctx || yz_blocks < 0 || x_blocks < 0 || lds_blocks_per_cu < 0 || lds_blocks_per_cu > 8)
        return set_error(ctx, BFVMB_ERR_ARG, "bfvmb_set_launch: bad argument");
    ctx->yz_blocks = yz_blocks;
    ctx->x_blocks = x_blocks;
    ctx->lds_blocks_per_cu = lds_blocks_per_cu > 0 ? lds_blocks_per_cu : 2;
    return BFVMB_OK;
}

int bfvmb_device_info(int device, int *lds_per_block, int *cu_count)
{
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess)
        return BFVMB_ERR_NODEV;
    if (lds_per_block)
        *lds_per_block = (int)prop.sharedMemPerBlock;
    if (cu_count)
        *cu_count = prop.multiProcessorCount;
    return BFVMB_OK;
}

int bfvmb_density_device_ptr(bfvmb_ctx *ctx, void **ptr)
{
    if (!ctx || !ptr)
        return set_error(ctx, BFVMB_ERR_ARG, "bfvmb_density_device_ptr: null argument");
    *ptr = ctx->field[ctx->cur];
    return BFVMB_OK;
}

}  // extern "C"
