// capi.hip -- the extern "C" boundary declared in include/biofvmb.h: the single-domain
// context (bfvmb_ctx, a group of one) and batches of uneven domains (bfvmb_batch).
#include <cmath>
#include <cstring>
#include <new>

#include "bfvmb_internal.hpp"

using namespace bfvmb;

namespace {
thread_local std::string g_create_error;

template <class G>
int make_group(G **out, int ndom, const int *nx, const int *ny, const int *nz, const int *nsub, const double *dx,
               int precision, int device)
{
    if (!out)
        return BFVMB_ERR_ARG;
    *out = nullptr;
    G *g = new (std::nothrow) G;
    if (!g)
        return BFVMB_ERR_NOMEM;
    const int rc = group_init(g, ndom, nx, ny, nz, nsub, dx, precision, device);
    if (rc) {
        g_create_error = g->err;
        group_free(g);
        delete g;
        return rc;
    }
    *out = g;
    return BFVMB_OK;
}
}  // namespace

extern "C" {

const char *bfvmb_version(void) { return "biofvmb-hip 0.2.0 (gfx950)"; }

// ------------------------------------------------------------------------- one domain
const char *bfvmb_last_error(const bfvmb_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int bfvmb_create(bfvmb_ctx **out, int nx, int ny, int nz, int nsub, double dx, int precision, int device)
{
    return make_group(out, 1, &nx, &ny, &nz, &nsub, &dx, precision, device);
}

int bfvmb_destroy(bfvmb_ctx *ctx)
{
    if (!ctx)
        return BFVMB_OK;
    group_free(ctx);
    delete ctx;
    return BFVMB_OK;
}

int bfvmb_set_stream(bfvmb_ctx *ctx, void *stream) { return ctx ? group_set_stream(ctx, stream) : BFVMB_ERR_ARG; }

int bfvmb_set_coefficients(bfvmb_ctx *ctx, const double *diffusion, const double *decay, double dt)
{
    return ctx ? domain_set_coefficients(ctx, 0, diffusion, decay, dt) : BFVMB_ERR_ARG;
}

int bfvmb_set_dirichlet(bfvmb_ctx *ctx, const double *value, const uint8_t *active)
{
    return ctx ? domain_set_dirichlet(ctx, 0, value, active) : BFVMB_ERR_ARG;
}

int bfvmb_set_dirichlet_nodes(bfvmb_ctx *ctx, int n, const int *voxel, const double *values)
{
    return ctx ? domain_set_dirichlet_nodes(ctx, 0, n, voxel, values) : BFVMB_ERR_ARG;
}

int bfvmb_upload_density(bfvmb_ctx *ctx, const double *rho) { return ctx ? domain_upload(ctx, 0, rho) : BFVMB_ERR_ARG; }

int bfvmb_download_density(bfvmb_ctx *ctx, double *rho) { return ctx ? domain_download(ctx, 0, rho) : BFVMB_ERR_ARG; }

int bfvmb_upload_density_device(bfvmb_ctx *ctx, const double *rho)
{
    return ctx ? domain_upload_device(ctx, 0, rho) : BFVMB_ERR_ARG;
}

int bfvmb_download_density_device(bfvmb_ctx *ctx, double *rho)
{
    return ctx ? domain_download_device(ctx, 0, rho) : BFVMB_ERR_ARG;
}

int bfvmb_set_agents(bfvmb_ctx *ctx, int n, const int *voxel, const double *temp1, const double *temp2,
                     const uint8_t *active)
{
    return ctx ? domain_set_agents(ctx, 0, n, voxel, temp1, temp2, active) : BFVMB_ERR_ARG;
}

int bfvmb_set_agent_rates(bfvmb_ctx *ctx, int n, const int *voxel, const double *volume, const double *secretion,
                          const double *uptake, const double *saturation, const uint8_t *active, double dt)
{
    if (!ctx || n < 0 || (n > 0 && (!voxel || !volume || !secretion || !uptake || !saturation)))
        return ctx ? set_error(ctx, BFVMB_ERR_ARG, "bfvmb_set_agent_rates: bad argument") : BFVMB_ERR_ARG;
    const Domain &d = ctx->doms[0];
    const int ns = d.nsub;
    const double dV = (d.dx * d.dx) * d.dx;
    std::vector<double> t1((size_t)n * ns), t2((size_t)n * ns);
    for (int a = 0; a < n; ++a) {
        // BioFVM_basic_agent.cpp:155-176
        const double k = dt * volume[a] / dV;
        for (int s = 0; s < ns; ++s) {
            const size_t i = (size_t)a * ns + s;
            t1[i] = ((0.0 + secretion[i]) * saturation[i]) * k;
            t2[i] = std::fma(k, uptake[i], std::fma(k, secretion[i], 1.0));
        }
    }
    return domain_set_agents(ctx, 0, n, voxel, t1.data(), t2.data(), active);
}

int bfvmb_voxel_indices(int nx, int ny, int nz, double dx, const double *bbox_min, int n, const double *pos, int *voxel)
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

int bfvmb_densities_at(bfvmb_ctx *ctx, int n, const int *voxel, double *out, int on_device)
{
    return ctx ? domain_densities_at(ctx, 0, n, voxel, out, on_device) : BFVMB_ERR_ARG;
}

int bfvmb_densities_at_noalloc(bfvmb_ctx *ctx, int n, const int *voxel, double *out)
{
    return ctx ? domain_densities_at_noalloc(ctx, 0, n, voxel, out) : BFVMB_ERR_ARG;
}

int bfvmb_simulate_cell_sources_and_sinks(bfvmb_ctx *ctx)
{
    if (!ctx)
        return BFVMB_ERR_ARG;
    ctx->epoch++;
    return launch_agents(ctx);
}

int bfvmb_set_bulk_rates(bfvmb_ctx *ctx, const double *supply, const double *target, const double *uptake)
{
    return ctx ? domain_set_bulk(ctx, 0, supply, target, uptake) : BFVMB_ERR_ARG;
}

int bfvmb_simulate_bulk_sources_and_sinks(bfvmb_ctx *ctx, double dt)
{
    return ctx ? group_bulk(ctx, dt) : BFVMB_ERR_ARG;
}

int bfvmb_simulate_diffusion_decay(bfvmb_ctx *ctx, int nsteps)
{
    return ctx ? group_steps(ctx, nsteps, false, false) : BFVMB_ERR_ARG;
}

int bfvmb_run(bfvmb_ctx *ctx, int nsteps) { return ctx ? group_steps(ctx, nsteps, true, false) : BFVMB_ERR_ARG; }

int bfvmb_run_graph(bfvmb_ctx *ctx, int nsteps) { return ctx ? group_steps(ctx, nsteps, true, true) : BFVMB_ERR_ARG; }

int bfvmb_sync(bfvmb_ctx *ctx)
{
    return ctx ? hip_check(ctx, hipStreamSynchronize(ctx->stream), "bfvmb_sync") : BFVMB_ERR_ARG;
}

int bfvmb_launch_phase(bfvmb_ctx *ctx, int phase, int count)
{
    return ctx ? group_launch_phase(ctx, phase, count) : BFVMB_ERR_ARG;
}

int bfvmb_set_launch(bfvmb_ctx *ctx, int sweep, int grid_blocks, int lds_blocks_per_cu, int ring_depth)
{
    return ctx ? group_set_launch(ctx, sweep, grid_blocks, lds_blocks_per_cu, ring_depth) : BFVMB_ERR_ARG;
}

int bfvmb_autotune(bfvmb_ctx *ctx, int reps, int *chosen)
{
    return ctx ? group_autotune(ctx, reps, chosen) : BFVMB_ERR_ARG;
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
        return BFVMB_ERR_ARG;
    *ptr = ctx->doms[0].field[ctx->cur];
    return BFVMB_OK;
}

// ------------------------------------------------------------------------- batches
const char *bfvmb_batch_last_error(const bfvmb_batch *b) { return b ? b->err.c_str() : g_create_error.c_str(); }

int bfvmb_batch_create(bfvmb_batch **out, int n, const int *nx, const int *ny, const int *nz, const int *nsub,
                       const double *dx, int precision, int device)
{
    return make_group(out, n, nx, ny, nz, nsub, dx, precision, device);
}

int bfvmb_batch_destroy(bfvmb_batch *b)
{
    if (!b)
        return BFVMB_OK;
    group_free(b);
    delete b;
    return BFVMB_OK;
}

int bfvmb_batch_set_stream(bfvmb_batch *b, void *stream) { return b ? group_set_stream(b, stream) : BFVMB_ERR_ARG; }

int bfvmb_batch_set_coefficients(bfvmb_batch *b, int i, const double *diffusion, const double *decay, double dt)
{
    return b ? domain_set_coefficients(b, i, diffusion, decay, dt) : BFVMB_ERR_ARG;
}

int bfvmb_batch_set_dirichlet(bfvmb_batch *b, int i, const double *value, const uint8_t *active)
{
    return b ? domain_set_dirichlet(b, i, value, active) : BFVMB_ERR_ARG;
}

int bfvmb_batch_set_dirichlet_nodes(bfvmb_batch *b, int i, int n, const int *voxel, const double *values)
{
    return b ? domain_set_dirichlet_nodes(b, i, n, voxel, values) : BFVMB_ERR_ARG;
}

int bfvmb_batch_upload_density(bfvmb_batch *b, int i, const double *rho)
{
    return b ? domain_upload(b, i, rho) : BFVMB_ERR_ARG;
}

int bfvmb_batch_download_density(bfvmb_batch *b, int i, double *rho)
{
    return b ? domain_download(b, i, rho) : BFVMB_ERR_ARG;
}

int bfvmb_batch_set_agents(bfvmb_batch *b, int i, int n, const int *voxel, const double *temp1, const double *temp2,
                           const uint8_t *active)
{
    return b ? domain_set_agents(b, i, n, voxel, temp1, temp2, active) : BFVMB_ERR_ARG;
}

int bfvmb_batch_set_bulk_rates(bfvmb_batch *b, int i, const double *supply, const double *target,
                               const double *uptake)
{
    return b ? domain_set_bulk(b, i, supply, target, uptake) : BFVMB_ERR_ARG;
}

int bfvmb_batch_run(bfvmb_batch *b, int nsteps, int use_graph)
{
    return b ? group_steps(b, nsteps, true, use_graph != 0) : BFVMB_ERR_ARG;
}

int bfvmb_batch_sync(bfvmb_batch *b)
{
    return b ? hip_check(b, hipStreamSynchronize(b->stream), "bfvmb_batch_sync") : BFVMB_ERR_ARG;
}

int bfvmb_batch_set_launch(bfvmb_batch *b, int sweep, int grid_blocks, int lds_blocks_per_cu, int ring_depth)
{
    return b ? group_set_launch(b, sweep, grid_blocks, lds_blocks_per_cu, ring_depth) : BFVMB_ERR_ARG;
}

int bfvmb_batch_autotune(bfvmb_batch *b, int reps, int *chosen)
{
    return b ? group_autotune(b, reps, chosen) : BFVMB_ERR_ARG;
}

int bfvmb_batch_launch_phase(bfvmb_batch *b, int phase, int count)
{
    return b ? group_launch_phase(b, phase, count) : BFVMB_ERR_ARG;
}

}  // extern "C"
