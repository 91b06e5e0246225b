// agents.hip -- moving agents: the voxel of every agent from its position, and the
// regrouping of the source/sink list by voxel, on the device.
//
// The two ways the reference moves an agent (world size 1):
//  * position set, then Basic_Agent::update_voxel_index (BioFVM_basic_agent.cpp:137-153): a
//    position outside the bounding box (BioFVM_mesh.cpp:204-213) deactivates the agent;
//  * Basic_Agent::assign_position (BioFVM_basic_agent.cpp:112-128): a position outside the box
//    is rejected, the agent keeps its voxel and stays active (keep_outside).
// Inside the box both give Cartesian_Mesh::nearest_voxel_local_index (BioFVM_parallel.cpp:
// 418-522): floor((p - min)/dx) per axis, clamped into the grid, voxel x*ny*nz + y*nz + z,
// clamped into [0, nvox).  An inactive agent stays inactive (the reference never sets
// is_active again).  The agents active when the list was set form a compact list (ascending
// caller index); each update sorts them stably by device offset z*nx*ny + y*nx + x (64-bit
// keys; a radix sort keeps equal keys in input order, so agents sharing a voxel stay in list
// order, as the reference's loop applies them), dead agents last, then rebuilds the CSR of
// occupied voxels and reorders temp1/temp2.  The caller's arithmetic for the mapping is repeated operation for operation
// (bfvmb_voxel_indices on the host), so both give the same voxels.
#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "bfvmb_internal.hpp"

namespace bfvmb {

struct AgentRemap {
    int n0 = 0;                          // agents active when the list was set
    int *ids = nullptr;                  // [n0] caller index of compact agent i (ascending)
    uint8_t *alive = nullptr;            // [n0] still active
    unsigned long long *key = nullptr;   // [n0] device offset of each compact agent's current voxel
    double *tab = nullptr;               // [n0][nsub][2] temp1, temp2 in compact order (agent-major:
                                         // the regrouping gathers whole agents)
    double *pos = nullptr;               // [given][3] staging of host positions
    unsigned long long *kin = nullptr, *kout = nullptr;  // [n0] sort keys (device offset, nvox = dead)
    int *vin = nullptr, *vout = nullptr;       // [n0] compact indices
    int *head = nullptr, *sid = nullptr;       // [n0] segment heads, segment ids
    int *counts = nullptr;               // [2] segments, agents alive
    int *seg_start_new = nullptr;        // CSR arrays with room for a segment per agent, adopted
    long long *seg_off_new = nullptr;    // by the domain after the first successful regrouping
    void *tmp = nullptr;                 // hipcub scratch
    size_t tmp_bytes = 0;
    std::vector<int> host_ids;
    bool order_stale = false;            // Domain::agent_order not yet read back from vout
};

void free_remap(Domain &d)
{
    AgentRemap *r = d.remap;
    if (!r)
        return;
    for (void *p : {(void *)r->ids, (void *)r->alive, (void *)r->key, (void *)r->tab, (void *)r->pos, (void *)r->kin,
                    (void *)r->kout, (void *)r->vin, (void *)r->vout, (void *)r->head, (void *)r->sid,
                    (void *)r->counts, r->tmp, (void *)r->seg_start_new, (void *)r->seg_off_new})
        (void)hipFree(p);
    delete r;
    d.remap = nullptr;
}

namespace {

struct MapGrid {
    double lo[3], hi[3], dx;
    int g[3];         // voxels per axis as the reference counts them: (int)((hi - lo)/dx)
    long long nv;     // voxels of the grid
    int nx, ny, nz;
};

// sort key of compact agent i: its voxel's device offset, or nvox once it is inactive.  A
// position outside the box deactivates the agent, or with keep_outside leaves it where it was.
__global__ __launch_bounds__(256) void k_agent_keys(const double *__restrict__ pos, const int *__restrict__ ids,
                                                    uint8_t *__restrict__ alive, unsigned long long *__restrict__ cur,
                                                    int n0, MapGrid m, int keep_outside,
                                                    unsigned long long *__restrict__ key, int *__restrict__ val)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n0)
        return;
    const double *p = pos + 3 * (size_t)ids[i];
    bool inside = true;
    long long c[3];
    for (int a = 0; a < 3; ++a) {
        const double v = p[a];
        if (!(v >= m.lo[a] && v <= m.hi[a]))  // outside the bounding box (or not a number)
            inside = false;
        long long q = inside ? (long long)floor((v - m.lo[a]) / m.dx) : 0;
        if (q >= m.g[a])
            q = m.g[a] - 1;
        if (q < 0)
            q = 0;
        c[a] = q;
    }
    unsigned long long k = (unsigned long long)m.nv;
    if (alive[i]) {
        if (inside) {
            long long v = c[0] * m.g[1] * m.g[2] + c[1] * m.g[2] + c[2];  // reference numbering
            if (v < 0)
                v = 0;
            if (v >= m.nv)
                v = m.nv - 1;
            const long long nyz = (long long)m.ny * m.nz;
            const long long x = v / nyz, rem = v - x * nyz, y = rem / m.nz, z = rem - y * m.nz;
            k = (unsigned long long)((z * m.ny + y) * m.nx + x);
            cur[i] = k;
        } else if (keep_outside) {
            k = cur[i];
        } else {
            alive[i] = 0;
        }
    }
    key[i] = k;
    val[i] = i;
}

__global__ __launch_bounds__(256) void k_agent_heads(const unsigned long long *__restrict__ key, int n0,
                                                     unsigned long long dead, int *__restrict__ head)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n0)
        return;
    head[k] = key[k] != dead && (k == 0 || key[k - 1] != key[k]) ? 1 : 0;
}

// CSR of occupied voxels and the agent tables in sorted order (dead agents sit after the
// last segment, in no segment)
__global__ __launch_bounds__(256) void k_agent_build(const unsigned long long *__restrict__ key,
                                                     const int *__restrict__ val, const int *__restrict__ head,
                                                     const int *__restrict__ sid, int n0, int nsub,
                                                     unsigned long long dead,
                                                     const double *__restrict__ tab,
                                                     int *__restrict__ seg_start, long long *__restrict__ seg_off,
                                                     double *__restrict__ t1, double *__restrict__ t2,
                                                     int *__restrict__ counts)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n0)
        return;
    const double *ta = tab + (size_t)val[k] * nsub * 2;
    for (int s = 0; s < nsub; ++s) {
        t1[(size_t)s * n0 + k] = ta[2 * s];
        t2[(size_t)s * n0 + k] = ta[2 * s + 1];
    }
    if (key[k] == dead)
        return;
    // sid is the exclusive scan of head: a segment's first agent holds the segment's index, the
    // others that index + 1
    const int seg = sid[k] + head[k] - 1;
    if (head[k]) {
        seg_start[seg] = k;
        seg_off[seg] = (long long)key[k];
    }
    if (k + 1 == n0 || key[k + 1] == dead) {  // the last live agent closes the list
        seg_start[seg + 1] = k + 1;
        counts[0] = seg + 1;
        counts[1] = k + 1;
    }
}

// compact-order tables from the tables set_agents sorted: tab[cidx[k]][s] = sorted[s][k]
__global__ __launch_bounds__(256) void k_agent_unsort(const double *__restrict__ t1, const double *__restrict__ t2,
                                                      const int *__restrict__ cidx, int n0, int nsub,
                                                      double *__restrict__ tab)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n0)
        return;
    double *ta = tab + (size_t)cidx[k] * nsub * 2;
    for (int s = 0; s < nsub; ++s) {
        ta[2 * s] = t1[(size_t)s * n0 + k];
        ta[2 * s + 1] = t2[(size_t)s * n0 + k];
    }
}

int bits_for(unsigned long long v)
{
    int b = 1;
    while (b < 64 && (v >> b) != 0)
        ++b;
    return b;
}

// first update after set_agents: compact list, tables in compact order, scratch
int remap_setup(Group *g, Domain &d)
{
    const int n0 = d.nact, ns = d.nsub;
    AgentRemap *r = new (std::nothrow) AgentRemap;
    if (!r)
        return set_error(g, BFVMB_ERR_NOMEM, "update_agent_positions: host memory");
    d.remap = r;
    r->n0 = n0;
    r->host_ids = d.agent_order;
    std::sort(r->host_ids.begin(), r->host_ids.end());
    std::vector<int> cidx(n0);
    std::vector<unsigned long long> key0(n0);  // current voxel (device offset) in compact order
    for (int k = 0; k < n0; ++k) {
        cidx[k] = (int)(std::lower_bound(r->host_ids.begin(), r->host_ids.end(), d.agent_order[k]) - r->host_ids.begin());
        key0[cidx[k]] = (unsigned long long)d.agent_off[k];
    }
    // the CSR may now need a segment per agent: new arrays, adopted once they are filled
    int *d_cidx = nullptr;
    int rc;
    const size_t tab = (size_t)ns * n0 * sizeof(double);
    if ((rc = hip_check(g, hipMalloc((void **)&r->ids, n0 * sizeof(int)), "agent ids")) ||
        (rc = hip_check(g, hipMalloc((void **)&r->alive, n0), "agent flags")) ||
        (rc = hip_check(g, hipMalloc((void **)&r->key, n0 * sizeof(unsigned long long)), "agent voxels")) ||
        (rc = hip_check(g, hipMalloc((void **)&r->tab, 2 * tab), "agent tables")) ||
        (rc = hip_check(g, hipMalloc((void **)&r->kin, n0 * sizeof(unsigned long long)), "agent keys")) ||
        (rc = hip_check(g, hipMalloc((void **)&r->kout, n0 * sizeof(unsigned long long)), "agent keys")) ||
        (rc = hip_check(g, hipMalloc((void **)&r->vin, n0 * sizeof(int)), "agent order")) ||
        (rc = hip_check(g, hipMalloc((void **)&r->vout, n0 * sizeof(int)), "agent order")) ||
        (rc = hip_check(g, hipMalloc((void **)&r->head, n0 * sizeof(int)), "agent heads")) ||
        (rc = hip_check(g, hipMalloc((void **)&r->sid, n0 * sizeof(int)), "agent segments")) ||
        (rc = hip_check(g, hipMalloc((void **)&r->counts, 2 * sizeof(int)), "agent counts")) ||
        (rc = hip_check(g, hipMalloc((void **)&r->seg_start_new, (n0 + 1) * sizeof(int)), "agent segments")) ||
        (rc = hip_check(g, hipMalloc((void **)&r->seg_off_new, n0 * sizeof(long long)), "agent voxels")) ||
        (rc = hip_check(g, hipMalloc((void **)&d_cidx, n0 * sizeof(int)), "agent ranks")))
        goto fail;
    {
        size_t b1 = 0, b2 = 0;
        if ((rc = hip_check(g, hipcub::DeviceRadixSort::SortPairs(nullptr, b1, r->kin, r->kout, r->vin, r->vout, n0, 0,
                                                                   64, g->stream), "agent sort size")) ||
            (rc = hip_check(g, hipcub::DeviceScan::ExclusiveSum(nullptr, b2, r->head, r->sid, n0, g->stream),
                            "agent scan size")))
            goto fail;
        r->tmp_bytes = std::max<size_t>(std::max(b1, b2), 1);
        if ((rc = hip_check(g, hipMalloc(&r->tmp, r->tmp_bytes), "agent sort scratch")))
            goto fail;
    }
    if ((rc = hip_check(g, hipMemcpyAsync(r->ids, r->host_ids.data(), n0 * sizeof(int), hipMemcpyHostToDevice,
                                          g->stream), "agent ids")) ||
        (rc = hip_check(g, hipMemsetAsync(r->alive, 1, n0, g->stream), "agent flags")) ||
        (rc = hip_check(g, hipMemcpyAsync(r->key, key0.data(), n0 * sizeof(unsigned long long), hipMemcpyHostToDevice,
                                          g->stream), "agent voxels")) ||
        (rc = hip_check(g, hipMemcpyAsync(d_cidx, cidx.data(), n0 * sizeof(int), hipMemcpyHostToDevice, g->stream),
                        "agent ranks")))
        goto fail;
    hipLaunchKernelGGL(k_agent_unsort, dim3((n0 + 255) / 256), dim3(256), 0, g->stream, d.t1, d.t2, d_cidx, n0, ns,
                       r->tab);
    if ((rc = hip_check(g, hipGetLastError(), "agent tables")) ||
        (rc = hip_check(g, hipStreamSynchronize(g->stream), "agent set-up")))
        goto fail;
    (void)hipFree(d_cidx);
    return BFVMB_OK;
fail:
    (void)hipFree(d_cidx);
    free_remap(d);
    return rc;
}

}  // namespace

int domain_update_agent_positions(Group *g, int i, int n, const double *pos, const double *bbox_min, int on_device,
                                  bool keep_outside)
{
    if (i < 0 || (size_t)i >= g->doms.size() || n < 0 || (n > 0 && !pos))
        return set_error(g, BFVMB_ERR_ARG, "update_agent_positions: bad argument");
    Domain &d = g->doms[i];
    if (n != d.agents_given)
        return set_error(g, BFVMB_ERR_ARG, "update_agent_positions: n differs from the agent list that was set");
    int rc = hip_check(g, hipStreamSynchronize(g->stream), "agents sync");
    if (rc)
        return rc;
    if (d.nact == 0)
        return BFVMB_OK;  // nobody active: nothing to move
    if (!d.remap && (rc = remap_setup(g, d)))
        return rc;
    AgentRemap *r = d.remap;
    const int n0 = r->n0;
    MapGrid m;
    const int nn[3] = {d.nx, d.ny, d.nz};
    for (int a = 0; a < 3; ++a) {
        m.lo[a] = bbox_min ? bbox_min[a] : 0.0;
        m.hi[a] = m.lo[a] + nn[a] * d.dx;
        m.g[a] = (int)((m.hi[a] - m.lo[a]) / d.dx);
    }
    m.dx = d.dx;
    m.nv = (long long)d.nvox;
    m.nx = d.nx;
    m.ny = d.ny;
    m.nz = d.nz;
    const double *src = pos;
    if (!on_device) {
        if (!r->pos && (rc = hip_check(g, hipMalloc((void **)&r->pos, (size_t)n * 3 * sizeof(double)), "positions")))
            return rc;
        if ((rc = hip_check(g, hipMemcpyAsync(r->pos, pos, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice,
                                              g->stream), "positions")))
            return rc;
        src = r->pos;
    }
    const unsigned long long dead = (unsigned long long)d.nvox;
    const dim3 grid((n0 + 255) / 256), block(256);
    hipLaunchKernelGGL(k_agent_keys, grid, block, 0, g->stream, src, r->ids, r->alive, r->key, n0, m,
                       keep_outside ? 1 : 0, r->kin, r->vin);
    size_t bytes = r->tmp_bytes;
    if ((rc = hip_check(g, hipGetLastError(), "agent keys")) ||
        (rc = hip_check(g, hipcub::DeviceRadixSort::SortPairs(r->tmp, bytes, r->kin, r->kout, r->vin, r->vout, n0, 0,
                                                               bits_for(d.nvox), g->stream), "agent sort")))
        return rc;
    hipLaunchKernelGGL(k_agent_heads, grid, block, 0, g->stream, r->kout, n0, dead, r->head);
    bytes = r->tmp_bytes;
    if ((rc = hip_check(g, hipGetLastError(), "agent heads")) ||
        (rc = hip_check(g, hipcub::DeviceScan::ExclusiveSum(r->tmp, bytes, r->head, r->sid, n0, g->stream),
                        "agent scan")) ||
        (rc = hip_check(g, hipMemsetAsync(r->counts, 0, 2 * sizeof(int), g->stream), "agent counts")))
        return rc;
    // up to here the domain's agent list is untouched; from the build on it is rewritten, and a
    // failure drops it (no descriptor may keep pointing at a half-built list)
    int *seg_start = r->seg_start_new ? r->seg_start_new : d.seg_start;
    long long *seg_off = r->seg_off_new ? r->seg_off_new : d.seg_off;
    int counts[2] = {0, 0};
    rc = hip_check(g, hipMemsetAsync(seg_start, 0, sizeof(int), g->stream), "agent segments");
    if (!rc) {
        hipLaunchKernelGGL(k_agent_build, grid, block, 0, g->stream, r->kout, r->vout, r->head, r->sid, n0, d.nsub,
                           dead, r->tab, seg_start, seg_off, d.t1, d.t2, r->counts);
        rc = hip_check(g, hipGetLastError(), "agent build");
    }
    if (!rc)
        rc = hip_check(g, hipMemcpyAsync(counts, r->counts, sizeof(counts), hipMemcpyDeviceToHost, g->stream),
                       "agent counts");
    if (!rc)
        rc = hip_check(g, hipStreamSynchronize(g->stream), "agent update");
    if (rc) {
        (void)hipStreamSynchronize(g->stream);
        free_agents(d);
        d.agents_given = 0;
        d.agent_order.clear();
        g->desc_dirty = true;
        g->epoch++;
        return rc;
    }
    if (r->seg_start_new) {  // first regrouping: the new arrays replace the set_agents ones
        (void)hipFree(d.seg_start);
        (void)hipFree(d.seg_off);
        d.seg_start = r->seg_start_new;
        d.seg_off = r->seg_off_new;
        d.seg_cap = (size_t)n0;
        r->seg_start_new = nullptr;
        r->seg_off_new = nullptr;
    }
    r->order_stale = true;  // read back when someone needs it (the adjoint's rate gradients)
    d.nseg = counts[0];
    g->desc_dirty = true;
    g->epoch++;
    return BFVMB_OK;
}

int domain_agent_order(Group *g, Domain &d)
{
    AgentRemap *r = d.remap;
    if (!r || !r->order_stale)
        return BFVMB_OK;
    std::vector<int> order(r->n0);
    int rc;
    if ((rc = hip_check(g, hipMemcpyAsync(order.data(), r->vout, r->n0 * sizeof(int), hipMemcpyDeviceToHost, g->stream),
                        "agent order")) ||
        (rc = hip_check(g, hipStreamSynchronize(g->stream), "agent order")))
        return rc;
    for (int &k : order)
        k = r->host_ids[k];
    d.agent_order.swap(order);
    r->order_stale = false;
    return BFVMB_OK;
}

}  // namespace bfvmb

extern "C" int bfvmb_update_agent_positions(bfvmb_ctx *ctx, int n, const double *pos, const double *bbox_min,
                                            int pos_on_device)
{
    return ctx ? bfvmb::domain_update_agent_positions(ctx, 0, n, pos, bbox_min, pos_on_device) : BFVMB_ERR_ARG;
}

extern "C" int bfvmb_batch_update_agent_positions(bfvmb_batch *b, int i, int n, const double *pos,
                                                  const double *bbox_min, int pos_on_device)
{
    return b ? bfvmb::domain_update_agent_positions(b, i, n, pos, bbox_min, pos_on_device) : BFVMB_ERR_ARG;
}

extern "C" int bfvmb_assign_agent_positions(bfvmb_ctx *ctx, int n, const double *pos, const double *bbox_min,
                                            int pos_on_device)
{
    return ctx ? bfvmb::domain_update_agent_positions(ctx, 0, n, pos, bbox_min, pos_on_device, true) : BFVMB_ERR_ARG;
}

extern "C" int bfvmb_batch_assign_agent_positions(bfvmb_batch *b, int i, int n, const double *pos,
                                                  const double *bbox_min, int pos_on_device)
{
    return b ? bfvmb::domain_update_agent_positions(b, i, n, pos, bbox_min, pos_on_device, true) : BFVMB_ERR_ARG;
}
