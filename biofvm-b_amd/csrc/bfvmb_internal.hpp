// bfvmb_internal.hpp -- device descriptors, groups of microenvironments and the kernel
// launchers shared by the HIP translation units.
//
// Every context is a Group: one or more microenvironments ("domains") stepped together
// on one stream, all on the same GPU and in the same precision.  Each phase of a step
// (sources/sinks, x, y, z sweep) is one kernel launch covering every domain of the group:
// the kernels read a device array of per-domain descriptors and map a global tile index
// to its domain through per-phase prefix sums.  A single-domain context (bfvmb_ctx) and a
// batch of uneven domains (bfvmb_batch) therefore run the same kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "biofvmb.h"

namespace bfvmb {

constexpr int kWave = 64;

// Per-axis Thomas tables, FP64 on both storage paths, [nsub][n].
struct AxisTables {
    const double *den;   // denominators  (thomas_denom{x,y,z})
    const double *cc;    // upper factors (thomas_c{x,y,z})
    const double *rden;  // RN(1/den)
};

// Device descriptor of one domain (plain data, copied to the GPU).
template <class T>
struct DomainDesc {
    T *field[2];                 // ping-pong buffers [nsub][nz][ny][nx]
    const double *c1;            // [nsub]  D*dt/dx^2
    AxisTables tab[3];           // x, y, z tables
    const T *dval;               // [nsub]  Dirichlet value
    const uint8_t *dact;         // [nsub]  Dirichlet activation
    const T *dface;              // per-voxel Dirichlet values of the six faces (face_x/y/z), or
                                 // nullptr: dval for every voxel of a substrate
    const int *seg_start;        // agents: [nseg+1] offsets into the voxel-sorted list
    const long long *seg_off;    // agents: [nseg] voxel offset z*nx*ny + y*nx + x
    const double *t1, *t2;       // agents: [nsub][nact] temp1 / temp2, sorted order
    const int *xseg;             // agents: [x-sweep tiles per substrate + 1] first segment of
                                 // each tile of 64 rows (z*ny + y), or nullptr: no agents
    const double *bulk_ts;       // bulk: [nsub][nz][ny][nx] target*supply (nullptr: none)
    const double *bulk_us;       // bulk: [nsub][nz][ny][nx] uptake+supply
    double dt;                   // time step frozen by set_coefficients
    int nseg, nact;
    int nx, ny, nz, nsub;
};

// Per-voxel Dirichlet values of a domain's six faces, in one array of the field's precision:
// the x faces as planes [2][nsub][nz][ny], then the y faces [2][nsub][nz][nx], then the z
// faces [2][nsub][ny][nx] (e = 0 the low face, 1 the high one).  A voxel on an edge or corner
// appears in several planes, with the same value.
__host__ __device__ __forceinline__ size_t face_x(int ny, int nz, int ns, int e, int s, int y, int z)
{
    return (((size_t)e * ns + s) * nz + z) * ny + y;
}
__host__ __device__ __forceinline__ size_t face_y(int nx, int ny, int nz, int ns, int e, int s, int x, int z)
{
    return 2 * (size_t)ns * nz * ny + (((size_t)e * ns + s) * nz + z) * nx + x;
}
__host__ __device__ __forceinline__ size_t face_z(int nx, int ny, int nz, int ns, int e, int s, int x, int y)
{
    return 2 * (size_t)ns * nz * ((size_t)ny + nx) + (((size_t)e * ns + s) * ny + y) * nx + x;
}
__host__ __device__ __forceinline__ size_t face_total(int nx, int ny, int nz, int ns)
{
    return 2 * (size_t)ns * ((size_t)nz * ny + (size_t)nz * nx + (size_t)ny * nx);
}

// Launch shape of one sweep as set through bfvmb_set_launch (0 = library default).
struct LaunchShape {
    long long grid = 0;          // blocks of the (strided) grid
    int lds_blocks_per_cu = 0;   // sweep wavefronts sharing one CU's LDS line store
    int ring = 0;                // y/z sweeps: rows in flight per lane (16, 32 or 64)
};

// Host bookkeeping of one domain.
struct Domain {
    int nx = 0, ny = 0, nz = 0, nsub = 0;
    double dx = 0.0;
    size_t nvox = 0;
    void *field[2] = {nullptr, nullptr};
    double *coef = nullptr;  // c1[nsub] | per axis den, cc, rden [nsub][n]
    bool have_coef = false;
    double dt = 0.0;
    std::vector<double> D, decay;
    std::vector<double> dval;
    std::vector<uint8_t> dact;
    void *d_dval = nullptr;
    uint8_t *d_dact = nullptr;
    std::vector<double> dface;     // per-voxel face values (face_x/y/z layout); empty: dval
    void *d_dface = nullptr;       // their device copy in the field's precision
    // agents (device CSR)
    int nseg = 0, nact = 0;
    int *seg_start = nullptr;
    long long *seg_off = nullptr;
    double *t1 = nullptr, *t2 = nullptr;
    size_t seg_cap = 0, act_cap = 0;  // segments / agents the device arrays have room for
    int *xseg = nullptr;           // first segment of each x-sweep tile (rebuilt with the descriptors)
    std::vector<int> agent_order;  // sorted position -> index in the caller's agent list (after
                                   // moves: current once domain_agent_order has run)
    std::vector<long long> agent_off;  // device offset of each sorted agent's voxel, as set
    int agents_given = 0;          // length of the caller's agent list
    struct AgentRemap *remap = nullptr;  // device regrouping state of moving agents (agents.hip)
    double *bulk = nullptr;        // bulk sources/sinks: T*S then U+S, each [nsub][nz][ny][nx]
    void *grad = nullptr;          // gradients [nsub*3 + axis][nz][ny][nx], field precision
    bool grad_valid = false;       // grad holds the gradients of some field
};

// Offsets inside a domain's coefficient block, in doubles.
struct CoefLayout {
    size_t c1, den[3], cc[3], rden[3], total;
};
CoefLayout coef_layout(const Domain &d);

struct Group {
    int precision = BFVMB_FP64;
    int device = 0;
    size_t elem = 8;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    std::vector<Domain> doms;
    int cur = 0;  // which buffer of every domain holds the current field
    LaunchShape launch[3];
    size_t max_lds_per_block = 0;
    int cu_count = 0;
    // device descriptors and per-phase prefix sums (0 = agent work items, 1..3 = x,y,z
    // tiles, 4 = bulk elements)
    void *d_desc = nullptr;
    long long *d_prefix = nullptr;
    std::vector<long long> prefix[5];
    bool desc_dirty = true;
    // FP64 steps apply the cell sources and sinks inside the x sweep's launch when no x-sweep
    // tile holds more than kFuseMaxSegsPerTile occupied voxels (set with the descriptors)
    bool fuse_agents = false;
    void *stage = nullptr;  // staging for AoS<->SoA transfers (double)
    size_t stage_bytes = 0;
    // captured step graphs, keyed by (steps, parity, with_agents)
    std::map<std::tuple<int, int, int>, hipGraphExec_t> graphs;
    std::string err;
    // bumped by every change of field, coefficients, Dirichlet data or agents; an adjoint
    // tape is valid only in the epoch it was recorded in
    unsigned long long epoch = 0;
    struct Tape *tape = nullptr;  // adjoint checkpoints (adjoint.hip)
};

// Launchers (lod.hip, field.hip): one kernel over all domains of the group.
int launch_sweep(Group *g, int axis);   // field[cur] -> field[cur^1] for every domain
int launch_sweep_x_agents(Group *g);    // cell sources/sinks, then the x sweep, in one launch
int launch_xseg(Group *g, const Domain &d);  // xseg from the domain's segment list
// Occupied voxels per x-sweep tile above which the agent update runs as its own kernel: the
// fused update is one wavefront per tile working through its segments 256 at a time, so a
// dense tile serialises where the separate kernel spreads the work over the whole GPU.
// Measured, FP64 step at 400^3 x 4 with agents in a 3 mm sphere (densest tile; fused against
// separate, ms): 1M agents (1369) 3.36 against 3.41; 2M (2579) 3.329 against 3.348; 3M
// (3788) 3.398 against 3.389; 5M (5785) 3.573 against 3.430.  640^3 x 8 with 10M agents
// (4639): fused x sweep 12.30 ms against 9.72 + 2.26 separate.
constexpr int kFuseMaxSegsPerTile = 3072;
int launch_agents(Group *g);            // in place on field[cur]
int launch_bulk(Group *g, double dt);   // in place on field[cur], domains with bulk rates;
                                        // dt <= 0: each domain's own time step
int launch_aos_to_soa(Group *g, Domain &d, const double *d_aos, void *d_soa);
int launch_aos_to_soa_f64(Group *g, Domain &d, const double *d_aos, double *d_soa);  // any precision
int launch_soa_to_aos(Group *g, Domain &d, const void *d_soa, double *d_aos);
int sweep_lds_attributes(size_t bytes);
int sync_descriptors(Group *g);         // upload descriptors/prefixes if stale

// Host FP64 Thomas set-up (thomas.cpp).
void lod_constants(int nsub, const double *D, const double *decay, double dt, double dx,
                   double *c1, double *c2, double *c3, double *c3a);
void thomas_axis(int n, int nsub, const double *c1, const double *c2, const double *c3,
                 const double *c3a, double *den, double *cc);

// Group operations behind both ABIs (group.hip).
int set_error(Group *g, int code, const std::string &msg);
int hip_check(Group *g, hipError_t e, const char *what);
int group_init(Group *g, int ndom, const int *nx, const int *ny, const int *nz, const int *nsub,
               const double *dx, int precision, int device);
void group_free(Group *g);
int domain_set_coefficients(Group *g, int i, const double *D, const double *decay, double dt);
int domain_set_dirichlet(Group *g, int i, const double *value, const uint8_t *active);
int domain_set_dirichlet_nodes(Group *g, int i, int n, const int *voxel, const double *values);
int domain_upload(Group *g, int i, const double *rho);
int domain_download(Group *g, int i, double *rho);
int domain_set_agents(Group *g, int i, int n, const int *voxel, const double *temp1, const double *temp2,
                      const uint8_t *active);
int domain_set_bulk(Group *g, int i, const double *supply, const double *target, const double *uptake);
// moving agents (agents.hip): new positions of the agents that were set, voxels and
// regrouping on the device; free_remap releases that state
// keep_outside: assign_position's rule (a position outside the box is rejected: the agent keeps
// its voxel and stays active) instead of update_voxel_index's (the agent is deactivated)
int domain_update_agent_positions(Group *g, int i, int n, const double *pos, const double *bbox_min, int on_device,
                                  bool keep_outside = false);
int domain_agent_order(Group *g, Domain &d);  // brings Domain::agent_order up to date
// densities at listed voxels (field.hip); on_device: voxel and out are device pointers
int domain_densities_at(Group *g, int i, int n, const int *voxel, double *out, int on_device);
// the same from host arrays, staged in the library's static device scratch (no allocation);
// scratch_setup finds that array's address on the group's device (group_init calls it)
int domain_densities_at_noalloc(Group *g, int i, int n, const int *voxel, double *out);
int scratch_setup(Group *g);
int domain_upload_device(Group *g, int i, const double *d_rho);   // group.hip, stream-ordered
int domain_download_device(Group *g, int i, double *d_rho);
void free_remap(Domain &d);
void free_agents(Domain &d);  // group.hip: releases a domain's agent list (and its remap state)
int group_bulk(Group *g, double dt);
int group_steps(Group *g, int nsteps, bool with_agents, bool use_graph);
int group_set_stream(Group *g, void *stream);
int group_launch_phase(Group *g, int phase, int count);
int group_set_launch(Group *g, int sweep, long long grid, int bpc, int ring);
int group_autotune(Group *g, int reps, int *chosen);
// Adjoint (adjoint.hip): forward run recording checkpoints, then the reverse pass.
int adjoint_forward(Group *g, int nsteps, int every);
int adjoint_backward(Group *g, const double *seed, double *grad_rho0, double *grad_t1, double *grad_t2,
                     bool on_device = false);
void adjoint_free(Group *g);
// Density gradients of a single-domain group (gradient.hip).
int domain_compute_gradients(Group *g);
int domain_download_gradients(Group *g, double *out);
int domain_gradients_at(Group *g, int n, const int *voxel, double *out);

// MAT v4 files of a single-domain group (matlab.hip).
int domain_write_matlab(Group *g, const char *path, const double *origin);
int domain_read_matlab(Group *g, const char *path, const double *origin);

}  // namespace bfvmb

struct bfvmb_ctx : bfvmb::Group {};
struct bfvmb_batch : bfvmb::Group {};
