// bfvmb_internal.hpp -- context and kernel interfaces shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "biofvmb.h"

namespace bfvmb {

// Per-axis Thomas tables, FP64 on both storage paths, [nsub][n] (substrate-major so that
// one substrate's table is contiguous and uniform across a wavefront).
struct AxisTables {
    const double *den;   // denominators  (thomas_denom{x,y,z})
    const double *cc;    // upper factors (thomas_c{x,y,z})
    const double *rden;  // RN(1/den)
};

template <class T>
struct SweepArgs {
    const T *__restrict__ in;   // field before the sweep   [nsub][nz][ny][nx]
    T *__restrict__ out;        // field after the sweep (a different buffer)
    const double *c1;           // [nsub]  D*dt/dx^2
    AxisTables tab;             // tables of the swept axis
    const T *dval;              // [nsub]  Dirichlet value
    const uint8_t *dact;        // [nsub]  Dirichlet activation
    int nx, ny, nz, nsub;
};

struct AgentCSR {
    int nseg = 0;                 // voxels holding at least one active agent
    int nact = 0;                 // active agents
    int *seg_start = nullptr;     // [nseg+1] offsets into the sorted agent list
    long long *seg_off = nullptr; // [nseg] voxel offset z*nx*ny + y*nx + x
    double *t1 = nullptr;         // [nsub][nact] temp1 in sorted order
    double *t2 = nullptr;         // [nsub][nact] temp2 in sorted order
};

}  // namespace bfvmb

struct bfvmb_ctx {
    int nx = 0, ny = 0, nz = 0, nsub = 0;
    double dx = 0.0;
    int precision = BFVMB_FP64;
    int device = 0;
    size_t nvox = 0;
    size_t elem = 8;  // bytes per stored real
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;

    void *field[2] = {nullptr, nullptr};  // ping-pong field buffers
    int cur = 0;                          // which buffer holds the current field

    // coefficient block (FP64): c1[nsub] | per axis den, cc, rden [nsub][n]
    double *coef = nullptr;
    bool have_coef = false;
    double dt = 0.0;
    std::vector<double> D, decay;

    std::vector<double> dval;
    std::vector<uint8_t> dact;
    void *d_dval = nullptr;   // [nsub] storage type
    uint8_t *d_dact = nullptr;

    bfvmb::AgentCSR agents;

    void *stage = nullptr;  // staging for AoS<->SoA transfers (double)
    int yz_blocks = 0, x_blocks = 0;
    int lds_blocks_per_cu = 2;        // sweep wavefronts sharing one CU's LDS
    size_t max_lds_per_block = 0;     // device limit on LDS per workgroup

    std::string err;
};

namespace bfvmb {

// Launchers (lod.hip).
int launch_sweep(bfvmb_ctx *ctx, int axis);  // one sweep, field[cur] -> field[cur^1]
// Launchers (field.hip).
int launch_agents(bfvmb_ctx *ctx);
int launch_aos_to_soa(bfvmb_ctx *ctx, const double *d_aos, void *d_soa);
int launch_soa_to_aos(bfvmb_ctx *ctx, const void *d_soa, double *d_aos);

// Host FP64 Thomas set-up (thomas.cpp).
void lod_constants(int nsub, const double *D, const double *decay, double dt, double dx,
                   double *c1, double *c2, double *c3, double *c3a);
void thomas_axis(int n, int nsub, const double *c1, const double *c2, const double *c3,
                 const double *c3a, double *den, double *cc);

// Opt the sweep kernels into the device's full LDS size (lod.hip).
int sweep_lds_attributes(size_t bytes);

// Offsets inside the coefficient block, in doubles.
struct CoefLayout {
    size_t c1, den[3], cc[3], rden[3], total;
};
CoefLayout coef_layout(const bfvmb_ctx *ctx);

int set_error(bfvmb_ctx *ctx, int code, const std::string &msg);
int hip_check(bfvmb_ctx *ctx, hipError_t e, const char *what);

}  // namespace bfvmb
