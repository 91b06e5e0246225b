/*
 * biofvmb.h -- C ABI of the MI355X (gfx950) BioFVM-B hot path: the constant-coefficient
 * LOD diffusion-decay step and the agent source/sink update.
 *
 * One context holds one microenvironment resident on one GPU, as [substrate][z][y][x]
 * arrays (x contiguous) in FP64 or FP32.  Host arrays cross the ABI in the reference's
 * own voxel-major layout, rho[((x*ny + y)*nz + z)*nsub + d]  (BioFVM_mesh.cpp:659-665,
 * BioFVM_parallel.cpp:34-39), and are transposed on the device.
 *
 * Conventions
 *   - every entry point returns an int status (BFVMB_OK == 0); no C++ exception crosses
 *     the ABI; bfvmb_last_error() gives the message of the last failure of a context;
 *   - host pointers are borrowed for the duration of the call only;
 *   - device memory is owned by the context; work is ordered on the context's stream
 *     (its own, or the one given to bfvmb_set_stream); one host thread per context;
 *   - the host copy of the field is stale after a step until bfvmb_download_density.
 */
#ifndef BIOFVMB_H
#define BIOFVMB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BFVMB_OK 0
#define BFVMB_ERR_ARG 1     /* bad argument or size */
#define BFVMB_ERR_STATE 2   /* call out of order (e.g. step before coefficients) */
#define BFVMB_ERR_HIP 3     /* HIP runtime failure */
#define BFVMB_ERR_NOMEM 4   /* device allocation failed */
#define BFVMB_ERR_NODEV 5   /* no usable gfx950 device */
#define BFVMB_ERR_IO 6      /* file could not be opened, read or written, or is malformed */

#define BFVMB_FP64 0        /* drop-in path: FP64 storage and sweeps */
#define BFVMB_FP32 1        /* bulk path: FP32 storage and sweeps, FP64 Thomas set-up */

typedef struct bfvmb_ctx bfvmb_ctx;

/* Library version string. */
const char *bfvmb_version(void);

/* Create a context for an nx*ny*nz grid with nsub substrates and voxel edge dx on HIP
 * device `device`.  Replaces the sizing done by Microenvironment::resize_densities +
 * resize_space_uniform (BioFVM_microenvironment.cpp:709-749, BioFVM_parallel.cpp:18-59):
 * the field starts at 0; Dirichlet starts at value 100.0, all substrates active. */
int bfvmb_create(bfvmb_ctx **out, int nx, int ny, int nz, int nsub, double dx, int precision,
                 int device);
int bfvmb_destroy(bfvmb_ctx *ctx);
const char *bfvmb_last_error(const bfvmb_ctx *ctx);

/* Use `stream` (a hipStream_t; NULL = the context's own stream) for all later work. */
int bfvmb_set_stream(bfvmb_ctx *ctx, void *stream);

/* Thomas set-up of diffusion_decay_solver__constant_coefficients_LOD_3D's first call
 * (BioFVM_parallel.cpp:714-918), computed on the host in FP64 and uploaded.  As in the
 * reference, dt, D and decay are frozen here; call again to change them. */
int bfvmb_set_coefficients(bfvmb_ctx *ctx, const double *diffusion, const double *decay,
                           double dt);

/* Dirichlet value and activation per substrate: set_substrate_dirichlet_value /
 * set_substrate_dirichlet_activation (BioFVM_microenvironment.cpp:328-359).  Applied to all
 * six faces before and between the sweeps, as apply_dirichlet_conditions (:381-484).
 * value[s] becomes the value of every voxel of substrate s, replacing values set per voxel
 * with bfvmb_set_dirichlet_nodes (set_substrate_dirichlet_value writes every voxel); value ==
 * NULL changes the activations only and keeps the values. */
int bfvmb_set_dirichlet(bfvmb_ctx *ctx, const double *value, const uint8_t *active);

/* Per-voxel Dirichlet values: Microenvironment::add_dirichlet_node / update_dirichlet_node
 * (BioFVM_microenvironment.cpp:226-290, declared BioFVM_microenvironment.h:253-256) for n
 * voxels (reference numbering) with values[k][substrate], i.e. dirichlet_value_vectors at those
 * voxels.  apply_dirichlet_conditions (:381-484) applies dirichlet_value_vectors on the six
 * faces only, for the substrates whose activation is on, so face voxels take these values in
 * every later step; interior voxels are accepted and leave the step unchanged.  Voxels not
 * listed keep their values (initially the per-substrate ones). */
int bfvmb_set_dirichlet_nodes(bfvmb_ctx *ctx, int n, const int *voxel, const double *values);

/* Host <-> device field transfer in the reference layout (p_density_vectors). */
int bfvmb_upload_density(bfvmb_ctx *ctx, const double *rho_voxel_major);
int bfvmb_download_density(bfvmb_ctx *ctx, double *rho_voxel_major);
/* The same transposes from / into a device buffer (FP64, voxel-major), in stream order on the
 * context's stream: no host transfer and no wait (a caller on another stream synchronizes
 * first).  For fields that GPU code produces or consumes, e.g. torch CUDA tensors. */
int bfvmb_upload_density_device(bfvmb_ctx *ctx, const double *rho_voxel_major);
int bfvmb_download_density_device(bfvmb_ctx *ctx, double *rho_voxel_major);

/* Agents for the source/sink update: voxel index (reference numbering), per-agent
 * cell_source_sink_solver_temp1/temp2 ([n][nsub], BioFVM_basic_agent.cpp:155-176) and the
 * active flag.  Agents are stably sorted by voxel inside, so that agents sharing a voxel are
 * applied in list order, as simulate_cell_sources_and_sinks does on one thread. */
int bfvmb_set_agents(bfvmb_ctx *ctx, int n, const int *voxel, const double *temp1,
                     const double *temp2, const uint8_t *active);

/* Same, from rates: temp1/temp2 are formed as set_internal_uptake_constants does, with
 * k = (dt*volume[a])/dV:  temp1 = (S*T)*k,  temp2 = fma(k,U,fma(k,S,1)). */
int bfvmb_set_agent_rates(bfvmb_ctx *ctx, int n, const int *voxel, const double *volume,
                          const double *secretion, const double *uptake,
                          const double *saturation, const uint8_t *active, double dt);

/* Host helper: voxel of each agent position ([n][3], micrometres) in a grid whose lower
 * corner is bbox_min, as Basic_Agent::assign_position + update_voxel_index compute it
 * (BioFVM_basic_agent.cpp:112-153, BioFVM_parallel.cpp:418-522 at world size 1):
 * floor((p - min)/dx) clamped into the grid.  A position outside the bounding box
 * (BioFVM_mesh.cpp:204-213) gets -1; the reference then leaves that agent at voxel 0. */
int bfvmb_voxel_indices(int nx, int ny, int nz, double dx, const double *bbox_min, int n,
                        const double *pos, int *voxel);

/* Densities at n voxels (reference numbering), out[k][substrate] in FP64: what
 * Microenvironment::nearest_density_vector(voxel) / density_vector(n) point at
 * (BioFVM_microenvironment.cpp:984-1015) and Basic_Agent::nearest_density_vector reads
 * (BioFVM_basic_agent.cpp:250-253), gathered from the resident field without downloading
 * it.  on_device != 0: voxel and out are device pointers (stream-ordered, no wait; a voxel
 * outside the grid reads as NaN); otherwise host arrays, indices checked, blocks until out is
 * filled. */
int bfvmb_densities_at(bfvmb_ctx *ctx, int n, const int *voxel, double *out, int on_device);

/* bfvmb_densities_at from host arrays (voxel[n], out[n][nsub]) that allocates no device memory:
 * the voxel list and the gathered values are staged in a static device array of
 * BFVMB_NOALLOC_SCRATCH_BYTES that the library keeps on every GPU it runs on until the process
 * ends.  No counterpart in the reference: it is for callers that cannot allocate (tight
 * per-step sensing loops, code that must not call hipMalloc).  Requires
 * 256*ceil(4n/256) + 8*n*nsub <= BFVMB_NOALLOC_SCRATCH_BYTES, else BFVMB_ERR_ARG.  All calls on
 * one GPU share that array: calls that overlap on one GPU (other host threads, other streams)
 * race.  Blocks until out is filled. */
#define BFVMB_NOALLOC_SCRATCH_BYTES (16u << 20)
int bfvmb_densities_at_noalloc(bfvmb_ctx *ctx, int n, const int *voxel, double *out);

/* Moving agents: new positions ([n][3], micrometres; grid lower corner bbox_min, NULL = the
 * origin) of the n agents last given to bfvmb_set_agents / bfvmb_set_agent_rates, as
 * Basic_Agent::update_voxel_index computes their voxels (BioFVM_basic_agent.cpp:137-153 with
 * BioFVM_parallel.cpp:418-522 at world size 1; the same arithmetic as bfvmb_voxel_indices).
 * The mapping, the stable regrouping by voxel and the reordering of temp1/temp2 run on the
 * device; the rates are kept.  An agent whose position lies outside the bounding box is
 * deactivated, as update_voxel_index does, and stays inactive until the agents are set again.
 * pos_on_device != 0: pos is a device pointer, read in stream order.  Blocks until the new
 * grouping is in place (one small read-back). */
int bfvmb_update_agent_positions(bfvmb_ctx *ctx, int n, const double *pos, const double *bbox_min,
                                 int pos_on_device);

/* The same move by Basic_Agent::assign_position (BioFVM_basic_agent.cpp:112-128): a position
 * outside the bounding box is rejected, so that agent keeps its voxel and stays active; inside
 * the box the voxel is update_voxel_index's.  Inactive agents stay inactive. */
int bfvmb_assign_agent_positions(bfvmb_ctx *ctx, int n, const double *pos, const double *bbox_min,
                                 int pos_on_device);

/* Microenvironment::simulate_cell_sources_and_sinks (BioFVM_microenvironment.cpp:1155-1169). */
int bfvmb_simulate_cell_sources_and_sinks(bfvmb_ctx *ctx);

/* Bulk sources and sinks: the per-voxel supply rate S, supply target T and uptake rate U
 * that the reference's bulk_supply_rate_function / bulk_supply_target_densities_function /
 * bulk_uptake_rate_function slots write (BioFVM_microenvironment.h:165-167), given as
 * voxel-major [voxel][substrate] arrays (NULL = the zero_function default).  T*S and U+S
 * are formed once here (update_rates, BioFVM_microenvironment.cpp:1171-1198) and kept on
 * the device; all three NULL removes the bulk phase.  Once set, each step of bfvmb_run /
 * bfvmb_run_graph starts with the bulk update (time step of bfvmb_set_coefficients). */
int bfvmb_set_bulk_rates(bfvmb_ctx *ctx, const double *supply, const double *target, const double *uptake);

/* Microenvironment::simulate_bulk_sources_and_sinks (BioFVM_microenvironment.cpp:1110-1153):
 * rho = (rho + dt*T*S) / (1 + dt*(U + S)) per voxel and substrate, with the reference's
 * rounding. */
int bfvmb_simulate_bulk_sources_and_sinks(bfvmb_ctx *ctx, double dt);

/* nsteps calls of the diffusion_decay_solver slot (BioFVM_parallel.cpp:693-1395): per step
 * Dirichlet, x sweep, Dirichlet, y sweep, Dirichlet, z sweep, Dirichlet. */
int bfvmb_simulate_diffusion_decay(bfvmb_ctx *ctx, int nsteps);

/* nsteps of the demo time loop (complete_diffusion_demo.cpp:489-493): per step, cell
 * sources/sinks then one LOD step (preceded by the bulk update when bulk rates are set).
 * In FP64 the sources/sinks run inside the x sweep's launch unless some x-sweep tile (64
 * rows y,z) holds more than 3072 occupied voxels; either way the results are those of the
 * two separate calls, bit for bit. */
int bfvmb_run(bfvmb_ctx *ctx, int nsteps);

/* bfvmb_run captured once as a HIP graph (nsteps x 4 kernel nodes) and replayed as one
 * launch; the graph is cached per (nsteps, buffer parity) until agents, coefficients
 * pointers or launch shapes change.  Same results as bfvmb_run. */
int bfvmb_run_graph(bfvmb_ctx *ctx, int nsteps);

/* Block the host until the context's stream is idle. */
int bfvmb_sync(bfvmb_ctx *ctx);

/* One kernel of the step, for measurement: phase 0 = sources/sinks, 1 = x, 2 = y, 3 = z
 * sweep, 4 = sources/sinks and x sweep as bfvmb_run launches them (FP64 with no x-sweep tile
 * of more than 3072 occupied voxels: one fused launch; otherwise phases 0 and 1); enqueued
 * `count` times on the context's stream. */
int bfvmb_launch_phase(bfvmb_ctx *ctx, int phase, int count);

/* Launch shape of sweep `sweep` (0 = x, 1 = y, 2 = z, -1 = all three); 0 keeps the
 * library default: grid size (a grid smaller than the number of line tiles strides over
 * them; the default is one block per tile, except for the FP32 x-sweep variants with
 * register chunks and the FP32 z sweep: one block per SIMD), how many sweep
 * wavefronts share one CU's LDS line store, and rows in flight per lane (y/z: 16/32/64) or
 * the x-sweep kernel variant (FP32: 1 = one column chunk in flight and 17 register chunks,
 * 3 = three chunks in flight and none, each with its Thomas tables read by scalar loads;
 * 2 and 4 = the same with the tables staged in LDS).  Used by the autotuner; results do
 * not depend on it. */
int bfvmb_set_launch(bfvmb_ctx *ctx, int sweep, int grid_blocks, int lds_blocks_per_cu, int ring_depth);

/* Autotuner: times the candidate launch shapes of each sweep (LDS wavefronts per CU 2, 3,
 * 4, 6, 8; ring depth 32/64 for y/z; the FP32 x variants; one block per tile or a grid of
 * one block per wavefront slot striding over the tiles) over `reps` launches with HIP
 * events and keeps the fastest.
 * The field is saved and restored, so tuning does not advance the simulation.  `chosen`
 * (may be NULL) receives 3 x {wavefronts per CU, ring, microseconds per launch}. */
int bfvmb_autotune(bfvmb_ctx *ctx, int reps, int *chosen);

/* ---- Adjoint.  The reference has no adjoint: these differentiate its time loop
 * (complete_diffusion_demo.cpp:489-493) through the LOD step (BioFVM_parallel.cpp:929-1392)
 * and the agent update (BioFVM_basic_agent.cpp:279-300), with the agent tables of
 * bfvmb_set_agents / bfvmb_set_agent_rates held fixed.
 *
 * bfvmb_adjoint_forward: `nsteps` of bfvmb_run from the current field rho_0 (same results),
 * recording a field checkpoint every `checkpoint_every` steps; the field is rho_T after it.
 * bfvmb_adjoint_backward: for the recorded run and a seed w = dL/drho_T (voxel-major
 * [voxel][substrate], like the densities), the gradients dL/drho_0 (same layout) and
 * dL/dtemp1, dL/dtemp2 ([agent][substrate] in the order given to set_agents, zero for
 * inactive agents).  Any gradient may be NULL.  It recomputes the forward states segment
 * by segment from the checkpoints and leaves the context as the forward run left it, so it
 * may be called again with another seed.  Any change of field, coefficients, Dirichlet
 * data or agents after the forward run (including further steps) makes it return
 * BFVMB_ERR_STATE. */
int bfvmb_adjoint_forward(bfvmb_ctx *ctx, int nsteps, int checkpoint_every);
int bfvmb_adjoint_backward(bfvmb_ctx *ctx, const double *seed, double *grad_rho0, double *grad_temp1,
                           double *grad_temp2);
/* bfvmb_adjoint_backward with device pointers for the seed and the three gradients (same
 * layouts), in stream order on the context's stream: no host transfer, no wait at the end
 * (a caller on another stream synchronizes first). */
int bfvmb_adjoint_backward_device(bfvmb_ctx *ctx, const double *seed, double *grad_rho0,
                                  double *grad_temp1, double *grad_temp2);

/* ---- Density gradients (Microenvironment::compute_all_gradient_vectors and
 * compute_gradient_vector, BioFVM_microenvironment.cpp:1243-1533): per voxel, substrate and
 * axis, (rho[+1] - rho[-1])/(2 dx) inside, one-sided (/dx) at the first and last voxel, 0
 * along an axis of length 1.  FP64 arithmetic; FP64 contexts match the reference bit for bit.
 * bfvmb_compute_gradients fills a device buffer from the current field (stream-ordered);
 * bfvmb_download_gradients copies it out as [voxel][substrate][3] (gradient_vectors'
 * layout); bfvmb_gradient_device_ptr gives its address ([nsub*3 + axis][z][y][x] of the
 * field's real type).  bfvmb_gradients_at computes them only at the listed voxels
 * (reference numbering) from the current field, out[k][substrate][3], as
 * gradient_vector(n) does for the voxel of one cell. */
int bfvmb_compute_gradients(bfvmb_ctx *ctx);
int bfvmb_download_gradients(bfvmb_ctx *ctx, double *grad_voxel_major);
int bfvmb_gradients_at(bfvmb_ctx *ctx, int n, const int *voxel, double *grad);
int bfvmb_gradient_device_ptr(bfvmb_ctx *ctx, void **ptr);

/* ---- MAT v4 output and input of the field.
 * bfvmb_write_matlab writes the file Microenvironment::write_to_matlab writes
 * (BioFVM_microenvironment.cpp:1079-1104; MPI-IO variant BioFVM_parallel.cpp:610-678 at
 * world size 1), byte for byte: variable "multiscale_microenvironment", one column per
 * voxel in mesh order holding center x, y, z, volume and the densities.  The columns are
 * assembled on the GPU from the resident field (FP32 fields are widened), so no host copy
 * of the field is needed.  `bbox_min` is the mesh's lower corner (x, y, z; NULL = origin).
 * bfvmb_read_matlab loads such a file into the field, which the reference leaves unwritten
 * (Microenvironment::read_from_matlab, BioFVM_microenvironment.h:284): the matrix must be
 * (4 + nsub) x voxels of doubles (read_matlab's checks, BioFVM_matlab.cpp:71-200) and every
 * column's center and volume must lie on the context's grid, else the field is unchanged
 * and BFVMB_ERR_ARG / BFVMB_ERR_IO is returned. */
int bfvmb_write_matlab(bfvmb_ctx *ctx, const char *filename, const double *bbox_min);
int bfvmb_read_matlab(bfvmb_ctx *ctx, const char *filename, const double *bbox_min);

/* LDS bytes per workgroup and CU count of HIP device `device`. */
int bfvmb_device_info(int device, int *lds_per_block, int *cu_count);

/* Device address of the current field, [nsub][nz][ny][nx] of 8- or 4-byte reals. */
int bfvmb_density_device_ptr(bfvmb_ctx *ctx, void **ptr);

/* ---- Batches: many independent microenvironments of uneven sizes and substrate counts,
 * advanced together.  Each phase of a step (sources/sinks, x, y, z sweep) is one kernel
 * launch over all members, and bfvmb_batch_run can capture the whole run as one HIP graph.
 * Member i has its own coefficients, Dirichlet, field and agents; the per-member calls
 * mean what the single-context calls above mean.  No counterpart in the reference, whose
 * drivers advance one Microenvironment per process (examples/test_VS.cpp:284-291). */
typedef struct bfvmb_batch bfvmb_batch;
int bfvmb_batch_create(bfvmb_batch **out, int n, const int *nx, const int *ny, const int *nz,
                       const int *nsub, const double *dx, int precision, int device);
int bfvmb_batch_destroy(bfvmb_batch *b);
const char *bfvmb_batch_last_error(const bfvmb_batch *b);
int bfvmb_batch_set_stream(bfvmb_batch *b, void *stream);
int bfvmb_batch_set_coefficients(bfvmb_batch *b, int i, const double *diffusion, const double *decay,
                                 double dt);
int bfvmb_batch_set_dirichlet(bfvmb_batch *b, int i, const double *value, const uint8_t *active);
/* bfvmb_set_dirichlet_nodes for member i. */
int bfvmb_batch_set_dirichlet_nodes(bfvmb_batch *b, int i, int n, const int *voxel, const double *values);
int bfvmb_batch_upload_density(bfvmb_batch *b, int i, const double *rho_voxel_major);
int bfvmb_batch_download_density(bfvmb_batch *b, int i, double *rho_voxel_major);
int bfvmb_batch_set_agents(bfvmb_batch *b, int i, int n, const int *voxel, const double *temp1,
                           const double *temp2, const uint8_t *active);
/* bfvmb_update_agent_positions / bfvmb_assign_agent_positions for member i. */
int bfvmb_batch_update_agent_positions(bfvmb_batch *b, int i, int n, const double *pos,
                                       const double *bbox_min, int pos_on_device);
int bfvmb_batch_assign_agent_positions(bfvmb_batch *b, int i, int n, const double *pos,
                                       const double *bbox_min, int pos_on_device);
/* bfvmb_set_bulk_rates for member i. */
int bfvmb_batch_set_bulk_rates(bfvmb_batch *b, int i, const double *supply, const double *target,
                               const double *uptake);
/* nsteps of (bulk sources/sinks where set, cell sources/sinks, LOD step) for every member;
 * use_graph != 0 replays a captured graph of the whole run. */
int bfvmb_batch_run(bfvmb_batch *b, int nsteps, int use_graph);
int bfvmb_batch_sync(bfvmb_batch *b);
int bfvmb_batch_set_launch(bfvmb_batch *b, int sweep, int grid_blocks, int lds_blocks_per_cu,
                           int ring_depth);
int bfvmb_batch_autotune(bfvmb_batch *b, int reps, int *chosen);
/* bfvmb_launch_phase for a batch: one phase for every member, `count` times. */
int bfvmb_batch_launch_phase(bfvmb_batch *b, int phase, int count);

#ifdef __cplusplus
}
#endif
#endif /* BIOFVMB_H */
