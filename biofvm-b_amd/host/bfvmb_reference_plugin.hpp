// bfvmb_reference_plugin.hpp -- drop-in entry points for an application built on the
// reference BioFVM-B library.  Include after the application's own "BioFVM.h".
//
//   M.diffusion_decay_solver = bfvmb_hip_LOD_3D;        // replaces
//       diffusion_decay_solver__constant_coefficients_LOD_3D (BioFVM_parallel.cpp:693),
//       same slot signature (BioFVM_microenvironment.h:164)
//   bfvmb_hip_cell_sources_and_sinks(M, all_basic_agents, dt);   // replaces
//       Microenvironment::simulate_cell_sources_and_sinks (BioFVM_microenvironment.cpp:1155-1169)
//
// Both run on one MI355X through the C ABI of include/biofvmb.h.  By default every call
// uploads p_density_vectors, runs, and downloads the result, so the host field is always
// current (exactly the reference's contract).  bfvmb_plugin_set_resident(M, true) keeps the
// field on the device between calls instead; bfvmb_plugin_sync(M) then brings the host
// copy up to date and bfvmb_plugin_invalidate(M) uploads host-side edits on the next call.
// Dirichlet data: activations are read at every call; the values (dirichlet_value_vectors on
// the six faces, per-voxel values of add_dirichlet_node / update_dirichlet_node included) at
// every call while the field is synced each call, and in resident mode at the first call and
// after bfvmb_plugin_invalidate(M).  The agent list goes to the device only when a voxel,
// active flag or source/sink constant differs from the last upload.  In resident mode the
// cell update waits for the next bfvmb_hip_LOD_3D call, which runs both together (FP64: the
// update inside the x sweep's launch); every other call here runs a waiting update first.
#pragma once

#include <string>
#include <vector>

#include "biofvmb.h"  // BFVMB_FP64 / BFVMB_FP32

void bfvmb_hip_LOD_3D(BioFVM::Microenvironment &M, double dt, int mpi_Size, int mpi_Rank, int *mpi_Coords,
                      int *mpi_Dims, MPI_Comm mpi_Cart_comm);
void bfvmb_hip_cell_sources_and_sinks(BioFVM::Microenvironment &M, std::vector<BioFVM::Basic_Agent *> &agents,
                                      double dt);

// replaces Microenvironment::simulate_bulk_sources_and_sinks (BioFVM_microenvironment.cpp:1110-1153)
void bfvmb_hip_bulk_sources_and_sinks(BioFVM::Microenvironment &M, double dt);

void bfvmb_plugin_set_resident(BioFVM::Microenvironment &M, bool resident);
void bfvmb_plugin_sync(BioFVM::Microenvironment &M);        // device -> p_density_vectors
void bfvmb_plugin_invalidate(BioFVM::Microenvironment &M);  // host field edited: upload next call
void bfvmb_plugin_finish(BioFVM::Microenvironment &M);      // wait until the GPU work queued so far is done
// Agents kept on the device too: bfvmb_hip_cell_sources_and_sinks stops reading the agent list
// at every call (for a million agents that walk costs ~2.6 ms on 16 host threads) and
// reuses the last upload until bfvmb_plugin_invalidate_agents(M) reports that agents moved,
// were added or removed, or changed volume, rates or activity.
void bfvmb_plugin_set_agents_resident(BioFVM::Microenvironment &M, bool resident);
void bfvmb_plugin_invalidate_agents(BioFVM::Microenvironment &M);
void bfvmb_plugin_release(BioFVM::Microenvironment &M);     // free the device context
// The next bfvmb_hip_LOD_3D call first times the sweeps' launch shapes on this
// microenvironment's field (bfvmb_autotune) and keeps the fastest for the calls after it; the
// field is restored after the timing and results do not depend on the shapes.  The library's
// defaults are tuned at 400^3 x 4; other grids can run faster at other shapes.
void bfvmb_plugin_autotune(BioFVM::Microenvironment &M);
// BFVMB_FP32: the field is stored and swept in FP32 (FP64 recurrence, one rounding per store;
// within 1e-5 of the reference instead of bit-identical, about 1.9x the FP64 throughput).
// Call before the other plugin calls for M: a context of the other precision is released
// (its field synced to p_density_vectors first) and the settings made for it are dropped.
void bfvmb_plugin_set_precision(BioFVM::Microenvironment &M, int precision);
// Microenvironment::write_to_matlab, from the device when it holds the newer field
void bfvmb_plugin_write_to_matlab(BioFVM::Microenvironment &M, const std::string &filename);
