// bfvmb_reference_plugin.cpp -- the reference application's hot-path entry points, backed by
// the MI355X library (see bfvmb_reference_plugin.hpp).  Compiled against the application's
// BioFVM-B headers; it reads only members those headers make public, plus the agents'
// protected source/sink constants through a pointer to member of a derived class (legal
// C++, no header change needed).
//
// Dirichlet values and activations are private to Microenvironment and have no getter, so
// they are read back through the public apply_dirichlet_conditions (BioFVM_microenvironment
// .cpp:381-484): the plugin marks one face voxel with NaN, lets the reference overwrite it,
// and restores it.  The first call checks every face voxel this way (the library takes one
// value per substrate, which is what every public setter produces).
#include <cmath>
#include <cstdio>
#include <limits>
#include <map>
#include <vector>

#include "BioFVM.h"
#include "biofvmb.h"
#include "bfvmb_reference_plugin.hpp"

using namespace BioFVM;

namespace {

struct AgentAccess : Basic_Agent {
    static std::vector<double> Basic_Agent::*t1() { return &AgentAccess::cell_source_sink_solver_temp1; }
    static std::vector<double> Basic_Agent::*t2() { return &AgentAccess::cell_source_sink_solver_temp2; }
    static bool Basic_Agent::*active() { return &AgentAccess::is_active; }
};

struct PluginState {
    bfvmb_ctx *ctx = nullptr;
    int nx = 0, ny = 0, nz = 0, nsub = 0;
    bool coef_done = false;
    bool resident = false;
    bool device_current = false;  // the device holds the newest field
    bool host_current = true;     // p_density_vectors holds the newest field
    std::vector<double> dval;
    std::vector<uint8_t> dact;
    std::vector<double> agent_volume;  // volumes seen at the last upload
    const std::vector<Basic_Agent *> *agents_key = nullptr;
    std::vector<double> bulk[3];       // bulk supply / target / uptake at the last upload
};

std::map<Microenvironment *, PluginState> g_state;

bool fail(const char *what, bfvmb_ctx *ctx)
{
    std::printf("Error: bfvmb plugin: %s: %s\n", what, bfvmb_last_error(ctx));
    return false;
}

// Dirichlet value/activation per substrate as apply_dirichlet_conditions would apply them.
bool probe_dirichlet(Microenvironment &M, PluginState &S, bool full)
{
    const int ns = S.nsub;
    std::vector<double> &rho = *M.p_density_vectors;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    std::vector<size_t> vox;
    if (full) {
        for (int i = 0; i < S.nx; ++i)
            for (int j = 0; j < S.ny; ++j)
                for (int k = 0; k < S.nz; ++k)
                    if (i == 0 || j == 0 || k == 0 || i == S.nx - 1 || j == S.ny - 1 || k == S.nz - 1)
                        vox.push_back(((size_t)i * S.ny + j) * S.nz + k);
    } else {
        vox.push_back(0);
    }
    std::vector<double> saved(vox.size() * ns);
    for (size_t v = 0; v < vox.size(); ++v)
        for (int d = 0; d < ns; ++d) {
            saved[v * ns + d] = rho[vox[v] * ns + d];
            rho[vox[v] * ns + d] = nan;
        }
    M.apply_dirichlet_conditions(0, 1);
    S.dval.assign(ns, 0.0);
    S.dact.assign(ns, 0);
    bool ok = true;
    for (int d = 0; d < ns; ++d) {
        const double v0 = rho[vox[0] * ns + d];
        S.dact[d] = std::isnan(v0) ? 0 : 1;
        S.dval[d] = S.dact[d] ? v0 : 0.0;
        for (size_t v = 1; v < vox.size() && ok; ++v) {
            const double w = rho[vox[v] * ns + d];
            if (std::isnan(w) != std::isnan(v0) || (!std::isnan(w) && w != v0))
                ok = false;
        }
    }
    for (size_t v = 0; v < vox.size(); ++v)
        for (int d = 0; d < ns; ++d)
            rho[vox[v] * ns + d] = saved[v * ns + d];
    if (!ok)
        std::printf("Error: bfvmb plugin: Dirichlet values differ between face voxels of one substrate\n");
    return ok;
}

PluginState *state_for(Microenvironment &M)
{
    PluginState &S = g_state[&M];
    const int nx = (int)M.mesh.x_coordinates.size(), ny = (int)M.mesh.y_coordinates.size();
    const int nz = (int)M.mesh.z_coordinates.size(), ns = M.number_of_densities();
    if (S.ctx && (S.nx != nx || S.ny != ny || S.nz != nz || S.nsub != ns)) {
        bfvmb_destroy(S.ctx);
        S = PluginState();
    }
    if (!S.ctx) {
        if (bfvmb_create(&S.ctx, nx, ny, nz, ns, M.mesh.dx, BFVMB_FP64, 0) != BFVMB_OK) {
            fail("bfvmb_create", nullptr);
            S.ctx = nullptr;
            return nullptr;
        }
        S.nx = nx;
        S.ny = ny;
        S.nz = nz;
        S.nsub = ns;
        S.host_current = true;
        S.device_current = false;
    }
    return &S;
}

bool to_device(Microenvironment &M, PluginState &S)
{
    if (S.resident && S.device_current)
        return true;
    if (bfvmb_upload_density(S.ctx, M.p_density_vectors->data()) != BFVMB_OK)
        return fail("upload", S.ctx);
    S.device_current = true;
    return true;
}

bool to_host(Microenvironment &M, PluginState &S)
{
    if (bfvmb_download_density(S.ctx, M.p_density_vectors->data()) != BFVMB_OK)
        return fail("download", S.ctx);
    S.host_current = true;
    return true;
}

}  // namespace

void bfvmb_hip_LOD_3D(Microenvironment &M, double dt, int mpi_Size, int mpi_Rank, int *mpi_Coords, int *mpi_Dims,
                      MPI_Comm mpi_Cart_comm)
{
    (void)mpi_Rank;
    (void)mpi_Coords;
    (void)mpi_Dims;
    (void)mpi_Cart_comm;
    // same guards as the reference solver (BioFVM_parallel.cpp:705-710)
    if (M.mesh.uniform_mesh == false || M.mesh.Cartesian_mesh == false) {
        std::printf("Error: This algorithm is written for uniform Cartesian meshes. Try: other solvers!\n\n");
        return;
    }
    if (mpi_Size != 1) {
        std::printf("Error: bfvmb plugin: one GPU per microenvironment (MPI world size %d)\n", mpi_Size);
        return;
    }
    PluginState *S = state_for(M);
    if (!S)
        return;
    if (!S->coef_done) {  // first call freezes dt, D and decay, as the reference's set-up does
        if (bfvmb_set_coefficients(S->ctx, M.diffusion_coefficients.data(), M.decay_rates.data(), dt) != BFVMB_OK) {
            fail("set_coefficients", S->ctx);
            return;
        }
        if (!probe_dirichlet(M, *S, true))
            return;
        S->coef_done = true;
    } else if (!S->resident) {
        if (!probe_dirichlet(M, *S, false))  // values may change between calls
            return;
    }
    if (bfvmb_set_dirichlet(S->ctx, S->dval.data(), S->dact.data()) != BFVMB_OK) {
        fail("set_dirichlet", S->ctx);
        return;
    }
    if (!to_device(M, *S))
        return;
    if (bfvmb_simulate_diffusion_decay(S->ctx, 1) != BFVMB_OK) {
        fail("simulate_diffusion_decay", S->ctx);
        return;
    }
    S->host_current = false;
    if (!S->resident)
        to_host(M, *S);
}

void bfvmb_hip_cell_sources_and_sinks(Microenvironment &M, std::vector<Basic_Agent *> &agents, double dt)
{
    PluginState *S = state_for(M);
    if (!S)
        return;
    // An agent whose volume changed since the last upload gets fresh constants first, as
    // simulate_secretion_and_uptake does for volume_is_changed (BioFVM_basic_agent.cpp:285-289).
    const size_t n = agents.size();
    const bool same_list = S->agents_key == &agents && S->agent_volume.size() == n;
    const int ns = S->nsub;
    std::vector<int> voxel(n);
    std::vector<double> t1(n * ns), t2(n * ns);
    std::vector<uint8_t> act(n);
    for (size_t i = 0; i < n; ++i) {
        Basic_Agent *a = agents[i];
        if (same_list && a->get_total_volume() != S->agent_volume[i])
            a->set_internal_uptake_constants(dt);
        const std::vector<double> &a1 = a->*AgentAccess::t1();
        const std::vector<double> &a2 = a->*AgentAccess::t2();
        voxel[i] = a->get_current_voxel_index();
        act[i] = (a->*AgentAccess::active()) && voxel[i] >= 0 ? 1 : 0;
        if (voxel[i] < 0)
            voxel[i] = 0;
        for (int d = 0; d < ns; ++d) {
            t1[i * ns + d] = a1[d];
            t2[i * ns + d] = a2[d];
        }
    }
    S->agents_key = &agents;
    S->agent_volume.resize(n);
    for (size_t i = 0; i < n; ++i)
        S->agent_volume[i] = agents[i]->get_total_volume();
    if (bfvmb_set_agents(S->ctx, (int)n, voxel.data(), t1.data(), t2.data(), act.data()) != BFVMB_OK) {
        fail("set_agents", S->ctx);
        return;
    }
    if (!to_device(M, *S))
        return;
    if (bfvmb_simulate_cell_sources_and_sinks(S->ctx) != BFVMB_OK) {
        fail("simulate_cell_sources_and_sinks", S->ctx);
        return;
    }
    S->host_current = false;
    if (!S->resident)
        to_host(M, *S);
}

void bfvmb_hip_bulk_sources_and_sinks(Microenvironment &M, double dt)
{
    PluginState *S = state_for(M);
    if (!S)
        return;
    // the rates the bulk_*_function slots give now, voxel by voxel, as the reference
    // evaluates them on every call (BioFVM_microenvironment.cpp:1122-1128); uploaded when
    // they differ from the last upload
    const size_t nv = (size_t)S->nx * S->ny * S->nz, ns = (size_t)S->nsub;
    std::vector<double> r[3];
    for (auto &v : r)
        v.assign(nv * ns, 0.0);
    for (size_t i = 0; i < nv; ++i) {
        M.bulk_supply_rate_function(&M, (int)i, &r[0][i * ns]);
        M.bulk_supply_target_densities_function(&M, (int)i, &r[1][i * ns]);
        M.bulk_uptake_rate_function(&M, (int)i, &r[2][i * ns]);
    }
    if (r[0] != S->bulk[0] || r[1] != S->bulk[1] || r[2] != S->bulk[2]) {
        if (bfvmb_set_bulk_rates(S->ctx, r[0].data(), r[1].data(), r[2].data()) != BFVMB_OK) {
            fail("set_bulk_rates", S->ctx);
            return;
        }
        for (int k = 0; k < 3; ++k)
            S->bulk[k].swap(r[k]);
    }
    if (!to_device(M, *S))
        return;
    if (bfvmb_simulate_bulk_sources_and_sinks(S->ctx, dt) != BFVMB_OK) {
        fail("simulate_bulk_sources_and_sinks", S->ctx);
        return;
    }
    S->host_current = false;
    if (!S->resident)
        to_host(M, *S);
}

void bfvmb_plugin_set_resident(Microenvironment &M, bool resident)
{
    PluginState *S = state_for(M);
    if (!S)
        return;
    if (!resident && !S->host_current && S->device_current)
        to_host(M, *S);
    S->resident = resident;
}

void bfvmb_plugin_sync(Microenvironment &M)
{
    auto it = g_state.find(&M);
    if (it != g_state.end() && it->second.ctx && !it->second.host_current)
        to_host(M, it->second);
}

void bfvmb_plugin_invalidate(Microenvironment &M)
{
    auto it = g_state.find(&M);
    if (it != g_state.end())
        it->second.device_current = false;
}

void bfvmb_plugin_write_to_matlab(Microenvironment &M, const std::string &filename)
{
    // a field newer on the device is written from there (same bytes as
    // Microenvironment::write_to_matlab, BioFVM_microenvironment.cpp:1079-1104)
    auto it = g_state.find(&M);
    if (it != g_state.end() && it->second.ctx && !it->second.host_current && it->second.device_current) {
        const double lo[3] = {M.mesh.bounding_box[0], M.mesh.bounding_box[1], M.mesh.bounding_box[2]};
        if (bfvmb_write_matlab(it->second.ctx, filename.c_str(), lo) != BFVMB_OK)
            fail("bfvmb_write_matlab", it->second.ctx);
        return;
    }
    M.write_to_matlab(filename);
}

void bfvmb_plugin_release(Microenvironment &M)
{
    auto it = g_state.find(&M);
    if (it == g_state.end())
        return;
    bfvmb_plugin_sync(M);
    if (it->second.ctx)
        bfvmb_destroy(it->second.ctx);
    g_state.erase(it);
}
