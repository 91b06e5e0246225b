// bfvmb_reference_plugin.cpp -- the reference application's hot-path entry points, backed by
// the MI355X library (see bfvmb_reference_plugin.hpp).  Compiled against the application's
// BioFVM-B headers without changing them.  It reads the members those headers make public,
// the agents' protected source/sink constants through a pointer to member of a derived
// class, and three private members -- Microenvironment::dirichlet_value_vectors and
// dirichlet_activation_vector (BioFVM_microenvironment.h:137-139), which no getter exposes, and
// Basic_Agent::volume_is_changed (BioFVM_basic_agent.h:68) -- through pointers to member that
// an explicit template instantiation hands out (the usual access checks do not apply to names
// used in an explicit instantiation, [temp.explicit]).
#include <cmath>
#include <cstdio>
#include <cstring>
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

template <class Tag, typename Tag::type P>
struct Expose {
    friend typename Tag::type member(Tag) { return P; }
};
struct DirichletValues {
    typedef std::vector<double> Microenvironment::*type;
    friend type member(DirichletValues);
};
struct DirichletActivation {
    typedef std::vector<bool> Microenvironment::*type;
    friend type member(DirichletActivation);
};
struct VolumeChanged {
    typedef bool Basic_Agent::*type;
    friend type member(VolumeChanged);
};
template struct Expose<DirichletValues, &Microenvironment::dirichlet_value_vectors>;
template struct Expose<DirichletActivation, &Microenvironment::dirichlet_activation_vector>;
template struct Expose<VolumeChanged, &Basic_Agent::volume_is_changed>;

struct PluginState {
    bfvmb_ctx *ctx = nullptr;
    int nx = 0, ny = 0, nz = 0, nsub = 0;
    int precision = BFVMB_FP64;  // bfvmb_plugin_set_precision
    bool coef_done = false;
    bool resident = false;
    bool device_current = false;  // the device holds the newest field
    bool host_current = true;     // p_density_vectors holds the newest field
    bool dirichlet_pushed = false;
    bool rescan_dirichlet = true;  // read dirichlet_value_vectors again at the next call
    std::vector<uint8_t> dact;     // activations last pushed
    std::vector<int> face_vox;     // the face voxels, reference numbering
    std::vector<double> face_val;  // dirichlet_value_vectors there ([face voxel][substrate]), last pushed
    // agents as last uploaded: voxel, active flag, temp1/temp2 [agent][substrate]
    std::vector<int> agent_voxel;
    std::vector<uint8_t> agent_active;
    std::vector<double> agent_t1, agent_t2;
    bool agents_uploaded = false;
    bool agents_resident = false;  // the caller reports agent changes (bfvmb_plugin_invalidate_agents)
    bool agents_dirty = true;
    const std::vector<Basic_Agent *> *agents_list = nullptr;
    std::vector<double> bulk[3];  // bulk supply / target / uptake at the last upload
    // With the field resident, a cell update is queued until the next solver call, which then
    // runs it and the LOD step together (bfvmb_run: in FP64 one launch for the update and the
    // x sweep); every other entry point runs a queued update first.  Not with bulk rates set,
    // since bfvmb_run would apply them again.
    bool agents_pending = false;
    bool tune_pending = false;  // bfvmb_plugin_autotune: tune at the next solver call
};

std::map<Microenvironment *, PluginState> g_state;

bool fail(const char *what, bfvmb_ctx *ctx)
{
    std::printf("Error: bfvmb plugin: %s: %s\n", what, bfvmb_last_error(ctx));
    return false;
}

bool same_value(double a, double b) { return a == b || (std::isnan(a) && std::isnan(b)); }

// The Dirichlet data apply_dirichlet_conditions (BioFVM_microenvironment.cpp:381-484) applies:
// the activation per substrate, and dirichlet_value_vectors on the six faces (values of
// interior voxels never reach the field).  A substrate whose face voxels all hold one value
// goes to the library as that value; otherwise every face voxel goes as a Dirichlet node.
// Values are read when scan_values is set; data equal to the last push are not pushed again.
bool push_dirichlet(Microenvironment &M, PluginState &S, bool scan_values)
{
    const int ns = S.nsub;
    const size_t nvox = (size_t)S.nx * S.ny * S.nz;
    const std::vector<bool> &act = M.*member(DirichletActivation());
    const std::vector<double> &val = M.*member(DirichletValues());
    if ((int)act.size() < ns || val.size() < nvox * ns) {
        std::printf("Error: bfvmb plugin: Dirichlet vectors are not sized for the mesh and substrates\n");
        return false;
    }
    std::vector<uint8_t> a(ns);
    for (int d = 0; d < ns; ++d)
        a[d] = act[d] ? 1 : 0;
    bool values_changed = false;
    if (scan_values || !S.dirichlet_pushed) {
        if (S.face_vox.empty())
            for (int i = 0; i < S.nx; ++i)
                for (int j = 0; j < S.ny; ++j)
                    for (int k = 0; k < S.nz; ++k)
                        if (i == 0 || j == 0 || k == 0 || i == S.nx - 1 || j == S.ny - 1 || k == S.nz - 1)
                            S.face_vox.push_back((i * S.ny + j) * S.nz + k);
        const size_t nf = S.face_vox.size();
        std::vector<double> fv(nf * ns);
        for (size_t v = 0; v < nf; ++v)
            std::memcpy(&fv[v * ns], &val[(size_t)S.face_vox[v] * ns], ns * sizeof(double));
        values_changed = !S.dirichlet_pushed || fv.size() != S.face_val.size() ||
                         std::memcmp(fv.data(), S.face_val.data(), fv.size() * sizeof(double)) != 0;
        if (values_changed)
            S.face_val.swap(fv);
    }
    if (!values_changed) {
        if (a == S.dact)
            return true;
        if (bfvmb_set_dirichlet(S.ctx, nullptr, a.data()) != BFVMB_OK)
            return fail("set_dirichlet", S.ctx);
        S.dact = a;
        return true;
    }
    const size_t nf = S.face_vox.size();
    std::vector<double> dval(ns);
    bool uniform = true;
    for (int d = 0; d < ns; ++d) {
        dval[d] = S.face_val[d];
        for (size_t v = 1; v < nf && uniform; ++v)
            uniform = same_value(S.face_val[v * ns + d], dval[d]);
    }
    S.dirichlet_pushed = false;
    if (bfvmb_set_dirichlet(S.ctx, dval.data(), a.data()) != BFVMB_OK)
        return fail("set_dirichlet", S.ctx);
    if (!uniform && bfvmb_set_dirichlet_nodes(S.ctx, (int)nf, S.face_vox.data(), S.face_val.data()) != BFVMB_OK)
        return fail("set_dirichlet_nodes", S.ctx);
    S.dact = a;
    S.dirichlet_pushed = true;
    return true;
}

PluginState *state_for(Microenvironment &M)
{
    PluginState &S = g_state[&M];
    const int nx = (int)M.mesh.x_coordinates.size(), ny = (int)M.mesh.y_coordinates.size();
    const int nz = (int)M.mesh.z_coordinates.size(), ns = M.number_of_densities();
    if (S.ctx && (S.nx != nx || S.ny != ny || S.nz != nz || S.nsub != ns)) {
        bfvmb_destroy(S.ctx);
        const int precision = S.precision;
        S = PluginState();
        S.precision = precision;
    }
    if (!S.ctx) {
        if (bfvmb_create(&S.ctx, nx, ny, nz, ns, M.mesh.dx, S.precision, 0) != BFVMB_OK) {
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

// runs a queued cell update (bfvmb_simulate_cell_sources_and_sinks, as if it had not waited)
bool flush_agents(PluginState &S)
{
    if (!S.agents_pending)
        return true;
    S.agents_pending = false;
    if (bfvmb_simulate_cell_sources_and_sinks(S.ctx) != BFVMB_OK)
        return fail("simulate_cell_sources_and_sinks", S.ctx);
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
        S->coef_done = true;
    }
    // activations every call; values every call while the field is synced each call, and
    // after bfvmb_plugin_invalidate when it is resident
    if (!push_dirichlet(M, *S, !S->resident || S->rescan_dirichlet))
        return;
    S->rescan_dirichlet = false;
    if (!to_device(M, *S))
        return;
    if (S->tune_pending) {  // times the shapes on the current field, then restores it
        S->tune_pending = false;
        int chosen[9];
        if (bfvmb_autotune(S->ctx, 3, chosen) != BFVMB_OK) {
            fail("autotune", S->ctx);
            return;
        }
    }
    if (S->agents_pending) {  // the queued cell update and this step, as bfvmb_run launches them
        S->agents_pending = false;
        if (bfvmb_run(S->ctx, 1) != BFVMB_OK) {
            fail("run", S->ctx);
            return;
        }
    } else if (bfvmb_simulate_diffusion_decay(S->ctx, 1) != BFVMB_OK) {
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
    if (!S || !flush_agents(*S))
        return;
    // The agents' voxels, flags and constants, gathered in parallel straight into the copy of
    // the last upload (no per-call allocation), noting whether anything changed.  An active
    // agent whose volume changed gets fresh constants first, as simulate_secretion_and_uptake
    // does (BioFVM_basic_agent.cpp:285-289).  The list goes to the device only when it differs
    // from the last upload.
    const long n = (long)agents.size();
    const int ns = S->nsub;
    const bool skip = S->agents_resident && !S->agents_dirty && S->agents_uploaded && S->agents_list == &agents &&
                      S->agent_voxel.size() == (size_t)n;
    int changed = S->agents_uploaded && S->agent_voxel.size() == (size_t)n ? 0 : 1;
    if (S->agent_voxel.size() != (size_t)n) {
        S->agent_voxel.resize(n);
        S->agent_active.resize(n);
        S->agent_t1.resize((size_t)n * ns);
        S->agent_t2.resize((size_t)n * ns);
    }
    int *vox = S->agent_voxel.data();
    uint8_t *act = S->agent_active.data();
    double *t1 = S->agent_t1.data(), *t2 = S->agent_t2.data();
#pragma omp parallel for schedule(static) reduction(| : changed) if (!skip)
    for (long i = 0; i < (skip ? 0 : n); ++i) {
        Basic_Agent *a = agents[i];
        const bool live = a->*AgentAccess::active();
        if (live && a->*member(VolumeChanged()))
            a->set_internal_uptake_constants(dt);
        const double *a1 = (a->*AgentAccess::t1()).data();
        const double *a2 = (a->*AgentAccess::t2()).data();
        const int v = a->get_current_voxel_index();
        const int vv = v < 0 ? 0 : v;
        const uint8_t on = live && v >= 0 ? 1 : 0;
        int c = vox[i] != vv || act[i] != on;
        vox[i] = vv;
        act[i] = on;
        double *d1 = t1 + (size_t)i * ns, *d2 = t2 + (size_t)i * ns;
        c |= std::memcmp(d1, a1, ns * sizeof(double)) != 0 || std::memcmp(d2, a2, ns * sizeof(double)) != 0;
        std::memcpy(d1, a1, ns * sizeof(double));
        std::memcpy(d2, a2, ns * sizeof(double));
        changed |= c;
    }
    if (changed) {
        S->agents_uploaded = false;
        if (bfvmb_set_agents(S->ctx, (int)n, vox, t1, t2, act) != BFVMB_OK) {
            fail("set_agents", S->ctx);
            return;
        }
        S->agents_uploaded = true;
    }
    S->agents_list = &agents;
    S->agents_dirty = false;
    if (!to_device(M, *S))
        return;
    if (S->resident && S->bulk[0].empty()) {  // queued for the next solver call
        S->agents_pending = true;
        S->host_current = false;
        return;
    }
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
    if (!S || !flush_agents(*S))
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
    if (!S || !flush_agents(*S))
        return;
    if (!resident && !S->host_current && S->device_current)
        to_host(M, *S);
    S->resident = resident;
}

void bfvmb_plugin_sync(Microenvironment &M)
{
    auto it = g_state.find(&M);
    if (it != g_state.end() && it->second.ctx && flush_agents(it->second) && !it->second.host_current)
        to_host(M, it->second);
}

void bfvmb_plugin_set_agents_resident(Microenvironment &M, bool resident)
{
    PluginState *S = state_for(M);
    if (!S || !flush_agents(*S))
        return;
    S->agents_resident = resident;
    S->agents_dirty = true;
}

void bfvmb_plugin_invalidate_agents(Microenvironment &M)
{
    auto it = g_state.find(&M);
    if (it != g_state.end() && flush_agents(it->second))
        it->second.agents_dirty = true;
}

void bfvmb_plugin_finish(Microenvironment &M)
{
    auto it = g_state.find(&M);
    if (it != g_state.end() && it->second.ctx && flush_agents(it->second) && bfvmb_sync(it->second.ctx) != BFVMB_OK)
        fail("bfvmb_sync", it->second.ctx);
}

void bfvmb_plugin_invalidate(Microenvironment &M)
{
    auto it = g_state.find(&M);
    if (it != g_state.end() && flush_agents(it->second)) {
        it->second.device_current = false;
        it->second.rescan_dirichlet = true;
    }
}

void bfvmb_plugin_write_to_matlab(Microenvironment &M, const std::string &filename)
{
    // a field newer on the device is written from there (same bytes as
    // Microenvironment::write_to_matlab, BioFVM_microenvironment.cpp:1079-1104)
    auto it = g_state.find(&M);
    if (it != g_state.end() && it->second.ctx && !flush_agents(it->second))
        return;
    if (it != g_state.end() && it->second.ctx && !it->second.host_current && it->second.device_current) {
        const double lo[3] = {M.mesh.bounding_box[0], M.mesh.bounding_box[1], M.mesh.bounding_box[2]};
        if (bfvmb_write_matlab(it->second.ctx, filename.c_str(), lo) != BFVMB_OK)
            fail("bfvmb_write_matlab", it->second.ctx);
        return;
    }
    M.write_to_matlab(filename);
}

void bfvmb_plugin_set_precision(Microenvironment &M, int precision)
{
    if (precision != BFVMB_FP64 && precision != BFVMB_FP32) {
        std::printf("Error: bfvmb plugin: set_precision: BFVMB_FP64 or BFVMB_FP32\n");
        return;
    }
    PluginState &S = g_state[&M];
    if (S.ctx && S.precision != precision) {  // the field goes back to the host, then a new context
        bfvmb_plugin_sync(M);
        bfvmb_destroy(S.ctx);
        S = PluginState();
    }
    S.precision = precision;
}

void bfvmb_plugin_autotune(Microenvironment &M)
{
    PluginState *S = state_for(M);
    if (S)
        S->tune_pending = true;
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
