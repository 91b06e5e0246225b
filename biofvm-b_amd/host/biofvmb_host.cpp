// biofvmb_host.cpp -- implementation of the C++ host layer (see biofvmb_host.hpp).
// Built with -ffp-contract=off: the agent constants reproduce the reference's rounding
// (one explicit fma per axpy of BioFVM_basic_agent.cpp:155-176, nothing else fused).
#include "biofvmb_host.hpp"

#include <cmath>
#include <cstdio>

namespace biofvmb {

std::vector<Basic_Agent *> all_basic_agents;

namespace {
Microenvironment *g_default_micro = nullptr;  // default_microenvironment of the reference
unsigned long long g_agent_epoch = 1;          // bumped by every agent mutation
long long g_volume_pending = 0;                // agents whose volume changed since their constants
}  // namespace

// ------------------------------------------------------------------------- Basic_Agent
Basic_Agent::Basic_Agent()
    : position(3, 0.0)
{
    static int next_id = 0;
    ID = next_id++;
    if (g_default_micro)
        register_microenvironment(g_default_micro);
}

void Basic_Agent::register_microenvironment(Microenvironment *m)
{
    micro_ = m;
    const size_t n = m ? (size_t)m->number_of_densities() : 0;
    secretion_rates.resize(n, 0.0);
    uptake_rates.resize(n, 0.0);
    saturation_densities.resize(n, 0.0);
    temp1_.resize(n, 0.0);
    temp2_.resize(n, 1.0);
    ++g_agent_epoch;
}

bool Basic_Agent::assign_position(double x, double y, double z)
{
    // BioFVM_basic_agent.cpp:112-128: a position outside the domain is rejected and the
    // agent keeps its voxel (0 for a new agent) and stays active
    const int v = micro_ ? micro_->nearest_voxel_index(x, y, z) : -1;
    if (v < 0)
        return false;
    position = {x, y, z};
    voxel_ = v;
    ++g_agent_epoch;
    return true;
}

void Basic_Agent::update_voxel_index()
{
    const int v = micro_ ? micro_->nearest_voxel_index(position[0], position[1], position[2]) : -1;
    if (v < 0) {
        voxel_ = -1;
        active_ = false;
    } else {
        voxel_ = v;
    }
    ++g_agent_epoch;
}

void Basic_Agent::set_total_volume(double v)
{
    volume_ = v;
    if (!volume_is_changed_)
        ++g_volume_pending;
    volume_is_changed_ = true;
    ++g_agent_epoch;
}

void Basic_Agent::set_active(bool a)
{
    active_ = a;
    ++g_agent_epoch;
}

void Basic_Agent::set_internal_uptake_constants(double dt)
{
    // k = dt*V_cell/V_voxel; temp1 = ((0+S)*T)*k; temp2 = fma(k,U,fma(k,S,1))
    const double k = dt * volume_ / micro_->voxel_volume();
    const size_t n = secretion_rates.size();
    temp1_.assign(n, 0.0);
    temp2_.assign(n, 1.0);
    for (size_t d = 0; d < n; ++d) {
        double t1 = 0.0;
        t1 += secretion_rates[d];
        t1 *= saturation_densities[d];
        t1 *= k;
        temp1_[d] = t1;
        temp2_[d] = std::fma(k, uptake_rates[d], std::fma(k, secretion_rates[d], 1.0));
    }
    if (volume_is_changed_)
        --g_volume_pending;
    volume_is_changed_ = false;
    ++g_agent_epoch;
}

Basic_Agent *create_basic_agent()
{
    Basic_Agent *a = new Basic_Agent;
    all_basic_agents.push_back(a);
    a->index = (int)all_basic_agents.size() - 1;
    return a;
}

void delete_all_basic_agents()
{
    for (Basic_Agent *a : all_basic_agents) {
        if (a->volume_is_changed_)
            --g_volume_pending;
        delete a;
    }
    all_basic_agents.clear();
    ++g_agent_epoch;
}

// ------------------------------------------------------------------------- Microenvironment
Microenvironment::Microenvironment(int precision, int device)
    : density_names(1, "unnamed"), density_units(1, "none"), diffusion_coefficients(1, 0.0), decay_rates(1, 0.0),
      precision_(precision), device_(device), dval_(1, 1.0), dact_(1, 1), host_(1, 0.0)
{
    if (!g_default_micro)
        g_default_micro = this;
}

Microenvironment::~Microenvironment()
{
    if (ctx_)
        bfvmb_destroy(ctx_);
    if (g_default_micro == this)
        g_default_micro = nullptr;
}

void Microenvironment::check(int rc, const char *what) const
{
    if (rc != BFVMB_OK)
        throw Error(rc, std::string(what) + ": " + bfvmb_last_error(ctx_));
}

void Microenvironment::resize_densities(int n)
{
    // BioFVM_microenvironment.cpp:709-749: field zeroed, Dirichlet values 1.0, all active
    density_names.assign(n, "unnamed");
    density_units.assign(n, "none");
    diffusion_coefficients.assign(n, 0.0);
    decay_rates.assign(n, 0.0);
    dval_.assign(n, 1.0);
    dact_.assign(n, 1);
    dnodes_.clear();
    host_.assign((size_t)number_of_voxels() * n, 0.0);
    if (ctx_) {
        bfvmb_destroy(ctx_);
        ctx_ = nullptr;
    }
    agents_pending_ = false;
    host_stale_ = false;
    host_dirty_ = true;
    dirichlet_dirty_ = true;
    solver_setup_done_ = false;
    agents_uploaded_ = false;
}

void Microenvironment::set_density(int index, const std::string &name, const std::string &units,
                                   double diffusion_constant, double decay_rate)
{
    density_names[index] = name;
    density_units[index] = units;
    diffusion_coefficients[index] = diffusion_constant;
    decay_rates[index] = decay_rate;
}

void Microenvironment::resize_space_uniform(double x_start, double x_end, double y_start, double y_end,
                                            double z_start, double z_end, double dx)
{
    // node counts as Cartesian_Mesh::resize (BioFVM_parallel.cpp:100-110); field zeroed and
    // every Dirichlet value reset to 100.0 (BioFVM_parallel.cpp:34-56)
    const double eps = 1e-16;
    nx_ = (int)std::ceil(eps + (x_end - x_start) / dx);
    ny_ = (int)std::ceil(eps + (y_end - y_start) / dx);
    nz_ = (int)std::ceil(eps + (z_end - z_start) / dx);
    dx_ = dx;
    x0_ = x_start;
    y0_ = y_start;
    z0_ = z_start;
    host_.assign((size_t)number_of_voxels() * number_of_densities(), 0.0);
    dval_.assign(number_of_densities(), 100.0);
    dnodes_.clear();
    if (ctx_) {
        bfvmb_destroy(ctx_);
        ctx_ = nullptr;
    }
    agents_pending_ = false;
    host_stale_ = false;
    host_dirty_ = true;
    dirichlet_dirty_ = true;
    solver_setup_done_ = false;
    agents_uploaded_ = false;
}

int Microenvironment::nearest_voxel_index(double x, double y, double z) const
{
    const double p[3] = {x, y, z}, lo[3] = {x0_, y0_, z0_};
    int v = -1;
    check(bfvmb_voxel_indices(nx_, ny_, nz_, dx_, lo, 1, p, &v), "bfvmb_voxel_indices");
    return v;
}

void Microenvironment::set_substrate_dirichlet_value(int substrate, double value)
{
    if (substrate < 0 || substrate >= number_of_densities()) {
        std::printf("BioFVM Warning: Invalid substrate index %d for Dirichlet value assignment.\n", substrate);
        return;
    }
    dval_[substrate] = value;
    for (auto &kv : dnodes_)  // the reference writes the value to every voxel
        kv.second[substrate] = value;
    dirichlet_dirty_ = true;
}

void Microenvironment::add_dirichlet_node(int voxel, const std::vector<double> &value)
{
    if (voxel < 0 || voxel >= number_of_voxels())
        throw Error(BFVMB_ERR_ARG, "add_dirichlet_node: voxel index out of range");
    auto it = dnodes_.find(voxel);
    if (it == dnodes_.end())
        it = dnodes_.emplace(voxel, dval_).first;
    for (size_t d = 0; d < value.size() && d < it->second.size(); ++d)
        it->second[d] = value[d];
    dirichlet_dirty_ = true;
}

void Microenvironment::set_substrate_dirichlet_activation(int substrate, bool active)
{
    dact_[substrate] = active ? 1 : 0;
    dirichlet_dirty_ = true;
}

void Microenvironment::ensure_context()
{
    if (ctx_)
        return;
    int rc = bfvmb_create(&ctx_, nx_, ny_, nz_, number_of_densities(), dx_, precision_, device_);
    if (rc != BFVMB_OK)
        throw Error(rc, std::string("bfvmb_create: ") + bfvmb_last_error(nullptr));
    host_dirty_ = true;
    dirichlet_dirty_ = true;
    agents_uploaded_ = false;
}

void Microenvironment::push_dirichlet()
{
    if (!dirichlet_dirty_)
        return;
    check(bfvmb_set_dirichlet(ctx_, dval_.data(), dact_.data()), "bfvmb_set_dirichlet");
    if (!dnodes_.empty()) {
        const size_t ns = dval_.size();
        std::vector<int> voxel;
        std::vector<double> values;
        voxel.reserve(dnodes_.size());
        values.reserve(dnodes_.size() * ns);
        for (const auto &kv : dnodes_) {
            voxel.push_back(kv.first);
            values.insert(values.end(), kv.second.begin(), kv.second.end());
        }
        check(bfvmb_set_dirichlet_nodes(ctx_, (int)voxel.size(), voxel.data(), values.data()),
              "bfvmb_set_dirichlet_nodes");
    }
    dirichlet_dirty_ = false;
}

void Microenvironment::flush_agents()
{
    if (!agents_pending_)
        return;
    agents_pending_ = false;
    check(bfvmb_simulate_cell_sources_and_sinks(ctx_), "bfvmb_simulate_cell_sources_and_sinks");
}

void Microenvironment::sync_to_device()
{
    ensure_context();
    flush_agents();
    if (host_dirty_) {
        check(bfvmb_upload_density(ctx_, host_.data()), "bfvmb_upload_density");
        host_dirty_ = false;
        host_stale_ = false;
    }
}

void Microenvironment::sync_to_host()
{
    flush_agents();
    if (ctx_ && host_stale_) {
        check(bfvmb_download_density(ctx_, host_.data()), "bfvmb_download_density");
        host_stale_ = false;
    }
}

void zero_function(Microenvironment *m, int, double *write_destination)
{
    for (int i = 0; i < m->number_of_densities(); ++i)
        write_destination[i] = 0.0;
}

void Microenvironment::update_rates()
{
    ensure_context();
    const size_t nv = (size_t)number_of_voxels(), ns = (size_t)number_of_densities();
    std::vector<double> r[3];
    for (auto &v : r)
        v.assign(nv * ns, 0.0);
    for (size_t i = 0; i < nv; ++i) {
        bulk_supply_rate_function(this, (int)i, &r[0][i * ns]);
        bulk_supply_target_densities_function(this, (int)i, &r[1][i * ns]);
        bulk_uptake_rate_function(this, (int)i, &r[2][i * ns]);
    }
    if (r[0] == bulk_[0] && r[1] == bulk_[1] && r[2] == bulk_[2])
        return;
    check(bfvmb_set_bulk_rates(ctx_, r[0].data(), r[1].data(), r[2].data()), "bfvmb_set_bulk_rates");
    for (int k = 0; k < 3; ++k)
        bulk_[k].swap(r[k]);
}

void Microenvironment::simulate_bulk_sources_and_sinks(double dt)
{
    sync_to_device();  // runs a waiting cell update first
    update_rates();
    check(bfvmb_simulate_bulk_sources_and_sinks(ctx_, dt), "bfvmb_simulate_bulk_sources_and_sinks");
    host_stale_ = true;
}

void Microenvironment::write_to_matlab(const std::string &filename)
{
    sync_to_device();  // runs a waiting cell update first
    const double lo[3] = {x0_, y0_, z0_};
    check(bfvmb_write_matlab(ctx_, filename.c_str(), lo), "bfvmb_write_matlab");
}

void Microenvironment::read_from_matlab(const std::string &filename)
{
    sync_to_device();  // host edits (and a waiting cell update) first, so that a refused file
                       // leaves a consistent field
    const double lo[3] = {x0_, y0_, z0_};
    check(bfvmb_read_matlab(ctx_, filename.c_str(), lo), "bfvmb_read_matlab");
    host_stale_ = true;
}

double *Microenvironment::density_data()
{
    sync_to_host();
    host_dirty_ = true;  // the caller may write through the pointer
    return host_.data();
}

double *Microenvironment::density_vector(int n) { return density_data() + (size_t)n * number_of_densities(); }

std::vector<double> Microenvironment::nearest_density_vectors(const std::vector<int> &voxels)
{
    const size_t ns = (size_t)number_of_densities();
    std::vector<double> out(voxels.size() * ns);
    if (!host_stale_) {  // the host copy is current (or newer): read it
        for (size_t k = 0; k < voxels.size(); ++k) {
            const int v = voxels[k];
            if (v < 0 || v >= number_of_voxels())
                throw Error(BFVMB_ERR_ARG, "nearest_density_vectors: voxel index out of range");
            for (size_t s = 0; s < ns; ++s)
                out[k * ns + s] = host_[(size_t)v * ns + s];
        }
        return out;
    }
    flush_agents();
    check(bfvmb_densities_at(ctx_, (int)voxels.size(), voxels.data(), out.data(), 0), "bfvmb_densities_at");
    return out;
}

void Microenvironment::ensure_solver(double dt)
{
    ensure_context();
    if (!solver_setup_done_) {
        check(bfvmb_set_coefficients(ctx_, diffusion_coefficients.data(), decay_rates.data(), dt),
              "bfvmb_set_coefficients");
        solver_setup_done_ = true;
    }
    push_dirichlet();
    sync_to_device();
    if (tune_pending_) {  // on the current field, which the autotuner restores
        tune_pending_ = false;
        int chosen[9];
        check(bfvmb_autotune(ctx_, 3, chosen), "bfvmb_autotune");
    }
}

void Microenvironment::simulate_diffusion_decay(double dt)
{
    const bool agents = agents_pending_;  // ensure_solver would run it on its own
    agents_pending_ = false;
    ensure_solver(dt);
    if (agents)  // the waiting cell update and this step, as bfvmb_run launches them
        check(bfvmb_run(ctx_, 1), "bfvmb_run");
    else
        check(bfvmb_simulate_diffusion_decay(ctx_, 1), "bfvmb_simulate_diffusion_decay");
    host_stale_ = true;
}

void Microenvironment::upload_agents(std::vector<Basic_Agent *> &agents, double dt)
{
    flush_agents();  // a waiting update uses the list on the device
    // an agent whose volume changed gets new constants first (BioFVM_basic_agent.cpp:285-289);
    // the list is walked only while some agent has such a change pending
    if (g_volume_pending > 0)
        for (Basic_Agent *a : agents)
            if (a->active_ && a->volume_is_changed_)
                a->set_internal_uptake_constants(dt);
    if (agents_uploaded_ && agents_key_ == &agents && agents_n_ == agents.size() && agents_epoch_ == g_agent_epoch)
        return;
    const int ns = number_of_densities();
    const size_t n = agents.size();
    std::vector<int> voxel(n);
    std::vector<double> t1(n * ns), t2(n * ns);
    std::vector<uint8_t> active(n);
    for (size_t i = 0; i < n; ++i) {
        const Basic_Agent *a = agents[i];
        voxel[i] = a->voxel_;
        active[i] = a->active_ ? 1 : 0;
        for (int d = 0; d < ns; ++d) {
            t1[i * ns + d] = d < (int)a->temp1_.size() ? a->temp1_[d] : 0.0;
            t2[i * ns + d] = d < (int)a->temp2_.size() ? a->temp2_[d] : 1.0;
        }
    }
    check(bfvmb_set_agents(ctx_, (int)n, voxel.data(), t1.data(), t2.data(), active.data()), "bfvmb_set_agents");
    agents_key_ = &agents;
    agents_n_ = n;
    agents_uploaded_ = true;
    agents_epoch_ = g_agent_epoch;
}

void Microenvironment::simulate_cell_sources_and_sinks(std::vector<Basic_Agent *> &agents, double dt)
{
    ensure_context();
    sync_to_device();
    upload_agents(agents, dt);
    host_stale_ = true;
    if (bulk_[0].empty()) {  // waits for the next simulate_diffusion_decay
        agents_pending_ = true;
        return;
    }
    check(bfvmb_simulate_cell_sources_and_sinks(ctx_), "bfvmb_simulate_cell_sources_and_sinks");
}

void Microenvironment::run(std::vector<Basic_Agent *> &agents, double dt, int nsteps)
{
    ensure_solver(dt);  // runs a waiting cell update first
    upload_agents(agents, dt);
    check(bfvmb_run(ctx_, nsteps), "bfvmb_run");
    host_stale_ = true;
}

}  // namespace biofvmb
