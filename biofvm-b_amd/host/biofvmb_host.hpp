// biofvmb_host.hpp -- C++ host layer over the C ABI (include/biofvmb.h): the part of
// BioFVM-B's Microenvironment / Basic_Agent interface that drives the LOD diffusion-decay
// step and the agent source/sink update, with the densities resident on one MI355X.
//
// Names, argument meaning and call order follow the reference for this path
// (BioFVM_microenvironment.h:66-356, BioFVM_basic_agent.h:60-119), so a driver written
// against the reference (examples/complete_diffusion_demo.cpp:379-493,
// examples/test_VS.cpp:118-291) reads the same, minus the MPI arguments (one GPU, world
// size 1).  Errors from the device library are raised as biofvmb::Error.
//
// Residency: after a step the device holds the current field.  density_vector(),
// operator() and density_data() download it on demand and mark the host copy as possibly
// modified, so the next step uploads it again; sync_to_host() / sync_to_device() make the
// transfers explicit.
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "biofvmb.h"

namespace biofvmb {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

class Microenvironment;

// Basic_Agent (BioFVM_basic_agent.h:60-119): a point source/sink at one voxel.
class Basic_Agent {
public:
    Basic_Agent();
    bool assign_position(double x, double y, double z);  // false outside the domain
    bool assign_position(const std::vector<double> &p) { return assign_position(p[0], p[1], p[2]); }
    // BioFVM_basic_agent.cpp:137-153: the voxel of `position` as it stands; a position outside
    // the domain deactivates the agent for good (voxel -1)
    void update_voxel_index();
    void set_total_volume(double v);
    double get_total_volume() const { return volume_; }
    void set_internal_uptake_constants(double dt);  // BioFVM_basic_agent.cpp:155-176
    void register_microenvironment(Microenvironment *m);
    Microenvironment *get_microenvironment() const { return micro_; }
    int get_current_voxel_index() const { return voxel_; }
    bool is_active() const { return active_; }
    void set_active(bool a);

    std::vector<double> position;
    std::vector<double> secretion_rates, uptake_rates, saturation_densities;
    const std::vector<double> &temp1() const { return temp1_; }
    const std::vector<double> &temp2() const { return temp2_; }
    int ID = 0;
    int index = 0;

private:
    friend class Microenvironment;
    friend void delete_all_basic_agents();
    Microenvironment *micro_ = nullptr;
    double volume_ = 1.0;
    bool volume_is_changed_ = false;
    int voxel_ = 0;
    bool active_ = true;
    std::vector<double> temp1_, temp2_;
};

// Global agent list and factory (BioFVM_basic_agent.cpp:55,201-208).
extern std::vector<Basic_Agent *> all_basic_agents;
Basic_Agent *create_basic_agent();
void delete_all_basic_agents();

class Microenvironment;
// Default of the bulk rate slots (BioFVM_microenvironment.cpp:101-108): zeros.
void zero_function(Microenvironment *m, int voxel_index, double *write_destination);

class Microenvironment {
public:
    // precision: BFVMB_FP64 (drop-in) or BFVMB_FP32 (FP32 storage, FP64 recurrence)
    explicit Microenvironment(int precision = BFVMB_FP64, int device = 0);
    ~Microenvironment();
    Microenvironment(const Microenvironment &) = delete;
    Microenvironment &operator=(const Microenvironment &) = delete;

    // ---- sizing, as resize_densities / set_density / resize_space_uniform
    void resize_densities(int n);
    void set_density(int index, const std::string &name, const std::string &units, double diffusion_constant,
                     double decay_rate);
    void resize_space_uniform(double x_start, double x_end, double y_start, double y_end, double z_start,
                              double z_end, double dx);
    int number_of_densities() const { return (int)density_names.size(); }
    int number_of_voxels() const { return nx_ * ny_ * nz_; }
    int x_size() const { return nx_; }
    int y_size() const { return ny_; }
    int z_size() const { return nz_; }
    double dx() const { return dx_; }
    double voxel_volume() const { return (dx_ * dx_) * dx_; }
    int voxel_index(int i, int j, int k) const { return (i * ny_ + j) * nz_ + k; }  // BioFVM_mesh.cpp:659-665
    // agent position -> voxel (-1 outside the bounding box)
    int nearest_voxel_index(double x, double y, double z) const;

    std::vector<std::string> density_names, density_units;
    std::vector<double> diffusion_coefficients, decay_rates;

    // ---- Dirichlet (BioFVM_microenvironment.cpp:328-373)
    void set_substrate_dirichlet_value(int substrate, double value);
    void set_substrate_dirichlet_activation(int substrate, bool active);
    double dirichlet_value(int substrate) const { return dval_[substrate]; }
    bool dirichlet_activation(int substrate) const { return dact_[substrate] != 0; }
    // per-voxel values (BioFVM_microenvironment.cpp:226-290): dirichlet_value_vectors at one
    // voxel, its first value.size() substrates; apply_dirichlet_conditions applies them on
    // the six faces only, so values of interior voxels never reach the field
    void add_dirichlet_node(int voxel_index, const std::vector<double> &value);
    void add_dirichlet_node(int x, int y, int z, const std::vector<double> &value)
    {
        add_dirichlet_node(voxel_index(x, y, z), value);
    }
    void update_dirichlet_node(int voxel_index, const std::vector<double> &value) { add_dirichlet_node(voxel_index, value); }
    void update_dirichlet_node(int x, int y, int z, const std::vector<double> &value)
    {
        add_dirichlet_node(voxel_index(x, y, z), value);
    }
    bool is_dirichlet_node(int voxel_index) const { return dnodes_.count(voxel_index) != 0; }

    // ---- densities in the reference layout [x][y][z][substrate]
    double *density_vector(int n);
    double *density_vector(int i, int j, int k) { return density_vector(voxel_index(i, j, k)); }
    double *operator()(int n) { return density_vector(n); }
    double *operator()(int i, int j, int k) { return density_vector(i, j, k); }
    double *density_data();  // the whole host field
    // densities ([k][substrate]) of the given voxels, read-only: what nearest_density_vector
    // (BioFVM_microenvironment.cpp:984-992) points at, gathered from the device field when it
    // is the newer copy, so that sampling a few voxels does not download the whole field
    std::vector<double> nearest_density_vectors(const std::vector<int> &voxels);
    void sync_to_host();
    void sync_to_device();

    // ---- bulk sources and sinks (BioFVM_microenvironment.h:165-167, .cpp:1110-1198): the
    // slots are evaluated for every voxel and the rates go to the device when they changed;
    // run() then starts each step with the bulk update too
    void (*bulk_supply_rate_function)(Microenvironment *, int, double *) = zero_function;
    void (*bulk_supply_target_densities_function)(Microenvironment *, int, double *) = zero_function;
    void (*bulk_uptake_rate_function)(Microenvironment *, int, double *) = zero_function;
    void update_rates();
    void simulate_bulk_sources_and_sinks(double dt);

    // ---- MAT v4 files of the field
    // Microenvironment::write_to_matlab (BioFVM_microenvironment.cpp:1079-1104): written
    // from the device, where the field lives, after any pending host edits are uploaded
    void write_to_matlab(const std::string &filename);
    // loads a multiscale_microenvironment file written for this mesh into the field (the
    // reference declares read_from_matlab, BioFVM_microenvironment.h:284, without a body)
    void read_from_matlab(const std::string &filename);

    // ---- the hot path
    // diffusion_decay_solver slot, LOD_3D: the first call freezes dt, D and decay
    // (BioFVM_parallel.cpp:714-918), as the reference does.
    void simulate_diffusion_decay(double dt);
    void simulate_cell_sources_and_sinks(std::vector<Basic_Agent *> &agents, double dt);
    void simulate_cell_sources_and_sinks(double dt) { simulate_cell_sources_and_sinks(all_basic_agents, dt); }
    // nsteps of: sources/sinks then LOD (complete_diffusion_demo.cpp:489-493), on device
    void run(std::vector<Basic_Agent *> &agents, double dt, int nsteps);
    // the next solver call (simulate_diffusion_decay or run) first times the sweeps' launch
    // shapes on the current field (bfvmb_autotune, which restores the field) and keeps the
    // fastest; results do not depend on the shapes.  The defaults are tuned at 400^3 x 4.
    void autotune() { tune_pending_ = true; }

    // the context, for direct C-ABI calls (a waiting cell update runs first)
    bfvmb_ctx *context()
    {
        flush_agents();
        return ctx_;
    }
    int precision() const { return precision_; }

private:
    void check(int rc, const char *what) const;
    void ensure_context();
    void ensure_solver(double dt);
    void upload_agents(std::vector<Basic_Agent *> &agents, double dt);
    void push_dirichlet();
    void flush_agents();

    int precision_, device_;
    int nx_ = 1, ny_ = 1, nz_ = 1;
    double dx_ = 1.0, x0_ = 0.0, y0_ = 0.0, z0_ = 0.0;
    std::vector<double> dval_;
    std::vector<uint8_t> dact_;
    std::map<int, std::vector<double>> dnodes_;  // voxel -> dirichlet_value_vectors there
    std::vector<double> host_;  // p_density_vectors
    bool host_stale_ = false;   // device is newer than host_
    bool host_dirty_ = true;    // host_ may be newer than the device
    bool solver_setup_done_ = false;
    bool dirichlet_dirty_ = true;
    const std::vector<Basic_Agent *> *agents_key_ = nullptr;
    size_t agents_n_ = 0;
    bool agents_uploaded_ = false;
    unsigned long long agents_epoch_ = 0;
    std::vector<double> bulk_[3];  // rates at the last upload
    // simulate_cell_sources_and_sinks leaves its update waiting for the next
    // simulate_diffusion_decay, which runs both as bfvmb_run does (FP64: the update inside the
    // x sweep's launch); every other method that reaches the device runs a waiting update
    // first.  Not once bulk rates are on the device (bfvmb_run would apply them again).
    bool agents_pending_ = false;
    bool tune_pending_ = false;  // autotune() at the next solver call
    bfvmb_ctx *ctx_ = nullptr;
};

}  // namespace biofvmb
