// ref_driver.cpp -- runs the reference BioFVM-B library itself on one case file.
//
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile against the reference sources
// (never copied here) into oracle/_ref/, it writes the golden vectors under tests/golden/
// and serves as the "reference" CPU baseline.  Nothing in the product links it.
//
// The set-up follows the harness order the reference's own drivers use
// (examples/test_VS.cpp:118-245, examples/complete_diffusion_demo.cpp:223-240,430-493):
//   resize_densities + set_density, resize_space_uniform (MPI overload, world size 1),
//   Dirichlet value/activation per substrate, initial field copied into p_density_vectors,
//   agents via create_basic_agent / register_microenvironment / assign_position /
//   set_total_volume / rates / set_internal_uptake_constants, then per step
//   simulate_cell_sources_and_sinks(dt) followed by simulate_diffusion_decay(...).
//
// usage: ref_driver <case.bin> <out.bin> [lod|avx] [granularity] [--time|--time-pipelined] [--mat out.mat]
//                   [--bulk rates.bin] [--grad grad.bin] [--dnodes nodes.bin] [--serial-agents]
//                   [--moves moves.bin moves_out.bin] [--autotune] [--fp32]
//        ref_driver_hip <case.bin> <out.bin> [hip|hipres|hipall] ...   (BFVMB_PLUGIN build: the
//        GPU solver in the diffusion_decay_solver slot; hipres keeps the field resident, hipall
//        the field and the agents)
#include <mpi.h>
#include <omp.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "BioFVM.h"
#ifdef BFVMB_PLUGIN
// the same driver with the MI355X entry points plugged into the reference's slots
#include "bfvmb_reference_plugin.hpp"
#endif

using namespace BioFVM;

namespace {

// Basic_Agent keeps the source/sink constants protected; a derived class may name them
// through a pointer to member of the base, which is then applied to any Basic_Agent.
struct AgentPeek : Basic_Agent {
    static std::vector<double> Basic_Agent::*temp1() { return &AgentPeek::cell_source_sink_solver_temp1; }
    static std::vector<double> Basic_Agent::*temp2() { return &AgentPeek::cell_source_sink_solver_temp2; }
    static bool Basic_Agent::*active() { return &AgentPeek::is_active; }
};

struct Reader {
    FILE *f;
    template <class T> void get(T *p, size_t n) {
        if (n && fread(p, sizeof(T), n, f) != n) {
            fprintf(stderr, "ref_driver: short read\n");
            exit(2);
        }
    }
    template <class T> T one() { T v; get(&v, 1); return v; }
};

// bulk supply / target / uptake per voxel ([nvox][nsub] each), served to the reference
// through its bulk_*_function slots (BioFVM_microenvironment.h:165-167)
std::vector<double> g_bulk[3];
int g_bulk_nsub = 0;
void bulk_supply(Microenvironment *, int v, double *dst)
{
    for (int p = 0; p < g_bulk_nsub; ++p)
        dst[p] = g_bulk[0][(size_t)v * g_bulk_nsub + p];
}
void bulk_target(Microenvironment *, int v, double *dst)
{
    for (int p = 0; p < g_bulk_nsub; ++p)
        dst[p] = g_bulk[1][(size_t)v * g_bulk_nsub + p];
}
void bulk_uptake(Microenvironment *, int v, double *dst)
{
    for (int p = 0; p < g_bulk_nsub; ++p)
        dst[p] = g_bulk[2][(size_t)v * g_bulk_nsub + p];
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: %s case.bin out.bin [lod|avx] [granularity] [--time|--time-pipelined]\n", argv[0]);
        return 2;
    }
    std::string solver = argc > 3 ? argv[3] : "lod";
    int gran = argc > 4 ? atoi(argv[4]) : 1;
    // --time: wall time per step, each GPU step finished inside it; --time-pipelined: the GPU
    // work is finished after the first and the last step only, so the host queues later
    // steps while earlier ones run (ms_per_later_step is then the throughput of steps 2..n)
    const bool pipelined = argc > 5 && std::string(argv[5]) == "--time-pipelined";
    bool timing = pipelined || (argc > 5 && std::string(argv[5]) == "--time");
    std::string mat_path;  // --mat PATH: Microenvironment::write_to_matlab after the last step
    std::string grad_path;  // --grad PATH: compute_all_gradient_vectors after the last step,
                            // written as [voxel][substrate][3] doubles
    std::string bulk_path;  // --bulk PATH: supply, target, uptake [nvox][nsub] doubles; per step
                            // simulate_bulk_sources_and_sinks runs before the cell update
    bool serial_agents = false;  // --serial-agents: simulate_cell_sources_and_sinks on one thread (its
                                 // omp loop races where agents share a voxel), the LOD on all
    bool autotune = false;  // --autotune (GPU slot): bfvmb_plugin_autotune before the first step
    bool fp32 = false;      // --fp32 (GPU slot): bfvmb_plugin_set_precision(M, BFVMB_FP32) first
    for (int i = 3; i < argc; ++i) {
        if (std::string(argv[i]) == "--serial-agents")
            serial_agents = true;
        if (std::string(argv[i]) == "--autotune")
            autotune = true;
        if (std::string(argv[i]) == "--fp32")
            fp32 = true;
    }
    // --moves IN OUT: agent moves, IN = nmoves, then per move: step, route, positions [n][3];
    // before the cell update of that step every agent moves by assign_position (route 0) or by
    // setting its position and calling update_voxel_index (route 1); OUT gets, per move, the
    // agents' voxels [n] (int) and active flags [n] (int) after it
    std::string moves_path, moves_out;
    for (int i = 3; i + 2 < argc; ++i)
        if (std::string(argv[i]) == "--moves") {
            moves_path = argv[i + 1];
            moves_out = argv[i + 2];
        }
    std::string cells_path;  // --cells-out PATH: p_density_vectors right after every step's cell
                             // update, before its LOD step (the GPU slot: after bfvmb_plugin_sync)
    std::string dnodes_path;  // --dnodes PATH: n, voxel[n], values[n][nsub]: per-voxel Dirichlet
                              // values set through the four public node setters in turn
    for (int i = 3; i + 1 < argc; ++i) {
        if (std::string(argv[i]) == "--dnodes")
            dnodes_path = argv[i + 1];
        if (std::string(argv[i]) == "--mat")
            mat_path = argv[i + 1];
        if (std::string(argv[i]) == "--bulk")
            bulk_path = argv[i + 1];
        if (std::string(argv[i]) == "--grad")
            grad_path = argv[i + 1];
        if (std::string(argv[i]) == "--cells-out")
            cells_path = argv[i + 1];
    }

    int provided = 0;
    MPI_Init_thread(nullptr, nullptr, MPI_THREAD_FUNNELED, &provided);
    int size = 0, rank = 0;
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    int dims[3] = {1, size, 1}, periodic[3] = {0, 0, 0}, coords[3] = {0, 0, 0};
    MPI_Comm cart;
    MPI_Cart_create(MPI_COMM_WORLD, 3, dims, periodic, 0, &cart);
    MPI_Cart_coords(cart, rank, 3, coords);

    Reader r{fopen(argv[1], "rb")};
    if (!r.f) { perror(argv[1]); return 2; }
    char magic[8];
    r.get(magic, 8);
    if (memcmp(magic, "BFVMCAS1", 8) != 0) { fprintf(stderr, "bad case magic\n"); return 2; }
    int nx = r.one<int>(), ny = r.one<int>(), nz = r.one<int>(), nsub = r.one<int>();
    int nsteps = r.one<int>(), with_agents = r.one<int>(), nagents = r.one<int>(), nsnap = r.one<int>();
    double x0 = r.one<double>(), y0 = r.one<double>(), z0 = r.one<double>();
    double dx = r.one<double>(), dt = r.one<double>();
    std::vector<double> D(nsub), decay(nsub), dval(nsub);
    std::vector<int> dact(nsub);
    r.get(D.data(), nsub); r.get(decay.data(), nsub); r.get(dval.data(), nsub); r.get(dact.data(), nsub);
    size_t nvox = (size_t)nx * ny * nz;
    std::vector<double> init(nvox * nsub);
    r.get(init.data(), init.size());
    std::vector<double> pos(3 * (size_t)nagents), vol(nagents), S((size_t)nagents * nsub),
        U((size_t)nagents * nsub), T((size_t)nagents * nsub);
    r.get(pos.data(), pos.size()); r.get(vol.data(), vol.size());
    r.get(S.data(), S.size()); r.get(U.data(), U.size()); r.get(T.data(), T.size());
    std::vector<int> snaps(nsnap);
    r.get(snaps.data(), nsnap);
    fclose(r.f);

    Microenvironment M;
    M.resize_densities(nsub);
    for (int d = 0; d < nsub; ++d)
        M.set_density(d, "substrate" + std::to_string(d), "dimensionless", D[d], decay[d]);
    M.resize_space_uniform(x0, x0 + nx * dx, y0, y0 + ny * dx, z0, z0 + nz * dx, dx, dims, coords);
    if (M.mesh.x_size != nx || M.mesh.y_size != ny || M.mesh.z_size != nz) {
        fprintf(stderr, "mesh size mismatch %d %d %d\n", M.mesh.x_size, M.mesh.y_size, M.mesh.z_size);
        return 3;
    }
    for (int d = 0; d < nsub; ++d) {
        if (dact[d] < 0)
            continue;  // keep the defaults resize_space leaves (value 100.0, active)
        M.set_substrate_dirichlet_value(d, dval[d]);
        M.set_substrate_dirichlet_activation(d, dact[d] != 0);
    }
    if (!dnodes_path.empty()) {
        // node k goes through add_dirichlet_node(voxel), add_dirichlet_node(x, y, z),
        // update_dirichlet_node(x, y, z) or update_dirichlet_node(voxel) for k % 4 = 0..3
        // (BioFVM_microenvironment.cpp:226-290); the last only for a voxel not yet marked, where
        // it adds the node (the marked case searches dirichlet_indices, which nothing fills)
        Reader q{fopen(dnodes_path.c_str(), "rb")};
        if (!q.f) { perror(dnodes_path.c_str()); return 2; }
        const int n = q.one<int>();
        std::vector<int> vox(n);
        std::vector<double> val((size_t)n * nsub);
        q.get(vox.data(), vox.size());
        q.get(val.data(), val.size());
        fclose(q.f);
        for (int k = 0; k < n; ++k) {
            std::vector<double> v(val.begin() + (size_t)k * nsub, val.begin() + (size_t)(k + 1) * nsub);
            const int x = vox[k] / (ny * nz), y = (vox[k] / nz) % ny, z = vox[k] % nz;
            int route = k % 4;
            if (route == 3 && M.is_dirichlet_node(vox[k]))
                route = 2;
            if (route == 0)
                M.add_dirichlet_node(vox[k], v);
            else if (route == 1)
                M.add_dirichlet_node(x, y, z, v);
            else if (route == 2)
                M.update_dirichlet_node(x, y, z, v);
            else
                M.update_dirichlet_node(vox[k], v);
        }
    }
    typedef void (*slot_t)(Microenvironment &, double, int, int, int *, int *, MPI_Comm);
    slot_t lod = diffusion_decay_solver__constant_coefficients_LOD_3D;
    slot_t avx = diffusion_decay_solver__constant_coefficients_LOD_3D_AVX256D;
    M.diffusion_decay_solver = solver == "avx" ? avx : lod;
    const bool hip = solver == "hip" || solver == "hipres" || solver == "hipall";
#ifdef BFVMB_PLUGIN
    if (hip)
        M.diffusion_decay_solver = bfvmb_hip_LOD_3D;
    if (hip && fp32)
        bfvmb_plugin_set_precision(M, BFVMB_FP32);
#else
    if (hip) {
        fprintf(stderr, "built without the GPU plugin\n");
        return 2;
    }
#endif
    M.granurality = gran;
    if (!bulk_path.empty()) {
        Reader b{fopen(bulk_path.c_str(), "rb")};
        if (!b.f) { perror(bulk_path.c_str()); return 2; }
        g_bulk_nsub = nsub;
        for (auto &v : g_bulk) {
            v.resize(nvox * nsub);
            b.get(v.data(), v.size());
        }
        fclose(b.f);
        M.bulk_supply_rate_function = bulk_supply;
        M.bulk_supply_target_densities_function = bulk_target;
        M.bulk_uptake_rate_function = bulk_uptake;
    }
    std::copy(init.begin(), init.end(), M.p_density_vectors->begin());
    std::vector<double>().swap(init);  // large cases: keep only the library's copy

    std::vector<int> voxel(nagents);
    std::vector<double> t1((size_t)nagents * nsub), t2((size_t)nagents * nsub);
    for (int a = 0; a < nagents; ++a) {
        Basic_Agent *ag = create_basic_agent();
        ag->register_microenvironment(&M);
        ag->assign_position(pos[3 * a], pos[3 * a + 1], pos[3 * a + 2], rank, dims);
        ag->set_total_volume(vol[a]);
        for (int d = 0; d < nsub; ++d) {
            (*ag->secretion_rates)[d] = S[(size_t)a * nsub + d];
            (*ag->uptake_rates)[d] = U[(size_t)a * nsub + d];
            (*ag->saturation_densities)[d] = T[(size_t)a * nsub + d];
        }
        ag->set_internal_uptake_constants(dt);
        voxel[a] = ag->get_current_voxel_index();
        const std::vector<double> &a1 = ag->*AgentPeek::temp1();
        const std::vector<double> &a2 = ag->*AgentPeek::temp2();
        for (int d = 0; d < nsub; ++d) {
            t1[(size_t)a * nsub + d] = a1[d];
            t2[(size_t)a * nsub + d] = a2[d];
        }
    }

    struct Move {
        int step, route;
        std::vector<double> pos;
    };
    std::vector<Move> moves;
    if (!moves_path.empty()) {
        Reader q{fopen(moves_path.c_str(), "rb")};
        if (!q.f) { perror(moves_path.c_str()); return 2; }
        const int nm = q.one<int>();
        for (int k = 0; k < nm; ++k) {
            Move mv;
            mv.step = q.one<int>();
            mv.route = q.one<int>();
            mv.pos.resize(3 * (size_t)nagents);
            q.get(mv.pos.data(), mv.pos.size());
            moves.push_back(mv);
        }
        fclose(q.f);
    }
    FILE *mout = moves.empty() ? nullptr : fopen(moves_out.c_str(), "wb");
    FILE *out = fopen(argv[2], "wb");
    if (!out) { perror(argv[2]); return 2; }
    fwrite("BFVMOUT1", 1, 8, out);
    fwrite(&nagents, sizeof(int), 1, out);
    fwrite(voxel.data(), sizeof(int), nagents, out);
    fwrite(t1.data(), sizeof(double), t1.size(), out);
    fwrite(t2.data(), sizeof(double), t2.size(), out);
    fwrite(&nsnap, sizeof(int), 1, out);

    int next = 0;
    double total_ms = 0.0, first_ms = 0.0;  // first_ms: step 1 alone (it carries set-up work)
    for (int s = 1; s <= nsteps; ++s) {
        auto t_a = std::chrono::steady_clock::now();
        for (const Move &mv : moves) {
            if (mv.step != s)
                continue;
            for (int a = 0; a < nagents; ++a) {
                Basic_Agent *ag = all_basic_agents[a];
                const double *p = &mv.pos[3 * (size_t)a];
                if (mv.route == 0) {
                    ag->assign_position(p[0], p[1], p[2], rank, dims);
                } else {
                    ag->position[0] = p[0];
                    ag->position[1] = p[1];
                    ag->position[2] = p[2];
                    ag->update_voxel_index(rank, dims);
                }
            }
#ifdef BFVMB_PLUGIN
            if (hip)
                bfvmb_plugin_invalidate_agents(M);  // the agents changed (hipall keeps them resident)
#endif
            for (int a = 0; a < nagents; ++a) {
                const int v = all_basic_agents[a]->get_current_voxel_index();
                fwrite(&v, sizeof(int), 1, mout);
            }
            for (int a = 0; a < nagents; ++a) {
                const int act = all_basic_agents[a]->*AgentPeek::active() ? 1 : 0;
                fwrite(&act, sizeof(int), 1, mout);
            }
        }
        if (!bulk_path.empty()) {
#ifdef BFVMB_PLUGIN
            if (hip)
                bfvmb_hip_bulk_sources_and_sinks(M, dt);
            else
#endif
                M.simulate_bulk_sources_and_sinks(dt);
        }
        if (with_agents) {
#ifdef BFVMB_PLUGIN
            if (hip)
                bfvmb_hip_cell_sources_and_sinks(M, all_basic_agents, dt);
            else
#endif
            {
                const int threads = omp_get_max_threads();
                if (serial_agents)
                    omp_set_num_threads(1);
                M.simulate_cell_sources_and_sinks(dt);
                omp_set_num_threads(threads);
            }
        }
        if (!cells_path.empty() && with_agents) {
#ifdef BFVMB_PLUGIN
            if (hip)
                bfvmb_plugin_sync(M);  // runs a cell update the plugin holds back, then downloads
#endif
            FILE *cf = fopen(cells_path.c_str(), s == 1 ? "wb" : "ab");
            if (!cf) { perror(cells_path.c_str()); return 2; }
            fwrite(M.p_density_vectors->data(), sizeof(double), M.p_density_vectors->size(), cf);
            fclose(cf);
        }
#ifdef BFVMB_PLUGIN
        if ((solver == "hipres" || solver == "hipall") && s == 1)
            bfvmb_plugin_set_resident(M, true);
        if (solver == "hipall" && s == 1)
            bfvmb_plugin_set_agents_resident(M, true);
        if (hip && autotune && s == 1)
            bfvmb_plugin_autotune(M);  // at this step's solver call
#endif
        M.simulate_diffusion_decay(dt, size, rank, coords, dims, cart);
#ifdef BFVMB_PLUGIN
        if (hip && timing && (!pipelined || s == 1 || s == nsteps))
            bfvmb_plugin_finish(M);  // the step's GPU work inside the timed region
#endif
        auto t_b = std::chrono::steady_clock::now();
        total_ms += std::chrono::duration<double, std::milli>(t_b - t_a).count();
        if (s == 1)
            first_ms = total_ms;
        while (next < nsnap && snaps[next] == s) {
            fwrite(&s, sizeof(int), 1, out);
#ifdef BFVMB_PLUGIN
            if (hip)
                bfvmb_plugin_sync(M);
#endif
            fwrite(M.p_density_vectors->data(), sizeof(double), nvox * nsub, out);
            ++next;
        }
    }
    fclose(out);
    if (mout)
        fclose(mout);
    if (!grad_path.empty()) {
#ifdef BFVMB_PLUGIN
        if (hip)
            bfvmb_plugin_sync(M);
#endif
        M.compute_all_gradient_vectors();
        FILE *g = fopen(grad_path.c_str(), "wb");
        if (!g) { perror(grad_path.c_str()); return 2; }
        for (size_t v = 0; v < nvox; ++v)
            for (int d = 0; d < nsub; ++d)
                fwrite(M.gradient_vector((int)v)[d].data(), sizeof(double), 3, g);
        fclose(g);
    }
    if (!mat_path.empty()) {
#ifdef BFVMB_PLUGIN
        if (hip)
            bfvmb_plugin_write_to_matlab(M, mat_path);  // from the device when resident
        else
#endif
            M.write_to_matlab(mat_path);
    }
    if (timing)
        printf("{\"steps\": %d, \"ms_total\": %.6f, \"ms_per_step\": %.6f, \"ms_first_step\": %.6f, "
               "\"ms_per_later_step\": %.6f, \"threads\": %d}\n", nsteps, total_ms, total_ms / nsteps, first_ms,
               nsteps > 1 ? (total_ms - first_ms) / (nsteps - 1) : 0.0, omp_get_max_threads());
    MPI_Finalize();
    return 0;
}
