// bfvmb_driver.cpp -- runs one case through the C++ host layer on the GPU.
//
// Reads the same case files as the reference driver used for the golden vectors and
// writes the same output format, so the two are compared directly.  The set-up and the
// time loop follow the reference's demo drivers (examples/complete_diffusion_demo.cpp:
// 223-240, 430-493; examples/test_VS.cpp:215-291): resize_densities + set_density,
// resize_space_uniform, Dirichlet per substrate, initial field, agents through
// create_basic_agent / assign_position / set_total_volume / rates /
// set_internal_uptake_constants, then per step simulate_cell_sources_and_sinks(dt)
// followed by simulate_diffusion_decay(dt).
//
// usage: bfvmb_driver <case.bin> <out.bin> [fp64|fp32] [--fused] [--time] [--mat out.mat]
//                     [--bulk rates.bin] [--sample samples.bin] [--dnodes nodes.bin]
//                     [--moves moves.bin moves_out.bin] [--autotune]
//        bfvmb_driver --bench-dropin <n> <nagents> <steps> [fp64|fp32]
// --bench-dropin: the drop-in workload (n^3 voxels x 4 substrates, agents in a centred sphere
// of radius 3 mm * n/400, lognormal(0, 0.5) fields; drawn with std::mt19937_64, for timing)
// through the host layer, timing `steps` steps called one at a time
// (simulate_cell_sources_and_sinks + simulate_diffusion_decay per step, as the reference's
// drivers call the slots) against one run(agents, dt, steps); prints one JSON line.
// --sample: at every snapshot, before the field is written, the densities at the agents'
// voxels through Microenvironment::nearest_density_vectors (read from the device field),
// [nagents][nsub] doubles per snapshot.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <random>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "biofvmb_host.hpp"

namespace {
template <class T>
void get(FILE *f, T *p, size_t n)
{
    if (n && std::fread(p, sizeof(T), n, f) != n) {
        std::fprintf(stderr, "bfvmb_driver: short read\n");
        std::exit(2);
    }
}
template <class T>
T one(FILE *f)
{
    T v;
    get(f, &v, 1);
    return v;
}

// per-voxel bulk rates served through the bulk_*_function slots
std::vector<double> g_bulk[3];
int g_bulk_nsub = 0;
template <int K>
void bulk_rate(biofvmb::Microenvironment *, int v, double *dst)
{
    for (int p = 0; p < g_bulk_nsub; ++p)
        dst[p] = g_bulk[K][(size_t)v * g_bulk_nsub + p];
}
}  // namespace

int bench_dropin(int n, int nagents, int steps, int precision)
{
    const double dx = 20.0, dt = 0.01;
    const double D[4] = {1e5, 1e3, 2.5e3, 5e4}, decay[4] = {0.1, 0.01, 0.4, 1e-3};
    biofvmb::Microenvironment M(precision);
    M.resize_densities(4);
    for (int d = 0; d < 4; ++d)
        M.set_density(d, "substrate" + std::to_string(d), "dimensionless", D[d], decay[d]);
    M.resize_space_uniform(0.0, n * dx, 0.0, n * dx, 0.0, n * dx, dx);
    for (int d = 0; d < 4; ++d) {
        M.set_substrate_dirichlet_value(d, d == 0 ? 38.0 : 0.0);
        M.set_substrate_dirichlet_activation(d, d == 0);
    }
    std::mt19937_64 gen(2024);
    std::lognormal_distribution<double> logn(0.0, 0.5);
    double *rho = M.density_data();
    for (size_t i = 0; i < (size_t)n * n * n * 4; ++i)
        rho[i] = logn(gen);
    std::normal_distribution<double> normal(0.0, 1.0);
    std::uniform_real_distribution<double> unit(0.0, 1.0);
    const double c = n * dx / 2.0, radius = 3000.0 * n / 400.0;
    for (int a = 0; a < nagents; ++a) {
        double v[3], r2 = 0.0;
        for (double &x : v) {
            x = normal(gen);
            r2 += x * x;
        }
        const double r = radius * std::cbrt(unit(gen)) / std::sqrt(r2);
        biofvmb::Basic_Agent *ag = biofvmb::create_basic_agent();
        ag->register_microenvironment(&M);
        ag->assign_position(c + v[0] * r, c + v[1] * r, c + v[2] * r);
        ag->set_total_volume(4.0 / 3.0 * M_PI * 125.0);
        for (int d = 0; d < 4; ++d) {
            ag->secretion_rates[d] = 5.0 * unit(gen);
            ag->uptake_rates[d] = 20.0 * unit(gen);
            ag->saturation_densities[d] = 1.0;
        }
        ag->set_internal_uptake_constants(dt);
    }
    auto now = [] { return std::chrono::steady_clock::now(); };
    auto ms = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
    };
    for (int k = 0; k < 3; ++k) {  // warm-up: uploads, agent list, coefficients
        M.simulate_cell_sources_and_sinks(dt);
        M.simulate_diffusion_decay(dt);
    }
    M.run(biofvmb::all_basic_agents, dt, 2);
    bfvmb_sync(M.context());
    double per_call = 1e30, fused = 1e30;
    for (int rep = 0; rep < 2; ++rep) {
        auto t0 = now();
        for (int k = 0; k < steps; ++k) {
            M.simulate_cell_sources_and_sinks(dt);
            M.simulate_diffusion_decay(dt);
        }
        bfvmb_sync(M.context());
        auto t1 = now();
        M.run(biofvmb::all_basic_agents, dt, steps);
        bfvmb_sync(M.context());
        auto t2 = now();
        per_call = std::min(per_call, ms(t0, t1) / steps);
        fused = std::min(fused, ms(t1, t2) / steps);
    }
    std::printf("{\"workload\": \"host layer, %d^3 x 4, %d agents\", \"steps\": %d, \"precision\": \"%s\", "
                "\"per_call_ms_per_step\": %.6f, \"run_ms_per_step\": %.6f, \"per_call_over_run\": %.6f}\n",
                n, nagents, steps, precision == BFVMB_FP64 ? "fp64" : "fp32", per_call, fused, per_call / fused);
    biofvmb::delete_all_basic_agents();
    return 0;
}

int main(int argc, char **argv)
{
    if (argc >= 5 && !std::strcmp(argv[1], "--bench-dropin")) {
        try {
            return bench_dropin(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]),
                                argc > 5 && !std::strcmp(argv[5], "fp32") ? BFVMB_FP32 : BFVMB_FP64);
        } catch (const biofvmb::Error &e) {
            std::fprintf(stderr, "bfvmb_driver: %s (code %d)\n", e.what(), e.code);
            return 4;
        }
    }
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s case.bin out.bin [fp64|fp32] [--fused] [--time] [--mat out.mat]\n", argv[0]);
        return 2;
    }
    int precision = BFVMB_FP64;
    bool fused = false, timing = false, autotune = false;  // --autotune: M.autotune() before the steps
    const char *mat_path = nullptr;   // --mat PATH: write_to_matlab after the last step
    const char *bulk_path = nullptr;  // --bulk PATH: supply, target, uptake [nvox][nsub] doubles
    const char *sample_path = nullptr;  // --sample PATH: densities at the agents' voxels per snapshot
    // --cells-sample PATH (per-call mode): the same densities right after every step's cell
    // update, before its LOD step (the host layer then runs the waiting update on its own)
    const char *cells_sample_path = nullptr;
    const char *dnodes_path = nullptr;  // --dnodes PATH: n, voxel[n], values[n][nsub] Dirichlet nodes
    const char *moves_path = nullptr, *moves_out = nullptr;  // --moves IN OUT (the reference driver's format)
    for (int i = 3; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--moves") && i + 2 < argc) {
            moves_path = argv[++i];
            moves_out = argv[++i];
            continue;
        }
        if (!std::strcmp(argv[i], "--dnodes") && i + 1 < argc) {
            dnodes_path = argv[++i];
            continue;
        }
        if (!std::strcmp(argv[i], "--mat") && i + 1 < argc)
            mat_path = argv[++i];
        else if (!std::strcmp(argv[i], "--bulk") && i + 1 < argc)
            bulk_path = argv[++i];
        else if (!std::strcmp(argv[i], "--sample") && i + 1 < argc)
            sample_path = argv[++i];
        else if (!std::strcmp(argv[i], "--cells-sample") && i + 1 < argc)
            cells_sample_path = argv[++i];
        else if (!std::strcmp(argv[i], "fp32"))
            precision = BFVMB_FP32;
        else if (!std::strcmp(argv[i], "--fused"))
            fused = true;
        else if (!std::strcmp(argv[i], "--time"))
            timing = true;
        else if (!std::strcmp(argv[i], "--autotune"))
            autotune = true;
    }
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) {
        std::perror(argv[1]);
        return 2;
    }
    char magic[8];
    get(f, magic, 8);
    if (std::memcmp(magic, "BFVMCAS1", 8) != 0) {
        std::fprintf(stderr, "bad case magic\n");
        return 2;
    }
    const int nx = one<int>(f), ny = one<int>(f), nz = one<int>(f), nsub = one<int>(f);
    const int nsteps = one<int>(f), with_agents = one<int>(f), nagents = one<int>(f), nsnap = one<int>(f);
    const double x0 = one<double>(f), y0 = one<double>(f), z0 = one<double>(f), dx = one<double>(f), dt = one<double>(f);
    std::vector<double> D(nsub), decay(nsub), dval(nsub);
    std::vector<int> dact(nsub);
    get(f, D.data(), nsub);
    get(f, decay.data(), nsub);
    get(f, dval.data(), nsub);
    get(f, dact.data(), nsub);
    const size_t nvox = (size_t)nx * ny * nz;
    std::vector<double> init(nvox * nsub);
    get(f, init.data(), init.size());
    std::vector<double> pos(3 * (size_t)nagents), vol(nagents), S((size_t)nagents * nsub), U((size_t)nagents * nsub),
        T((size_t)nagents * nsub);
    get(f, pos.data(), pos.size());
    get(f, vol.data(), vol.size());
    get(f, S.data(), S.size());
    get(f, U.data(), U.size());
    get(f, T.data(), T.size());
    std::vector<int> snaps(nsnap);
    get(f, snaps.data(), nsnap);
    std::fclose(f);

    try {
        biofvmb::Microenvironment M(precision);
        M.resize_densities(nsub);
        for (int d = 0; d < nsub; ++d)
            M.set_density(d, "substrate" + std::to_string(d), "dimensionless", D[d], decay[d]);
        M.resize_space_uniform(x0, x0 + nx * dx, y0, y0 + ny * dx, z0, z0 + nz * dx, dx);
        if (M.x_size() != nx || M.y_size() != ny || M.z_size() != nz) {
            std::fprintf(stderr, "mesh size mismatch\n");
            return 3;
        }
        for (int d = 0; d < nsub; ++d) {
            if (dact[d] < 0)
                continue;  // keep resize_space's defaults (100.0, active)
            M.set_substrate_dirichlet_value(d, dval[d]);
            M.set_substrate_dirichlet_activation(d, dact[d] != 0);
        }
        if (dnodes_path) {  // the four node setters in turn, as the reference driver does
            FILE *q = std::fopen(dnodes_path, "rb");
            if (!q) {
                std::perror(dnodes_path);
                return 2;
            }
            const int n = one<int>(q);
            std::vector<int> vox(n);
            std::vector<double> val((size_t)n * nsub);
            get(q, vox.data(), vox.size());
            get(q, val.data(), val.size());
            std::fclose(q);
            for (int k = 0; k < n; ++k) {
                std::vector<double> v(val.begin() + (size_t)k * nsub, val.begin() + (size_t)(k + 1) * nsub);
                const int x = vox[k] / (ny * nz), y = (vox[k] / nz) % ny, z = vox[k] % nz;
                switch (k % 4) {
                case 0: M.add_dirichlet_node(vox[k], v); break;
                case 1: M.add_dirichlet_node(x, y, z, v); break;
                case 2: M.update_dirichlet_node(x, y, z, v); break;
                default: M.update_dirichlet_node(vox[k], v); break;
                }
            }
        }
        std::copy(init.begin(), init.end(), M.density_data());

        std::vector<int> voxel(nagents);
        std::vector<double> t1((size_t)nagents * nsub), t2((size_t)nagents * nsub);
        for (int a = 0; a < nagents; ++a) {
            biofvmb::Basic_Agent *ag = biofvmb::create_basic_agent();
            ag->register_microenvironment(&M);
            ag->assign_position(pos[3 * a], pos[3 * a + 1], pos[3 * a + 2]);
            ag->set_total_volume(vol[a]);
            for (int d = 0; d < nsub; ++d) {
                ag->secretion_rates[d] = S[(size_t)a * nsub + d];
                ag->uptake_rates[d] = U[(size_t)a * nsub + d];
                ag->saturation_densities[d] = T[(size_t)a * nsub + d];
            }
            ag->set_internal_uptake_constants(dt);
            voxel[a] = ag->get_current_voxel_index();
            for (int d = 0; d < nsub; ++d) {
                t1[(size_t)a * nsub + d] = ag->temp1()[d];
                t2[(size_t)a * nsub + d] = ag->temp2()[d];
            }
        }

        FILE *out = std::fopen(argv[2], "wb");
        if (!out) {
            std::perror(argv[2]);
            return 2;
        }
        std::fwrite("BFVMOUT1", 1, 8, out);
        std::fwrite(&nagents, sizeof(int), 1, out);
        std::fwrite(voxel.data(), sizeof(int), nagents, out);
        std::fwrite(t1.data(), sizeof(double), t1.size(), out);
        std::fwrite(t2.data(), sizeof(double), t2.size(), out);
        std::fwrite(&nsnap, sizeof(int), 1, out);
        if (bulk_path) {
            FILE *b = std::fopen(bulk_path, "rb");
            if (!b) {
                std::perror(bulk_path);
                return 2;
            }
            g_bulk_nsub = nsub;
            for (auto &v : g_bulk) {
                v.resize(nvox * nsub);
                get(b, v.data(), v.size());
            }
            std::fclose(b);
            M.bulk_supply_rate_function = bulk_rate<0>;
            M.bulk_supply_target_densities_function = bulk_rate<1>;
            M.bulk_uptake_rate_function = bulk_rate<2>;
            if (fused)
                M.update_rates();  // constant rates: run() applies them every step
        }
        struct Move {
            int step, route;
            std::vector<double> pos;
        };
        std::vector<Move> moves;
        if (moves_path) {
            FILE *q = std::fopen(moves_path, "rb");
            if (!q) {
                std::perror(moves_path);
                return 2;
            }
            const int nm = one<int>(q);
            for (int k = 0; k < nm; ++k) {
                Move mv;
                mv.step = one<int>(q);
                mv.route = one<int>(q);
                mv.pos.resize(3 * (size_t)nagents);
                get(q, mv.pos.data(), mv.pos.size());
                moves.push_back(mv);
            }
            std::fclose(q);
        }
        FILE *mout = moves.empty() ? nullptr : std::fopen(moves_out, "wb");
        auto move_agents = [&](int s) {  // before the cell update of step s, as the reference driver
            for (const Move &mv : moves) {
                if (mv.step != s)
                    continue;
                for (int a = 0; a < nagents; ++a) {
                    biofvmb::Basic_Agent *ag = biofvmb::all_basic_agents[a];
                    const double *p = &mv.pos[3 * (size_t)a];
                    if (mv.route == 0) {
                        ag->assign_position(p[0], p[1], p[2]);
                    } else {
                        ag->position = {p[0], p[1], p[2]};
                        ag->update_voxel_index();
                    }
                }
                for (int a = 0; a < nagents; ++a) {
                    const int v = biofvmb::all_basic_agents[a]->get_current_voxel_index();
                    std::fwrite(&v, sizeof(int), 1, mout);
                }
                for (int a = 0; a < nagents; ++a) {
                    const int act = biofvmb::all_basic_agents[a]->is_active() ? 1 : 0;
                    std::fwrite(&act, sizeof(int), 1, mout);
                }
            }
        };
        if (autotune)
            M.autotune();  // at the first solver call
        int next = 0, done = 0;
        double ms = 0.0;
        for (int s = 1; s <= nsteps; ++s) {
            const bool snap = next < nsnap && snaps[next] == s;
            const bool move_next = !moves.empty() && std::any_of(moves.begin(), moves.end(),
                                                                 [&](const Move &m) { return m.step == s + 1; });
            if (fused && !snap && s < nsteps && !move_next)
                continue;  // batch steps up to the next snapshot (or move) into one call
            const int k = s - done;
            auto ta = std::chrono::steady_clock::now();
            if (fused) {
                move_agents(done + 1);
                std::vector<biofvmb::Basic_Agent *> none;
                if (with_agents || bulk_path)
                    M.run(with_agents ? biofvmb::all_basic_agents : none, dt, k);
                else
                    for (int i = 0; i < k; ++i)
                        M.simulate_diffusion_decay(dt);
            } else {
                for (int i = 0; i < k; ++i) {
                    move_agents(done + 1 + i);
                    if (bulk_path)
                        M.simulate_bulk_sources_and_sinks(dt);
                    if (with_agents)
                        M.simulate_cell_sources_and_sinks(dt);
                    if (cells_sample_path) {
                        FILE *cf = std::fopen(cells_sample_path, done + i == 0 ? "wb" : "ab");
                        if (!cf) {
                            std::perror(cells_sample_path);
                            return 3;
                        }
                        const std::vector<double> d = M.nearest_density_vectors(voxel);
                        std::fwrite(d.data(), sizeof(double), d.size(), cf);
                        std::fclose(cf);
                    }
                    M.simulate_diffusion_decay(dt);
                }
            }
            bfvmb_sync(M.context());
            ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ta).count();
            done = s;
            while (next < nsnap && snaps[next] == s) {
                if (sample_path) {
                    FILE *sf = std::fopen(sample_path, next == 0 ? "wb" : "ab");
                    if (!sf) {
                        std::perror(sample_path);
                        return 3;
                    }
                    const std::vector<double> d = M.nearest_density_vectors(voxel);
                    std::fwrite(d.data(), sizeof(double), d.size(), sf);
                    std::fclose(sf);
                }
                std::fwrite(&s, sizeof(int), 1, out);
                std::fwrite(M.density_data(), sizeof(double), nvox * nsub, out);
                ++next;
            }
        }
        std::fclose(out);
        if (mout)
            std::fclose(mout);
        if (mat_path)
            M.write_to_matlab(mat_path);
        if (timing)
            std::printf("{\"steps\": %d, \"ms_total\": %.6f}\n", nsteps, ms);
        biofvmb::delete_all_basic_agents();
    } catch (const biofvmb::Error &e) {
        std::fprintf(stderr, "bfvmb_driver: %s (code %d)\n", e.what(), e.code);
        return 4;
    }
    return 0;
}
