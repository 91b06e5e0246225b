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
// --sample: at every snapshot, before the field is written, the densities at the agents'
// voxels through Microenvironment::nearest_density_vectors (read from the device field),
// [nagents][nsub] doubles per snapshot.
#include <chrono>
#include <cstdio>
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

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s case.bin out.bin [fp64|fp32] [--fused] [--time] [--mat out.mat]\n", argv[0]);
        return 2;
    }
    int precision = BFVMB_FP64;
    bool fused = false, timing = false;
    const char *mat_path = nullptr;   // --mat PATH: write_to_matlab after the last step
    const char *bulk_path = nullptr;  // --bulk PATH: supply, target, uptake [nvox][nsub] doubles
    const char *sample_path = nullptr;  // --sample PATH: densities at the agents' voxels per snapshot
    const char *dnodes_path = nullptr;  // --dnodes PATH: n, voxel[n], values[n][nsub] Dirichlet nodes
    for (int i = 3; i < argc; ++i) {
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
        else if (!std::strcmp(argv[i], "fp32"))
            precision = BFVMB_FP32;
        else if (!std::strcmp(argv[i], "--fused"))
            fused = true;
        else if (!std::strcmp(argv[i], "--time"))
            timing = true;
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
        int next = 0, done = 0;
        double ms = 0.0;
        for (int s = 1; s <= nsteps; ++s) {
            const bool snap = next < nsnap && snaps[next] == s;
            if (fused && !snap && s < nsteps)
                continue;  // batch steps up to the next snapshot into one call
            const int k = s - done;
            auto ta = std::chrono::steady_clock::now();
            if (fused) {
                std::vector<biofvmb::Basic_Agent *> none;
                if (with_agents || bulk_path)
                    M.run(with_agents ? biofvmb::all_basic_agents : none, dt, k);
                else
                    for (int i = 0; i < k; ++i)
                        M.simulate_diffusion_decay(dt);
            } else {
                for (int i = 0; i < k; ++i) {
                    if (bulk_path)
                        M.simulate_bulk_sources_and_sinks(dt);
                    if (with_agents)
                        M.simulate_cell_sources_and_sinks(dt);
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
