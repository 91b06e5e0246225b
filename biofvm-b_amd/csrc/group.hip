// group.hip -- operations on a group of microenvironments: device state, Thomas set-up,
// transfers, agents, stepping (eager or as a captured HIP graph) and the autotuner.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "bfvmb_internal.hpp"

namespace bfvmb {

int set_error(Group *g, int code, const std::string &msg)
{
    if (g)
        g->err = msg;
    return code;
}

int hip_check(Group *g, hipError_t e, const char *what)
{
    if (e == hipSuccess)
        return BFVMB_OK;
    return set_error(g, e == hipErrorOutOfMemory ? BFVMB_ERR_NOMEM : BFVMB_ERR_HIP,
                     std::string(what) + ": " + hipGetErrorString(e));
}

CoefLayout coef_layout(const Domain &d)
{
    CoefLayout L;
    const size_t s = (size_t)d.nsub;
    const int n[3] = {d.nx, d.ny, d.nz};
    size_t off = 0;
    L.c1 = off;
    off += s;
    for (int a = 0; a < 3; ++a) {
        L.den[a] = off;
        off += s * n[a];
        L.cc[a] = off;
        off += s * n[a];
        L.rden[a] = off;
        off += s * n[a];
    }
    L.total = off;
    return L;
}

namespace {

void clear_graphs(Group *g)
{
    for (auto &kv : g->graphs)
        (void)hipGraphExecDestroy(kv.second);
    g->graphs.clear();
}

int upload_dirichlet(Group *g, Domain &d)
{
    std::vector<unsigned char> vals(d.nsub * g->elem);
    for (int s = 0; s < d.nsub; ++s) {
        if (g->elem == 8) {
            const double v = d.dval[s];
            std::memcpy(vals.data() + s * 8, &v, 8);
        } else {
            const float v = (float)d.dval[s];
            std::memcpy(vals.data() + s * 4, &v, 4);
        }
    }
    int rc = hip_check(g, hipMemcpyAsync(d.d_dval, vals.data(), vals.size(), hipMemcpyHostToDevice, g->stream),
                       "dirichlet values");
    if (!rc)
        rc = hip_check(g, hipMemcpyAsync(d.d_dact, d.dact.data(), d.nsub, hipMemcpyHostToDevice, g->stream),
                       "dirichlet flags");
    if (!rc)
        rc = hip_check(g, hipStreamSynchronize(g->stream), "dirichlet sync");
    return rc;
}


template <class T>
std::vector<DomainDesc<T>> build_descs(Group *g)
{
    std::vector<DomainDesc<T>> v(g->doms.size());
    for (size_t i = 0; i < g->doms.size(); ++i) {
        Domain &d = g->doms[i];
        DomainDesc<T> &e = v[i];
        std::memset(&e, 0, sizeof(e));
        e.field[0] = static_cast<T *>(d.field[0]);
        e.field[1] = static_cast<T *>(d.field[1]);
        if (d.coef) {
            const CoefLayout L = coef_layout(d);
            e.c1 = d.coef + L.c1;
            for (int a = 0; a < 3; ++a)
                e.tab[a] = AxisTables{d.coef + L.den[a], d.coef + L.cc[a], d.coef + L.rden[a]};
        }
        e.dval = static_cast<const T *>(d.d_dval);
        e.dact = d.d_dact;
        e.dface = d.dface.empty() ? nullptr : static_cast<const T *>(d.d_dface);
        e.seg_start = d.seg_start;
        e.seg_off = d.seg_off;
        e.t1 = d.t1;
        e.t2 = d.t2;
        e.xseg = d.nseg > 0 ? d.xseg : nullptr;
        e.bulk_ts = d.bulk;
        e.bulk_us = d.bulk ? d.bulk + d.nvox * d.nsub : nullptr;
        e.dt = d.dt;
        e.nseg = d.nseg;
        e.nact = d.nact;
        e.nx = d.nx;
        e.ny = d.ny;
        e.nz = d.nz;
        e.nsub = d.nsub;
    }
    return v;
}

}  // namespace

void free_agents(Domain &d)
{
    free_remap(d);
    (void)hipFree(d.seg_start);
    (void)hipFree(d.seg_off);
    (void)hipFree(d.t1);
    (void)hipFree(d.t2);
    d.seg_start = nullptr;
    d.seg_off = nullptr;
    d.t1 = d.t2 = nullptr;
    d.nseg = d.nact = 0;
    d.seg_cap = d.act_cap = 0;
}

int sync_descriptors(Group *g)
{
    if (!g->desc_dirty)
        return BFVMB_OK;
    const size_t nd = g->doms.size();
    for (int p = 0; p < 5; ++p)
        g->prefix[p].assign(nd + 1, 0);
    for (size_t i = 0; i < nd; ++i) {
        const Domain &d = g->doms[i];
        const long long tx = (long long)((d.ny * (long long)d.nz + kWave - 1) / kWave) * d.nsub;
        const long long ty = (long long)((d.nx * (long long)d.nz + kWave - 1) / kWave) * d.nsub;
        const long long tz = (long long)((d.nx * (long long)d.ny + kWave - 1) / kWave) * d.nsub;
        g->prefix[0][i + 1] = g->prefix[0][i] + (long long)d.nseg * d.nsub;
        g->prefix[1][i + 1] = g->prefix[1][i] + tx;
        g->prefix[2][i + 1] = g->prefix[2][i] + ty;
        g->prefix[3][i + 1] = g->prefix[3][i] + tz;
        g->prefix[4][i + 1] = g->prefix[4][i] + (d.bulk ? (long long)d.nvox * d.nsub : 0);
    }
    std::vector<long long> flat;
    for (int p = 0; p < 5; ++p)
        flat.insert(flat.end(), g->prefix[p].begin(), g->prefix[p].end());
    int rc;
    const size_t desc_bytes = nd * (g->precision == BFVMB_FP64 ? sizeof(DomainDesc<double>) : sizeof(DomainDesc<float>));
    if (!g->d_desc && (rc = hip_check(g, hipMalloc(&g->d_desc, desc_bytes), "descriptors")))
        return rc;
    if (!g->d_prefix && (rc = hip_check(g, hipMalloc((void **)&g->d_prefix, flat.size() * sizeof(long long)), "prefixes")))
        return rc;
    // the host copies live until the stream synchronisation below
    std::vector<DomainDesc<double>> vd;
    std::vector<DomainDesc<float>> vf;
    if (g->precision == BFVMB_FP64) {
        vd = build_descs<double>(g);
        rc = hip_check(g, hipMemcpyAsync(g->d_desc, vd.data(), desc_bytes, hipMemcpyHostToDevice, g->stream), "descriptors");
    } else {
        vf = build_descs<float>(g);
        rc = hip_check(g, hipMemcpyAsync(g->d_desc, vf.data(), desc_bytes, hipMemcpyHostToDevice, g->stream), "descriptors");
    }
    if (!rc)
        rc = hip_check(g, hipMemcpyAsync(g->d_prefix, flat.data(), flat.size() * sizeof(long long), hipMemcpyHostToDevice,
                                         g->stream), "prefixes");
    for (size_t i = 0; i < nd && !rc; ++i)  // the x sweep's tile index of the agent lists
        rc = launch_xseg(g, g->doms[i]);
    // the densest tile decides whether the step fuses the agent update into the x sweep
    std::vector<std::vector<int>> xs(nd);
    for (size_t i = 0; i < nd && !rc; ++i) {
        const Domain &d = g->doms[i];
        if (d.nseg == 0)
            continue;
        xs[i].resize((d.ny * (size_t)d.nz + kWave - 1) / kWave + 1);
        rc = hip_check(g, hipMemcpyAsync(xs[i].data(), d.xseg, xs[i].size() * sizeof(int), hipMemcpyDeviceToHost,
                                         g->stream), "agent tile index");
    }
    if (!rc)
        rc = hip_check(g, hipStreamSynchronize(g->stream), "descriptor sync");
    else
        (void)hipStreamSynchronize(g->stream);  // copies queued before the error use the host vectors
    if (rc)
        return rc;
    int densest = 0;
    for (const auto &v : xs)
        for (size_t t = 0; t + 1 < v.size(); ++t)
            densest = std::max(densest, v[t + 1] - v[t]);
    g->fuse_agents = g->elem == 8 && densest <= kFuseMaxSegsPerTile;
    g->desc_dirty = false;
    clear_graphs(g);  // captured launches hold the old pointers
    return BFVMB_OK;
}

int group_init(Group *g, int ndom, const int *nx, const int *ny, const int *nz, const int *nsub, const double *dx,
               int precision, int device)
{
    if (ndom < 1 || !nx || !ny || !nz || !nsub || !dx || (precision != BFVMB_FP64 && precision != BFVMB_FP32))
        return set_error(g, BFVMB_ERR_ARG, "bad group arguments");
    for (int i = 0; i < ndom; ++i) {
        if (nx[i] < 1 || ny[i] < 1 || nz[i] < 1 || nsub[i] < 1 || !(dx[i] > 0.0))
            return set_error(g, BFVMB_ERR_ARG, "bad grid, substrate count or dx");
        if ((long long)ny[i] * nz[i] >= (1LL << 31) || (long long)nx[i] * ny[i] >= (1LL << 31) ||
            (long long)nx[i] * nz[i] >= (1LL << 31))
            return set_error(g, BFVMB_ERR_ARG, "plane too large");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0 || device < 0 || device >= ndev)
        return set_error(g, BFVMB_ERR_NODEV, "no HIP device");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess || std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_error(g, BFVMB_ERR_NODEV, std::string("device is not gfx950: ") + prop.gcnArchName);
    g->precision = precision;
    g->device = device;
    g->elem = precision == BFVMB_FP64 ? 8 : 4;
    g->max_lds_per_block = prop.sharedMemPerBlock;
    g->cu_count = prop.multiProcessorCount;
    int rc;
    if ((rc = hip_check(g, hipSetDevice(device), "hipSetDevice")) ||
        (rc = hip_check(g, hipStreamCreate(&g->own_stream), "hipStreamCreate")))
        return rc;
    g->stream = g->own_stream;
    if ((rc = hip_check(g, (hipError_t)sweep_lds_attributes(g->max_lds_per_block), "LDS attributes")) ||
        (rc = scratch_setup(g)))
        return rc;
    g->doms.resize(ndom);
    for (int i = 0; i < ndom; ++i) {
        Domain &d = g->doms[i];
        d.nx = nx[i];
        d.ny = ny[i];
        d.nz = nz[i];
        d.nsub = nsub[i];
        d.dx = dx[i];
        d.nvox = (size_t)nx[i] * ny[i] * nz[i];
        d.dval.assign(d.nsub, 100.0);  // resize_space's default (BioFVM_parallel.cpp:56)
        d.dact.assign(d.nsub, 1);      // resize_densities' default (BioFVM_microenvironment.cpp:743)
        const size_t bytes = d.nvox * d.nsub * g->elem;
        if ((rc = hip_check(g, hipMalloc(&d.field[0], bytes), "field buffer")) ||
            (rc = hip_check(g, hipMalloc(&d.field[1], bytes), "field buffer")) ||
            (rc = hip_check(g, hipMalloc(&d.d_dval, d.nsub * g->elem), "dirichlet values")) ||
            (rc = hip_check(g, hipMalloc((void **)&d.d_dact, d.nsub), "dirichlet flags")) ||
            (rc = hip_check(g, hipMalloc((void **)&d.xseg, ((d.ny * (size_t)d.nz + kWave - 1) / kWave + 1) * sizeof(int)),
                            "agent tile index")) ||
            (rc = hip_check(g, hipMemsetAsync(d.field[0], 0, bytes, g->stream), "zero field")) ||
            (rc = upload_dirichlet(g, d)))
            return rc;
    }
    g->desc_dirty = true;
    return BFVMB_OK;
}

void group_free(Group *g)
{
    (void)hipSetDevice(g->device);
    if (g->stream)
        (void)hipStreamSynchronize(g->stream);
    clear_graphs(g);
    for (Domain &d : g->doms) {
        free_agents(d);
        (void)hipFree(d.bulk);
        (void)hipFree(d.grad);
        (void)hipFree(d.field[0]);
        (void)hipFree(d.field[1]);
        (void)hipFree(d.coef);
        (void)hipFree(d.d_dval);
        (void)hipFree(d.d_dact);
        (void)hipFree(d.xseg);
        (void)hipFree(d.d_dface);
    }
    adjoint_free(g);
    g->doms.clear();
    (void)hipFree(g->d_desc);
    (void)hipFree(g->d_prefix);
    (void)hipFree(g->stage);
    if (g->own_stream)
        (void)hipStreamDestroy(g->own_stream);
    g->d_desc = nullptr;
    g->d_prefix = nullptr;
    g->stage = nullptr;
    g->own_stream = g->stream = nullptr;
}

int group_set_stream(Group *g, void *stream)
{
    int rc = hip_check(g, hipStreamSynchronize(g->stream), "stream switch");
    g->stream = stream ? static_cast<hipStream_t>(stream) : g->own_stream;
    return rc;
}

static bool bad_index(Group *g, int i)
{
    return i < 0 || i >= (int)g->doms.size();
}

int domain_set_coefficients(Group *g, int i, const double *D, const double *decay, double dt)
{
    if (bad_index(g, i) || !D || !decay || !(dt > 0.0))
        return set_error(g, BFVMB_ERR_ARG, "set_coefficients: bad argument");
    Domain &d = g->doms[i];
    const int ns = d.nsub;
    const int n[3] = {d.nx, d.ny, d.nz};
    std::vector<double> c1(ns), c2(ns), c3(ns), c3a(ns);
    lod_constants(ns, D, decay, dt, d.dx, c1.data(), c2.data(), c3.data(), c3a.data());
    const CoefLayout L = coef_layout(d);
    std::vector<double> blob(L.total);
    for (int s = 0; s < ns; ++s)
        blob[L.c1 + s] = c1[s];
    for (int a = 0; a < 3; ++a) {
        std::vector<double> den((size_t)ns * n[a]), cc((size_t)ns * n[a]);
        thomas_axis(n[a], ns, c1.data(), c2.data(), c3.data(), c3a.data(), den.data(), cc.data());
        for (size_t k = 0; k < den.size(); ++k) {
            blob[L.den[a] + k] = den[k];
            blob[L.cc[a] + k] = cc[k];
            blob[L.rden[a] + k] = 1.0 / den[k];
        }
    }
    const size_t bytes = blob.size() * sizeof(double);
    int rc;
    if (!d.coef) {
        if ((rc = hip_check(g, hipMalloc((void **)&d.coef, bytes), "coefficient block")))
            return rc;
        g->desc_dirty = true;
    }
    if ((rc = hip_check(g, hipMemcpyAsync(d.coef, blob.data(), bytes, hipMemcpyHostToDevice, g->stream), "coefficients")) ||
        (rc = hip_check(g, hipStreamSynchronize(g->stream), "coefficient sync")))
        return rc;
    d.D.assign(D, D + ns);
    d.decay.assign(decay, decay + ns);
    d.dt = dt;
    d.have_coef = true;
    g->desc_dirty = true;  // the descriptor carries dt
    g->epoch++;
    return BFVMB_OK;
}

int domain_set_dirichlet(Group *g, int i, const double *value, const uint8_t *active)
{
    if (bad_index(g, i) || !active)
        return set_error(g, BFVMB_ERR_ARG, "set_dirichlet: bad argument");
    Domain &d = g->doms[i];
    if (value) {
        // one value per substrate on every face voxel: set_substrate_dirichlet_value writes it
        // to every voxel (BioFVM_microenvironment.cpp:345-351), replacing per-voxel values
        int rc = hip_check(g, hipStreamSynchronize(g->stream), "dirichlet sync");
        if (rc)
            return rc;
        d.dval.assign(value, value + d.nsub);
        if (!d.dface.empty()) {
            std::vector<double>().swap(d.dface);
            (void)hipFree(d.d_dface);
            d.d_dface = nullptr;
            g->desc_dirty = true;
        }
    }
    for (int s = 0; s < d.nsub; ++s)
        d.dact[s] = active[s] ? 1 : 0;
    g->epoch++;
    return upload_dirichlet(g, d);
}

// add_dirichlet_node / update_dirichlet_node (BioFVM_microenvironment.cpp:226-290): values of
// listed voxels.  apply_dirichlet_conditions (:381-484) reads dirichlet_value_vectors on the six
// faces only, so the face planes hold everything the LOD step can see; interior voxels are
// accepted and change nothing.
int domain_set_dirichlet_nodes(Group *g, int i, int n, const int *voxel, const double *values)
{
    if (bad_index(g, i) || n < 0 || (n > 0 && (!voxel || !values)))
        return set_error(g, BFVMB_ERR_ARG, "set_dirichlet_nodes: bad argument");
    Domain &d = g->doms[i];
    const int nx = d.nx, ny = d.ny, nz = d.nz, ns = d.nsub;
    for (int k = 0; k < n; ++k)
        if (voxel[k] < 0 || (size_t)voxel[k] >= d.nvox)
            return set_error(g, BFVMB_ERR_ARG, "set_dirichlet_nodes: voxel index out of range");
    std::vector<double> f = d.dface;
    if (f.empty()) {  // start from the per-substrate values
        f.resize(face_total(nx, ny, nz, ns));
        for (int e = 0; e < 2; ++e)
            for (int s = 0; s < ns; ++s) {
                const double v = d.dval[s];
                std::fill_n(f.begin() + face_x(ny, nz, ns, e, s, 0, 0), (size_t)nz * ny, v);
                std::fill_n(f.begin() + face_y(nx, ny, nz, ns, e, s, 0, 0), (size_t)nz * nx, v);
                std::fill_n(f.begin() + face_z(nx, ny, nz, ns, e, s, 0, 0), (size_t)ny * nx, v);
            }
    }
    const long long nyz = (long long)ny * nz;
    for (int k = 0; k < n; ++k) {
        const long long v = voxel[k];  // reference numbering (x*ny + y)*nz + z
        const int x = (int)(v / nyz), y = (int)((v - x * nyz) / nz), z = (int)(v - x * nyz - (long long)y * nz);
        const double *val = values + (size_t)k * ns;
        for (int s = 0; s < ns; ++s) {
            for (int e = 0; e < 2; ++e) {
                if (x == (e ? nx - 1 : 0))
                    f[face_x(ny, nz, ns, e, s, y, z)] = val[s];
                if (y == (e ? ny - 1 : 0))
                    f[face_y(nx, ny, nz, ns, e, s, x, z)] = val[s];
                if (z == (e ? nz - 1 : 0))
                    f[face_z(nx, ny, nz, ns, e, s, x, y)] = val[s];
            }
        }
    }
    int rc = hip_check(g, hipStreamSynchronize(g->stream), "dirichlet sync");
    if (rc)
        return rc;
    if (!d.d_dface && (rc = hip_check(g, hipMalloc(&d.d_dface, f.size() * g->elem), "dirichlet faces")))
        return rc;
    if (g->elem == 8) {
        rc = hip_check(g, hipMemcpy(d.d_dface, f.data(), f.size() * 8, hipMemcpyHostToDevice), "dirichlet faces");
    } else {
        std::vector<float> h(f.begin(), f.end());
        rc = hip_check(g, hipMemcpy(d.d_dface, h.data(), h.size() * 4, hipMemcpyHostToDevice), "dirichlet faces");
    }
    if (rc) {  // back to the per-substrate values rather than half-written planes
        (void)hipFree(d.d_dface);
        d.d_dface = nullptr;
        std::vector<double>().swap(d.dface);
        g->desc_dirty = true;
        g->epoch++;
        return rc;
    }
    const bool was_on = !d.dface.empty();
    d.dface.swap(f);
    if (!was_on)
        g->desc_dirty = true;
    g->epoch++;
    return BFVMB_OK;
}

static int ensure_stage(Group *g, size_t bytes)
{
    if (g->stage && g->stage_bytes >= bytes)
        return BFVMB_OK;
    (void)hipFree(g->stage);
    g->stage = nullptr;
    g->stage_bytes = 0;
    int rc = hip_check(g, hipMalloc(&g->stage, bytes), "staging buffer");
    if (!rc)
        g->stage_bytes = bytes;
    return rc;
}

int domain_upload(Group *g, int i, const double *rho)
{
    if (bad_index(g, i) || !rho)
        return set_error(g, BFVMB_ERR_ARG, "upload_density: bad argument");
    Domain &d = g->doms[i];
    const size_t bytes = d.nvox * d.nsub * sizeof(double);
    int rc;
    g->epoch++;
    if ((rc = ensure_stage(g, bytes)) ||
        (rc = hip_check(g, hipMemcpyAsync(g->stage, rho, bytes, hipMemcpyHostToDevice, g->stream), "density upload")) ||
        (rc = launch_aos_to_soa(g, d, static_cast<const double *>(g->stage), d.field[g->cur])))
        return rc;
    return hip_check(g, hipStreamSynchronize(g->stream), "upload sync");
}

// the same transposes from / into a caller's device buffer, in stream order (no wait)
int domain_upload_device(Group *g, int i, const double *d_rho)
{
    if (bad_index(g, i) || !d_rho)
        return set_error(g, BFVMB_ERR_ARG, "upload_density_device: bad argument");
    Domain &d = g->doms[i];
    g->epoch++;
    return launch_aos_to_soa(g, d, d_rho, d.field[g->cur]);
}

int domain_download_device(Group *g, int i, double *d_rho)
{
    if (bad_index(g, i) || !d_rho)
        return set_error(g, BFVMB_ERR_ARG, "download_density_device: bad argument");
    Domain &d = g->doms[i];
    return launch_soa_to_aos(g, d, d.field[g->cur], d_rho);
}

int domain_download(Group *g, int i, double *rho)
{
    if (bad_index(g, i) || !rho)
        return set_error(g, BFVMB_ERR_ARG, "download_density: bad argument");
    Domain &d = g->doms[i];
    const size_t bytes = d.nvox * d.nsub * sizeof(double);
    int rc;
    if ((rc = ensure_stage(g, bytes)) ||
        (rc = launch_soa_to_aos(g, d, d.field[g->cur], static_cast<double *>(g->stage))) ||
        (rc = hip_check(g, hipMemcpyAsync(rho, g->stage, bytes, hipMemcpyDeviceToHost, g->stream), "density download")))
        return rc;
    return hip_check(g, hipStreamSynchronize(g->stream), "download sync");
}

int domain_set_agents(Group *g, int i, int n, const int *voxel, const double *temp1, const double *temp2,
                      const uint8_t *active)
{
    if (bad_index(g, i) || n < 0 || (n > 0 && (!voxel || !temp1 || !temp2)))
        return set_error(g, BFVMB_ERR_ARG, "set_agents: bad argument");
    Domain &d = g->doms[i];
    const int ns = d.nsub;
    std::vector<int> order;
    std::vector<long long> off(n);  // device offset z*nx*ny + y*nx + x of each agent's voxel
    order.reserve(n);
    const long long nyz = (long long)d.ny * d.nz, nxy = (long long)d.nx * d.ny;
    for (int a = 0; a < n; ++a) {
        if (active && !active[a])
            continue;
        if (voxel[a] < 0 || (size_t)voxel[a] >= d.nvox)
            return set_error(g, BFVMB_ERR_ARG, "set_agents: voxel index out of range");
        const long long v = voxel[a];
        const long long x = v / nyz, rem = v - x * nyz, y = rem / d.nz, z = rem - y * d.nz;
        off[a] = z * nxy + y * d.nx + x;
        order.push_back(a);
    }
    // Voxels in device order, so that neighbouring threads of the agent kernel touch
    // neighbouring field elements; list order is kept inside a voxel, as the reference's loop
    // applies them.
    std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return off[p] < off[q]; });
    const int nact = (int)order.size();
    std::vector<int> seg_start;
    std::vector<long long> seg_off;
    std::vector<double> t1((size_t)ns * nact), t2((size_t)ns * nact);
    for (int k = 0; k < nact; ++k) {
        const int a = order[k];
        if (k == 0 || off[a] != off[order[k - 1]]) {
            seg_start.push_back(k);
            seg_off.push_back(off[a]);
        }
        for (int s = 0; s < ns; ++s) {
            t1[(size_t)s * nact + k] = temp1[(size_t)a * ns + s];
            t2[(size_t)s * nact + k] = temp2[(size_t)a * ns + s];
        }
    }
    seg_start.push_back(nact);
    int rc = hip_check(g, hipStreamSynchronize(g->stream), "agents sync");
    if (rc)
        return rc;
    free_remap(d);  // the regrouping state belongs to the previous list
    const size_t nseg = seg_off.size();
    // device arrays are reused while they have room (a list of the same size every call
    // allocates nothing); the descriptor carries the counts, t1/t2 are strided by nact
    if (nseg > d.seg_cap || (size_t)nact > d.act_cap) {
        free_agents(d);
        if (nact > 0 &&
            ((rc = hip_check(g, hipMalloc((void **)&d.seg_start, (nseg + 1) * sizeof(int)), "agent segments")) ||
             (rc = hip_check(g, hipMalloc((void **)&d.seg_off, nseg * sizeof(long long)), "agent voxels")) ||
             (rc = hip_check(g, hipMalloc((void **)&d.t1, t1.size() * sizeof(double)), "agent temp1")) ||
             (rc = hip_check(g, hipMalloc((void **)&d.t2, t2.size() * sizeof(double)), "agent temp2")))) {
            free_agents(d);
            d.agents_given = 0;
            d.agent_order.clear();
            g->desc_dirty = true;
            g->epoch++;
            return rc;
        }
        d.seg_cap = nact > 0 ? nseg : 0;
        d.act_cap = (size_t)nact;
    }
    d.agent_order = order;
    d.agents_given = n;
    d.agent_off.resize(nact);
    for (int k = 0; k < nact; ++k)
        d.agent_off[k] = off[order[k]];
    g->desc_dirty = true;
    g->epoch++;
    d.nseg = (int)nseg;
    d.nact = nact;
    if (nact == 0)
        return BFVMB_OK;
    if ((rc = hip_check(g, hipMemcpyAsync(d.seg_start, seg_start.data(), seg_start.size() * sizeof(int),
                                          hipMemcpyHostToDevice, g->stream), "agent upload")) ||
        (rc = hip_check(g, hipMemcpyAsync(d.seg_off, seg_off.data(), nseg * sizeof(long long), hipMemcpyHostToDevice,
                                          g->stream), "agent upload")) ||
        (rc = hip_check(g, hipMemcpyAsync(d.t1, t1.data(), t1.size() * sizeof(double), hipMemcpyHostToDevice, g->stream),
                        "agent upload")) ||
        (rc = hip_check(g, hipMemcpyAsync(d.t2, t2.data(), t2.size() * sizeof(double), hipMemcpyHostToDevice, g->stream),
                        "agent upload")) ||
        (rc = hip_check(g, hipStreamSynchronize(g->stream), "agent upload"))) {
        (void)hipStreamSynchronize(g->stream);
        free_agents(d);  // no descriptor may keep pointing at a half-written list
        d.agents_given = 0;
        d.agent_order.clear();
        return rc;
    }
    return BFVMB_OK;
}

int domain_set_bulk(Group *g, int i, const double *supply, const double *target, const double *uptake)
{
    if (bad_index(g, i))
        return set_error(g, BFVMB_ERR_ARG, "set_bulk_rates: bad domain index");
    Domain &d = g->doms[i];
    int rc = hip_check(g, hipStreamSynchronize(g->stream), "bulk sync");
    if (rc)
        return rc;
    g->desc_dirty = true;
    g->epoch++;
    if (!supply && !target && !uptake) {  // back to the zero_function defaults: no bulk phase
        (void)hipFree(d.bulk);
        d.bulk = nullptr;
        return BFVMB_OK;
    }
    // T*S and U+S per element, as the reference's loop rounds them (missing rates are 0)
    const size_t n = d.nvox * d.nsub;
    std::vector<double> ts(n), us(n);
    for (size_t k = 0; k < n; ++k) {
        const double S = supply ? supply[k] : 0.0, T = target ? target[k] : 0.0, U = uptake ? uptake[k] : 0.0;
        ts[k] = T * S;
        us[k] = U + S;
    }
    if (!d.bulk && (rc = hip_check(g, hipMalloc((void **)&d.bulk, 2 * n * sizeof(double)), "bulk rates")))
        return rc;
    if ((rc = ensure_stage(g, n * sizeof(double))))
        return rc;
    const std::vector<double> *src[2] = {&ts, &us};
    for (int k = 0; k < 2; ++k)
        if ((rc = hip_check(g, hipMemcpyAsync(g->stage, src[k]->data(), n * sizeof(double), hipMemcpyHostToDevice,
                                              g->stream), "bulk upload")) ||
            (rc = launch_aos_to_soa_f64(g, d, static_cast<const double *>(g->stage), d.bulk + k * n)) ||
            (rc = hip_check(g, hipStreamSynchronize(g->stream), "bulk upload"))) {
            (void)hipStreamSynchronize(g->stream);
            (void)hipFree(d.bulk);  // no bulk phase with half-written rates
            d.bulk = nullptr;
            return rc;
        }
    return BFVMB_OK;
}

int group_bulk(Group *g, double dt)
{
    if (!(dt > 0.0))
        return set_error(g, BFVMB_ERR_ARG, "simulate_bulk_sources_and_sinks: dt must be positive");
    g->epoch++;
    return launch_bulk(g, dt);
}

// The step's cell sources and sinks and its x sweep: one launch (fuse_agents) or two.
static int launch_agents_and_x(Group *g)
{
    if (g->fuse_agents)
        return launch_sweep_x_agents(g);
    int rc = launch_agents(g);
    return rc ? rc : launch_sweep(g, 0);
}

static int eager_steps(Group *g, int nsteps, bool with_agents)
{
    const bool bulk = g->prefix[4].back() > 0;
    for (int s = 0; s < nsteps; ++s) {
        int rc;
        // time loop of a step: bulk sources/sinks, cell sources/sinks, LOD
        // (BioFVM_microenvironment.cpp:1110-1169, BioFVM_parallel.cpp:693-1395)
        if (with_agents && bulk && (rc = launch_bulk(g, 0.0)))
            return rc;
        // FP64: the cell sources and sinks run inside the x sweep, tile by tile, before it
        // reads each tile (results identical to a separate pass; one launch fewer), unless a
        // tile is dense (kFuseMaxSegsPerTile).  FP32 keeps the separate pass: its x sweep runs
        // one wavefront per SIMD, which leaves no other wavefront to cover the update's latency.
        if ((rc = with_agents ? launch_agents_and_x(g) : launch_sweep(g, 0)))
            return rc;
        for (int axis = 1; axis < 3; ++axis)
            if ((rc = launch_sweep(g, axis)))
                return rc;
    }
    return BFVMB_OK;
}

int group_steps(Group *g, int nsteps, bool with_agents, bool use_graph)
{
    if (nsteps < 0)
        return set_error(g, BFVMB_ERR_ARG, "nsteps < 0");
    for (const Domain &d : g->doms)
        if (!d.have_coef)
            return set_error(g, BFVMB_ERR_STATE, "set_coefficients has not been called for every domain");
    int rc = sync_descriptors(g);
    if (rc || nsteps == 0)
        return rc;
    g->epoch++;
    if (!use_graph || g->stream == nullptr)
        return eager_steps(g, nsteps, with_agents);
    // the whole run as one graph: nsteps x (agents, x, y, z) kernel nodes, replayed as one launch
    const int start = g->cur;
    const auto key = std::make_tuple(nsteps, start, with_agents ? 1 : 0);
    auto it = g->graphs.find(key);
    if (it == g->graphs.end()) {
        hipGraph_t graph = nullptr;
        if ((rc = hip_check(g, hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal), "graph capture")))
            return rc;
        const int rc_steps = eager_steps(g, nsteps, with_agents);
        const hipError_t e = hipStreamEndCapture(g->stream, &graph);
        g->cur = start;
        if (rc_steps)
            return rc_steps;
        if ((rc = hip_check(g, e, "graph capture end")))
            return rc;
        hipGraphExec_t exec = nullptr;
        rc = hip_check(g, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
        (void)hipGraphDestroy(graph);
        if (rc)
            return rc;
        it = g->graphs.emplace(key, exec).first;
    }
    if ((rc = hip_check(g, hipGraphLaunch(it->second, g->stream), "graph launch")))
        return rc;
    g->cur = start ^ (nsteps & 1);  // three buffer swaps per step
    return BFVMB_OK;
}

int group_launch_phase(Group *g, int phase, int count)
{
    if (phase < 0 || phase > 4 || count < 0)
        return set_error(g, BFVMB_ERR_ARG, "launch_phase: bad argument");
    if (phase > 0)
        for (const Domain &d : g->doms)
            if (!d.have_coef)
                return set_error(g, BFVMB_ERR_STATE, "set_coefficients has not been called for every domain");
    g->epoch++;
    for (int i = 0; i < count; ++i) {
        int rc = phase == 0 ? launch_agents(g) : phase == 4 ? launch_agents_and_x(g) : launch_sweep(g, phase - 1);
        if (rc)
            return rc;
    }
    return BFVMB_OK;
}

int group_set_launch(Group *g, int sweep, long long grid, int bpc, int ring)
{
    if (sweep < -1 || sweep > 2 || grid < 0 || bpc < 0 || bpc > 8 || ring < 0 || ring > 64)
        return set_error(g, BFVMB_ERR_ARG, "set_launch: bad argument");
    for (int a = 0; a < 3; ++a)
        if (sweep < 0 || sweep == a) {
            g->launch[a].grid = grid;
            g->launch[a].lds_blocks_per_cu = bpc;
            g->launch[a].ring = ring;
        }
    clear_graphs(g);
    return BFVMB_OK;
}

// Autotuner: for each sweep, time the candidate launch shapes (LDS wavefronts per CU; ring
// depth for y/z; the variant for the FP32 x sweep; one block per tile or a strided grid)
// over `reps` launches with HIP events and keep the fastest.  The fields are
// saved before and restored after, so tuning does not advance the simulation.
int group_autotune(Group *g, int reps, int *chosen)
{
    if (reps < 1)
        reps = 3;
    for (const Domain &d : g->doms)
        if (!d.have_coef)
            return set_error(g, BFVMB_ERR_STATE, "set_coefficients has not been called for every domain");
    int rc = sync_descriptors(g);
    if (rc)
        return rc;
    const int start = g->cur;
    std::vector<void *> saved(g->doms.size(), nullptr);
    for (size_t i = 0; i < g->doms.size(); ++i) {
        const Domain &d = g->doms[i];
        const size_t bytes = d.nvox * d.nsub * g->elem;
        if ((rc = hip_check(g, hipMalloc(&saved[i], bytes), "autotune save")) ||
            (rc = hip_check(g, hipMemcpyAsync(saved[i], d.field[start], bytes, hipMemcpyDeviceToDevice, g->stream),
                            "autotune save")))
            break;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc)
        rc = hip_check(g, hipEventCreate(&e0), "event");
    if (!rc)
        rc = hip_check(g, hipEventCreate(&e1), "event");
    for (int axis = 0; axis < 3 && !rc; ++axis) {
        float best = 1e30f;
        LaunchShape best_shape;
        // y/z: rows in flight; x: the FP32 variant by its chunks in flight (FP64 has one)
        const int yz_rings[2] = {32, 64}, x_rings[4] = {1, 2, 3, 4};
        const int *rings = axis == 0 ? x_rings : yz_rings;
        const int nring = axis == 0 ? (g->elem == 4 ? 4 : 1) : 2;
        const int bpcs[5] = {2, 3, 4, 6, 8};
        // each sweep tries one block per tile (a grid request above the tile count) and a grid
        // of one block per wavefront slot (up to four per CU) striding over the tiles
        const int ngrid = 2;
        for (int b = 0; b < 5 && !rc; ++b)
            for (int r = 0; r < nring && !rc; ++r)
                for (int gi = 0; gi < ngrid && !rc; ++gi) {
                    g->launch[axis] = LaunchShape();
                    g->launch[axis].grid =
                        gi == 0 ? (1LL << 40) : (long long)std::max(g->cu_count, 1) * std::min(bpcs[b], 4);
                    g->launch[axis].lds_blocks_per_cu = bpcs[b];
                    g->launch[axis].ring = axis == 0 && g->elem == 8 ? 0 : rings[r];
                    if ((rc = launch_sweep(g, axis)))  // warm-up
                        break;
                    (void)hipEventRecord(e0, g->stream);
                    for (int k = 0; k < reps && !rc; ++k)
                        rc = launch_sweep(g, axis);
                    (void)hipEventRecord(e1, g->stream);
                    if (!rc)
                        rc = hip_check(g, hipEventSynchronize(e1), "autotune timing");
                    float ms = 0.f;
                    if (!rc && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms < best) {
                        best = ms;
                        best_shape = g->launch[axis];
                    }
                }
        g->launch[axis] = best_shape;
        if (chosen) {
            chosen[3 * axis] = best_shape.lds_blocks_per_cu;
            chosen[3 * axis + 1] = best_shape.ring;
            chosen[3 * axis + 2] = (int)(best * 1000.f / reps);  // microseconds per launch
        }
    }
    if (e0)
        (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    // restore whatever was saved, also after a failure part-way through the timing (the sweeps
    // that ran have advanced the field); a field that could not be restored is a new state
    bool restored = true;
    for (size_t i = 0; i < g->doms.size(); ++i) {
        if (!saved[i])
            continue;
        const Domain &d = g->doms[i];
        const int rr = hip_check(g, hipMemcpyAsync(d.field[start], saved[i], d.nvox * d.nsub * g->elem,
                                                   hipMemcpyDeviceToDevice, g->stream), "autotune restore");
        if (rr) {
            restored = false;
            if (!rc)
                rc = rr;
        }
    }
    int rc2 = hip_check(g, hipStreamSynchronize(g->stream), "autotune sync");
    for (void *p : saved)
        (void)hipFree(p);
    g->cur = start;
    if (!restored || rc2)
        g->epoch++;
    clear_graphs(g);
    return rc ? rc : rc2;
}

}  // namespace bfvmb
