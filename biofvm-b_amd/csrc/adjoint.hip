// adjoint.hip -- reverse-mode derivative of nsteps of (sources/sinks, LOD step).
//
// Given a seed w = dL/drho_T, the reverse pass returns dL/drho_0 and dL/dtemp1, dL/dtemp2
// of every agent for the run recorded by adjoint_forward.
//
// Forward step n:  rho <- B(rho) (bulk sources/sinks when rates are set,
//                  BioFVM_microenvironment.cpp:1110-1153), rho <- AG(rho) (agents,
//                  BioFVM_basic_agent.cpp:279-300), then
//                  rho <- D Z D Y D X D rho   (BioFVM_parallel.cpp:929-1392)
// with D the Dirichlet overwrite (Jacobian M: 0 on active faces, 1 elsewhere) and X, Y, Z
// the 1-D solves A^-1 of symmetric tridiagonal operators, so X^T = X etc.  The adjoint
// step is therefore
//     lam <- M X M Y M Z M lam   (the same sweep kernels, reverse order, Dirichlet value 0)
//     lam <- AG^T lam            (per voxel, agents in reverse list order)
//     lam <- B^T lam = lam / (1 + dt (U + S))   (B is affine in rho, with that rounded factor)
// AG at one voxel composes r_a = (r_{a-1} + t1_a)/t2_a; its adjoint accumulates
//     dL/dt1_a += lam_a / t2_a,  dL/dt2_a -= (lam_a / t2_a) r_a,  lam_{a-1} = lam_a / t2_a.
// The states before each agent update are recomputed from checkpoints of the field taken
// every `every` steps, one segment at a time, last segment first; each segment (recompute
// and reverse steps) is replayed as one captured HIP graph, kept with the tape.
#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>

#include "bfvmb_internal.hpp"

namespace bfvmb {
namespace {

// the first M of an adjoint step, on the two z faces (the z sweep's face lines and its
// stores apply the rest)
template <class T>
__global__ __launch_bounds__(256) void k_zero_zfaces(T *f, const uint8_t *dact, int nx, int ny, int nz, int nsub)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nxy = (long long)nx * ny;
    if (i >= nxy * nsub)
        return;
    const int s = (int)(i / nxy);
    if (!dact[s])
        return;
    const long long xy = i - (long long)s * nxy;
    T *p = f + (size_t)s * nxy * nz;
    p[xy] = T(0);
    p[(size_t)(nz - 1) * nxy + xy] = T(0);
}

// field values at the occupied voxels (the inputs of the agent update), [nsub][nseg]
template <class T>
__global__ __launch_bounds__(256) void k_gather(const T *f, const long long *seg_off, int nseg, int nsub, size_t nvox,
                                                double *out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nseg * nsub)
        return;
    const int s = (int)(i / nseg);
    const int k = (int)(i - (long long)s * nseg);
    out[i] = (double)f[(size_t)s * nvox + seg_off[k]];
}

// adjoint of the agent update at every occupied voxel; rbuf is per-agent scratch
template <class T>
__global__ __launch_bounds__(256) void k_agents_adjoint(T *lam, const double *r0, const int *seg_start,
                                                        const long long *seg_off, const double *t1, const double *t2,
                                                        int nseg, int nact, int nsub, size_t nvox, double *rbuf,
                                                        double *g1, double *g2)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nseg * nsub)
        return;
    const int s = (int)(i / nseg);
    const int k = (int)(i - (long long)s * nseg);
    const double *a1 = t1 + (size_t)s * nact, *a2 = t2 + (size_t)s * nact;
    double *rb = rbuf + (size_t)s * nact;
    double r = r0[i];
    const int a0 = seg_start[k], a1e = seg_start[k + 1];
    for (int a = a0; a < a1e; ++a) {  // forward states r_a, as the agent kernel computes them
        r = r + a1[a];
        r = r / a2[a];
        rb[a] = r;
    }
    T *p = lam + (size_t)s * nvox + seg_off[k];
    double l = (double)*p;
    for (int a = a1e - 1; a >= a0; --a) {
        const double q = l / a2[a];
        g1[(size_t)s * nact + a] += q;
        g2[(size_t)s * nact + a] -= q * rb[a];
        l = q;
    }
    *p = (T)l;
}

// adjoint of the bulk update: lam /= (U + S)*dt + 1, the forward denominator
template <class T>
__global__ __launch_bounds__(256) void k_bulk_adjoint(T *lam, const double *us, size_t nel, double dt)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nel)
        return;
    lam[e] = (T)((double)lam[e] / (us[e] * dt + 1.0));
}

// rate gradients from the sorted [nsub][nact] accumulators to the caller's [agent][nsub]
// order (order[k] = caller index of sorted agent k); inactive agents keep the zeros the
// output was cleared to
__global__ __launch_bounds__(256) void k_scatter_rates(const double *__restrict__ g, const int *__restrict__ order,
                                                       int nact, int nsub, double *__restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nact * nsub)
        return;
    const int s = (int)(i / nact), k = (int)(i - (long long)s * nact);
    out[(size_t)order[k] * nsub + s] = g[i];
}

}  // namespace

// Device state of the reverse pass, created by the first adjoint_backward of a tape and
// kept with it: adjoint fields, agent inputs per step, gradient accumulators, the adjoint
// group (its descriptors) and the captured segment graphs.
// What the reverse pass's buffers and graphs depend on: a new record with the same stamp
// keeps them.
struct RevStamp {
    const void *ptr[7];  // agent CSR and tables, coefficients, bulk rates, Dirichlet flags
    double dt;           // time step of the bulk update
    int nseg, nact, every;
    size_t nel;
    int launch[3][2];
    bool operator==(const RevStamp &o) const { return std::memcmp(this, &o, sizeof(RevStamp)) == 0; }
};

struct Reverse {
    RevStamp stamp;
    void *lam[2] = {nullptr, nullptr};
    void *stage = nullptr;  // FP64 transfers
    void *zeros = nullptr;  // Dirichlet value 0 of the adjoint sweeps
    double *r0 = nullptr;   // agent inputs per step of a segment [every][nsub][nseg]
    double *rbuf = nullptr, *g1 = nullptr, *g2 = nullptr;  // [nsub][nact]
    int *order = nullptr;   // device copy of the agent order, for device-side gradient outputs
    Group adj;
    std::map<std::tuple<int, int, int>, hipGraphExec_t> graphs;  // (segment steps, adjoint parity,
                                                                  // primal parity)
};

struct Tape {
    int nsteps = 0, every = 1, nck = 0;
    int start = 0;                 // buffer parity holding rho_0
    void *ck = nullptr;            // nck fields: rho at steps 0, every, 2 every, ...
    void *fin = nullptr;           // rho_T
    unsigned long long epoch = 0;  // Group::epoch right after the forward run
    Reverse *rev = nullptr;        // reverse-pass state (adjoint_backward)
    int nck_cap = 0;               // checkpoints the ck buffer holds
    size_t fbytes = 0;             // bytes of one field
};

namespace {

void free_reverse(Reverse *r)
{
    if (!r)
        return;
    for (auto &kv : r->graphs)
        (void)hipGraphExecDestroy(kv.second);
    for (void *p : {r->lam[0], r->lam[1], r->stage, r->zeros, (void *)r->r0, (void *)r->rbuf, (void *)r->g1,
                    (void *)r->g2, (void *)r->order, r->adj.d_desc, (void *)r->adj.d_prefix})
        (void)hipFree(p);
    delete r;
}

void free_tape(Group *g)
{
    if (!g->tape)
        return;
    if (g->stream)
        (void)hipStreamSynchronize(g->stream);
    free_reverse(g->tape->rev);
    (void)hipFree(g->tape->ck);
    (void)hipFree(g->tape->fin);
    delete g->tape;
    g->tape = nullptr;
}

RevStamp stamp_of(const Group *g, int every)
{
    const Domain &d = g->doms[0];
    RevStamp st;
    std::memset(&st, 0, sizeof(st));
    const void *p[7] = {d.seg_start, d.seg_off, d.t1, d.t2, d.coef, d.bulk, d.d_dact};
    for (int k = 0; k < 7; ++k)
        st.ptr[k] = p[k];
    st.dt = d.dt;
    st.nseg = d.nseg;
    st.nact = d.nact;
    st.every = every;
    st.nel = d.nvox * d.nsub;
    for (int a = 0; a < 3; ++a) {
        st.launch[a][0] = g->launch[a].lds_blocks_per_cu;
        st.launch[a][1] = g->launch[a].ring;
    }
    return st;
}

int rev_alloc(Group *g, void **p, size_t bytes)
{
    return hip_check(g, hipMalloc(p, bytes ? bytes : 8), "adjoint buffers");
}

// the reverse-pass state of the tape (first backward call): buffers and the adjoint group,
// one domain on the lam buffers with the primal's tables and flags and Dirichlet value 0
int make_reverse(Group *g, Tape &tp)
{
    Domain &d = g->doms[0];
    const size_t nel = d.nvox * d.nsub;
    const int nseg = d.nseg, nact = d.nact, ns = d.nsub;
    Reverse *r = new Reverse;
    tp.rev = r;
    r->stamp = stamp_of(g, tp.every);
    int rc;
    const size_t abytes = (size_t)nact * ns * sizeof(double);
    if ((rc = rev_alloc(g, &r->lam[0], nel * g->elem)) || (rc = rev_alloc(g, &r->lam[1], nel * g->elem)) ||
        (rc = rev_alloc(g, &r->stage, nel * sizeof(double))) || (rc = rev_alloc(g, &r->zeros, ns * g->elem)) ||
        (rc = hip_check(g, hipMemsetAsync(r->zeros, 0, ns * g->elem, g->stream), "adjoint zeros")))
        return rc;
    if (nseg > 0 &&
        ((rc = rev_alloc(g, (void **)&r->r0, (size_t)tp.every * nseg * ns * sizeof(double))) ||
         (rc = rev_alloc(g, (void **)&r->rbuf, abytes)) || (rc = rev_alloc(g, (void **)&r->g1, abytes)) ||
         (rc = rev_alloc(g, (void **)&r->g2, abytes))))
        return rc;
    Group &adj = r->adj;
    adj.precision = g->precision;
    adj.device = g->device;
    adj.elem = g->elem;
    adj.stream = g->stream;
    adj.max_lds_per_block = g->max_lds_per_block;
    adj.cu_count = g->cu_count;
    for (int a = 0; a < 3; ++a)
        adj.launch[a] = g->launch[a];
    adj.doms.resize(1);
    Domain &ad = adj.doms[0];
    ad.nx = d.nx;
    ad.ny = d.ny;
    ad.nz = d.nz;
    ad.nsub = d.nsub;
    ad.dx = d.dx;
    ad.nvox = d.nvox;
    ad.field[0] = r->lam[0];
    ad.field[1] = r->lam[1];
    ad.coef = d.coef;
    ad.have_coef = true;
    ad.d_dval = r->zeros;
    ad.d_dact = d.d_dact;
    if ((rc = sync_descriptors(&adj)))
        return set_error(g, rc, adj.err);
    return BFVMB_OK;
}

// on_device: seed and the three outputs are device pointers; the outputs are written in
// stream order (no wait at the end)
template <class T>
int backward_t(Group *g, const double *seed, double *grad_rho0, double *grad_t1, double *grad_t2, bool on_device)
{
    Tape &tp = *g->tape;
    Domain &d = g->doms[0];
    const size_t nel = d.nvox * d.nsub;
    const size_t fbytes = nel * sizeof(T);
    const int nseg = d.nseg, nact = d.nact, ns = d.nsub, every = tp.every, nsteps = tp.nsteps;
    const bool bulk = d.bulk != nullptr;
    int rc;
    if (!tp.rev && (rc = make_reverse(g, tp)))
        return rc;
    Reverse &rv = *tp.rev;
    Group &adj = rv.adj;
    Domain &ad = adj.doms[0];
    const size_t abytes = (size_t)nact * ns * sizeof(double);
    if (!on_device &&
        (rc = hip_check(g, hipMemcpyAsync(rv.stage, seed, nel * sizeof(double), hipMemcpyHostToDevice, g->stream),
                        "seed upload")))
        return rc;
    adj.cur = 0;
    if ((rc = launch_aos_to_soa(g, d, on_device ? seed : static_cast<const double *>(rv.stage), rv.lam[0])))
        return rc;
    if (nseg > 0 && ((rc = hip_check(g, hipMemsetAsync(rv.g1, 0, abytes, g->stream), "gradient init")) ||
                     (rc = hip_check(g, hipMemsetAsync(rv.g2, 0, abytes, g->stream), "gradient init"))))
        return rc;
    if ((rc = sync_descriptors(g)))
        return rc;

    auto field = [&](int parity) { return static_cast<T *>(d.field[parity]); };
    auto ckp = [&](int j) { return static_cast<T *>(tp.ck) + (size_t)j * nel; };
    const unsigned zgrid = (unsigned)(((long long)d.nx * d.ny * ns + 255) / 256);
    const unsigned agrid = (unsigned)(((long long)nseg * ns + 255) / 256);
    const unsigned bgrid = (unsigned)((nel + 255) / 256);
    const double *us = bulk ? d.bulk + nel : nullptr;
    // one segment of L steps from the checkpoint in field(tp.start): recompute it, recording
    // each step's agent inputs (after the bulk update), then the reverse steps, last first
    auto segment = [&](int L) -> int {
        for (int j = 0; j < L; ++j) {
            int e;
            if (bulk && (e = launch_bulk(g, 0.0)))
                return e;
            if (nseg > 0) {
                hipLaunchKernelGGL(k_gather<T>, dim3(agrid), dim3(256), 0, g->stream, field(g->cur), d.seg_off, nseg,
                                   ns, d.nvox, rv.r0 + (size_t)j * nseg * ns);
                if ((e = launch_agents(g)))
                    return e;
            }
            for (int axis = 0; axis < 3; ++axis)
                if ((e = launch_sweep(g, axis)))
                    return e;
        }
        for (int j = L - 1; j >= 0; --j) {
            hipLaunchKernelGGL(k_zero_zfaces<T>, dim3(zgrid), dim3(256), 0, g->stream,
                               static_cast<T *>(ad.field[adj.cur]), d.d_dact, d.nx, d.ny, d.nz, ns);
            int e;
            if ((e = launch_sweep(&adj, 2)) || (e = launch_sweep(&adj, 1)) || (e = launch_sweep(&adj, 0)))
                return set_error(g, e, adj.err);
            if (nseg > 0)
                hipLaunchKernelGGL(k_agents_adjoint<T>, dim3(agrid), dim3(256), 0, g->stream,
                                   static_cast<T *>(ad.field[adj.cur]), rv.r0 + (size_t)j * nseg * ns, d.seg_start,
                                   d.seg_off, d.t1, d.t2, nseg, nact, ns, d.nvox, rv.rbuf, rv.g1, rv.g2);
            if (bulk)
                hipLaunchKernelGGL(k_bulk_adjoint<T>, dim3(bgrid), dim3(256), 0, g->stream,
                                   static_cast<T *>(ad.field[adj.cur]), us, nel, d.dt);
        }
        return hip_check(g, hipGetLastError(), "adjoint launch");
    };
    for (int seg = (nsteps - 1) / every; nsteps > 0 && seg >= 0; --seg) {
        const int n0 = seg * every, L = std::min(nsteps, n0 + every) - n0;
        if ((rc = hip_check(g, hipMemcpyAsync(field(tp.start), ckp(seg), fbytes, hipMemcpyDeviceToDevice, g->stream),
                            "checkpoint restore")))
            return rc;
        const int a0 = adj.cur;
        g->cur = tp.start;
        if (g->stream == nullptr) {  // the legacy stream cannot be captured: launch eagerly
            if ((rc = segment(L)))
                return rc;
        } else {
            const auto key = std::make_tuple(L, a0, tp.start);
            auto it = rv.graphs.find(key);
            if (it == rv.graphs.end()) {
                hipGraph_t graph = nullptr;
                if ((rc = hip_check(g, hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal),
                                    "adjoint capture")))
                    return rc;
                const int rc_seg = segment(L);
                const hipError_t e = hipStreamEndCapture(g->stream, &graph);
                if (rc_seg) {
                    if (graph)
                        (void)hipGraphDestroy(graph);
                    return rc_seg;
                }
                if ((rc = hip_check(g, e, "adjoint capture end")))
                    return rc;
                hipGraphExec_t exec = nullptr;
                rc = hip_check(g, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "adjoint graph");
                (void)hipGraphDestroy(graph);
                if (rc)
                    return rc;
                it = rv.graphs.emplace(key, exec).first;
            }
            if ((rc = hip_check(g, hipGraphLaunch(it->second, g->stream), "adjoint graph launch")))
                return rc;
        }
        // three buffer swaps per step on both sides
        g->cur = tp.start ^ (L & 1);
        adj.cur = a0 ^ (L & 1);
    }
    // outputs
    if (on_device) {
        const size_t obytes = (size_t)d.agents_given * ns * sizeof(double);
        if (grad_rho0 && (rc = launch_soa_to_aos(g, d, ad.field[adj.cur], grad_rho0)))
            return rc;
        if (obytes && (grad_t1 || grad_t2)) {
            if (nseg > 0 && !rv.order) {  // the order is fixed for the record's agent tables
                if ((rc = domain_agent_order(g, d)) ||
                    (rc = rev_alloc(g, (void **)&rv.order, (size_t)nact * sizeof(int))) ||
                    (rc = hip_check(g, hipMemcpyAsync(rv.order, d.agent_order.data(), (size_t)nact * sizeof(int),
                                                      hipMemcpyHostToDevice, g->stream), "agent order")) ||
                    (rc = hip_check(g, hipStreamSynchronize(g->stream), "agent order")))
                    return rc;
            }
            const unsigned sgrid = (unsigned)(((long long)nact * ns + 255) / 256);
            double *outs[2] = {grad_t1, grad_t2};
            const double *accs[2] = {rv.g1, rv.g2};
            for (int k = 0; k < 2; ++k) {
                if (!outs[k])
                    continue;
                if ((rc = hip_check(g, hipMemsetAsync(outs[k], 0, obytes, g->stream), "gradient clear")))
                    return rc;
                if (nseg > 0 && nact > 0)
                    hipLaunchKernelGGL(k_scatter_rates, dim3(sgrid), dim3(256), 0, g->stream, accs[k], rv.order, nact,
                                       ns, outs[k]);
            }
            if ((rc = hip_check(g, hipGetLastError(), "gradient scatter")))
                return rc;
        }
        const int end = tp.start ^ (nsteps & 1);
        if ((rc = hip_check(g, hipMemcpyAsync(field(end), tp.fin, fbytes, hipMemcpyDeviceToDevice, g->stream),
                            "restore")))
            return rc;
        g->cur = end;
        g->epoch = tp.epoch;
        return BFVMB_OK;
    }
    if (grad_rho0 &&
        ((rc = launch_soa_to_aos(g, d, ad.field[adj.cur], static_cast<double *>(rv.stage))) ||
         (rc = hip_check(g, hipMemcpyAsync(grad_rho0, rv.stage, nel * sizeof(double), hipMemcpyDeviceToHost,
                                           g->stream), "gradient download"))))
        return rc;
    std::vector<double> h1, h2;
    if (nseg > 0 && (grad_t1 || grad_t2)) {
        h1.resize((size_t)nact * ns);
        h2.resize((size_t)nact * ns);
        if ((rc = hip_check(g, hipMemcpyAsync(h1.data(), rv.g1, abytes, hipMemcpyDeviceToHost, g->stream),
                            "gradient download")) ||
            (rc = hip_check(g, hipMemcpyAsync(h2.data(), rv.g2, abytes, hipMemcpyDeviceToHost, g->stream),
                            "gradient download")))
            return rc;
    }
    // the context is left as the forward run left it: rho_T current, tape still valid
    const int end = tp.start ^ (nsteps & 1);
    if ((rc = hip_check(g, hipMemcpyAsync(field(end), tp.fin, fbytes, hipMemcpyDeviceToDevice, g->stream), "restore")) ||
        (rc = hip_check(g, hipStreamSynchronize(g->stream), "adjoint sync")))
        return rc;
    g->cur = end;
    g->epoch = tp.epoch;
    if (!h1.empty() && (rc = domain_agent_order(g, d)))
        return rc;
    const size_t nag = (size_t)d.agents_given;
    if (grad_t1)
        std::fill(grad_t1, grad_t1 + nag * ns, 0.0);
    if (grad_t2)
        std::fill(grad_t2, grad_t2 + nag * ns, 0.0);
    if (!h1.empty())
        for (int k = 0; k < nact; ++k) {
            const size_t a = (size_t)d.agent_order[k];
            for (int s = 0; s < ns; ++s) {
                if (grad_t1)
                    grad_t1[a * ns + s] = h1[(size_t)s * nact + k];
                if (grad_t2)
                    grad_t2[a * ns + s] = h2[(size_t)s * nact + k];
            }
        }
    return BFVMB_OK;
}

}  // namespace

void adjoint_free(Group *g)
{
    free_tape(g);
}

int adjoint_forward(Group *g, int nsteps, int every)
{
    if (g->doms.size() != 1 || nsteps < 0 || every < 1)
        return set_error(g, BFVMB_ERR_ARG, "adjoint_forward: bad argument");
    Domain &d = g->doms[0];
    if (!d.have_coef)
        return set_error(g, BFVMB_ERR_STATE, "set_coefficients has not been called");
    const size_t fbytes = d.nvox * d.nsub * g->elem;
    const int nck = std::max(1, (nsteps + every - 1) / every);
    // a previous record's buffers are reused when large enough, and its reverse-pass state
    // (buffers and captured graphs) when nothing it depends on has changed
    Tape *tp = g->tape;
    if (tp && (tp->nck_cap < nck || tp->fbytes != fbytes)) {
        free_tape(g);
        tp = nullptr;
    }
    if (tp && tp->rev && !(tp->rev->stamp == stamp_of(g, every))) {
        (void)hipStreamSynchronize(g->stream);
        free_reverse(tp->rev);
        tp->rev = nullptr;
    }
    int rc;
    if (!tp) {
        tp = new Tape;
        g->tape = tp;
        if ((rc = hip_check(g, hipMalloc(&tp->ck, nck * fbytes), "adjoint checkpoints")) ||
            (rc = hip_check(g, hipMalloc(&tp->fin, fbytes), "adjoint final field"))) {
            free_tape(g);
            return rc;
        }
        tp->nck_cap = nck;
        tp->fbytes = fbytes;
    }
    tp->nsteps = nsteps;
    tp->every = every;
    tp->nck = nck;
    tp->start = g->cur;
    tp->epoch = 0;
    char *ck = static_cast<char *>(tp->ck);
    for (int n = 0; n < nsteps || n == 0; n += every) {
        if ((rc = hip_check(g, hipMemcpyAsync(ck + (size_t)(n / every) * fbytes, d.field[g->cur], fbytes,
                                              hipMemcpyDeviceToDevice, g->stream), "checkpoint")))
            break;
        // a segment is replayed as one captured graph when the context has a stream
        if (nsteps > 0 && (rc = group_steps(g, std::min(every, nsteps - n), true, true)))
            break;
        if (nsteps == 0)
            break;
    }
    if (!rc)
        rc = hip_check(g, hipMemcpyAsync(tp->fin, d.field[g->cur], fbytes, hipMemcpyDeviceToDevice, g->stream),
                       "final field");
    if (!rc)
        rc = hip_check(g, hipStreamSynchronize(g->stream), "adjoint forward sync");
    if (rc) {
        free_tape(g);
        return rc;
    }
    tp->epoch = g->epoch;
    return BFVMB_OK;
}

int adjoint_backward(Group *g, const double *seed, double *grad_rho0, double *grad_t1, double *grad_t2,
                     bool on_device)
{
    if (!seed)
        return set_error(g, BFVMB_ERR_ARG, "adjoint_backward: seed is NULL");
    if (!g->tape || g->tape->epoch != g->epoch)
        return set_error(g, BFVMB_ERR_STATE,
                         "adjoint_backward: no adjoint_forward record for the current field, coefficients and agents");
    if (g->precision == BFVMB_FP64)
        return backward_t<double>(g, seed, grad_rho0, grad_t1, grad_t2, on_device);
    return backward_t<float>(g, seed, grad_rho0, grad_t1, grad_t2, on_device);
}

}  // namespace bfvmb

extern "C" int bfvmb_adjoint_forward(bfvmb_ctx *ctx, int nsteps, int checkpoint_every)
{
    if (!ctx)
        return BFVMB_ERR_ARG;
    return bfvmb::adjoint_forward(ctx, nsteps, checkpoint_every);
}

extern "C" int bfvmb_adjoint_backward(bfvmb_ctx *ctx, const double *seed, double *grad_rho0, double *grad_temp1,
                                      double *grad_temp2)
{
    if (!ctx)
        return BFVMB_ERR_ARG;
    return bfvmb::adjoint_backward(ctx, seed, grad_rho0, grad_temp1, grad_temp2);
}

extern "C" int bfvmb_adjoint_backward_device(bfvmb_ctx *ctx, const double *seed, double *grad_rho0,
                                             double *grad_temp1, double *grad_temp2)
{
    if (!ctx)
        return BFVMB_ERR_ARG;
    return bfvmb::adjoint_backward(ctx, seed, grad_rho0, grad_temp1, grad_temp2, true);
}
