/*
 * bfvmb_oracle.c -- CPU restatement of BioFVM-B's FP64 constant-coefficient LOD
 * diffusion-decay step and its agent source/sink update.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle for the HIP library in
 * biofvm-b_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * load it; the product path never links, calls or falls back to it.
 *
 * Layout: the reference's own voxel-major / substrate-minor store,
 *     rho[((x*ny + y)*nz + z)*nsub + d]                 (BioFVM_mesh.cpp:659-665,
 *                                                         BioFVM_parallel.cpp:724-726)
 * Every floating-point operation is written out with the rounding the reference gets
 * from g++ -O3 with its default -ffp-contract=fast (Makefile:41): `a += b*c` and
 * axpy() become one fused multiply-add, `a -= b*c` becomes fma(-b, c, a), and every
 * other operation rounds on its own.  The file is compiled with -ffp-contract=off so
 * that the fma() calls below are the only fused operations.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   - bitwise against fixtures written by the reference itself (oracle/_ref, built
 *     from the reference sources by oracle/Makefile; vectors under tests/golden/);
 *   - against a dense numpy solve of each 1-D operator (brute force on tiny grids);
 *   - against closed forms: D=0 (three divisions by 1+c2), Neumann cosine
 *     eigenmodes, a Dirichlet-valued uniform field staying fixed, mass conservation
 *     with no Dirichlet and no decay; and the single/multiple-agent compositions.
 */
#include <math.h>
#include <stddef.h>

#define ORC_API __attribute__((visibility("default")))

static inline size_t orc_idx(int x, int y, int z, int d, int ny, int nz, int nsub)
{
    return (((size_t)x * ny + y) * nz + z) * nsub + d;
}

/* LOD constants, BioFVM_parallel.cpp:780-801:
 *   c1  = ((D*dt)/dx)/dx        c2  = (decay*dt)/3
 *   c3  = ((1+c1)+c1)+c2        c3a = (1+c1)+c2                                     */
ORC_API void orc_lod_constants(int nsub, const double *D, const double *decay, double dt,
                               double dx, double *c1, double *c2, double *c3, double *c3a)
{
    for (int d = 0; d < nsub; ++d) {
        double a = D[d];
        a *= dt;
        a /= dx;
        a /= dx;
        double b = decay[d];
        b *= dt;
        b /= 3.0;
        c1[d] = a;
        c2[d] = b;
        c3[d] = ((1.0 + a) + a) + b;
        c3a[d] = (1.0 + a) + b;
    }
}

/* Thomas coefficients of one axis, BioFVM_parallel.cpp:882-896 (y), :898-912 (z) and,
 * at MPI world size 1, :810-851 (x).  denom and c are [n][nsub].
 *   c[i] = -c1;  denom[i] = c3, ends c3a, n==1 -> 1+c2
 *   c[0] /= denom[0];  for i>=1: denom[i] = fma(c1, c[i-1], denom[i]); c[i] /= denom[i]  */
ORC_API void orc_thomas_axis(int n, int nsub, const double *c1, const double *c2,
                             const double *c3, const double *c3a, double *denom, double *c)
{
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < nsub; ++d) {
            c[(size_t)i * nsub + d] = c1[d] * -1.0;
            denom[(size_t)i * nsub + d] = c3[d];
        }
    for (int d = 0; d < nsub; ++d) {
        denom[d] = c3a[d];
        denom[(size_t)(n - 1) * nsub + d] = c3a[d];
        if (n == 1)
            denom[d] = 1.0 + c2[d];
        c[d] /= denom[d];
    }
    for (int i = 1; i < n; ++i)
        for (int d = 0; d < nsub; ++d) {
            size_t k = (size_t)i * nsub + d;
            denom[k] = fma(c1[d], c[k - nsub], denom[k]);
            c[k] /= denom[k];
        }
}

/* apply_dirichlet_conditions, BioFVM_microenvironment.cpp:381-484: overwrite every voxel
 * on all six faces (x faces, then y, then z), for each substrate whose activation flag is
 * set, with dirichlet_value_vectors[voxel*nsub + d].  `dvox` is that per-voxel array
 * ([nvox][nsub], as add_dirichlet_node / update_dirichlet_node write it, :226-290); NULL
 * means every voxel of substrate d holds value[d] (set_substrate_dirichlet_value, :334-359,
 * and the 100.0 / 1.0 defaults of resize_space / resize_densities).                      */
static inline double orc_dval(const double *value, const double *dvox, size_t i, int d)
{
    return dvox ? dvox[i] : value[d];
}

ORC_API void orc_apply_dirichlet_voxels(int nx, int ny, int nz, int nsub, double *rho,
                                        const double *value, const double *dvox,
                                        const unsigned char *active)
{
    size_t i;
    for (int y = 0; y < ny; ++y)
        for (int z = 0; z < nz; ++z)
            for (int d = 0; d < nsub; ++d)
                if (active[d]) {
                    i = orc_idx(0, y, z, d, ny, nz, nsub);
                    rho[i] = orc_dval(value, dvox, i, d);
                    i = orc_idx(nx - 1, y, z, d, ny, nz, nsub);
                    rho[i] = orc_dval(value, dvox, i, d);
                }
    for (int x = 0; x < nx; ++x)
        for (int z = 0; z < nz; ++z)
            for (int d = 0; d < nsub; ++d)
                if (active[d]) {
                    i = orc_idx(x, 0, z, d, ny, nz, nsub);
                    rho[i] = orc_dval(value, dvox, i, d);
                    i = orc_idx(x, ny - 1, z, d, ny, nz, nsub);
                    rho[i] = orc_dval(value, dvox, i, d);
                }
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int d = 0; d < nsub; ++d)
                if (active[d]) {
                    i = orc_idx(x, y, 0, d, ny, nz, nsub);
                    rho[i] = orc_dval(value, dvox, i, d);
                    i = orc_idx(x, y, nz - 1, d, ny, nz, nsub);
                    rho[i] = orc_dval(value, dvox, i, d);
                }
}

ORC_API void orc_apply_dirichlet(int nx, int ny, int nz, int nsub, double *rho,
                                 const double *value, const unsigned char *active)
{
    orc_apply_dirichlet_voxels(nx, ny, nz, nsub, rho, value, NULL, active);
}

/* One Thomas solve along a strided line: forward elimination
 *   u[0] /= denom[0];  u[i] = fma(c1, u[i-1], u[i]); u[i] /= denom[i]
 * and back substitution  u[i] = fma(-c[i], u[i+1], u[i]),  i = n-2 .. 0
 * (BioFVM_parallel.cpp:954-979 and :1131-1149 for x, :1274-1319 for y, :1334-1378 for z). */
static void orc_line(double *u, size_t stride, int n, int nsub, int d, double c1,
                     const double *denom, const double *c)
{
    u[0] /= denom[d];
    for (int i = 1; i < n; ++i) {
        double v = fma(c1, u[(size_t)(i - 1) * stride], u[(size_t)i * stride]);
        u[(size_t)i * stride] = v / denom[(size_t)i * nsub + d];
    }
    for (int i = n - 2; i >= 0; --i)
        u[(size_t)i * stride] = fma(-c[(size_t)i * nsub + d], u[(size_t)(i + 1) * stride],
                                    u[(size_t)i * stride]);
}

/* One sweep along axis 0 (x), 1 (y) or 2 (z), every line and every substrate. */
ORC_API void orc_sweep(int axis, int nx, int ny, int nz, int nsub, double *rho,
                       const double *c1, const double *denom, const double *c)
{
    size_t sx = (size_t)ny * nz * nsub, sy = (size_t)nz * nsub, sz = (size_t)nsub;
    if (axis == 0) {
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z)
                for (int d = 0; d < nsub; ++d)
                    orc_line(rho + orc_idx(0, y, z, d, ny, nz, nsub), sx, nx, nsub, d, c1[d],
                             denom, c);
    } else if (axis == 1) {
        for (int x = 0; x < nx; ++x)
            for (int z = 0; z < nz; ++z)
                for (int d = 0; d < nsub; ++d)
                    orc_line(rho + orc_idx(x, 0, z, d, ny, nz, nsub), sy, ny, nsub, d, c1[d],
                             denom, c);
    } else {
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y)
                for (int d = 0; d < nsub; ++d)
                    orc_line(rho + orc_idx(x, y, 0, d, ny, nz, nsub), sz, nz, nsub, d, c1[d],
                             denom, c);
    }
}

/* diffusion_decay_solver__constant_coefficients_LOD_3D, one call after setup
 * (BioFVM_parallel.cpp:929-1392): Dirichlet, x, Dirichlet, y, Dirichlet, z, Dirichlet.  */
static void orc_lod_step_dv(int nx, int ny, int nz, int nsub, double *rho, const double *c1,
                            const double *denomx, const double *cx, const double *denomy,
                            const double *cy, const double *denomz, const double *cz,
                            const double *dvalue, const double *dvox, const unsigned char *dactive)
{
    orc_apply_dirichlet_voxels(nx, ny, nz, nsub, rho, dvalue, dvox, dactive);
    orc_sweep(0, nx, ny, nz, nsub, rho, c1, denomx, cx);
    orc_apply_dirichlet_voxels(nx, ny, nz, nsub, rho, dvalue, dvox, dactive);
    orc_sweep(1, nx, ny, nz, nsub, rho, c1, denomy, cy);
    orc_apply_dirichlet_voxels(nx, ny, nz, nsub, rho, dvalue, dvox, dactive);
    orc_sweep(2, nx, ny, nz, nsub, rho, c1, denomz, cz);
    orc_apply_dirichlet_voxels(nx, ny, nz, nsub, rho, dvalue, dvox, dactive);
}

ORC_API void orc_lod_step(int nx, int ny, int nz, int nsub, double *rho, const double *c1,
                          const double *denomx, const double *cx, const double *denomy,
                          const double *cy, const double *denomz, const double *cz,
                          const double *dvalue, const unsigned char *dactive)
{
    orc_lod_step_dv(nx, ny, nz, nsub, rho, c1, denomx, cx, denomy, cy, denomz, cz, dvalue, NULL, dactive);
}

/* Basic_Agent::set_internal_uptake_constants, BioFVM_basic_agent.cpp:155-176:
 *   k     = (dt*V_cell)/V_voxel
 *   temp1 = ((0+S)*T)*k
 *   temp2 = fma(k, U, fma(k, S, 1))      (two axpy calls, S first)                      */
ORC_API void orc_agent_constants(int nsub, double dt, double volume, double voxel_volume,
                                 const double *S, const double *U, const double *T,
                                 double *temp1, double *temp2)
{
    double k = dt * volume / voxel_volume;
    for (int d = 0; d < nsub; ++d) {
        double t1 = 0.0;
        t1 += S[d];
        t1 *= T[d];
        t1 *= k;
        temp1[d] = t1;
        double t2 = 1.0;
        t2 = fma(k, S[d], t2);
        t2 = fma(k, U[d], t2);
        temp2[d] = t2;
    }
}

/* Microenvironment::simulate_cell_sources_and_sinks, BioFVM_microenvironment.cpp:1155-1169,
 * with Basic_Agent::simulate_secretion_and_uptake, BioFVM_basic_agent.cpp:279-300, run in
 * list order (one OpenMP thread; the reference's omp loop races when agents share a voxel):
 *   rho[v][d] += temp1[d]  then  rho[v][d] /= temp2[d]                                   */
ORC_API void orc_sources_sinks(int nagents, int nsub, const int *voxel,
                               const unsigned char *active, const double *temp1,
                               const double *temp2, double *rho)
{
    for (int a = 0; a < nagents; ++a) {
        if (!active[a])
            continue;
        double *r = rho + (size_t)voxel[a] * nsub;
        for (int d = 0; d < nsub; ++d)
            r[d] += temp1[(size_t)a * nsub + d];
        for (int d = 0; d < nsub; ++d)
            r[d] /= temp2[(size_t)a * nsub + d];
    }
}

/* Bulk sources and sinks, Microenvironment::simulate_bulk_sources_and_sinks
 * (BioFVM_microenvironment.cpp:1110-1153), with supply S, target T and uptake U per voxel
 * and substrate ([nvox][nsub], what the bulk_*_function slots write):
 *   temp2 = T*S;  rho += dt*temp2  (one fused multiply-add);
 *   temp3 = U + S;  temp3 *= dt;  temp3 += 1;  rho /= temp3  (each rounded on its own,
 *   the statements sit in separate loops).                                               */
ORC_API void orc_bulk_sources_sinks(size_t nvox, int nsub, double dt, const double *S, const double *T,
                                    const double *U, double *rho)
{
    for (size_t i = 0; i < nvox * (size_t)nsub; ++i) {
        const double ts = T[i] * S[i];
        const double num = fma(dt, ts, rho[i]);
        double den = U[i] + S[i];
        den = den * dt;
        den = den + 1.0;
        rho[i] = num / den;
    }
}

/* Density gradients, Microenvironment::compute_all_gradient_vectors
 * (BioFVM_microenvironment.cpp:1243-1386), grad[voxel][d][axis]:
 *   interior   (rho[+1] - rho[-1]) / (2*dx)      (two_dx = dx*2.0)
 *   first      (rho[1] - rho[0]) / dx
 *   last       (rho[n-1] - rho[n-2]) / dx
 * each subtraction and division rounded on its own.  An axis of length 1 gives 0 (the
 * reference skips z then, and reads outside the field for x or y).                     */
ORC_API void orc_gradients(int nx, int ny, int nz, int nsub, double dx, const double *rho, double *grad)
{
    const int n[3] = {nx, ny, nz};
    const size_t jump[3] = {(size_t)ny * nz * nsub, (size_t)nz * nsub, (size_t)nsub};
    const double two_dx = dx * 2.0;
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                const int at[3] = {x, y, z};
                const size_t v = ((size_t)x * ny + y) * nz + z;
                for (int d = 0; d < nsub; ++d) {
                    const double *r = rho + v * nsub + d;
                    double *g = grad + (v * nsub + d) * 3;
                    for (int a = 0; a < 3; ++a) {
                        const ptrdiff_t j = (ptrdiff_t)jump[a];
                        if (n[a] == 1)
                            g[a] = 0.0;
                        else if (at[a] == 0)
                            g[a] = (r[j] - r[0]) / dx;
                        else if (at[a] == n[a] - 1)
                            g[a] = (r[0] - r[-j]) / dx;
                        else
                            g[a] = (r[j] - r[-j]) / two_dx;
                    }
                }
            }
}

/* Agent -> voxel, Basic_Agent::assign_position + update_voxel_index
 * (BioFVM_basic_agent.cpp:112-153) with Cartesian_Mesh::nearest_voxel_local_index at world
 * size 1 (BioFVM_parallel.cpp:418-522) and is_position_valid (BioFVM_mesh.cpp:204-213).
 * bbox = {xmin, ymin, zmin, xmax, ymax, zmax}.  Returns the voxel index, or -1 when
 * assign_position rejects the position (the reference then leaves the agent at its
 * constructor default, voxel 0, still active).                                           */
ORC_API int orc_voxel_index(const double *pos, const double *bbox, double dx, double dy,
                            double dz, int nx, int ny, int nz)
{
    for (int a = 0; a < 3; ++a)
        if (pos[a] < bbox[a] || pos[a] > bbox[a + 3])
            return -1;
    int xv = (int)floor((pos[0] - bbox[0]) / dx);
    int yv = (int)floor((pos[1] - bbox[1]) / dy);
    int zv = (int)floor((pos[2] - bbox[2]) / dz);
    int gx = (int)((bbox[3] - bbox[0]) / dx);
    int gy = (int)((bbox[4] - bbox[1]) / dy);
    int gz = (int)((bbox[5] - bbox[2]) / dz);
    if (xv >= gx) xv = gx - 1;
    if (xv < 0) xv = 0;
    if (yv >= gy) yv = gy - 1;
    if (yv < 0) yv = 0;
    if (zv >= gz) zv = gz - 1;
    if (zv < 0) zv = 0;
    int ly = (int)((bbox[4] - bbox[1]) / (1 * dy));
    int lz = (int)((bbox[5] - bbox[2]) / (1 * dz));
    long v = (long)xv * ly * lz + (long)yv * lz + zv;
    long nv = (long)nx * ny * nz;
    if (v < 0) v = 0;
    if (v >= nv) v = nv - 1;
    return (int)v;
}

/* A whole run as the BioFVM-B demos drive it (complete_diffusion_demo.cpp:489-493):
 * per step, cell sources/sinks (when with_agents) then one LOD step.  The Thomas tables
 * are built once from D, decay, dt and dx, as the solver's first call does.  dvox: the
 * per-voxel Dirichlet values ([nvox][nsub]) or NULL (dvalue per substrate). */
ORC_API void orc_run_dv(int nx, int ny, int nz, int nsub, double dx, double dt, const double *D,
                        const double *decay, const double *dvalue, const double *dvox,
                        const unsigned char *dactive, int nagents, const int *voxel,
                        const unsigned char *active, const double *temp1, const double *temp2,
                        int nsteps, double *rho,
                        double *work /* >= (3*(nx+ny+nz) + 4) * nsub doubles */)
{
    double *c1 = work, *c2 = c1 + nsub, *c3 = c2 + nsub, *c3a = c3 + nsub;
    double *denx = c3a + nsub, *cx = denx + (size_t)nx * nsub;
    double *deny = cx + (size_t)nx * nsub, *cy = deny + (size_t)ny * nsub;
    double *denz = cy + (size_t)ny * nsub, *cz = denz + (size_t)nz * nsub;
    orc_lod_constants(nsub, D, decay, dt, dx, c1, c2, c3, c3a);
    orc_thomas_axis(nx, nsub, c1, c2, c3, c3a, denx, cx);
    orc_thomas_axis(ny, nsub, c1, c2, c3, c3a, deny, cy);
    orc_thomas_axis(nz, nsub, c1, c2, c3, c3a, denz, cz);
    for (int s = 0; s < nsteps; ++s) {
        if (nagents > 0)
            orc_sources_sinks(nagents, nsub, voxel, active, temp1, temp2, rho);
        orc_lod_step_dv(nx, ny, nz, nsub, rho, c1, denx, cx, deny, cy, denz, cz, dvalue, dvox, dactive);
    }
}

ORC_API void orc_run(int nx, int ny, int nz, int nsub, double dx, double dt, const double *D,
                     const double *decay, const double *dvalue, const unsigned char *dactive,
                     int nagents, const int *voxel, const unsigned char *active,
                     const double *temp1, const double *temp2, int nsteps, double *rho,
                     double *work)
{
    orc_run_dv(nx, ny, nz, nsub, dx, dt, D, decay, dvalue, NULL, dactive, nagents, voxel, active, temp1,
               temp2, nsteps, rho, work);
}
