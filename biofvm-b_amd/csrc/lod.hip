// lod.hip -- the three Thomas sweeps of one LOD diffusion-decay step on gfx950.
//
// The reference (BioFVM_parallel.cpp:693-1395) solves, for every grid line along x, then
// y, then z, the tridiagonal system
//     forward   u[0] /= den[0];   u[i] = fma(c1, u[i-1], u[i]) / den[i]
//     backward  u[i] = fma(-c[i], u[i+1], u[i]),   i = n-2 .. 0
// and overwrites all six domain faces with the Dirichlet value of each active substrate
// before the x sweep and after every sweep (BioFVM_microenvironment.cpp:381-484).
//
// Design (the field is [substrate][z][y][x], x contiguous; every sweep reads one buffer and
// writes the other):
//  * One wavefront owns 64 lines.  The forward results of the tail of every line are held
//    in a per-wavefront LDS "line store" (a share of the CU's 160 KB); the backward pass
//    reads them from there.  The part of a line that fits neither the store nor registers
//    is only swept in the forward pass, keeping the recurrence value at each segment end in
//    the output row the backward pass overwrites last; the backward pass then recomputes each
//    such segment from its checkpoint (re-reading the input, bit-identical arithmetic) into
//    the freed store.  HBM sees one read and one write per voxel, plus that re-read.
//  * y and z sweeps: lane = line, consecutive lanes on consecutive x (y sweep) or xy
//    (z sweep), so each step of the recurrence is a coalesced 512 B row access; a ring of
//    P rows per lane is kept in flight.
//  * x sweep: lines lie along the contiguous axis.  Chunks of 64 rows x 16 columns are
//    loaded with coalesced row segments (several chunks in flight), transposed through an
//    LDS staging tile, and each lane walks its own row.  Final values go back the same way.
//  * FP64 division: q = a*r, then two corrections e = fma(-q,d,a), q = fma(e,r,q) with
//    r = RN(1/d) from the host.  With q faithful after the first correction, Markstein's
//    theorem makes the second result the correctly rounded a/d, i.e. bit-identical to the
//    reference's division, with a 6-deep instead of a ~10-deep dependent chain.  Numerators
//    below 2^-960 (outside the theorem's no-underflow range), infinities and NaNs take a true
//    division.
//  * FP32 storage path: the recurrence runs in FP64 on values converted on load (with the
//    FP64 reciprocal), results are rounded to FP32 once when stored.
//  * Dirichlet is fused: the x sweep applies it on load (the reference's first call) and
//    every sweep applies it on store; a line lying entirely on an active face is written
//    with the value directly (the reference solves it and then overwrites all of it).
#include <type_traits>

#include "bfvmb_internal.hpp"

namespace bfvmb {
namespace {

// Thomas tables are read through the constant address space: the index is uniform, so
// every read becomes a scalar load that never waits on the vector-memory counter that
// tracks the in-flight field rows.
typedef const __attribute__((address_space(4))) double *ctab;
__device__ __forceinline__ ctab as_ctab(const double *p) { return (ctab)p; }

// Field rows are accessed through the global address space: a generic (flat) access also
// counts on the LDS counter, so every wait for the line store would drain the rows in flight.
template <class T>
using gmem = __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ gmem<T> as_global(T *p)
{
    return (gmem<T>)p;
}

// A sweep's results, which nothing in the sweep reads again, are stored non-temporally where
// each store instruction writes whole 128-byte lines.  The HBM bytes stay the same (FETCH_SIZE
// and WRITE_SIZE per launch unchanged at 400^3 x 4), but the sweeps move them faster (FP64 x
// with the agent update 5.08 -> 5.45 TB/s).  Measured at 400^3 x 4 (us per sweep, results
// bit-identical): FP64 x 1211 -> 1144, y 911 -> 858, z 962 -> 916; FP32 y 451 -> 432, z 474
// -> 448.  Where a store covers parts of lines (rows
// that do not start on a line: most members of a batch of uneven domains, FP64 11.6 ->
// 13.2 ms per batch step with non-temporal stores everywhere; the FP32 x sweep's 64-byte
// halves of lines: 648 -> 696 us), partial lines leave the L2 before their other part is
// written, so those stores stay plain.  NT is a kernel parameter (a run-time choice between
// the two stores is merged back into one plain store by the compiler); the host launches the
// NT kernels where every domain's rows (x, y) or planes (z) are whole lines (nt_rows).
template <bool NT, class T>
__device__ __forceinline__ void store_result(gmem<T> p, T v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

// Checkpoint of the forward recurrence at a recomputed segment's start, kept in output elements that
// the backward pass overwrites later.  FP64 storage keeps the value itself; FP32 storage
// keeps the double's two 32-bit halves in two elements (off and off2), so that a recomputed
// segment starts from exactly the value of the first pass, and results do not depend on how
// much of a line stays on chip (launch shape, kernel variant).
template <class T>
__device__ __forceinline__ void put_ck(gmem<T> p, size_t off, size_t off2, double y)
{
    if constexpr (sizeof(T) == 8) {
        (void)off2;
        p[off] = y;
    } else {
        const unsigned long long b = (unsigned long long)__double_as_longlong(y);
        p[off] = __uint_as_float((unsigned)b);
        p[off2] = __uint_as_float((unsigned)(b >> 32));
    }
}
template <class T>
__device__ __forceinline__ double get_ck(gmem<T> p, size_t off, size_t off2)
{
    if constexpr (sizeof(T) == 8) {
        (void)off2;
        return p[off];
    } else {
        const unsigned long long lo = __float_as_uint(p[off]), hi = __float_as_uint(p[off2]);
        return __longlong_as_double((long long)((hi << 32) | lo));
    }
}

// a/d correctly rounded from r = RN(1/d): q = a*r, two corrections (Markstein).  Valid for
// finite |a| >= 2^-960 (no underflow in the corrections; an infinite a would give NaN where
// the division gives +-inf); callers flag the other numerators.
__device__ __forceinline__ double div_corr(double a, double d, double r)
{
    double q = a * r;
    double e = fma(-q, d, a);
    q = fma(e, r, q);
    e = fma(-q, d, a);
    return fma(e, r, q);
}
// Numerator range check in one integer operation per division: with h = the upper 32 bits
// of a's pattern, key(a) = (h << 1) - (kTinyHi << 1) (mod 2^32; the shift drops the sign) is
// below kKeyBad exactly when 2^-960 <= |a| < inf, i.e. when div_corr is exact.  Tiny
// numerators wrap round to the top of the range, infinities and NaNs (h >= 0x7ff00000) land
// at or above kKeyBad.  A line keeps the largest key it met.
constexpr unsigned kTinyHi = (1023u - 960u) << 20;  // upper half of 2^-960
constexpr unsigned kKeyBad = (0x7ff00000u << 1) - (kTinyHi << 1);
__device__ __forceinline__ unsigned num_key(double a)
{
    return ((unsigned)((unsigned long long)__double_as_longlong(a) >> 32) << 1) - (kTinyHi << 1);
}

// One forward step in double: y = fma(c1, prev, v) / den.  With prev = -0.0 the same
// expression gives the first element's v / den[0] exactly, signed zeros included.
// FP64 storage: corrected reciprocal; `akey` keeps the largest num_key of the line's
// numerators (two integer operations) so that a numerator outside its range is detected once
// per line; FP32 storage: the FP64 reciprocal.
template <class T>
__device__ __forceinline__ double fwd(double c1, double prev, double v, double den, double rden, unsigned &akey)
{
    const double num = fma(c1, prev, v);
    if constexpr (sizeof(T) == 8) {
        akey = max(akey, num_key(num));
        return div_corr(num, den, rden);
    } else {
        (void)den;
        return num * rden;
    }
}

// First row of a line kept in the line store of L rows: the forward pass keeps [lo, n) in
// LDS; rows before lo go to registers or are recomputed; lo is a multiple of P (L >= P is
// required).
__host__ __device__ __forceinline__ int park_boundary(int n, int L, int P)
{
    if (L >= n)
        return 0;
    const int lo = (n - L + P - 1) / P * P;
    return lo < n ? lo : n;
}

// XCD-aware tile order (y/z sweeps).  Workgroups are dealt to the 8 XCDs round-robin (block b
// runs on XCD b % 8), each XCD with its own L2.  XCD x takes the contiguous tiles
// [x*q + min(x, r), ...) (q + 1 of them for x < r; q, r = n / 8, n % 8) and its blocks
// stride over them in dispatch order, so that neighbouring tiles, which share the cache
// lines at their borders wherever a row is not 128-byte aligned, meet in the same L2.
constexpr int kXcd = 8;

// Domain of global tile t: the largest i with prefix[i] <= t (prefix has ndom+1 entries).
__device__ __forceinline__ int find_domain(const long long *prefix, int ndom, long long t)
{
    int lo = 0, hi = ndom - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// y/z sweeps: Thomas tables read with scalar loads (true) or staged in LDS per substrate.
// Every lane of a wavefront is at the same row, so the table entries are wave-uniform: read
// with scalar loads they take no LDS and no vector registers, and the line store gets the
// 24 bytes per row the staged tables took (+19 rows of 64 lanes at 400 rows, FP64).
// Measured at 400^3 x 4 (ms, isolated launches): FP64 y 0.980 -> 0.932, z 1.023 -> 0.967
// (step 3.364 -> 3.231); FP32 y 0.483 -> 0.462, but z 0.479 -> 0.487, so the FP32 z sweep
// keeps its staged tables.  The gain is the line store's: it needs lines that overflow the
// on-chip storage.  Isolated y / z launches of n^3 domains (us, scalar against staged): FP64
// n = 128: 28.4 / 29.1 against 26.4 / 26.5; n = 320: 295.5 / 302.9 against 277.6 / 288.9;
// n = 400: 512 / 541 against 533 / 552; n = 512: 563 / 597 against 573 / 614; at 640^3 x 8
// staged tables win (y 8.37 against 8.74 ms, z 8.75 against 9.07: 15 KB of tables per
// substrate, which the scalar cache holds less well).  FP32 y: n = 128: 21.4 against 18.3;
// n = 400: 259.8 against 265.6; 640^3 x 8: 3.32 against 3.64 ms.  So a launch reads its
// tables with scalar loads where its longest line has at least kYzScalarMinRows rows (FP64:
// and at most kYzScalarMaxRows).
constexpr int kYzScalarMinRows = 352, kYzScalarMaxRows = 512;
template <class T, int AXIS>
__host__ __device__ constexpr bool yz_scalar_ok() { return sizeof(T) == 8 || AXIS == 1; }
// LDS bytes per row of a y/z line taken by its staged tables ({den, rden}, cc)
__host__ __device__ constexpr long long yz_table_bytes(bool st) { return st ? 0 : 24; }

// Elements per line kept in the y/z line store for a line of n with `budget` LDS bytes
// (after the staged tables, if any).
template <class T>
__host__ __device__ __forceinline__ int yz_line_store(int n, long long budget, bool st)
{
    long long L = (budget - yz_table_bytes(st) * n) / (kWave * (long long)sizeof(T));
    if (L > n)
        L = n;
    return L < 0 ? 0 : (int)L;
}

// x-sweep chunk width: 16 columns (one 128-byte cache line of a row in FP64, half of one
// in FP32, where 32-column chunks measured slower: twice the registers per chunk).
template <class T>
constexpr int x_chunk() { return 16; }

// Columns per row kept in the x line store (a multiple of the chunk width, odd pitch),
// after the 64 x (W+1) staging tile and, with tables in LDS (tl), two FP64 tables of nx.
template <class T>
__host__ __device__ __forceinline__ int x_line_store(int nx, long long budget, bool tl)
{
    constexpr int W = x_chunk<T>();
    if (tl)
        budget -= 16LL * nx;
    const long long stg = (long long)kWave * (W + 1) * sizeof(T);
    long long L = budget > stg ? (budget - stg) / (kWave * (long long)sizeof(T)) - 1 : 0;
    L = L / W * W;
    const long long need = (nx + W - 1) / W * W;
    if (L > need)
        L = need;
    return L < 0 ? 0 : (int)L;
}

// ------------------------------------------------------------------------------ y / z
// Block = one wavefront = 64 lines of one substrate.  LDS: the line store [L][64] (staged
// tables {den, rden}[n], cc[n] of the block's substrate before it, in launches that do not
// read them with scalar loads, ST false).  P = rows in flight per lane (ring depth); table entries are read in batches of TB.  Before the line store, up to
// reg_rows<T>() forward results per line are kept in registers (the wavefront has a SIMD to
// itself at the default four per CU, and its register file is larger than its LDS share).
// Register rows per lane: 160 in FP64 (320 registers, room made by table batches of 8), 320 in
// FP32 (one register per row, table batches of 16), so that with the line store an FP32
// 400-long line stays on chip.  Measured at 400^3 x 4: FP32 y 0.522 -> 0.499 ms, z 0.542 ->
// 0.505 ms against 128 FP32 rows; FP64 y 0.972 -> 0.961, z 1.029 -> 1.012 against 128 rows and
// batches of 16.
#ifndef BFVMB_YZ_REG_ROWS_FP64
#define BFVMB_YZ_REG_ROWS_FP64 160
#endif
template <class T>
constexpr int reg_rows() { return sizeof(T) == 8 ? BFVMB_YZ_REG_ROWS_FP64 : 320; }
template <class T>
constexpr int table_batch() { return sizeof(T) == 8 ? 8 : 16; }

// Per-voxel Dirichlet values (descriptor dface) of a y/z line's Dirichlet voxels, stored over
// what the sweep stored there; the lane itself wrote those elements, so program order makes
// these stores the last.  The line (x, o) runs along y (AXIS 1, o = z) or z (AXIS 2, o = y):
// its two ends lie on that axis' faces, and a line with x or o on a face lies wholly on it.
template <class T, int AXIS>
__device__ __attribute__((noinline)) void yz_face_values(const DomainDesc<T> &D, int s, gmem<T> dst, size_t stride,
                                                         int n, int x, int o, bool face)
{
    const int nx = D.nx, ny = D.ny, nz = D.nz, ns = D.nsub;
    const T *fv = D.dface;
    if (AXIS == 1) {
        dst[0] = fv[face_y(nx, ny, nz, ns, 0, s, x, o)];
        dst[(size_t)(n - 1) * stride] = fv[face_y(nx, ny, nz, ns, 1, s, x, o)];
    } else {
        dst[0] = fv[face_z(nx, ny, nz, ns, 0, s, x, o)];
        dst[(size_t)(n - 1) * stride] = fv[face_z(nx, ny, nz, ns, 1, s, x, o)];
    }
    if (!face)
        return;
    for (int k = 0; k < n; ++k) {
        size_t i;
        if (x == 0 || x == nx - 1)
            i = AXIS == 1 ? face_x(ny, nz, ns, x == 0 ? 0 : 1, s, k, o) : face_x(ny, nz, ns, x == 0 ? 0 : 1, s, o, k);
        else if (AXIS == 1)
            i = face_z(nx, ny, nz, ns, o == 0 ? 0 : 1, s, x, k);
        else
            i = face_y(nx, ny, nz, ns, o == 0 ? 0 : 1, s, x, k);
        dst[(size_t)k * stride] = fv[i];
    }
}

template <class T, int AXIS, int P, bool ST, bool NT = false>
__global__ __launch_bounds__(64) void k_sweep_yz(const DomainDesc<T> *__restrict__ descs,
                                                 const long long *__restrict__ prefix, int ndom, int parity,
                                                 int budget)
{
    constexpr int TB = table_batch<T>();
    static_assert(P % TB == 0, "ring depth must be a multiple of the table batch");
    constexpr int RC = (P == 64 ? reg_rows<T>() / 2 : reg_rows<T>()) / P;  // rings of rows kept in registers
                                                                 // (fewer beside a 64-row ring)
    extern __shared__ unsigned char smem[];
    const int lane = threadIdx.x;
    const long long ntiles = prefix[ndom];
    int tab_dom = -1, tab_s = -1;

    // XCD-aware tile order (one tile per block with one block per tile); a grid that would
    // leave an XCD with tiles but no blocks strides over all tiles in order instead
    const long long G = gridDim.x;
    const bool by_xcd = G >= kXcd || G == ntiles;
    const long long xq = ntiles / kXcd, xr = ntiles % kXcd, bx = blockIdx.x % kXcd;
    const long long t_first = by_xcd ? bx * xq + min(bx, xr) : 0;                 // XCD's first tile
    const long long t_count = by_xcd ? xq + (bx < xr ? 1 : 0) : ntiles;          // its tile count
    const long long t_step = by_xcd ? G / kXcd + (bx < G % kXcd ? 1 : 0) : G;    // its blocks
    for (long long k = by_xcd ? blockIdx.x / kXcd : blockIdx.x; k < t_count; k += t_step) {
        const long long tile = t_first + k;
        const int di = __builtin_amdgcn_readfirstlane(find_domain(prefix, ndom, tile));  // uniform
        const DomainDesc<T> &D = descs[di];
        const int nx = D.nx, ny = D.ny, nz = D.nz;
        const size_t nxy = (size_t)nx * ny;
        const size_t nvox = nxy * nz;
        const int n = AXIS == 1 ? ny : nz;
        double2 *tab2 = reinterpret_cast<double2 *>(smem);
        double *tabc = reinterpret_cast<double *>(tab2 + n);
        T *ls = ST ? reinterpret_cast<T *>(smem) : reinterpret_cast<T *>(tabc + n);
        const int L = yz_line_store<T>(n, budget, ST);
        const long long lines_per_s = AXIS == 1 ? (long long)nx * nz : (long long)nxy;
        const long long tiles_per_s = (lines_per_s + kWave - 1) / kWave;
        const size_t stride = AXIS == 1 ? (size_t)nx : nxy;
        // rows [lds_lo, n) go to the line store, the head [0, R) to registers (whole rings of
        // rows), the middle [R, lds_lo) is recomputed.  The middle is recomputed right after the
        // tail is substituted back, segment by segment towards the head, so each re-read follows
        // its first read at the shortest distance the order of the two passes allows.
        const int lds_lo = park_boundary(n, L, P);
        const int nregc = min(RC, lds_lo / P);
        const int R = nregc * P;
        const long long lt = tile - prefix[di];
        const int s = __builtin_amdgcn_readfirstlane((int)(lt / tiles_per_s));  // 64-bit division runs on
                                                                              // the vector unit
        long long line = (lt - (long long)s * tiles_per_s) * kWave + lane;
        if (line >= lines_per_s)
            line = lines_per_s - 1;  // idle lanes repeat the last line (same values, same stores)
        const int o = (int)(line / nx);
        const int x = (int)(line - (long long)o * nx);
        const size_t base = (size_t)s * nvox + (AXIS == 1 ? (size_t)o * nxy + x : (size_t)line);
        const int on = AXIS == 1 ? nz : ny;
        const bool act = __builtin_amdgcn_readfirstlane((int)D.dact[s]) != 0;
        const T dv = ((const __attribute__((address_space(4))) T *)D.dval)[s];
        const gmem<const T> src = as_global<const T>(D.field[parity] + base);
        const gmem<T> dst = as_global<T>(D.field[parity ^ 1] + base);
        // a line on an active face is computed along but stored as the Dirichlet value
        const bool face = act && (x == 0 || x == nx - 1 || o == 0 || o == on - 1);
        const double c1 = as_ctab(D.c1)[s];

        // ---- forward elimination; a ring of P rows is in flight.  The head [0, R) goes to
        // registers; the middle [R, lds_lo) is only swept here (its results are recomputed
        // segment by segment in the backward pass, from checkpoints kept at the segment
        // starts); the tail [lds_lo, n) goes to the line store.  Measured at 400^3 x 4 against
        // registers holding the rows just before the tail and the head recomputed last: FP64 y
        // 0.967 -> 0.946 ms, z 1.024 -> 1.019 ms, FP32 y 0.491 -> 0.475 ms.
        const int S = L / P * P;  // middle segment length (the host guarantees L >= P)
        const int nmid = lds_lo > R ? (lds_lo - R + S - 1) / S : 0;
        const int m0 = lds_lo - (nmid - 1) * S;  // end of the first middle segment
        T buf[P];
#pragma unroll
        for (int j = 0; j < P; ++j)
            buf[j] = src[(size_t)min(j, n - 1) * stride];
        // stage this substrate's Thomas tables after the first ring of rows is on its way, so
        // that the two load latencies overlap (a block starts with both at every new tile)
        const ctab sden = as_ctab(D.tab[AXIS].den + (size_t)s * n);
        const ctab srden = as_ctab(D.tab[AXIS].rden + (size_t)s * n);
        const ctab scc = as_ctab(D.tab[AXIS].cc + (size_t)s * n);
        auto tdr = [&](int k) -> double2 { return ST ? make_double2(sden[k], srden[k]) : tab2[k]; };
        auto tcc = [&](int k) -> double { return ST ? scc[k] : tabc[k]; };
        if (!ST && (di != tab_dom || s != tab_s)) {
            __syncthreads();
            const double *gd = D.tab[AXIS].den + (size_t)s * n, *gr = D.tab[AXIS].rden + (size_t)s * n;
            const double *gc = D.tab[AXIS].cc + (size_t)s * n;
            for (int i = lane; i < n; i += kWave) {
                tab2[i] = make_double2(gd[i], gr[i]);
                tabc[i] = gc[i];
            }
            __syncthreads();
            tab_dom = di;
            tab_s = s;
        }
        double y = -0.0;
        unsigned akey = 0u;
        T regs[RC * P];  // register rows [0, R)
        // one full chunk of P rows from k0: prefetches row min(k + P, lim) (PF), stores the
        // results (STORE: 1 = line store at row k - base, 2 = registers rp[j])
        auto chunk = [&](int k0, int base, int lim, auto store_tag, auto exact_tag, auto pf_tag, T *rp) {
            constexpr int STORE = decltype(store_tag)::value;
            constexpr bool EXACT = decltype(exact_tag)::value;
            constexpr bool PF = decltype(pf_tag)::value;
#pragma unroll
            for (int jb = 0; jb < P; jb += TB) {
                double2 t[TB];
#pragma unroll
                for (int jj = 0; jj < TB; ++jj)
                    t[jj] = tdr(k0 + jb + jj);
#pragma unroll
                for (int jj = 0; jj < TB; ++jj) {
                    const int j = jb + jj;
                    const int k = k0 + j;
                    const double v = (double)buf[j];
                    if constexpr (PF)
                        buf[j] = src[(size_t)min(k + P, lim) * stride];
                    if constexpr (EXACT)
                        y = fma(c1, y, v) / t[jj].x;
                    else
                        y = fwd<T>(c1, y, v, t[jj].x, t[jj].y, akey);
                    if constexpr (STORE == 1)
                        ls[(size_t)(k - base) * kWave + lane] = (T)y;
                    else if constexpr (STORE == 2)
                        rp[j] = (T)y;
                }
            }
        };
        const std::integral_constant<int, 0> no_store;
        const std::integral_constant<int, 1> store;
        const std::integral_constant<int, 2> reg_store;
        const std::false_type inexact, no_pf;
        const std::true_type exact_div, pf;
#pragma unroll
        for (int q = 0; q < RC; ++q)  // register rows
            if (q < nregc)
                chunk(q * P, 0, n - 1, reg_store, inexact, pf, regs + q * P);
        if (R > 0 && R < lds_lo)  // checkpoint of the first middle segment
            put_ck<T>(dst, (size_t)(R - 1) * stride, (size_t)(R - 2) * stride, y);
        int k0 = R;
        for (int next_ck = m0; k0 < lds_lo; k0 += P) {  // middle: checkpoints only
            chunk(k0, 0, n - 1, no_store, inexact, pf, nullptr);
            if (k0 + P == next_ck && next_ck < lds_lo) {
                put_ck<T>(dst, (size_t)(next_ck - 1) * stride, (size_t)(next_ck - 2) * stride, y);
                next_ck += S;
            }
        }
        for (k0 = lds_lo; k0 + P <= n; k0 += P)  // tail: line store
            chunk(k0, lds_lo, n - 1, store, inexact, pf, nullptr);
        if (k0 < n) {  // the last, partial chunk (uniform per-row conditions)
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const int k = k0 + j;
                if (k < n) {
                    const double2 t = tdr(k);
                    y = fwd<T>(c1, y, (double)buf[j], t.x, t.y, akey);
                    ls[(size_t)(k - lds_lo) * kWave + lane] = (T)y;
                }
            }
        }
        bool exact = false;
        if constexpr (sizeof(T) == 8) {
            // a numerator below 2^-960 or not finite somewhere in the tile: redo the forward
            // pass with true divisions (same results everywhere else); the middle recomputation below
            // then divides the same way
            if (__builtin_expect(__any(akey >= kKeyBad), 0)) {
                exact = true;
                y = -0.0;
#pragma unroll
                for (int q = 0; q < RC; ++q)
                    if (q < nregc) {
#pragma unroll
                        for (int j = 0; j < P; ++j) {
                            const int k = q * P + j;
                            y = fma(c1, y, (double)src[(size_t)k * stride]) / tdr(k).x;
                            regs[q * P + j] = (T)y;
                        }
                    }
                if (R > 0 && R < lds_lo)
                    dst[(size_t)(R - 1) * stride] = (T)y;
                for (int k = R; k < lds_lo; ++k) {
                    y = fma(c1, y, (double)src[(size_t)k * stride]) / tdr(k).x;
                    if (k + 1 >= m0 && k + 1 < lds_lo && (k + 1 - m0) % S == 0)
                        dst[(size_t)k * stride] = (T)y;
                }
                for (int k = lds_lo; k < n; ++k) {
                    y = fma(c1, y, (double)src[(size_t)k * stride]) / tdr(k).x;
                    ls[(size_t)(k - lds_lo) * kWave + lane] = (T)y;
                }
            }
        }

        // ---- back substitution; the last element is y[n-1] itself
        double xv = y;
        store_result<NT>(dst + (size_t)(n - 1) * stride, act ? dv : (T)xv);
        // rows [base, top] from the line store (row k at k - base), in batches of TB reads
        auto back = [&](int top, int base) {
            for (int kt = top; kt >= base; kt -= TB) {
                if (kt - (TB - 1) >= base) {
                    double v[TB], c[TB];
#pragma unroll
                    for (int j = 0; j < TB; ++j) {
                        v[j] = (double)ls[(size_t)(kt - j - base) * kWave + lane];
                        c[j] = tcc(kt - j);
                    }
#pragma unroll
                    for (int j = 0; j < TB; ++j) {
                        const int k = kt - j;
                        xv = fma(-c[j], xv, v[j]);
                        store_result<NT>(dst + (size_t)k * stride, (face || (act && k == 0)) ? dv : (T)xv);
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < TB; ++j) {
                        const int k = kt - j;
                        if (k >= base) {
                            xv = fma(-tcc(k), xv, (double)ls[(size_t)(k - base) * kWave + lane]);
                            store_result<NT>(dst + (size_t)k * stride, (face || (act && k == 0)) ? dv : (T)xv);
                        }
                    }
                }
            }
        };
        // middle segments, last first: recompute the forward results from the checkpoint at
        // the row before the segment (re-reading the input), then substitute back.  The ring
        // and checkpoint of the next segment are loaded while the current one substitutes back.
        double yck = -0.0;
        auto seg_begin = [&](int b) { return b > m0 ? b - S : R; };
        auto prefetch_segment = [&](int b) {
            const int a = seg_begin(b);
            yck = a == 0 ? -0.0 : get_ck<T>(dst, (size_t)(a - 1) * stride, (size_t)(a - 2) * stride);
#pragma unroll
            for (int j = 0; j < P; ++j)
                buf[j] = src[(size_t)min(a + j, b - 1) * stride];
        };
        if (nmid > 0)
            prefetch_segment(lds_lo);
        back(n - 2, lds_lo);
        for (int b = lds_lo; b > R;) {
            const int a = seg_begin(b);
            y = yck;
            for (int kk = a; kk < b; kk += P) {
                const bool last = kk + P >= b;
                if (exact) {
                    if (last)
                        chunk(kk, a, b - 1, store, exact_div, no_pf, nullptr);
                    else
                        chunk(kk, a, b - 1, store, exact_div, pf, nullptr);
                } else {
                    if (last)
                        chunk(kk, a, b - 1, store, inexact, no_pf, nullptr);
                    else
                        chunk(kk, a, b - 1, store, inexact, pf, nullptr);
                }
            }
            if (a > R)
                prefetch_segment(a);
            back(b - 1, a);
            b = a;
        }
#pragma unroll
        for (int q = RC - 1; q >= 0; --q)  // register rows, last first
            if (q < nregc) {
#pragma unroll
                for (int jb = P - TB; jb >= 0; jb -= TB) {
                    double c[TB];
#pragma unroll
                    for (int jj = 0; jj < TB; ++jj)
                        c[jj] = tcc(q * P + jb + jj);
#pragma unroll
                    for (int jj = TB - 1; jj >= 0; --jj) {
                        const int k = q * P + jb + jj;
                        xv = fma(-c[jj], xv, (double)regs[q * P + jb + jj]);
                        store_result<NT>(dst + (size_t)k * stride, (face || (act && k == 0)) ? dv : (T)xv);
                    }
                }
            }
        if (act && D.dface != nullptr)  // per-voxel face values (uniform branch)
            yz_face_values<T, AXIS>(D, s, dst, stride, n, x, o, face);
    }
}

// ------------------------------------------------------------------------------ x
// Block = one wavefront = 64 consecutive rows (y,z) of one substrate.  LDS: staging tile
// [64][W+1] and line store [64][L+1] (odd pitches keep the transposed accesses
// conflict-free).  Column chunks are W = 16 wide; PC chunks of loads are kept in flight.
// Variants (PC chunks in flight, REGC register chunks), chosen per launch by x_variant:
//  * FP64 (2, 0): at most 256 registers per lane, so that two wavefronts share a SIMD (the x
//    sweep gains more from a second wavefront hiding its latencies than from a deeper ring;
//    one wavefront per SIMD with register chunks measured no faster: FP64 register chunks
//    beside the FP64 working set spill, 86 to 372 registers per lane for 8 to 12 chunks).
//  * FP32 at up to four wavefronts per CU (1, 17): one wavefront per SIMD, one chunk in
//    flight, and the forward results of the first 17 chunks kept in registers (an FP32
//    chunk takes 16 registers), so that with the line store's 8 chunks a 400-long line is
//    recomputed nowhere.  Measured at 400^3 x 4: deeper
//    rings with fewer register chunks were slower at every step (3 and 15: -10%).
//  * FP32 (3, 0): two wavefronts per SIMD and no register chunks, three chunks in flight (the
//    kernel needs ~230 registers per lane without register chunks, so two wavefronts per SIMD
//    leave no room for them: 8 register chunks spilled 392 bytes per lane; a (2, 16) variant,
//    two chunks in flight beside 16 register chunks, measured no faster than (1, 17), with
//    its tables by scalar loads or in LDS).
// Each FP32 variant reads its Thomas tables (rden forward, cc backward) either with scalar
// loads or, staged once per tile, from LDS (tl): scalar loads return out of order, so every
// wait for an LDS access also waits for them, and the FP32 sweep, one wavefront per SIMD,
// has no other wavefront to cover that; the LDS copy costs 16 bytes per column of line store.
// The launch shape's ring value picks the FP32 variant: 1 = (1, 17), 2 = (1, 17) tl,
// 3 = (3, 0), 4 = (3, 0) tl; ring 0 takes 2 at up to four wavefronts per CU and 3 above.
// Measured at 400^3 x 4 (ms): 1 and 2 run equally fast with one block per tile (0.649 /
// 0.654: the tables cost a chunk of line store, and a 400-column line then recomputes one
// chunk), but with a grid of one block per SIMD striding over the tiles 2 stages its tables
// once per substrate and takes 0.620 (1: 0.639), so that grid is the default for the
// register-chunk variants.  On the batch of 64 uneven domains (lines of 32-256) 2 takes
// 2.41 against 2.65 for 1.
struct XVariant {
    int pc, regc;
    bool tl;
};
template <class T>
XVariant x_variant(int bpc, int ring)
{
    if (sizeof(T) == 8)
        return {2, 0, false};
    if (ring == 1)
        return {1, 17, false};
    if (ring == 2)
        return {1, 17, true};
    if (ring == 3)
        return {3, 0, false};
    if (ring == 4)
        return {3, 0, true};
    return bpc <= 4 ? XVariant{1, 17, true} : XVariant{3, 0, false};
}

// Wavefronts per SIMD an x-sweep variant is compiled for: two without register chunks (256
// registers per lane), otherwise one.
template <class T, int REGC>
constexpr int x_waves_per_simd()
{
    return REGC == 0 ? 2 : 1;
}

#ifndef BFVMB_XAGENT_BATCHES
#define BFVMB_XAGENT_BATCHES 4
#endif
constexpr int kAgentBatches = BFVMB_XAGENT_BATCHES;  // segment batches of 64 in flight (AG)

// VF: the kernel also applies per-voxel Dirichlet values (descriptor dface) where a domain has
// them; a separate instantiation, so that the per-substrate case keeps its registers.
// AG: the step's cell sources and sinks are applied first, tile by tile (see below).
template <class T, int PC, int REGC, bool TL, bool VF, bool AG, bool NT = false>
__global__ __launch_bounds__(64, (x_waves_per_simd<T, REGC>())) void k_sweep_x(const DomainDesc<T> *__restrict__ descs,
                                                const long long *__restrict__ prefix, int ndom, int parity,
                                                int budget)
{
    static_assert(REGC % PC == 0, "register chunks come in whole rings");
    static_assert(!TL || sizeof(T) == 4, "tables in LDS: FP32 variants only");
    static_assert(!NT || sizeof(T) == 8, "non-temporal x-sweep stores: FP64 (whole-line row segments)");
    constexpr int W = x_chunk<T>();  // columns per chunk
    constexpr int RPI = kWave / W;   // rows per load instruction (4)
    constexpr int NI = kWave / RPI;  // load instructions per chunk (W)
    extern __shared__ unsigned char smem[];
    // LDS: TL tables rden[nx], cc[nx]; staging tile [64][W+1]; line store [64][L+1]
    double *xtr = reinterpret_cast<double *>(smem);
    const int lane = threadIdx.x;
    const int lrow = lane / W, lcol = lane % W;  // position of this lane in a load instruction
    const long long ntiles = prefix[ndom];
    int tab_dom = -1, tab_s = -1;

    for (long long t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int di = __builtin_amdgcn_readfirstlane(find_domain(prefix, ndom, t));  // uniform
        const DomainDesc<T> &D = descs[di];
        const int nx = D.nx, ny = D.ny, nz = D.nz;
        const size_t nvox = (size_t)nx * ny * nz;
        const int nrows = ny * nz;
        const int tiles_per_s = (nrows + kWave - 1) / kWave;
        double *xtc = xtr + nx;
        T *stg = reinterpret_cast<T *>(smem + (TL ? 16 * (size_t)nx : 0));  // [64][W+1]
        T *ls = stg + kWave * (W + 1);                                       // [64][L+1]
        const int L = x_line_store<T>(nx, budget, TL);
        const int Lp = L + 1;
        const int nch = (nx + W - 1) / W;
        const int Lc = L / W;  // chunks the line store holds (L is a multiple of W)
        // chunks [cl, nch) live in the line store, the head [0, cr) in registers and the middle
        // [cr, cl) is recomputed (right after the line store's chunks are substituted back, so
        // that each re-read follows its first read as closely as the two passes allow).  With
        // register chunks, cr and cl are multiples of PC, so that chunk c always sits in ring
        // slot c % PC (the host guarantees Lc >= PC for a line longer than the store, which
        // keeps cl < nch).
        int cl = nch - Lc;
        if constexpr (REGC > 0)
            cl = (cl + PC - 1) / PC * PC;
        if (cl < 0)
            cl = 0;
        const int cr = min(REGC, cl);
        const long long lt = t - prefix[di];
        const int s = __builtin_amdgcn_readfirstlane((int)(lt / tiles_per_s));  // 64-bit division runs on
                                                                              // the vector unit
        const int r0 = __builtin_amdgcn_readfirstlane((int)(lt - (long long)s * tiles_per_s) * kWave);
        const int rows_here = min(kWave, nrows - r0);
        const int my_row = r0 + min(lane, rows_here - 1);
        const int yy = my_row % ny, zz = my_row / ny;
        // per-substrate constants as scalar (wave-uniform) values
        const bool act = __builtin_amdgcn_readfirstlane((int)D.dact[s]) != 0;
        const T dv = ((const __attribute__((address_space(4))) T *)D.dval)[s];
        const bool whole = act && (yy == 0 || yy == ny - 1 || zz == 0 || zz == nz - 1);
        // per-voxel Dirichlet values (descriptor dface): read where they are used, inside
        // uniform branches, so that the per-substrate case keeps no registers for them
        const bool vf = VF && act && D.dface != nullptr;
        const gmem<const T> src = as_global<const T>(D.field[parity] + (size_t)s * nvox + (size_t)r0 * nx);
        const gmem<T> dst = as_global<T>(D.field[parity ^ 1] + (size_t)s * nvox + (size_t)r0 * nx);
        const double c1 = as_ctab(D.c1)[s];
        const ctab den = as_ctab(D.tab[0].den + (size_t)s * nx);
        const ctab cc = as_ctab(D.tab[0].cc + (size_t)s * nx);
        const ctab rden = as_ctab(D.tab[0].rden + (size_t)s * nx);

        // chunk c of this tile <-> registers: instruction i covers rows RPI*i .. RPI*i+3,
        // columns c*W .. c*W+15; out-of-range lanes re-read an in-range element.  Whole
        // chunks (all 64 rows, all 16 columns) address with a uniform base and one lane offset.
        const bool full_rows = rows_here == kWave;
        const unsigned lane_off = (unsigned)(lrow * nx + lcol);
        auto load_chunk = [&](gmem<const T> from, int c, T *r) {
            if (full_rows && c * W + W <= nx) {
                const gmem<const T> b = from + c * W;
#pragma unroll
                for (int i = 0; i < NI; ++i)
                    r[i] = b[(size_t)(RPI * i) * nx + lane_off];
            } else {
                const int col = min(c * W + lcol, nx - 1);
#pragma unroll
                for (int i = 0; i < NI; ++i) {
                    const int rr = min(RPI * i + lrow, rows_here - 1);
                    r[i] = from[(size_t)rr * nx + col];
                }
            }
        };
        // registers -> staging tile; with Dirichlet active, the x-face columns read as the
        // Dirichlet value (the overwrite before the x sweep)
        auto put_chunk = [&](const T *r, int c) {
            const int col = c * W + lcol;
            const bool edge = act && (col == 0 || col == nx - 1);
#pragma unroll
            for (int i = 0; i < NI; ++i)
                stg[(RPI * i + lrow) * (W + 1) + lcol] = edge ? dv : r[i];
        };
        // this lane's row of the staging tile gets its x-face values (chunk 0: column 0, last
        // chunk: column nx - 1); stored after put_chunk / the substitution wrote that element,
        // by the same wavefront, so the LDS keeps this value
        auto put_xfaces = [&](int c) {
            if (c == 0)
                stg[lane * (W + 1)] = D.dface[face_x(ny, nz, D.nsub, 0, s, yy, zz)];
            if (c == nch - 1)
                stg[lane * (W + 1) + (nx - 1 - c * W)] = D.dface[face_x(ny, nz, D.nsub, 1, s, yy, zz)];
        };
        auto put_result = [&](gmem<T> p, T v) { store_result<NT>(p, v); };  // NT: FP64 only
        auto store_chunk = [&](int c) {  // staging tile -> dst, coalesced
            if (full_rows && c * W + W <= nx) {
                const gmem<T> b = dst + c * W;
#pragma unroll
                for (int i = 0; i < NI; ++i)
                    put_result(b + (size_t)(RPI * i) * nx + lane_off, stg[(RPI * i + lrow) * (W + 1) + lcol]);
            } else {
                const int col = c * W + lcol;
#pragma unroll
                for (int i = 0; i < NI; ++i) {
                    const int rr = RPI * i + lrow;
                    if (rr < rows_here && col < nx)
                        put_result(dst + (size_t)rr * nx + col, stg[rr * (W + 1) + lcol]);
                }
            }
        };

        // ---- forward.  Head chunks [0, cr) go to registers; the middle [cr, cl) is swept
        // without storing its results (the backward pass recomputes it segment by segment from
        // checkpoints kept at the segment starts); the tail [cl, nch) goes to the line store.
        const int Sc = Lc;  // middle segment length in chunks (the host guarantees L >= W)
        const int nmid = cl > cr ? (cl - cr + Sc - 1) / Sc : 0;
        const int r0c = cl - (nmid - 1) * Sc;  // end of the first middle segment
        const size_t my_off = (size_t)min(lane, rows_here - 1) * nx;  // this lane's row
        // Checkpoint of the middle segment starting at chunk ca, in the output tile where that
        // segment's back-substitution writes last: in a whole tile the 64 lanes fill the
        // first RPI rows of chunk ca (contiguous row segments: a single scattered element
        // per row would make L2 fetch a line for every 8-byte write); otherwise the lane's
        // own row at the column before the segment.
        auto ck_off = [&](int ca) -> size_t {
            return full_rows ? (size_t)(lane / W) * nx + (size_t)ca * W + (lane % W) : my_off + (size_t)ca * W - 1;
        };
        // second element of an FP32 checkpoint: the next RPI rows of chunk ca, or the column
        // before ck_off's
        auto ck_off2 = [&](int ca) -> size_t {
            return full_rows ? ck_off(ca) + (size_t)RPI * nx : my_off + (size_t)ca * W - 2;
        };
        T ring[PC][NI];
        T regs[REGC > 0 ? REGC * W : 1];  // forward results of chunks [0, cr)
        double y = -0.0;
        // chunk c: stage the ring slot, refill it with chunk `next`, then W recurrence steps;
        // results (STORE) 1: to the line store at chunk c - base, 2: to registers rp[j]
        auto fwd_chunk = [&](int c, T *slot, int next, int base, auto store_tag, auto exact_tag, unsigned &akey,
                             T *rp) __attribute__((always_inline)) {
            constexpr int STORE = decltype(store_tag)::value;
            constexpr bool EXACT = decltype(exact_tag)::value;
            put_chunk(slot, c);
            if (vf && (c == 0 || c == nch - 1))
                put_xfaces(c);
            if (next >= 0)  // next < 0: nothing left to prefetch into this slot
                load_chunk(src, next, slot);
            __syncthreads();
            const int w = min(W, nx - c * W);
            double dd[W], rd[W], v[W];
#pragma unroll
            for (int j = 0; j < W; ++j) {
                const int kk = min(c * W + j, nx - 1);
                if constexpr (TL) {
                    dd[j] = 1.0;  // unused: FP32 storage multiplies by rden
                    rd[j] = xtr[kk];
                } else {
                    dd[j] = den[kk];
                    rd[j] = rden[kk];
                }
                v[j] = (double)stg[lane * (W + 1) + j];
            }
            T *tgt = ls + lane * Lp + (c - base) * W;
            auto step = [&](int j) {
                if constexpr (EXACT)
                    y = fma(c1, y, v[j]) / dd[j];
                else
                    y = fwd<T>(c1, y, v[j], dd[j], rd[j], akey);
                if constexpr (STORE == 1)
                    tgt[j] = (T)y;
                else if constexpr (STORE == 2)
                    rp[j] = (T)y;
            };
            if (STORE == 2 || w == W) {  // register chunks lie before the last chunk: never partial
#pragma unroll
                for (int j = 0; j < W; ++j)
                    step(j);
            } else {
#pragma unroll
                for (int j = 0; j < W; ++j)
                    if (j < w)
                        step(j);
            }
            __syncthreads();
        };
        const std::integral_constant<int, 0> no_store;
        const std::integral_constant<int, 1> store;
        const std::integral_constant<int, 2> reg_store;
        const std::false_type inexact;
        const std::true_type exact_div;
        auto forward = [&](auto exact_tag) {
            unsigned akey = 0u;
#pragma unroll
            for (int p = 0; p < PC; ++p)
                load_chunk(src, min(p, nch - 1), ring[p]);
            if constexpr (TL) {
                // this substrate's tables, staged after the first chunks are on their way so
                // that the two load latencies overlap (FP32 only: forward runs once per tile)
                if (di != tab_dom || s != tab_s) {
                    __syncthreads();
                    const double *gr = D.tab[0].rden + (size_t)s * nx, *gc = D.tab[0].cc + (size_t)s * nx;
                    for (int i = lane; i < nx; i += kWave) {
                        xtr[i] = gr[i];
                        xtc[i] = gc[i];
                    }
                    __syncthreads();
                    tab_dom = di;
                    tab_s = s;
                }
            }
            y = -0.0;
            int next_ck = r0c;
            if constexpr (REGC == 0) {
                for (int c0 = 0; c0 < nch; c0 += PC) {
#pragma unroll
                    for (int p = 0; p < PC; ++p) {
                        const int c = c0 + p;
                        if (c < nch) {
                            if (c < cl)
                                fwd_chunk(c, ring[p], min(c + PC, nch - 1), 0, no_store, exact_tag, akey, nullptr);
                            else
                                fwd_chunk(c, ring[p], min(c + PC, nch - 1), cl, store, exact_tag, akey, nullptr);
                            if (c + 1 == next_ck && next_ck < cl) {  // checkpoint at a segment end
                                put_ck<T>(dst, ck_off(next_ck), ck_off2(next_ck), y);
                                next_ck += Sc;
                            }
                        }
                    }
                }
            } else {
#pragma unroll
                for (int q = 0; q < REGC; ++q)  // register chunks: chunk q in ring slot q % PC
                    if (q < cr)
                        fwd_chunk(q, ring[q % PC], min(q + PC, nch - 1), 0, reg_store, exact_tag, akey, regs + q * W);
                if (cr > 0 && cr < cl)  // checkpoint of the first middle segment
                    put_ck<T>(dst, ck_off(cr), ck_off2(cr), y);
                for (int c0 = cr; c0 < cl; c0 += PC) {  // middle (cr, cl multiples of PC)
#pragma unroll
                    for (int p = 0; p < PC; ++p) {
                        const int c = c0 + p;
                        fwd_chunk(c, ring[p], min(c + PC, nch - 1), 0, no_store, exact_tag, akey, nullptr);
                        if (c + 1 == next_ck && next_ck < cl) {  // checkpoint at a segment end
                            put_ck<T>(dst, ck_off(next_ck), ck_off2(next_ck), y);
                            next_ck += Sc;
                        }
                    }
                }
                for (int c0 = cl; c0 < nch; c0 += PC) {  // line store (cl is a multiple of PC)
#pragma unroll
                    for (int p = 0; p < PC; ++p) {
                        const int c = c0 + p;
                        if (c < nch)
                            fwd_chunk(c, ring[p], min(c + PC, nch - 1), cl, store, exact_tag, akey, nullptr);
                    }
                }
            }
            return akey;
        };
        // Cell sources and sinks of this tile's rows, before the sweep reads them (the step's
        // agent update, fused): the occupied voxels of rows [r0, r0 + 64) are the segments
        // [xseg[tile], xseg[tile + 1]) of the voxel-sorted list.  Each is updated in place in
        // the input, r = (r + temp1[a]) / temp2[a] over its agents in list order, as the
        // separate agent kernel does (FP64 arithmetic, one rounding to T); the input is dead
        // after this sweep and no other tile reads these rows.  kAgentBatches batches of 64
        // segments are in flight at a time.
        if constexpr (AG) {
            if (D.xseg != nullptr) {
                const int tt = r0 / kWave;
                const int lo = __builtin_amdgcn_readfirstlane(D.xseg[tt]);
                const int hi = __builtin_amdgcn_readfirstlane(D.xseg[tt + 1]);
                T *f = D.field[parity] + (size_t)s * nvox;
                const double *a1 = D.t1 + (size_t)s * D.nact;
                const double *a2 = D.t2 + (size_t)s * D.nact;
                constexpr int NB = kAgentBatches;
                for (int k0 = lo; k0 < hi; k0 += NB * kWave) {
                    long long off[NB];
                    int b0[NB], b1[NB];
                    double r[NB];
#pragma unroll
                    for (int b = 0; b < NB; ++b) {
                        const int k = k0 + b * kWave + lane;
                        const bool ok = k < hi;
                        off[b] = ok ? D.seg_off[k] : -1;
                        b0[b] = ok ? D.seg_start[k] : 0;
                        b1[b] = ok ? D.seg_start[k + 1] : 0;
                    }
#pragma unroll
                    for (int b = 0; b < NB; ++b)
                        r[b] = off[b] >= 0 ? (double)f[off[b]] : 0.0;
#pragma unroll
                    for (int b = 0; b < NB; ++b)
                        for (int a = b0[b]; a < b1[b]; ++a) {
                            r[b] = r[b] + a1[a];
                            r[b] = r[b] / a2[a];
                        }
#pragma unroll
                    for (int b = 0; b < NB; ++b)
                        if (off[b] >= 0)
                            f[off[b]] = (T)r[b];
                }
                __threadfence_block();  // the updates are stored before the sweep loads them
            }
        }

        bool exact = false;
        const unsigned akey = forward(inexact);
        if constexpr (sizeof(T) == 8) {
            if (__builtin_expect(__any(akey >= kKeyBad), 0)) {
                exact = true;
                forward(exact_div);
            }
        } else {
            (void)akey;
        }

        // ---- backward: chunk c's forward results sit in the line store at c - base, or in
        // the registers rp (FROM_REGS)
        double xv = y;
        auto back_chunk = [&](int c, int base, auto regs_tag, const T *rp) __attribute__((always_inline)) {
            constexpr bool FROM_REGS = decltype(regs_tag)::value;
            const int w = min(W, nx - c * W);
            double cw[W], v[W];
#pragma unroll
            for (int j = 0; j < W; ++j) {
                cw[j] = TL ? xtc[min(c * W + j, nx - 1)] : cc[min(c * W + j, nx - 1)];
                if constexpr (FROM_REGS)
                    v[j] = (double)rp[j];
                else
                    v[j] = (double)ls[lane * Lp + (c - base) * W + j];
            }
            auto bstep = [&](int j) {
                const int k = c * W + j;
                if (k != nx - 1)
                    xv = fma(-cw[j], xv, v[j]);
                stg[lane * (W + 1) + j] = (whole || (act && (k == 0 || k == nx - 1))) ? dv : (T)xv;
            };
            if (FROM_REGS || w == W) {
#pragma unroll
                for (int j = W - 1; j >= 0; --j)
                    bstep(j);
            } else {
#pragma unroll
                for (int j = W - 1; j >= 0; --j)
                    if (j < w)
                        bstep(j);
            }
            if (vf) {
                if (c == 0 || c == nch - 1)
                    put_xfaces(c);
                if (whole) {  // a row on a y or z face: its values from that face's plane
                    const T *row = zz == 0 || zz == nz - 1
                                       ? D.dface + face_z(nx, ny, nz, D.nsub, zz == 0 ? 0 : 1, s, 0, yy)
                                       : D.dface + face_y(nx, ny, nz, D.nsub, yy == 0 ? 0 : 1, s, 0, zz);
                    for (int j = 0; j < w; ++j)
                        stg[lane * (W + 1) + j] = row[c * W + j];
                }
            }
            __syncthreads();
            store_chunk(c);
            __syncthreads();
        };
        // middle segments, last first: recompute from the checkpoint (re-reading the input),
        // then substitute back; the next segment's ring and checkpoint are loaded while the
        // current one substitutes back.  The register chunks (the head) come last.
        double yck = -0.0;
        auto seg_begin = [&](int cb) { return cb > r0c ? cb - Sc : cr; };
        auto prefetch_segment = [&](int cb) {
            const int ca = seg_begin(cb);
            yck = ca == 0 ? -0.0 : get_ck<T>(dst, ck_off(ca), ck_off2(ca));
#pragma unroll
            for (int p = 0; p < PC; ++p)
                load_chunk(src, min(ca + p, cb - 1), ring[p]);
        };
        const std::false_type from_ls;
        const std::true_type from_regs;
        if (cl > cr)
            prefetch_segment(cl);
        for (int c = nch - 1; c >= cl; --c)
            back_chunk(c, cl, from_ls, nullptr);
        for (int cb = cl; cb > cr;) {
            const int ca = seg_begin(cb);
            y = yck;
            unsigned unused = 0u;
            for (int c0 = ca; c0 < cb; c0 += PC) {
#pragma unroll
                for (int p = 0; p < PC; ++p) {
                    const int c = c0 + p;
                    if (c < cb) {
                        const int next = c + PC < cb ? c + PC : -1;
                        if (exact)
                            fwd_chunk(c, ring[p], next, ca, store, exact_div, unused, nullptr);
                        else
                            fwd_chunk(c, ring[p], next, ca, store, inexact, unused, nullptr);
                    }
                }
            }
            if (ca > cr)
                prefetch_segment(ca);
            for (int c = cb - 1; c >= ca; --c)
                back_chunk(c, ca, from_ls, nullptr);
            cb = ca;
        }
        if constexpr (REGC > 0) {
#pragma unroll
            for (int q = REGC - 1; q >= 0; --q)
                if (q < cr)
                    back_chunk(q, 0, from_regs, regs + q * W);
        }
    }
}

constexpr size_t kLdsBytes = 160 * 1024;

struct Shape {
    long long grid;  // 0 = one block per tile
    int bpc;         // sweep wavefronts sharing one CU's LDS
    int ring;        // y/z: rows in flight per lane; x: chunks in flight (0 = by bpc)
};

Shape shape_of(const Group *g, int axis)
{
    const LaunchShape &u = g->launch[axis];
    Shape sh;
    sh.grid = u.grid;
    // defaults measured on MI355X at 400^3 x 4 (with the head recomputation): four y/z sweep
    // wavefronts share a CU's LDS, six FP64 x sweep wavefronts (two per SIMD on some SIMDs),
    // four FP32 x sweep wavefronts (one per SIMD, holding register chunks, tables in LDS, a
    // strided grid), and y/z keep 32 rows per lane in flight
    sh.bpc = u.lds_blocks_per_cu > 0 ? u.lds_blocks_per_cu : (axis == 0 && g->elem == 8 ? 6 : 4);
    sh.ring = u.ring > 0 ? u.ring : (axis == 0 ? 0 : 32);  // x: the variant's chunks in flight
    return sh;
}

// whether every domain's rows (x: FP64 only, whose 16-column row segments are whole lines;
// y: rows of nx elements) or planes (z: nx * ny elements) are whole 128-byte lines, so that
// each result store of the sweep writes whole lines (non-temporal stores, store_result).  The
// FP32 y/z sweeps (256-byte row segments) also gain where rows are whole 64-byte halves of
// lines: FP32 y at 400^3 x 4 (1600-byte rows) 451 -> 432 us.
template <class T>
bool nt_rows(const Group *g, int axis)
{
    if (axis == 0 && sizeof(T) != 8)
        return false;
    const size_t unit = axis != 0 && sizeof(T) == 4 ? 64 : 128;
    for (const Domain &dm : g->doms) {
        const size_t row = axis == 2 ? (size_t)dm.nx * dm.ny : (size_t)dm.nx;
        if (row * sizeof(T) % unit != 0)
            return false;
    }
    return true;
}

// NT kernels at the default ring only (32 rows); other rings store plainly
template <class T, int AXIS, bool ST>
void launch_yz_st(const Group *g, const DomainDesc<T> *d, const long long *pre, int ndom, unsigned grid, size_t lds,
                  int ring, bool nt)
{
    if (ring <= 16)
        hipLaunchKernelGGL((k_sweep_yz<T, AXIS, 16, ST>), dim3(grid), dim3(kWave), lds, g->stream, d, pre, ndom, g->cur, (int)lds);
    else if (ring <= 32 && nt)
        hipLaunchKernelGGL((k_sweep_yz<T, AXIS, 32, ST, true>), dim3(grid), dim3(kWave), lds, g->stream, d, pre, ndom, g->cur, (int)lds);
    else if (ring <= 32)
        hipLaunchKernelGGL((k_sweep_yz<T, AXIS, 32, ST>), dim3(grid), dim3(kWave), lds, g->stream, d, pre, ndom, g->cur, (int)lds);
    else
        hipLaunchKernelGGL((k_sweep_yz<T, AXIS, 64, ST>), dim3(grid), dim3(kWave), lds, g->stream, d, pre, ndom, g->cur, (int)lds);
}

// whether a y/z launch over the group reads its tables with scalar loads
template <class T, int AXIS>
bool yz_scalar(const Group *g)
{
    if (!yz_scalar_ok<T, AXIS>())
        return false;
    int longest = 0;
    for (const Domain &dm : g->doms)
        longest = std::max(longest, AXIS == 1 ? dm.ny : dm.nz);
    return longest >= kYzScalarMinRows && (sizeof(T) == 4 || longest <= kYzScalarMaxRows);
}

template <class T, int AXIS>
void launch_yz(const Group *g, const DomainDesc<T> *d, const long long *pre, int ndom, unsigned grid, size_t lds,
               int ring)
{
    const bool nt = nt_rows<T>(g, AXIS);
    if constexpr (yz_scalar_ok<T, AXIS>()) {
        if (yz_scalar<T, AXIS>(g)) {
            launch_yz_st<T, AXIS, true>(g, d, pre, ndom, grid, lds, ring, nt);
            return;
        }
    }
    launch_yz_st<T, AXIS, false>(g, d, pre, ndom, grid, lds, ring, nt);
}

template <class T, bool VF, bool AG>
void launch_x_vf(const Group *g, const DomainDesc<T> *d, const long long *pre, int ndom, unsigned grid, size_t lds,
                 XVariant xv)
{
    if constexpr (sizeof(T) == 8) {
        (void)xv;
        if (nt_rows<T>(g, 0))
            hipLaunchKernelGGL((k_sweep_x<T, 2, 0, false, VF, AG, true>), dim3(grid), dim3(kWave), lds, g->stream, d, pre, ndom, g->cur, (int)lds);
        else
            hipLaunchKernelGGL((k_sweep_x<T, 2, 0, false, VF, AG>), dim3(grid), dim3(kWave), lds, g->stream, d, pre, ndom, g->cur, (int)lds);
    } else if (xv.pc == 1) {
        if (xv.tl)
            hipLaunchKernelGGL((k_sweep_x<T, 1, 17, true, VF, AG>), dim3(grid), dim3(kWave), lds, g->stream, d, pre, ndom, g->cur, (int)lds);
        else
            hipLaunchKernelGGL((k_sweep_x<T, 1, 17, false, VF, AG>), dim3(grid), dim3(kWave), lds, g->stream, d, pre, ndom, g->cur, (int)lds);
    } else {
        if (xv.tl)
            hipLaunchKernelGGL((k_sweep_x<T, 3, 0, true, VF, AG>), dim3(grid), dim3(kWave), lds, g->stream, d, pre, ndom, g->cur, (int)lds);
        else
            hipLaunchKernelGGL((k_sweep_x<T, 3, 0, false, VF, AG>), dim3(grid), dim3(kWave), lds, g->stream, d, pre, ndom, g->cur, (int)lds);
    }
}

template <class T>
void launch_x(const Group *g, const DomainDesc<T> *d, const long long *pre, int ndom, unsigned grid, size_t lds,
              XVariant xv, bool agents)
{
    bool vf = false;  // some domain has per-voxel Dirichlet values
    for (const Domain &dm : g->doms)
        vf = vf || !dm.dface.empty();
    if constexpr (sizeof(T) == 8) {  // the agent update runs in FP64 steps only
        if (agents) {
            if (vf)
                launch_x_vf<T, true, true>(g, d, pre, ndom, grid, lds, xv);
            else
                launch_x_vf<T, false, true>(g, d, pre, ndom, grid, lds, xv);
            return;
        }
    }
    if (vf)
        launch_x_vf<T, true, false>(g, d, pre, ndom, grid, lds, xv);
    else
        launch_x_vf<T, false, false>(g, d, pre, ndom, grid, lds, xv);
}

template <bool VF, bool AG>
hipError_t x_attributes(int bytes)
{
    const void *k[6] = {(const void *)k_sweep_x<double, 2, 0, false, VF, AG>, (const void *)k_sweep_x<float, 1, 17, false, VF, AG>,
                        (const void *)k_sweep_x<float, 1, 17, true, VF, AG>, (const void *)k_sweep_x<float, 3, 0, false, VF, AG>,
                        (const void *)k_sweep_x<float, 3, 0, true, VF, AG>,
                        (const void *)k_sweep_x<double, 2, 0, false, VF, AG, true>};
    hipError_t e = hipSuccess;
    for (int i = 0; i < 6 && e == hipSuccess; ++i)
        e = hipFuncSetAttribute(k[i], hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    return e;
}

template <class T>
int launch_sweep_t(Group *g, int axis, bool agents)
{
    int rc = sync_descriptors(g);
    if (rc)
        return rc;
    const int ndom = (int)g->doms.size();
    const Shape sh = shape_of(g, axis);
    XVariant xv = x_variant<T>(sh.bpc, axis == 0 ? sh.ring : 0);
    if (axis == 0 && xv.tl) {
        // tables in LDS only where every domain's tables fit beside the smallest line store
        constexpr int W = x_chunk<T>();
        for (const Domain &d : g->doms)
            if (16 * (size_t)d.nx + (size_t)kWave * (W + 1) * sizeof(T) +
                    (size_t)kWave * ((size_t)xv.pc * W + 1) * sizeof(T) > g->max_lds_per_block)
                xv.tl = false;
    }
    size_t budget = std::min(kLdsBytes / sh.bpc, g->max_lds_per_block);
    // LDS actually needed by the largest domain, capped by the budget; the budget is raised
    // (fewer wavefronts per CU) where a line store would hold less than one 16-row ring or
    // one 16-column chunk
    size_t need = 0;
    for (const Domain &d : g->doms) {
        size_t full, floor;
        if (axis == 0) {
            constexpr int W = x_chunk<T>();
            const size_t stg = (size_t)kWave * (W + 1) * sizeof(T) + (xv.tl ? 16 * (size_t)d.nx : 0);
            full = stg + (size_t)kWave * ((d.nx + W - 1) / W * W + 1) * sizeof(T);
            // one chunk, or one ring of chunks with register chunks
            const size_t chunks = xv.regc > 0 ? xv.pc : 1;
            floor = std::min(full, stg + (size_t)kWave * (chunks * W + 1) * sizeof(T));
        } else {
            const int n = axis == 1 ? d.ny : d.nz;
            const size_t tb = (size_t)yz_table_bytes(axis == 1 ? yz_scalar<T, 1>(g) : yz_scalar<T, 2>(g));
            full = (size_t)n * tb + (size_t)kWave * n * sizeof(T);
            floor = std::min(full, (size_t)n * tb + (size_t)kWave * 16 * sizeof(T));
        }
        if (floor > g->max_lds_per_block)
            return set_error(g, BFVMB_ERR_ARG, "sweep: axis too long for the LDS line store");
        budget = std::max(budget, floor);
        need = std::max(need, full);
    }
    const size_t lds = std::min(budget, need);
    int ring = sh.ring;
    if (axis == 0) {
        constexpr int W = x_chunk<T>();
        for (const Domain &d : g->doms) {
            const int L = x_line_store<T>(d.nx, (long long)lds, xv.tl);
            if (L < W || (xv.regc > 0 && (d.nx + W - 1) / W > L / W && L / W < xv.pc))
                return set_error(g, BFVMB_ERR_ARG, "sweep: LDS budget too small for the x line store");
        }
    } else {
        // a line longer than its line store is recomputed in segments of whole rings, so the
        // store must hold at least one ring: take the deepest ring (<= the requested one)
        // that fits every such domain
        int minL = 1 << 30;
        for (const Domain &d : g->doms) {
            const int n = axis == 1 ? d.ny : d.nz;
            const int L = yz_line_store<T>(n, (long long)lds, axis == 1 ? yz_scalar<T, 1>(g) : yz_scalar<T, 2>(g));
            if (L < n)
                minL = std::min(minL, L);
        }
        ring = ring >= 64 ? 64 : ring >= 32 ? 32 : 16;
        while (ring > 16 && ring > minL)
            ring /= 2;
        if (minL < ring)
            return set_error(g, BFVMB_ERR_ARG, "sweep: LDS budget too small for the y/z line store");
    }
    const long long *pre = g->d_prefix + (size_t)(axis + 1) * (ndom + 1);
    const long long ntiles = g->prefix[axis + 1][ndom];
    if (ntiles == 0)
        return BFVMB_OK;
    // default grid: one block per tile, except for the FP32 x variants with register chunks
    // (one wavefront per SIMD) and the FP32 z sweep: one block per SIMD striding over the
    // tiles (measured at 400^3 x 4: x 0.651 -> 0.620 ms with the tables in LDS, z 0.504 ->
    // 0.489 ms; FP32 y and the FP64 sweeps are faster with one block per tile)
    long long grid = ntiles;
    if (sh.grid > 0)
        grid = std::min<long long>(sh.grid, ntiles);
    else if ((axis == 0 && xv.regc > 0) || (axis == 2 && sizeof(T) == 4))
        grid = std::min<long long>(ntiles, (long long)std::max(g->cu_count, 1) * std::min(sh.bpc, 4));
    grid = std::min<long long>(grid, 1LL << 30);
    const DomainDesc<T> *d = static_cast<const DomainDesc<T> *>(g->d_desc);
    if (axis == 0)
        launch_x<T>(g, d, pre, ndom, (unsigned)grid, lds, xv, agents);
    else if (axis == 1)
        launch_yz<T, 1>(g, d, pre, ndom, (unsigned)grid, lds, ring);
    else
        launch_yz<T, 2>(g, d, pre, ndom, (unsigned)grid, lds, ring);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return hip_check(g, e, "sweep launch");
    g->cur ^= 1;
    return BFVMB_OK;
}

template <class T, int AXIS, bool ST>
hipError_t yz_attributes_st(int bytes)
{
    hipError_t e = hipFuncSetAttribute((const void *)k_sweep_yz<T, AXIS, 16, ST>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_sweep_yz<T, AXIS, 32, ST>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_sweep_yz<T, AXIS, 64, ST>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_sweep_yz<T, AXIS, 32, ST, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                bytes);
    return e;
}

template <class T, int AXIS>
hipError_t yz_attributes(int bytes)
{
    hipError_t e = yz_attributes_st<T, AXIS, false>(bytes);
    if constexpr (yz_scalar_ok<T, AXIS>())
        if (e == hipSuccess)
            e = yz_attributes_st<T, AXIS, true>(bytes);
    return e;
}

}  // namespace

int launch_sweep(Group *g, int axis)
{
    if (g->precision == BFVMB_FP64)
        return launch_sweep_t<double>(g, axis, false);
    return launch_sweep_t<float>(g, axis, false);
}

int launch_sweep_x_agents(Group *g)
{
    if (g->precision != BFVMB_FP64)
        return set_error(g, BFVMB_ERR_STATE, "the fused agent update is an FP64 kernel");
    return launch_sweep_t<double>(g, 0, true);
}

int sweep_lds_attributes(size_t lds_bytes)
{
    // allow the device's full LDS for every sweep instantiation
    const int bytes = (int)lds_bytes;
    hipError_t e = yz_attributes<double, 1>(bytes);
    if (e == hipSuccess)
        e = yz_attributes<double, 2>(bytes);
    if (e == hipSuccess)
        e = yz_attributes<float, 1>(bytes);
    if (e == hipSuccess)
        e = yz_attributes<float, 2>(bytes);
    if (e == hipSuccess)
        e = x_attributes<false, false>(bytes);
    if (e == hipSuccess)
        e = x_attributes<true, false>(bytes);
    if (e == hipSuccess)  // FP64 x sweep with the agent update
        e = hipFuncSetAttribute((const void *)k_sweep_x<double, 2, 0, false, false, true>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_sweep_x<double, 2, 0, false, true, true>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_sweep_x<double, 2, 0, false, false, true, true>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_sweep_x<double, 2, 0, false, true, true, true>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    return e == hipSuccess ? 0 : (int)e;
}

}  // namespace bfvmb
