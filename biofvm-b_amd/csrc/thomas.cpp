// thomas.cpp -- host FP64 set-up of the LOD solver's Thomas tables.
//
// Follows the first-call block of diffusion_decay_solver__constant_coefficients_LOD_3D
// (BioFVM_parallel.cpp:714-918), whose FP64 denominator recursion both storage paths
// keep: its values must be exact for the FP64 path to match the reference bit for bit,
// and the FP32 path only rounds the finished tables.  Built with -ffp-contract=off; the
// one fused operation of the reference (axpy, BioFVM_vector.cpp:336-352, contracted by
// g++ -O3) is written as std::fma.
#include <cmath>

#include "bfvmb_internal.hpp"

namespace bfvmb {

// c1 = ((D*dt)/dx)/dx, c2 = (decay*dt)/3, c3 = ((1+c1)+c1)+c2, c3a = (1+c1)+c2
// (BioFVM_parallel.cpp:780-801).
void lod_constants(int nsub, const double *D, const double *decay, double dt, double dx,
                   double *c1, double *c2, double *c3, double *c3a)
{
    for (int d = 0; d < nsub; ++d) {
        double a = D[d] * dt;
        a = a / dx;
        a = a / dx;
        double b = decay[d] * dt;
        b = b / 3.0;
        c1[d] = a;
        c2[d] = b;
        c3[d] = ((1.0 + a) + a) + b;
        c3a[d] = (1.0 + a) + b;
    }
}

// One axis, tables [nsub][n] (BioFVM_parallel.cpp:882-896; x at world size 1 :810-851):
// c = -c1, den = c3 with c3a at both ends (1+c2 when n == 1); c[0] /= den[0];
// den[i] = fma(c1, c[i-1], den[i]); c[i] /= den[i].
void thomas_axis(int n, int nsub, const double *c1, const double *c2, const double *c3,
                 const double *c3a, double *den, double *cc)
{
    for (int d = 0; d < nsub; ++d) {
        double *dd = den + (size_t)d * n;
        double *c = cc + (size_t)d * n;
        for (int i = 0; i < n; ++i) {
            c[i] = c1[d] * -1.0;
            dd[i] = c3[d];
        }
        dd[0] = c3a[d];
        dd[n - 1] = c3a[d];
        if (n == 1)
            dd[0] = 1.0 + c2[d];
        c[0] = c[0] / dd[0];
        for (int i = 1; i < n; ++i) {
            dd[i] = std::fma(c1[d], c[i - 1], dd[i]);
            c[i] = c[i] / dd[i];
        }
    }
}

}  // namespace bfvmb
