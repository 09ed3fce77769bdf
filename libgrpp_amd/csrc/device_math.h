// Device special functions of the GRPP engine (FP64).
#pragma once
#include <hip/hip_runtime.h>

#include "tables.h"

namespace grpp {

struct DevTables {
    const double *r, *w, *rpow, *bes, *rfun, *boys;
    const double *bes_t;      // bes in order-major layout [l][node][5] (grid tabulation)
    const double *besp;       // Taylor coefficients over k! (Horner form) of K_l and K_{l-1} per node and l,
                              // 80 bytes each: steering-only evaluations
    const double *rsh;
    const int *rsh_off, *comb, *comb_off;
    const double *lmat;
    const int *lmat_off;
    const double *slm;
    const int *slm_off;
    const double *W;          // angular coupling blocks
    const int *wdir;          // [(L*(nmax+1)+n)] -> offset, -1 if absent
    const int *wstride;       // mu stride per (L, n)
    int w_nmax;
    const double *w1;         // type-1 angular couplings of the numeric nuclear-attraction path
};

// K_n(x) = exp(-x) i_n(x): series head for tiny x, 4th-order Taylor about the nearest 0.01
// node for x <= 16, asymptotic polynomial beyond (specfunc_scaled_mod_sph_bessel.c:67-122).
__device__ __forceinline__ double bessel_k(const double *__restrict__ bes, const double *__restrict__ rfun, int n,
                                           double x)
{
    if (x <= 1e-7) {
        if (n == 0) return 1.0 - x;
        if (n == 1) return (1.0 - x) * x * (1.0 / 3.0);
        if (n == 2) return (1.0 - x) * x * x * (1.0 / 15.0);
        return 0.0;
    }
    if (x <= 16.0) {
        int i0 = (int) (rint(x * 100) + 1e-5);
        double dx = x - i0 * 0.01;
        const double *d = bes + ((size_t) i0 * BES_NL + n) * 5;
        double dx2 = dx * dx, dx3 = dx2 * dx, dx4 = dx3 * dx;
        return d[0] + d[1] * dx + d[2] * dx2 * 0.5 + d[3] * dx3 * (1.0 / 6.0) + d[4] * dx4 * (1.0 / 24.0);
    }
    double inv = -1.0 / x, s = 0.0, den = 1.0;
    const double *c = rfun + n * BES_NL;
    for (int k = 0; k <= n; k++) {
        s += c[k] * den;
        den *= inv;
    }
    return s / (2.0 * x);
}

// F_n(x): 6th-order Taylor about the nearest 0.1 node (x <= 30), asymptotic beyond (specfunc_boys.c:21-46)
__device__ __forceinline__ double boys_f(const double *__restrict__ boys, int n, double x)
{
    if (x <= 30.0) {
        int idx = (int) (rint(x / 0.1) + 1e-6);
        double dx = x - idx * 0.1;
        double dx2 = dx * dx, dx3 = dx * dx2, dx4 = dx * dx3, dx5 = dx * dx4, dx6 = dx * dx5;
        const double *b = boys + idx * BOYS_NN;
        return +b[n] - b[n + 1] * dx + b[n + 2] * dx2 / 2.0 - b[n + 3] * dx3 / 6.0 + b[n + 4] * dx4 / 24.0 -
               b[n + 5] * dx5 / 120.0 + b[n + 6] * dx6 / 720.0;
    }
    double df = 1.0;
    for (int k = 2 * n - 1; k > 1; k -= 2) df *= k;
    return df / pow(2.0, n + 1) * sqrt(M_PI / pow(x, 2 * n + 1));
}

// Dawson's integral by Rybicki's sampling formula D(x) = pi^{-1/2} sum_{n odd} e^{-(x-nh)^2}/n,
// h = 0.2 (G. B. Rybicki, Computers in Physics 3, 85 (1989)).
__device__ inline double dawson(double x)
{
    const double h = 0.2;
    double ax = fabs(x), s = 0.0;
    long nc = (long) floor(ax / h);
    for (long n = nc - 40; n <= nc + 40; n++) {
        if ((n & 1L) == 0) continue;
        double d = ax - n * h;
        s += exp(-d * d) / (double) n;
    }
    double v = s * 0.56418958354775628695;  // 1/sqrt(pi)
    return x < 0 ? -v : v;
}

// G_n(x) = int_0^1 t^{2n} e^{-x(1-t^2)} dt, n = 0..nmax (specfunc_gfun.c:40-93)
__device__ inline void gfun_values(double x, int nmax, double *g)
{
    if (x <= 12.0) {
        double s = 0.0, fk = 1.0;
        for (int k = 0; k < 100; k++) {
            if (k > 0) fk *= k;
            double c = exp(-x) * pow(x, k) / fk / (2.0 * nmax + 2.0 * k + 1.0);
            s += c;
            if (fabs(c) < 1e-15) break;
        }
        g[nmax] = s;
        for (int n = nmax; n > 0; n--) g[n - 1] = (1.0 - 2.0 * x * g[n]) / (2.0 * n - 1.0);
    } else {
        double sx = sqrt(x);
        g[0] = dawson(sx) / sx;
        for (int n = 0; n < nmax; n++) g[n + 1] = (1.0 - (2 * n + 1) * g[n]) / (2.0 * x);
    }
}

}  // namespace grpp
