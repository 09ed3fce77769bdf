/*
 * Oracle (test infrastructure): special functions of the GRPP path.
 *
 * The value tables that libgrpp ships as generated source (bessel_table_0[1601][60],
 * boys_table[301][40], rfun_coeffs[50][50]) are NOT copied: they are regenerated here
 * from their defining series in 80-bit long double, then used with the same
 * interpolation schemes as the reference.
 */
#include <math.h>
#include <string.h>
#include "orc.h"

#define BES_NPT 1601   /* x = 0, 0.01, ..., 16 */
#define BES_NLAM 60
#define BES_NDER 5
#define BOYS_NPT 301   /* x = 0, 0.1, ..., 30 */
#define BOYS_NN 40

static double bes_tab[BES_NPT][BES_NLAM][BES_NDER];
static double boys_tab[BOYS_NPT][BOYS_NN];
static double rfun[BES_NLAM][BES_NLAM];
static int orc_initialized = 0;

/*
 * K_lambda(x) = exp(-x) i_lambda(x), i_lambda the modified spherical Bessel function of the
 * first kind, from its power series
 *   i_l(x) = x^l/(2l+1)!! * sum_k (x^2/2)^k / (k! (2l+3)(2l+5)...(2l+2k+1)),
 * all terms positive, so the long-double sum is accurate to ~1e-18 relative.
 */
static long double k_scaled_series(int l, long double x)
{
    if (x == 0.0L) return l == 0 ? 1.0L : 0.0L;
    long double pref = 1.0L;
    for (int j = 1; j <= l; j++) pref *= x / (2 * j + 1);
    long double h = 0.5L * x * x, t = 1.0L, s = 1.0L;
    for (int k = 1; k < 2000; k++) {
        t *= h / ((long double) k * (2 * l + 2 * k + 1));
        s += t;
        if (t < 1e-22L * s) break;
    }
    return expl(-x) * pref * s;
}

/*
 * Boys function F_n(x) = int_0^1 t^{2n} exp(-x t^2) dt from the positive series
 *   F_n(x) = exp(-x) sum_k (2x)^k / ((2n+1)(2n+3)...(2n+2k+1)).
 */
static long double boys_series(int n, long double x)
{
    long double t = 1.0L / (2 * n + 1), s = t;
    for (int k = 1; k < 5000; k++) {
        t *= 2.0L * x / (2 * n + 2 * k + 1);
        s += t;
        if (t < 1e-22L * s) break;
    }
    return expl(-x) * s;
}

/*
 * Derivatives of K_lambda from the recursion
 *   K^(k)_l = l/(2l+1) K^(k-1)_{l-1} + (l+1)/(2l+1) K^(k-1)_{l+1} - K^(k-1)_l
 * evaluated in the same order as the reference (specfunc_scaled_mod_sph_bessel.c:45-61,135-156),
 * orders l <= 54 only.
 */
static void build_bessel_table(void)
{
    for (int i = 0; i < BES_NPT; i++) {
        long double x = (long double) i / 100.0L;
        for (int l = 0; l < BES_NLAM; l++) {
            bes_tab[i][l][0] = (double) k_scaled_series(l, x);
            for (int k = 1; k < BES_NDER; k++) bes_tab[i][l][k] = 0.0;
        }
        for (int k = 1; k < BES_NDER; k++) {
            for (int l = 0; l <= BES_NLAM - 1 - 5; l++) {
                double d = 0.0;
                if (l > 0) d += l / (2.0 * l + 1.0) * bes_tab[i][l - 1][k - 1];
                d += (l + 1.0) / (2.0 * l + 1.0) * bes_tab[i][l + 1][k - 1];
                d -= bes_tab[i][l][k - 1];
                bes_tab[i][l][k] = d;
            }
        }
    }
    /* asymptotic coefficients (k+n)!/(k!(n-k)! 2^k) (specfunc_scaled_mod_sph_bessel.c:168-175) */
    for (int n = 0; n < BES_NLAM; n++)
        for (int k = 0; k < BES_NLAM; k++) {
            if (k > n) { rfun[n][k] = 0.0; continue; }
            long double v = 1.0L;
            for (int j = 1; j <= n + k; j++) v *= j;
            for (int j = 1; j <= k; j++) v /= j;
            for (int j = 1; j <= n - k; j++) v /= j;
            for (int j = 0; j < k; j++) v /= 2.0L;
            rfun[n][k] = (double) v;
        }
}

static void build_boys_table(void)
{
    for (int i = 0; i < BOYS_NPT; i++) {
        long double x = (long double) i / 10.0L;
        for (int n = 0; n < BOYS_NN; n++) boys_tab[i][n] = (double) boys_series(n, x);
    }
}

void orc_init(void)
{
#pragma omp critical(orc_init_lock)
    {
        if (!orc_initialized) {
            build_bessel_table();
            build_boys_table();
            orc_initialized = 1;
        }
    }
}

double orc_bessel_table_value(int i, int lambda, int k) { orc_init(); return bes_tab[i][lambda][k]; }
double orc_boys_table_value(int i, int n) { orc_init(); return boys_tab[i][n]; }

/* K_n(x) = exp(-x) i_n(x) (specfunc_scaled_mod_sph_bessel.c:67-122):
 * tiny x: first series term; x <= 16: 4th-order Taylor about the nearest 0.01 node;
 * x > 16: R_n(-x)/(2x) with the asymptotic polynomial (:176-190). */
double orc_bessel_scaled(int n, double x)
{
    if (x <= 1e-7) {
        switch (n) {
            case 0: return 1.0 - x;
            case 1: return (1.0 - x) * x * (1.0 / 3.0);
            case 2: return (1.0 - x) * x * x * (1.0 / 15.0);
            default: return 0.0;
        }
    }
    if (x <= 16.0) {
        int i0 = (int) (round(x * 100) + 1e-5);
        double dx = x - i0 * 0.01;
        const double *d = bes_tab[i0][n];
        double dx2 = dx * dx, dx3 = dx2 * dx, dx4 = dx3 * dx;
        return d[0] + d[1] * dx + d[2] * dx2 * 0.5 + d[3] * dx3 * (1.0 / 6.0) + d[4] * dx4 * (1.0 / 24.0);
    }
    double xm = -x, s = 0.0, den = 1.0, inv = 1.0 / xm;
    for (int k = 0; k <= n; k++) {
        s += rfun[n][k] * den;
        den *= inv;
    }
    return s / (2.0 * x);
}

/* F_n(x): 6th-order Taylor about the nearest 0.1 node for x <= 30, asymptotic beyond
 * (specfunc_boys.c:21-46). */
double orc_boys(int n, double x)
{
    if (x <= 30.0) {
        double step = 0.1;
        int idx = (int) (round(x / step) + 1e-6);
        double dx = x - idx * step;
        double dx2 = dx * dx, dx3 = dx * dx2, dx4 = dx * dx3, dx5 = dx * dx4, dx6 = dx * dx5;
        const double *b = boys_tab[idx];
        return +b[n] - b[n + 1] * dx + b[n + 2] * dx2 / 2.0 - b[n + 3] * dx3 / 6.0 + b[n + 4] * dx4 / 24.0
               - b[n + 5] * dx5 / 120.0 + b[n + 6] * dx6 / 720.0;
    }
    return orc_double_factorial(2 * n - 1) / pow(2.0, n + 1) * sqrt(M_PI / pow(x, 2 * n + 1));
}

/*
 * Dawson's integral D(x) = exp(-x^2) int_0^x exp(t^2) dt by Rybicki's sampling-theorem
 * formula (G. B. Rybicki, Computers in Physics 3, 85 (1989)):
 *   D(x) = pi^{-1/2} sum_{n odd} exp(-(x - n h)^2) / n,
 * whose error is ~exp(-(pi/(2h))^2); h = 0.2 gives ~1e-27.  Long double.
 * (The reference evaluates the same function with MyMathLib Chebyshev pieces,
 * specfunc_dawson.c:73-113; both are accurate to double rounding.)
 */
double orc_dawson(double xd)
{
    long double x = fabsl((long double) xd);
    if (x == 0.0L) return 0.0;
    const long double h = 0.2L;
    long double s = 0.0L;
    long long nc = (long long) floorl(x / h);
    for (long long n = nc - 80; n <= nc + 80; n++) {
        if ((n & 1LL) == 0) continue;
        long double d = x - n * h;
        s += expl(-d * d) / (long double) n;
    }
    long double v = s / sqrtl(3.14159265358979323846264338327950288L);
    return (double) (xd < 0 ? -v : v);
}

/* G_n(x) = int_0^1 t^{2n} exp(-x(1-t^2)) dt for n = 0..nmax (specfunc_gfun.c:40-93):
 * x <= 12: the reference's truncated Taylor series for G_nmax, then downward recursion;
 * x > 12: G_0 = D(sqrt x)/sqrt x and upward recursion. */
void orc_gfun_values(double x, int nmax, double *g)
{
    memset(g, 0, (nmax + 1) * sizeof(double));
    if (x <= 12.0) {
        double s = 0.0;
        for (int k = 0; k < 100; k++) {
            double c = exp(-x) * pow(x, k) / orc_factorial(k) / (2.0 * nmax + 2.0 * k + 1.0);
            s += c;
            if (fabs(c) < 1e-15) break;
        }
        g[nmax] = s;
        for (int n = nmax; n > 0; n--) g[n - 1] = (1.0 - 2.0 * x * g[n]) / (2.0 * n - 1.0);
    } else {
        double sx = sqrt(x);
        g[0] = orc_dawson(sx) / sx;
        for (int n = 0; n < nmax; n++) g[n + 1] = (1.0 - (2 * n + 1) * g[n]) / (2.0 * x);
    }
}
