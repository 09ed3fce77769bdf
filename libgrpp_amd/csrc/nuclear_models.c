/*
 * Nuclear charge distribution models (host code): the estimates, densities, radii and electrostatic
 * potentials libgrpp exports from nuclear_models.c (nuclear_models.h:13-39), and the potential of a
 * model tabulated on the quadrature grid for the numeric nuclear-attraction integrals.
 *
 * The Fermi models need S_k(x) = sum_{n>=1} (-1)^n e^{nx} / n^k (specfunc_fermi_sk.c:42-90): a
 * cubic Taylor expansion about the nodes x0 = -10 + 0.005 i of [-10, 0] (dS_k/dx = S_{k-1}), and
 * one to three exponential terms below -10.  The node values S_{-2} ... S_7 are computed here once:
 * S_0, S_-1, S_-2 in closed form from the logistic function, S_1 ... S_7 by the Cohen-Rodriguez
 * Villegas-Zagier acceleration of the alternating series (40 terms, long double).
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>

#include "engine.h"
#include "libgrpp.h"

#define FSK_NPT 2001
static double fsk[FSK_NPT][10];   /* column j holds S_{j-2}(x0) */
static pthread_once_t fsk_once = PTHREAD_ONCE_INIT;

static long double fsk_series(int m, long double x)
{
    const int N = 40;
    const long double z = expl(x);
    long double d = powl(3.0L + sqrtl(8.0L), N);
    d = (d + 1.0L / d) / 2.0L;
    long double b = -1.0L, c = -d, s = 0.0L, zn = 1.0L;
    for (int k = 0; k < N; k++) {
        zn *= z;
        c = b - c;
        s += c * zn / powl((long double) (k + 1), m);
        b = (k + N) * (long double) (k - N) * b / ((k + 0.5L) * (k + 1.0L));
    }
    return -s / d;
}

static void fsk_build(void)
{
    for (int i = 0; i < FSK_NPT; i++) {
        const long double x = -10.0L + 0.005L * i, e = expl(x), sg = e / (1.0L + e);
        fsk[i][2] = (double) (-sg);
        fsk[i][1] = (double) (-sg * (1.0L - sg));
        fsk[i][0] = (double) (-sg * (1.0L - sg) * (1.0L - 2.0L * sg));
        for (int m = 1; m <= 7; m++) fsk[i][m + 2] = (double) fsk_series(m, x);
    }
}

double libgrpp_amd_fermi_sk(int k, double x)
{
    pthread_once(&fsk_once, fsk_build);
    if (x <= -30) return -1.0 * exp(x);
    if (x <= -14) return -1.0 * exp(x) + 1.0 * exp(2 * x) / pow(2, k);
    if (x <= -10) return -1.0 * exp(x) + 1.0 * exp(2 * x) / pow(2, k) - 1.0 * exp(3 * x) / pow(3, k);
    const int i0 = (int) (floor((x + 10.0) * 200) + 1e-10);
    const double x0 = -10.0 + i0 * 0.005, dx = x - x0;
    if (i0 == 2000) return fsk[i0][k + 2];
    const double *d = fsk[i0];
    const double dx2 = dx * dx, dx3 = dx2 * dx;
    return d[k + 2] + d[k + 1] * dx + d[k] * dx2 * 0.5 + d[k - 1] * dx3 * (1.0 / 6.0);
}

double libgrpp_amd_fermi_sk_node(int i, int col)
{
    pthread_once(&fsk_once, fsk_build);
    return fsk[i][col];
}

#define SK libgrpp_amd_fermi_sk

/* ---------------------------------------------------------------- estimates (fm) */
double libgrpp_estimate_nuclear_rms_radius_johnson_1985(int A) { return 0.836 * pow(A, 1.0 / 3.0) + 0.570; }
double libgrpp_estimate_nuclear_rms_radius_golovko_2008(int A) { return 0.77 * pow(A, 1.0 / 3.0) + 0.98; }

int libgrpp_estimate_fermi_model_parameters(double R_rms, double *c, double *a)
{
    const double t = 2.3;
    *a = t / (4 * log(3));
    const double c2 = 5.0 / 3.0 * R_rms * R_rms - 7.0 / 3.0 * M_PI * M_PI * (*a) * (*a);
    if (c2 < 0) return -1;
    *c = sqrt(c2);
    return 1;
}

/* ---------------------------------------------------------------- normalisation factors */
static double fermi_norm(double c, double a)
{
    const double a2 = a * a, a3 = a2 * a, c2 = c * c, c3 = c2 * c;
    return 1.0 + M_PI * M_PI * a2 / c2 - 6.0 * a3 / c3 * SK(3, -c / a);
}

static double bubble_norm(double c, double a, double k)
{
    const double a2 = a * a, a4 = a2 * a2, a5 = a4 * a, c2 = c * c, c4 = c2 * c2, c5 = c4 * c;
    return fermi_norm(c, a) + 3.0 / 5.0 * k + 2.0 * M_PI * M_PI * a2 * k / c2 +
           7.0 * M_PI * M_PI * M_PI * M_PI * a4 * k / (5.0 * c4) - 72.0 * a5 * k / c5 * SK(5, -c / a);
}

/* ---------------------------------------------------------------- densities and radii */
double libgrpp_charge_density_ball(double r, double Z, double R_rms)
{
    const double R0 = sqrt(5.0 / 3.0) * R_rms;
    return r <= R0 ? 3.0 * Z / (4.0 * M_PI * R0 * R0 * R0) : 0.0;
}

double libgrpp_charge_density_gaussian(double r, double Z, double R_rms)
{
    const double xi = 3.0 / (2.0 * R_rms * R_rms);
    return Z * pow(xi / M_PI, 1.5) * exp(-xi * r * r);
}

double libgrpp_charge_density_fermi(double r, double Z, double c, double a)
{
    const double rho0 = 3.0 * Z / (4 * M_PI * c * c * c * fermi_norm(c, a));
    return rho0 / (1.0 + exp((r - c) / a));
}

double libgrpp_charge_density_fermi_bubble(double r, double Z, double c, double a, double k)
{
    const double rho0 = 3.0 * Z / (4 * M_PI * c * c * c * bubble_norm(c, a, k));
    return rho0 * (1 + k * pow(r / c, 2)) / (1.0 + exp((r - c) / a));
}

double libgrpp_rms_radius_fermi(int Z, double c, double a)
{
    const double rho0 = 3.0 * Z / (4 * M_PI * c * c * c * fermi_norm(c, a));
    const double r2 = 4 * M_PI * rho0 / Z * pow(c, 5) / 5.0 *
                      (1.0 + 10.0 / 3.0 * a * a * M_PI * M_PI / (c * c) + 7.0 / 3.0 * pow(M_PI, 4) * pow(a, 4) / pow(c, 4) -
                       120.0 * pow(a, 5) / pow(c, 5) * SK(5, -c / a));
    return sqrt(r2);
}

double libgrpp_rms_radius_fermi_bubble(int Z, double c, double a, double k)
{
    const double rho0 = 3.0 * Z / (4 * M_PI * c * c * c * bubble_norm(c, a, k));
    const double r4 = pow(c, 5) / 5.0 *
                      (1.0 + 10.0 / 3.0 * a * a * M_PI * M_PI / (c * c) + 7.0 / 3.0 * pow(M_PI, 4) * pow(a, 4) / pow(c, 4) -
                       120.0 * pow(a, 5) / pow(c, 5) * SK(5, -c / a));
    const double r6 = pow(c, 7) / 7.0 *
                      (1.0 + 7.0 * a * a * M_PI * M_PI / (c * c) + 49.0 / 3.0 * pow(M_PI, 4) * pow(a, 4) / pow(c, 4) +
                       31.0 / 3.0 * pow(M_PI, 6) * pow(a, 6) / pow(c, 6) - 5040.0 * pow(a, 7) / pow(c, 7) * SK(7, -c / a));
    return sqrt(4 * M_PI * rho0 / Z * (r4 + k / (c * c) * r6));
}

/* ---------------------------------------------------------------- potentials */
double libgrpp_coulomb_potential_point(double r, double Z) { return -Z / r; }

double libgrpp_coulomb_potential_ball(double r, double Z, double R_rms)
{
    const double R0 = sqrt(5.0 / 3.0) * R_rms;
    if (r <= R0) return -Z / (2.0 * R0) * (3.0 - r * r / (R0 * R0));
    return -Z / r;
}

double libgrpp_coulomb_potential_gaussian(double r, double Z, double R_rms)
{
    const double xi = 3.0 / (2.0 * R_rms * R_rms);
    return -Z / r * erf(sqrt(xi) * r);
}

double libgrpp_coulomb_potential_fermi(double r, double Z, double c, double a)
{
    const double a2 = a * a, a3 = a * a * a, c3 = c * c * c, N = fermi_norm(c, a);
    if (r > c) return -Z / (N * r) * (N + 3 * a2 * r / c3 * SK(2, (c - r) / a) + 6 * a3 / c3 * SK(3, (c - r) / a));
    const double P2 = SK(2, (r - c) / a), P3 = SK(3, (r - c) / a), S3 = SK(3, -c / a), r3 = r * r * r;
    return -Z / (N * r) *
           (1.5 * r / c - r3 / (2 * c3) + M_PI * M_PI * a2 * r / (2 * c3) - 3 * a2 * r / c3 * P2 + 6 * a3 / c3 * (P3 - S3));
}

double libgrpp_coulomb_potential_fermi_bubble(double r, double Z, double c, double a, double k)
{
    const double Nk = bubble_norm(c, a, k);
    double F0, F2;
    if (r < c) {
        const double S2 = SK(2, (r - c) / a), S3 = SK(3, (r - c) / a), S4 = SK(4, (r - c) / a), S5 = SK(5, (r - c) / a);
        F0 = -pow(r, 3) / 6.0 - r * a * a * S2 + 2.0 * pow(a, 3) * S3 + r * c * c / 2.0 + M_PI * M_PI / 6.0 * r * a * a -
             2.0 * pow(a, 3) * SK(3, -c / a);
        F2 = -pow(r, 5) / 20.0 - pow(r, 3) * pow(a, 2) * S2 + 6.0 * pow(r, 2) * pow(a, 3) * S3 - 18.0 * r * pow(a, 4) * S4 +
             24.0 * pow(a, 5) * S5 + r * pow(c, 4) / 4.0 + r * M_PI * M_PI * c * c * a * a / 2.0 +
             r * pow(a, 4) * pow(M_PI, 4) * 7.0 / 60.0 - 24.0 * pow(a, 5) * SK(5, -c / a);
    } else {
        const double S2 = SK(2, (c - r) / a), S3 = SK(3, (c - r) / a), S4 = SK(4, (c - r) / a), S5 = SK(5, (c - r) / a);
        F0 = pow(c, 3) / 3.0 + M_PI * M_PI / 3.0 * c * a * a - 2.0 * pow(a, 3) * SK(3, -c / a) + r * a * a * S2 +
             2.0 * pow(a, 3) * S3;
        F2 = pow(c, 5) / 5.0 + 2.0 * pow(c, 3) * a * a * M_PI * M_PI / 3.0 + 7.0 * pow(a, 4) * c * pow(M_PI, 4) / 15.0 -
             24.0 * pow(a, 5) * SK(5, -c / a) + pow(a, 2) * pow(r, 3) * S2 + 6.0 * pow(a, 3) * pow(r, 2) * S3 +
             18.0 * r * pow(a, 4) * S4 + 24.0 * pow(a, 5) * S5;
    }
    return -Z / (Nk * r) * 3.0 / pow(c, 3) * (F0 + k / (c * c) * F2);
}

/* the potential of model (1 ball, 2 Gaussian, 3 Fermi, 4 Fermi + bubble, 5 point) for charge Z and
 * model parameters p at the n radii r (nuclear_attraction.c:86-120 with the wrappers :258-321) */
int nuc_model_potential(int model, double Z, const double *p, int n, const double *r, double *U)
{
    for (int i = 0; i < n; i++) {
        switch (model) {
        case 1: U[i] = libgrpp_coulomb_potential_ball(r[i], Z, p[0]); break;
        case 2: U[i] = libgrpp_coulomb_potential_gaussian(r[i], Z, p[0]); break;
        case 3: U[i] = libgrpp_coulomb_potential_fermi(r[i], Z, p[0], p[1]); break;
        case 4: U[i] = libgrpp_coulomb_potential_fermi_bubble(r[i], Z, p[0], p[1], p[2]); break;
        case 5: U[i] = libgrpp_coulomb_potential_point(r[i], Z); break;
        default: return 2;
        }
    }
    return 0;
}
