/*
 * Oracle (test infrastructure): nuclear attraction integrals of a shell pair for a point or finite
 * nuclear charge distribution, restated from
 *   nuclear_attraction.c:54-175       driver over primitive pairs and nuclear models;
 *   type1_integrals.c:144-270         radially-local potential integral of a primitive pair
 *                                     (binomial expansion about C, radial table, angular sum);
 *   radial_type1_integral.c:50-325    Log3 quadrature from 2047 points, nested doubling to 8191;
 *   angular_integrals.c:35-56         type-1 angular integral;
 *   nuclear_models.c                  potentials, densities and radii of the nuclear models;
 *   specfunc_fermi_sk.c:42-90         S_k(x) for the Fermi models: cubic Taylor about 0.005 nodes on
 *                                     [-10, 0], exponential tails below; the node table is
 *                                     regenerated here from the series S_k(x) = sum_n (-1)^n e^{nx}/n^k.
 * The point-charge model is the analytic type-1 integral with an r^-1 term of exponent 0.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

/* ------------------------------------------------------------------ S_k(x) */
#define FSK_NPT 2001
static double fsk_tab[FSK_NPT][10];   /* column j: S_{j-2}(x0), x0 = -10 + 0.005 i */
static int fsk_ready = 0;

/* sum_{n>=1} (-1)^n z^n / n^m, z = e^x in (0, 1], m >= 1, by the Cohen-Rodriguez Villegas-Zagier
 * acceleration of alternating series (Experimental Math. 9 (2000) 3-12, algorithm 1) */
static long double fsk_series(int m, long double x)
{
    const int N = 40;
    long double z = expl(x);
    long double d = powl(3.0L + sqrtl(8.0L), N);
    d = (d + 1.0L / d) / 2.0L;
    long double b = -1.0L, c = -d, s = 0.0L;
    for (int k = 0; k < N; k++) {
        c = b - c;
        const long double a = powl(z, k + 1) / powl((long double) (k + 1), m);
        s += c * a;
        b = (k + N) * (long double) (k - N) * b / ((k + 0.5L) * (k + 1.0L));
    }
    return -s / d;   /* the series starts with -z: (-1)^{k+1} a_k */
}

static void fsk_build(void)
{
#pragma omp critical(orc_fsk_lock)
    {
        if (!fsk_ready) {
            for (int i = 0; i < FSK_NPT; i++) {
                long double x = -10.0L + 0.005L * i;
                long double sg = expl(x) / (1.0L + expl(x));
                fsk_tab[i][2] = (double) (-sg);                                       /* S_0  */
                fsk_tab[i][1] = (double) (-sg * (1.0L - sg));                         /* S_-1 */
                fsk_tab[i][0] = (double) (-sg * (1.0L - sg) * (1.0L - 2.0L * sg));    /* S_-2 */
                for (int m = 1; m <= 7; m++) fsk_tab[i][m + 2] = (double) fsk_series(m, x);
            }
            fsk_ready = 1;
        }
    }
}

double orc_fermi_sk_table_value(int i, int col)
{
    fsk_build();
    return fsk_tab[i][col];
}

double orc_fermi_sk(int k, double x)
{
    fsk_build();
    if (x <= -30) return -1.0 * exp(x);
    if (x <= -14) return -1.0 * exp(x) + 1.0 * exp(2 * x) / pow(2, k);
    if (x <= -10) return -1.0 * exp(x) + 1.0 * exp(2 * x) / pow(2, k) - 1.0 * exp(3 * x) / pow(3, k);
    int i0 = (int) (floor((x + 10.0) * 200) + 1e-10);
    double x0 = -10.0 + i0 * 0.005, dx = x - x0;
    if (i0 == 2000) return fsk_tab[i0][k + 2];
    const double *d = fsk_tab[i0];
    double dx2 = dx * dx, dx3 = dx2 * dx;
    return d[k + 2] + d[k + 1] * dx + d[k] * dx2 * 0.5 + d[k - 1] * dx3 * (1.0 / 6.0);
}

/* ------------------------------------------------------------------ nuclear models */
double orc_rms_johnson_1985(int A) { return 0.836 * pow(A, 1.0 / 3.0) + 0.570; }
double orc_rms_golovko_2008(int A) { return 0.77 * pow(A, 1.0 / 3.0) + 0.98; }

int orc_fermi_params(double R_rms, double *c, double *a)
{
    double t = 2.3;
    *a = t / (4 * log(3));
    double c2 = 5.0 / 3.0 * R_rms * R_rms - 7.0 / 3.0 * M_PI * M_PI * (*a) * (*a);
    if (c2 < 0) return -1;
    *c = sqrt(c2);
    return 1;
}

static double fermi_norm(double c, double a)
{
    double a2 = a * a, a3 = a * a * a, c2 = c * c, c3 = c * c * c;
    return 1.0 + M_PI * M_PI * a2 / c2 - 6.0 * a3 / c3 * orc_fermi_sk(3, -c / a);
}

/* nuclear_models.c:354-373 */
static double bubble_norm(double c, double a, double k)
{
    double a2 = a * a, a3 = a * a2, a4 = a * a3, a5 = a * a4, c2 = c * c, c3 = c * c2, c4 = c * c3, c5 = c * c4;
    (void) c3;
    return fermi_norm(c, a) + 3.0 / 5.0 * k + 2.0 * M_PI * M_PI * a2 * k / c2 +
           7.0 * M_PI * M_PI * M_PI * M_PI * a4 * k / (5.0 * c4) - 72.0 * a5 * k / c5 * orc_fermi_sk(5, -c / a);
}

double orc_potential_point(double r, double Z) { return -Z / r; }

double orc_potential_ball(double r, double Z, double R_rms)
{
    double R0 = sqrt(5.0 / 3.0) * R_rms;
    if (r <= R0) return -Z / (2.0 * R0) * (3.0 - r * r / (R0 * R0));
    return -Z / r;
}

double orc_potential_gaussian(double r, double Z, double R_rms)
{
    double xi = 3.0 / (2.0 * R_rms * R_rms);
    return -Z / r * erf(sqrt(xi) * r);
}

double orc_potential_fermi(double r, double Z, double c, double a)
{
    double a2 = a * a, a3 = a * a * a, c3 = c * c * c;
    double N = fermi_norm(c, a);
    if (r > c) {
        double S2 = orc_fermi_sk(2, (c - r) / a), S3 = orc_fermi_sk(3, (c - r) / a);
        return -Z / (N * r) * (N + 3 * a2 * r / c3 * S2 + 6 * a3 / c3 * S3);
    }
    double P2 = orc_fermi_sk(2, (r - c) / a), P3 = orc_fermi_sk(3, (r - c) / a), S3 = orc_fermi_sk(3, -c / a);
    double r3 = r * r * r;
    return -Z / (N * r) *
           (1.5 * r / c - r3 / (2 * c3) + M_PI * M_PI * a2 * r / (2 * c3) - 3 * a2 * r / c3 * P2 + 6 * a3 / c3 * (P3 - S3));
}

/* nuclear_models.c:290-349 (Parpia-Mohanty form of the Fermi term plus the k (r/c)^2 term) */
double orc_potential_fermi_bubble(double r, double Z, double c, double a, double k)
{
    double a2 = a * a, c2 = c * c;
    double Nk = bubble_norm(c, a, k);
    double F0, F2;
    if (r < c) {
        double S2 = orc_fermi_sk(2, (r - c) / a), S3 = orc_fermi_sk(3, (r - c) / a);
        double S4 = orc_fermi_sk(4, (r - c) / a), S5 = orc_fermi_sk(5, (r - c) / a);
        F0 = -pow(r, 3) / 6.0 - r * a * a * S2 + 2.0 * pow(a, 3) * S3 + r * c * c / 2.0 + M_PI * M_PI / 6.0 * r * a * a -
             2.0 * pow(a, 3) * orc_fermi_sk(3, -c / a);
        F2 = -pow(r, 5) / 20.0 - pow(r, 3) * pow(a, 2) * S2 + 6.0 * pow(r, 2) * pow(a, 3) * S3 - 18.0 * r * pow(a, 4) * S4 +
             24.0 * pow(a, 5) * S5 + r * pow(c, 4) / 4.0 + r * M_PI * M_PI * c * c * a * a / 2.0 +
             r * pow(a, 4) * pow(M_PI, 4) * 7.0 / 60.0 - 24.0 * pow(a, 5) * orc_fermi_sk(5, -c / a);
    } else {
        double S2 = orc_fermi_sk(2, (c - r) / a), S3 = orc_fermi_sk(3, (c - r) / a);
        double S4 = orc_fermi_sk(4, (c - r) / a), S5 = orc_fermi_sk(5, (c - r) / a);
        F0 = pow(c, 3) / 3.0 + M_PI * M_PI / 3.0 * c * a * a - 2.0 * pow(a, 3) * orc_fermi_sk(3, -c / a) + r * a * a * S2 +
             2.0 * pow(a, 3) * S3;
        F2 = pow(c, 5) / 5.0 + 2.0 * pow(c, 3) * a * a * M_PI * M_PI / 3.0 + 7.0 * pow(a, 4) * c * pow(M_PI, 4) / 15.0 -
             24.0 * pow(a, 5) * orc_fermi_sk(5, -c / a) + pow(a, 2) * pow(r, 3) * S2 + 6.0 * pow(a, 3) * pow(r, 2) * S3 +
             18.0 * r * pow(a, 4) * S4 + 24.0 * pow(a, 5) * S5;
    }
    (void) a2;
    return -Z / (Nk * r) * 3.0 / pow(c, 3) * (F0 + k / c2 * F2);
}

/* ------------------------------------------------------------------ numeric radial type-1 */
#define T1_MIN_GRID 2047
#define T1_MAX_GRID 10000
#define T1_TOTAL 8191

typedef struct {
    double k, aA, aB, CA2, CB2, pref;
    int model;
    const double *mp;   /* Z, then the model parameters */
} t1par_t;

static double model_potential(const t1par_t *p, double r)
{
    const double Z = p->mp[0];
    switch (p->model) {
    case 5: return orc_potential_point(r, Z);
    case 1: return orc_potential_ball(r, Z, p->mp[1]);
    case 2: return orc_potential_gaussian(r, Z, p->mp[1]);
    case 3: return orc_potential_fermi(r, Z, p->mp[1], p->mp[2]);
    default: return orc_potential_fermi_bubble(r, Z, p->mp[1], p->mp[2], p->mp[3]);
    }
}

/* radial_type1_integral.c:170-180 */
static double t1_gto(const t1par_t *p, double r)
{
    double power = p->k * r - (p->aA + p->aB) * r * r - p->aA * p->CA2 - p->aB * p->CB2;
    return p->pref * exp(power);
}

typedef struct {
    int nr, nfill, lmax, nmax;
    double r[T1_TOTAL], w[T1_TOTAL], pot[T1_TOTAL], gto[T1_TOTAL];
    double *rN, *bes;   /* [n][T1_TOTAL], [lambda][T1_TOTAL] */
} t1grid_t;

static void t1_point(t1grid_t *g, const t1par_t *p, int idx, int i, int nr)
{
    const double R = 5.0, R3 = R * R * R;
    double xi = i / (nr + 1.0), xi3 = xi * xi * xi, ln_xi = log(1 - xi3);
    double wi = 3 * R3 * xi * xi * ln_xi * ln_xi / ((1 - xi3) * (nr + 1.0)), ri = -R * ln_xi;
    g->r[idx] = ri;
    g->w[idx] = wi;
    g->pot[idx] = model_potential(p, ri);
    g->gto[idx] = t1_gto(p, ri);
    for (int l = 0; l <= g->lmax; l++) g->bes[l * T1_TOTAL + idx] = orc_bessel_scaled(l, ri * p->k);
    for (int n = 0; n <= g->nmax; n++) g->rN[n * T1_TOTAL + idx] = pow(ri, n);
}

/* radial_type1_integral.c:100-150, 240-325: table over lambda, N <= lmax of the quadrature from 2047
 * points, doubled until |S - S_prev| <= tol or beyond 10000 points */
static void t1_table(const t1par_t *p, int lmax, double tol, double *Q)
{
    t1grid_t *g = (t1grid_t *) calloc(1, sizeof(t1grid_t));
    g->lmax = lmax;
    g->nmax = lmax;
    g->rN = (double *) calloc((size_t) (lmax + 1) * T1_TOTAL, sizeof(double));
    g->bes = (double *) calloc((size_t) (lmax + 1) * T1_TOTAL, sizeof(double));
    for (int i = 1; i <= T1_MIN_GRID; i++) t1_point(g, p, i - 1, i, T1_MIN_GRID);
    g->nr = g->nfill = T1_MIN_GRID;
    for (int lam = 0; lam <= lmax; lam++)
        for (int n = 0; n <= lmax; n++) {
            const double *rn = g->rN + n * T1_TOTAL, *bs = g->bes + lam * T1_TOTAL;
            int nr = T1_MIN_GRID;
            double sum = 0.0, prev;
            for (int i = 0; i < nr; i++) sum += g->w[i] * g->pot[i] * g->gto[i] * rn[i] * bs[i];
            int converged = 0;
            do {
                int idx = nr;
                nr = 2 * nr + 1;
                if (nr > T1_MAX_GRID) break;
                prev = sum;
                sum = 0.5 * sum;
                if (nr > g->nr) {   /* expand_radial_type1_grid: odd points appended */
                    int k = g->nfill;
                    for (int i = 1; i <= nr; i += 2) t1_point(g, p, k++, i, nr);
                    g->nfill = k;
                    g->nr = nr;
                }
                for (int i = idx; i < nr; i++) sum += g->w[i] * g->pot[i] * g->gto[i] * rn[i] * bs[i];
                converged = fabs(sum - prev) <= tol;
            } while (!converged);
            Q[lam * (lmax + 1) + n] = sum;
        }
    free(g->rN);
    free(g->bes);
    free(g);
}

/* angular_integrals.c:35-56 */
static double angular_type1(int lambda, int I, int J, int K, const double *k)
{
    double y[64];
    orc_rsh_values(lambda, k, y);
    double sum = 0.0;
    for (int mu = -lambda; mu <= lambda; mu++) {
        double sum2 = 0.0;
        for (int r = 0; r <= lambda; r++)
            for (int s = 0; s <= lambda; s++)
                for (int t = 0; t <= lambda; t++)
                    if (r + s + t == lambda)
                        sum2 += orc_rsh_coef(lambda, mu, r, s, t) * orc_unit_sphere_integral(I + r, J + s, K + t);
        sum += sum2 * y[mu + lambda];
    }
    return sum;
}

/* type1_integrals.c:144-270: one primitive pair */
static void radially_local_prim(const orc_shell_t *SA, double aA, const orc_shell_t *SB, double aB, const double *C,
                                int model, const double *mp, double tol, double *M)
{
    const int nA = SA->cart_size, nB = SB->cart_size;
    memset(M, 0, sizeof(double) * nA * nB);
    double CA[3], CB[3], kv[3];
    for (int d = 0; d < 3; d++) {
        CA[d] = C[d] - SA->origin[d];
        CB[d] = C[d] - SB->origin[d];
    }
    double CA2 = CA[0] * CA[0] + CA[1] * CA[1] + CA[2] * CA[2], CB2 = CB[0] * CB[0] + CB[1] * CB[1] + CB[2] * CB[2];
    for (int d = 0; d < 3; d++) kv[d] = -2.0 * (aA * CA[d] + aB * CB[d]);
    double k = sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
    const int LA = SA->L, LB = SB->L, lmax = LA + LB;
    t1par_t p = {k, aA, aB, CA2, CB2, 4 * M_PI * orc_gaussian_norm_factor(LA, 0, 0, aA) * orc_gaussian_norm_factor(LB, 0, 0, aB),
                 model, mp};
    double *Q = (double *) calloc((size_t) (lmax + 1) * (lmax + 1), sizeof(double));
    t1_table(&p, lmax, tol, Q);
    for (int ic = 0; ic < nA; ic++)
        for (int jc = 0; jc < nB; jc++) {
            const int *ta = SA->cart_list + 3 * ic, *tb = SB->cart_list + 3 * jc;
            double chi = 0.0;
            for (int a = 0; a <= ta[0]; a++)
                for (int b = 0; b <= ta[1]; b++)
                    for (int c = 0; c <= ta[2]; c++)
                        for (int d = 0; d <= tb[0]; d++)
                            for (int e = 0; e <= tb[1]; e++)
                                for (int f = 0; f <= tb[2]; f++) {
                                    double factor = (double) orc_binomial(ta[0], a) * orc_binomial(ta[1], b) *
                                                    orc_binomial(ta[2], c) * orc_binomial(tb[0], d) *
                                                    orc_binomial(tb[1], e) * orc_binomial(tb[2], f) *
                                                    pow(CA[0], ta[0] - a) * pow(CA[1], ta[1] - b) * pow(CA[2], ta[2] - c) *
                                                    pow(CB[0], tb[0] - d) * pow(CB[1], tb[1] - e) * pow(CB[2], tb[2] - f);
                                    if (fabs(factor) < 1e-13) continue;
                                    int N = a + b + c + d + e + f;
                                    double s = 0.0;
                                    for (int lam = 0; lam <= lmax; lam++) {
                                        double q = Q[lam * (lmax + 1) + N];
                                        if (fabs(q) < 1e-16) continue;
                                        s += angular_type1(lam, a + d, b + e, c + f, kv) * q;
                                    }
                                    chi += factor * s;
                                }
            M[ic * nB + jc] = chi;
        }
    free(Q);
}

/* nuclear_attraction.c:54-133: model 0 point charge (analytic), 1 ball (R_rms), 2 Gaussian (R_rms),
 * 3 Fermi (c, a), 4 Fermi + bubble (c, a, k), 5 point charge by the numeric quadrature */
void orc_nuclear_attraction(const orc_shell_t *A, const orc_shell_t *B, const double *C, int charge, int model,
                            const double *params, double tol, double *M)
{
    const int sz = A->cart_size * B->cart_size;
    memset(M, 0, sizeof(double) * sz);
    double *buf = (double *) calloc(sz, sizeof(double));
    double mp[4] = {(double) charge, 0, 0, 0};
    for (int q = 0; q < 3 && params && model >= 1 && model <= 4; q++) mp[q + 1] = params[q];
    for (int i = 0; i < A->num_primitives; i++)
        for (int j = 0; j < B->num_primitives; j++) {
            double cij = A->coeffs[i] * B->coeffs[j];
            if (fabs(cij) < 1e-13) continue;
            if (model == 0) {
                orc_type1_prim_n1(A, A->alpha[i], B, B->alpha[j], C, 0.0, buf);
                for (int k = 0; k < sz; k++) buf[k] *= (-1) * charge;
            } else {
                radially_local_prim(A, A->alpha[i], B, B->alpha[j], C, model, mp, tol, buf);
            }
            for (int k = 0; k < sz; k++) M[k] += cij * buf[k];
        }
    free(buf);
}
