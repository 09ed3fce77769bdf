/*
 * Oracle (test infrastructure): type-2 radial integrals
 *   T^N_{l1 l2} = int_0^inf r^{2+N} U(r) F^A_{l1}(r) F^B_{l2}(r) dr,
 *   F^A_l(r) = sum_i c_i N_i exp(-a_i (r-|CA|)^2) K_l(2 a_i |CA| r),
 * on libgrpp's Log3 grid with nested doubling and Shaw-Hill screening
 * (radial_type2_integral.c:101-489, screening.c:193-378).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

#define MIN_GRID 31
#define MAX_GRID 10000
#define NLEVEL 9            /* 31, 63, ..., 8191 points */
#define GRID_TOTAL 8191

orc_radial_opts_t orc_radial_opts = {1e-16, 1, 0, -1, 0};

typedef struct {
    int nfill;              /* points tabulated so far (level-ordered) */
    int nr;                 /* current grid size n (= nfill) */
    int l1max, l2max, nmax;
    double CA, CB;
    const orc_potential_t *pot;
    const orc_shell_t *bra, *ket;   /* shrunk + normalised copies */
    double *r, *w, *U;
    double *rN, *F1, *F2;   /* [n][GRID_TOTAL], [l][GRID_TOTAL] */
} grid_t;

/* sum_i c_i exp(-a_i (r-CA)^2) K_l(2 a_i CA r) (radial_type2_integral.c:465-489) */
static double integrand_fun(double r, int lambda, double CA, const orc_shell_t *s)
{
    double F = 0.0, d2 = (r - CA) * (r - CA);
    for (int i = 0; i < s->num_primitives; i++) {
        double k = 2.0 * s->alpha[i] * CA;
        F += s->coeffs[i] * exp(-s->alpha[i] * d2) * orc_bessel_scaled(lambda, k * r);
    }
    return F;
}

static void fill_point(grid_t *g, int idx, int i, int nr)
{
    const double R = 5.0, R3 = R * R * R;
    double xi = i / (nr + 1.0);
    double xi3 = xi * xi * xi;
    double ln_xi = log(1 - xi3);
    double wi = 3 * R3 * xi * xi * ln_xi * ln_xi / ((1 - xi3) * (nr + 1.0));
    double ri = -R * ln_xi;
    g->r[idx] = ri;
    g->w[idx] = wi;
    g->U[idx] = orc_potential_value(g->pot, ri);
    for (int n = 0; n <= g->nmax; n++) g->rN[n * GRID_TOTAL + idx] = pow(ri, n);
    for (int l = 0; l <= g->l1max; l++) g->F1[l * GRID_TOTAL + idx] = integrand_fun(ri, l, g->CA, g->bra);
    for (int l = 0; l <= g->l2max; l++) g->F2[l * GRID_TOTAL + idx] = integrand_fun(ri, l, g->CB, g->ket);
}

/* first 31 points (radial_type2_integral.c:324-384) */
static void grid_create(grid_t *g)
{
    g->nr = MIN_GRID;
    for (int i = 1; i <= MIN_GRID; i++) fill_point(g, i - 1, i, MIN_GRID);
    g->nfill = MIN_GRID;
}

/* doubled grid: the new (odd-index) points are appended (radial_type2_integral.c:390-444) */
static void grid_expand(grid_t *g, int nr)
{
    if (nr > MAX_GRID || nr <= g->nr) return;
    int idx = g->nfill;
    for (int i = 1; i <= nr; i += 2) fill_point(g, idx++, i, nr);
    g->nfill = idx;
    g->nr = nr;
}

static int equation_count;

/* fixed-point map for the integrand maximum (screening.c:348-378) */
static double eq_for_max(double r, int n, int l1, int l2, double p, double k1, double k2)
{
    double k1r = k1 * r, k2r = k2 * r;
    double K1 = (l1 == 0) ? orc_bessel_scaled(1, k1r) / orc_bessel_scaled(0, k1r)
                          : orc_bessel_scaled(l1 - 1, k1r) / orc_bessel_scaled(l1, k1r);
    double K2 = (l2 == 0) ? orc_bessel_scaled(1, k2r) / orc_bessel_scaled(0, k2r)
                          : orc_bessel_scaled(l2 - 1, k2r) / orc_bessel_scaled(l2, k2r);
    double a = K1 * k1r + K2 * k2r + n;
    if (l1 > 0) a = a - l1 - 1;
    if (l2 > 0) a = a - l2 - 1;
    return sqrt(a / (2.0 * p));
}

/* envelope estimate for one primitive triple (screening.c:272-342); returns 0 on success */
static int screen_primitive(int l1, int l2, int n, double CA2, double CB2, double aA, double aB, double k1, double k2,
                            double eta, double *val)
{
    *val = 0.0;
    if (l1 >= 1 && fabs(k1) <= ORC_ZERO_THRESH) return 0;
    if (l2 >= 1 && fabs(k2) <= ORC_ZERO_THRESH) return 0;
    double p = aA + aB + eta;
    double CA = sqrt(CA2), CB = sqrt(CB2);
    if (l1 == 0 && l2 == 0 && fabs(k1) <= ORC_ZERO_THRESH && fabs(k2) <= ORC_ZERO_THRESH) {
        *val = exp(-aA * CA * CA - aB * CB * CB) * orc_gaussian_integral(n, p);
        return 0;
    }
    const double tol = 1e-2;
    double r0 = (aA * CA + aB * CB) / p, r0p;
    int nsteps = 0;
    do {
        nsteps++;
        if (nsteps == 5) { *val = 0.0; return 1; }
        r0p = r0;
        r0 = eq_for_max(r0, n, l1, l2, p, k1, k2);
    } while (fabs(r0 - r0p) > tol);
    *val = sqrt(M_PI / p) * pow(r0, n) * orc_bessel_scaled(l1, k1 * r0) * orc_bessel_scaled(l2, k2 * r0) *
           exp(-eta * r0 * r0 - aA * (r0 - CA) * (r0 - CA) - aB * (r0 - CB) * (r0 - CB)) * 0.5 *
           (1 + erf(sqrt(p) * r0));
    return 0;
}

/* contracted estimate (screening.c:193-248); returns 0 on success */
static int screen_contracted(int l1, int l2, int n, double CA2, double CB2, const orc_shell_t *bra,
                             const orc_shell_t *ket, const orc_potential_t *pot, double *val)
{
    *val = 0.0;
    double CA = sqrt(CA2), CB = sqrt(CB2);
    for (int i = 0; i < bra->num_primitives; i++) {
        double k1 = 2 * bra->alpha[i] * CA;
        for (int j = 0; j < ket->num_primitives; j++) {
            double k2 = 2 * ket->alpha[j] * CB;
            for (int k = 0; k < pot->num_primitives; k++) {
                double v;
                if (screen_primitive(l1, l2, n + pot->powers[k], CA2, CB2, bra->alpha[i], ket->alpha[j], k1, k2,
                                     pot->alpha[k], &v))
                    return 1;
                *val += bra->coeffs[i] * ket->coeffs[j] * pot->coeffs[k] * v;
            }
        }
    }
    return 0;
}

/* sum of w U F1 F2 r^n over the grid points [b, e), added to s (radial_type2_integral.c:285, :304) */
static inline double point_sum(const double *w, const double *U, const double *F1, const double *F2,
                               const double *rN, int b, int e, double s)
{
    for (int i = b; i < e; i++) s += w[i] * U[i] * F1[i] * F2[i] * rN[i];
    return s;
}

/* the same sum in another arithmetic order (orc_radial_opts.engine_order, diagnostics only): the
 * level's points in a fresh sum (bit 0) and/or in groups of four (bit 2), each point's product as
 * (F1 r^nA s)(F2 r^nB sign s), s = sqrt|w U|, N = nA + nB with nA the smallest allowed (bit 1) */
static double point_sum_alt(const grid_t *g, const double *F1, const double *F2, int N, int b, int e, double s,
                            int mode)
{
    int nb = N - g->ket->L < 0 ? 0 : N - g->ket->L, na = N - nb;
    if (na > g->bra->L) { na = g->bra->L; nb = N - na; }
    double acc = (mode & 1) ? 0.0 : s, four = 0.0;
    int cnt = 0;
    for (int i = b; i < e; i++) {
        double p;
        if (mode & 2) {
            double wu = g->w[i] * g->U[i], sq = sqrt(fabs(wu)), sg = wu < 0 ? -sq : sq;
            p = (F1[i] * (pow(g->r[i], na) * sq)) * (F2[i] * (pow(g->r[i], nb) * sg));
        } else {
            p = g->w[i] * g->U[i] * F1[i] * F2[i] * g->rN[N * GRID_TOTAL + i];
        }
        if (mode & 4) {
            four += p;
            if (++cnt == 4 || i == e - 1) { acc += four; four = 0.0; cnt = 0; }
        } else {
            acc += p;
        }
    }
    return (mode & 1) ? s + acc : acc;
}

/* adaptive nested-doubling quadrature of one entry (radial_type2_integral.c:242-318) */
static double integrate_entry(grid_t *g, int n, int l1, int l2, int *level_out)
{
    const orc_radial_opts_t *o = &orc_radial_opts;
    double tol = o->tol;
    double screened = 0.0;
    int screen_ok = 0;
    if (o->screening && o->fixed_level < 0) {
        screen_ok = (screen_contracted(l1, l2, n, g->CA * g->CA, g->CB * g->CB, g->bra, g->ket, g->pot, &screened)
                     == 0);
        if (screen_ok && fabs(screened) < tol) {
            if (level_out) *level_out = -1;
            return screened;
        }
    }
    const double *U = g->U, *w = g->w;
    const double *F1 = g->F1 + l1 * GRID_TOTAL, *F2 = g->F2 + l2 * GRID_TOTAL, *rN = g->rN + n * GRID_TOTAL;
    int nr = MIN_GRID, level = 0;
    double sum = 0.0, prev;
    sum = o->engine_order ? point_sum_alt(g, F1, F2, n, 0, nr, sum, o->engine_order) : point_sum(w, U, F1, F2, rN, 0, nr, sum);
    if (o->fixed_level == 0) { if (level_out) *level_out = 0; return sum; }
    int converged = 0;
    do {
        int idx = nr;
        nr = 2 * nr + 1;
        if (nr > MAX_GRID) break;
        prev = sum;
        sum = 0.5 * sum;
        grid_expand(g, nr);
        sum = o->engine_order ? point_sum_alt(g, F1, F2, n, idx, nr, sum, o->engine_order)
                              : point_sum(w, U, F1, F2, rN, idx, nr, sum);
        level++;
        if (o->fixed_level >= 0) {
            if (level >= o->fixed_level) break;
            continue;
        }
        if (screen_ok && (fabs(sum) / fabs(screened) < 0.001)) { converged = 0; continue; }
        if (level < o->min_level) { converged = 0; continue; }
        converged = fabs(sum - prev) <= tol;
    } while (!converged);
    if (level_out) *level_out = level;
    return sum;
}

static orc_shell_t *shrink_normalise(const orc_shell_t *s)
{
    orc_shell_t *c = orc_shell_copy(s);
    int np = 0;
    for (int i = 0; i < c->num_primitives; i++)
        if (fabs(c->coeffs[i]) > ORC_ZERO_THRESH) {
            c->coeffs[np] = c->coeffs[i];
            c->alpha[np] = c->alpha[i];
            np++;
        }
    c->num_primitives = np;
    for (int i = 0; i < np; i++) c->coeffs[i] *= orc_gaussian_norm_factor(c->L, 0, 0, c->alpha[i]);
    return c;
}

static orc_potential_t *shrink_potential(const orc_potential_t *p)
{
    return orc_new_potential(p->L, p->J, p->num_primitives, p->powers, p->coeffs, p->alpha);
}

static grid_t *grid_new(int l1max, int l2max, int nmax, double CA, double CB, const orc_potential_t *pot,
                        const orc_shell_t *bra, const orc_shell_t *ket)
{
    grid_t *g = (grid_t *) calloc(1, sizeof(grid_t));
    g->l1max = l1max;
    g->l2max = l2max;
    g->nmax = nmax;
    g->CA = CA;
    g->CB = CB;
    g->pot = pot;
    g->bra = bra;
    g->ket = ket;
    g->r = (double *) calloc(GRID_TOTAL, sizeof(double));
    g->w = (double *) calloc(GRID_TOTAL, sizeof(double));
    g->U = (double *) calloc(GRID_TOTAL, sizeof(double));
    g->rN = (double *) calloc((size_t) (nmax + 1) * GRID_TOTAL, sizeof(double));
    g->F1 = (double *) calloc((size_t) (l1max + 1) * GRID_TOTAL, sizeof(double));
    g->F2 = (double *) calloc((size_t) (l2max + 1) * GRID_TOTAL, sizeof(double));
    grid_create(g);
    return g;
}

static void grid_free(grid_t *g)
{
    free(g->r); free(g->w); free(g->U); free(g->rN); free(g->F1); free(g->F2);
    free(g);
}

/* whole table (radial_type2_integral.c:101-207), including the analytic one-centre branch */
orc_radial_table_t *orc_tabulate_radial(int l1max, int l2max, int nmax, double CA2, double CB2,
                                        const orc_potential_t *pot, const orc_shell_t *bra, const orc_shell_t *ket)
{
    orc_radial_table_t *t = (orc_radial_table_t *) calloc(1, sizeof(orc_radial_table_t));
    t->lambda1_max = l1max;
    t->lambda2_max = l2max;
    t->n_max = nmax;
    size_t nent = (size_t) (l1max + 1) * (l2max + 1) * (nmax + 1);
    t->Q = (double *) calloc(nent, sizeof(double));
    t->level = (int *) calloc(nent, sizeof(int));

    if (CA2 < 1e-14 && CB2 < 1e-14) {
        for (int i = 0; i < bra->num_primitives; i++) {
            double ci = bra->coeffs[i] * orc_gaussian_norm_factor(bra->L, 0, 0, bra->alpha[i]);
            for (int j = 0; j < ket->num_primitives; j++) {
                double cj = ket->coeffs[j] * orc_gaussian_norm_factor(ket->L, 0, 0, ket->alpha[j]);
                for (int k = 0; k < pot->num_primitives; k++) {
                    double p = bra->alpha[i] + ket->alpha[j] + pot->alpha[k];
                    double f = ci * cj * pot->coeffs[k];
                    for (int n = 0; n <= nmax; n++) t->Q[n] += f * orc_gaussian_integral(pot->powers[k] + n, p);
                }
            }
        }
        return t;
    }

    orc_potential_t *ps = shrink_potential(pot);
    orc_shell_t *bs = shrink_normalise(bra), *ks = shrink_normalise(ket);
    grid_t *g = grid_new(l1max, l2max, nmax, sqrt(CA2), sqrt(CB2), ps, bs, ks);
    for (int l1 = 0; l1 <= l1max; l1++)
        for (int l2 = 0; l2 <= l2max; l2++)
            for (int n = 0; n <= nmax; n++) {
                size_t idx = (size_t) (l2max + 1) * (nmax + 1) * l1 + (nmax + 1) * l2 + n;
                int lev;
                t->Q[idx] = integrate_entry(g, n, l1, l2, &lev);
                t->level[idx] = lev;
            }
    grid_free(g);
    orc_delete_potential(ps);
    orc_delete_shell(bs);
    orc_delete_shell(ks);
    return t;
}

void orc_free_radial(orc_radial_table_t *t)
{
    free(t->Q);
    free(t->level);
    free(t);
}

double orc_radial_integral(int lambda1, int lambda2, int n, double CA, double CB, const orc_potential_t *pot,
                           const orc_shell_t *bra, const orc_shell_t *ket, int *level_out)
{
    orc_potential_t *ps = shrink_potential(pot);
    orc_shell_t *bs = shrink_normalise(bra), *ks = shrink_normalise(ket);
    grid_t *g = grid_new(lambda1, lambda2, n, CA, CB, ps, bs, ks);
    double v = integrate_entry(g, n, lambda1, lambda2, level_out);
    grid_free(g);
    orc_delete_potential(ps);
    orc_delete_shell(bs);
    orc_delete_shell(ks);
    return v;
}
