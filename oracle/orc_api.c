/*
 * Oracle (test infrastructure): object constructors (shell.c, potential.c, grpp.c) and the
 * whole-matrix builders used as the checker and as the CPU baseline ("port").
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

#define ORC_MAX_POT 100   /* grpp.c:12 */

orc_shell_t *orc_new_shell(const double *origin, int L, int nprim, const double *coeffs, const double *alpha)
{
    orc_shell_t *s = (orc_shell_t *) calloc(1, sizeof(orc_shell_t));
    s->L = L;
    for (int k = 0; k < 3; k++) s->origin[k] = origin[k];
    s->cart_size = (L + 1) * (L + 2) / 2;
    s->cart_list = (int *) calloc(3 * s->cart_size, sizeof(int));
    orc_cartesian_generator(L, s->cart_list);
    s->num_primitives = nprim;
    s->coeffs = (double *) calloc(nprim > 0 ? nprim : 1, sizeof(double));
    s->alpha = (double *) calloc(nprim > 0 ? nprim : 1, sizeof(double));
    for (int i = 0; i < nprim; i++) {
        s->coeffs[i] = coeffs[i];
        s->alpha[i] = alpha[i];
    }
    return s;
}

orc_shell_t *orc_shell_copy(const orc_shell_t *s)
{
    return orc_new_shell(s->origin, s->L, s->num_primitives, s->coeffs, s->alpha);
}

void orc_delete_shell(orc_shell_t *s)
{
    if (!s) return;
    free(s->cart_list);
    free(s->coeffs);
    free(s->alpha);
    free(s);
}

/* drops terms with |c| < 1e-14 (potential.c:22-46) */
orc_potential_t *orc_new_potential(int L, int J, int nprim, const int *powers, const double *coeffs,
                                   const double *alpha)
{
    orc_potential_t *p = (orc_potential_t *) calloc(1, sizeof(orc_potential_t));
    p->L = L;
    p->J = J;
    p->powers = (int *) calloc(nprim > 0 ? nprim : 1, sizeof(int));
    p->coeffs = (double *) calloc(nprim > 0 ? nprim : 1, sizeof(double));
    p->alpha = (double *) calloc(nprim > 0 ? nprim : 1, sizeof(double));
    for (int i = 0; i < nprim; i++) {
        if (fabs(coeffs[i]) < ORC_ZERO_THRESH) continue;
        p->coeffs[p->num_primitives] = coeffs[i];
        p->powers[p->num_primitives] = powers[i];
        p->alpha[p->num_primitives] = alpha[i];
        p->num_primitives++;
    }
    return p;
}

void orc_delete_potential(orc_potential_t *p)
{
    if (!p) return;
    free(p->powers);
    free(p->coeffs);
    free(p->alpha);
    free(p);
}

/* sum_k c_k r^{n_k-2} exp(-z_k r^2) (potential.c:70-81) */
double orc_potential_value(const orc_potential_t *p, double r)
{
    double v = 0.0, r2 = r * r;
    for (int i = 0; i < p->num_primitives; i++) v += p->coeffs[i] * pow(r, p->powers[i] - 2) * exp(-p->alpha[i] * r2);
    return v;
}

orc_grpp_t *orc_new_grpp(void) { return (orc_grpp_t *) calloc(1, sizeof(orc_grpp_t)); }

void orc_grpp_set_local(orc_grpp_t *g, orc_potential_t *p)
{
    if (g->U_L) orc_delete_potential(g->U_L);
    g->U_L = p;
}

void orc_grpp_add_arep(orc_grpp_t *g, orc_potential_t *p)
{
    if (g->n_arep == 0) g->U_arep = calloc(ORC_MAX_POT, sizeof(orc_potential_t *));
    g->U_arep[g->n_arep++] = p;
}

void orc_grpp_add_esop(orc_grpp_t *g, orc_potential_t *p)
{
    if (g->n_esop == 0) g->U_esop = calloc(ORC_MAX_POT, sizeof(orc_potential_t *));
    g->U_esop[g->n_esop++] = p;
}

void orc_grpp_add_oc(orc_grpp_t *g, orc_potential_t *p, orc_shell_t *s)
{
    if (g->n_oc_shells == 0) {
        g->U_oc = calloc(ORC_MAX_POT, sizeof(orc_potential_t *));
        g->oc_shells = calloc(ORC_MAX_POT, sizeof(orc_shell_t *));
    }
    g->U_oc[g->n_oc_shells] = p;
    g->oc_shells[g->n_oc_shells] = s;
    g->n_oc_shells++;
}

void orc_delete_grpp(orc_grpp_t *g)
{
    if (!g) return;
    orc_delete_potential(g->U_L);
    for (int i = 0; i < g->n_arep; i++) orc_delete_potential(g->U_arep[i]);
    for (int i = 0; i < g->n_esop; i++) orc_delete_potential(g->U_esop[i]);
    for (int i = 0; i < g->n_oc_shells; i++) {
        orc_delete_potential(g->U_oc[i]);
        orc_delete_shell(g->oc_shells[i]);
    }
    free(g->U_arep);
    free(g->U_esop);
    free(g->U_oc);
    free(g->oc_shells);
    free(g);
}

static void add_block(int dim, double *m, int r0, int c0, int nr, int nc, const double *blk)
{
    for (int i = 0; i < nr; i++)
        for (int j = 0; j < nc; j++) m[(size_t) (r0 + i) * dim + c0 + j] += blk[i * nc + j];
}

/* Dense dim x dim matrices summed over GRPP centres for every ordered shell pair, as the
 * reference C driver assembles them (test_libgrpp_c/eval_integrals.c:42-130).
 * pair_list (npsel entries, p = i*nshell + j) selects ordered pairs, NULL = all; returns seconds in the loop. */
double orc_build_matrices(int nshell, orc_shell_t **shells, int ncenter, const double *xyz, orc_grpp_t **grpps,
                          double *arep, double *so_x, double *so_y, double *so_z, int nthreads, int npsel,
                          const int *pair_list)
{
    orc_init();
    int *off = malloc(sizeof(int) * (nshell + 1));
    off[0] = 0;
    for (int i = 0; i < nshell; i++) off[i + 1] = off[i] + shells[i]->cart_size;
    int dim = off[nshell];
    size_t n2 = (size_t) dim * dim;
    memset(arep, 0, n2 * sizeof(double));
    memset(so_x, 0, n2 * sizeof(double));
    memset(so_y, 0, n2 * sizeof(double));
    memset(so_z, 0, n2 * sizeof(double));
    int npairs = pair_list ? npsel : nshell * nshell;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    double t0 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic)
    for (int q = 0; q < npairs; q++) {
        int p = pair_list ? pair_list[q] : q;
        int i = p / nshell, j = p % nshell;
        orc_shell_t *a = shells[i], *b = shells[j];
        int sz = a->cart_size * b->cart_size;
        double *buf = malloc(sizeof(double) * 4 * sz);
        for (int c = 0; c < ncenter; c++) {
            if (!grpps[c]) continue;
            orc_full_grpp(a, b, grpps[c], xyz + 3 * c, buf, buf + sz, buf + 2 * sz, buf + 3 * sz);
            add_block(dim, arep, off[i], off[j], a->cart_size, b->cart_size, buf);
            add_block(dim, so_x, off[i], off[j], a->cart_size, b->cart_size, buf + sz);
            add_block(dim, so_y, off[i], off[j], a->cart_size, b->cart_size, buf + 2 * sz);
            add_block(dim, so_z, off[i], off[j], a->cart_size, b->cart_size, buf + 3 * sz);
        }
        free(buf);
    }
    double t1 = omp_get_wtime();
    free(off);
    return t1 - t0;
}

/* grad[op][3*atom+k] dim x dim (test_libgrpp_c/eval_integrals_grad.c:37-163); centres = atoms */
double orc_build_gradients(int nshell, orc_shell_t **shells, int natom, const double *xyz, orc_grpp_t **grpps,
                           double *out, int nthreads)
{
    orc_init();
    int *off = malloc(sizeof(int) * (nshell + 1));
    off[0] = 0;
    for (int i = 0; i < nshell; i++) off[i + 1] = off[i] + shells[i]->cart_size;
    int dim = off[nshell];
    size_t n2 = (size_t) dim * dim;
    memset(out, 0, sizeof(double) * 12 * natom * n2);
    if (nthreads > 0) omp_set_num_threads(nthreads);
    double t0 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic)
    for (int p = 0; p < nshell * nshell; p++) {
        int i = p / nshell, j = p % nshell;
        orc_shell_t *a = shells[i], *b = shells[j];
        int sz = a->cart_size * b->cart_size;
        double *g[4][3];
        for (int op = 0; op < 4; op++)
            for (int k = 0; k < 3; k++) g[op][k] = calloc(sz, sizeof(double));
        for (int d = 0; d < natom; d++)
            for (int c = 0; c < natom; c++) {
                if (!grpps[c]) continue;
                orc_full_grpp_gradient(a, b, grpps[c], xyz + 3 * c, xyz + 3 * d, g[0], g[1], g[2], g[3]);
                for (int op = 0; op < 4; op++)
                    for (int k = 0; k < 3; k++)
                        add_block(dim, out + (size_t) (op * 3 * natom + 3 * d + k) * n2, off[i], off[j], a->cart_size,
                                  b->cart_size, g[op][k]);
            }
        for (int op = 0; op < 4; op++)
            for (int k = 0; k < 3; k++) free(g[op][k]);
    }
    double t1 = omp_get_wtime();
    free(off);
    return t1 - t0;
}
