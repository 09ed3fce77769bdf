/*
 * Oracle (test infrastructure): analytic nuclear gradients of the full GRPP integrals
 * (grpp_integrals_gradient.c:280-1016, diff_gaussian.c:35-73).
 *
 * d/dA_k of x_A^n e^{-a x_A^2} = -n x_A^{n-1}(...) + 2a x_A^{n+1}(...): the bra (or ket) shell
 * is replaced by an L-1 shell and an L+1 shell whose coefficients absorb the norm-factor ratio
 * (and 2a for L+1); the -n weight is applied when the blocks are scattered back.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

static double nfac(double a, int L) { return pow(2 * a / M_PI, 0.75) * pow(4 * a, 0.5 * L); }

/* diff_gaussian.c:35-63 */
static void diff_shell(const orc_shell_t *s, orc_shell_t **down, orc_shell_t **up)
{
    *down = NULL;
    if (s->L > 0) {
        *down = orc_new_shell(s->origin, s->L - 1, s->num_primitives, s->coeffs, s->alpha);
        for (int i = 0; i < s->num_primitives; i++)
            (*down)->coeffs[i] *= nfac(s->alpha[i], s->L) / nfac(s->alpha[i], s->L - 1);
    }
    *up = orc_new_shell(s->origin, s->L + 1, s->num_primitives, s->coeffs, s->alpha);
    for (int i = 0; i < s->num_primitives; i++)
        (*up)->coeffs[i] *= 2.0 * s->alpha[i] * nfac(s->alpha[i], s->L) / nfac(s->alpha[i], s->L + 1);
}

/* position of (n,l,m) in the Cartesian order of L = n+l+m (overlap_gradient.c:219-241) */
static int nlm_index(int n, int l, int m)
{
    int L = n + l + m;
    return (L - n) * (L - n + 1) / 2 + (L - n - l);
}

/* grad += factor * <dA|U|B> (diff_bra) or factor * <A|U|dB> (grpp_integrals_gradient.c:786-1016) */
static void contribution(const orc_shell_t *A, const orc_shell_t *B, const orc_grpp_t *g, const double *C,
                         double **ga, double **gx, double **gy, double **gz, int diff_bra, double factor)
{
    const orc_shell_t *ds = diff_bra ? A : B;
    orc_shell_t *down, *up;
    diff_shell(ds, &down, &up);
    int nA = A->cart_size, nB = B->cart_size;
    double *dn[4] = {0, 0, 0, 0}, *upb[4];
    int ndown = down ? down->cart_size : 0, nup = up->cart_size;
    if (down) {
        size_t sz = (size_t) (diff_bra ? ndown * nB : nA * ndown);
        for (int k = 0; k < 4; k++) dn[k] = calloc(sz, sizeof(double));
        orc_full_grpp(diff_bra ? down : A, diff_bra ? B : down, g, C, dn[0], dn[1], dn[2], dn[3]);
    }
    size_t szu = (size_t) (diff_bra ? nup * nB : nA * nup);
    for (int k = 0; k < 4; k++) upb[k] = calloc(szu, sizeof(double));
    orc_full_grpp(diff_bra ? up : A, diff_bra ? B : up, g, C, upb[0], upb[1], upb[2], upb[3]);
    double **outs[4] = {ga, gx, gy, gz};
    for (int k = 0; k < 3; k++)
        for (int i = 0; i < nA; i++)
            for (int j = 0; j < nB; j++) {
                int bra[3] = {A->cart_list[3 * i], A->cart_list[3 * i + 1], A->cart_list[3 * i + 2]};
                int ket[3] = {B->cart_list[3 * j], B->cart_list[3 * j + 1], B->cart_list[3 * j + 2]};
                int *dv = diff_bra ? bra : ket;
                int idx = i * nB + j;
                int n = dv[k];
                if (down && n > 0) {
                    dv[k] -= 1;
                    int bi = nlm_index(bra[0], bra[1], bra[2]), ki = nlm_index(ket[0], ket[1], ket[2]);
                    dv[k] += 1;
                    int row = diff_bra ? nB : ndown;
                    for (int op = 0; op < 4; op++) outs[op][k][idx] -= factor * n * dn[op][row * bi + ki];
                }
                dv[k] += 1;
                int bi = nlm_index(bra[0], bra[1], bra[2]), ki = nlm_index(ket[0], ket[1], ket[2]);
                dv[k] -= 1;
                int row = diff_bra ? nB : nup;
                for (int op = 0; op < 4; op++) outs[op][k][idx] += factor * upb[op][row * bi + ki];
            }
    for (int k = 0; k < 4; k++) { free(dn[k]); free(upb[k]); }
    if (down) orc_delete_shell(down);
    orc_delete_shell(up);
}

/* geometry cases ACB / ACA / ACC / CCB (grpp_integrals_gradient.c:280-393) */
void orc_full_grpp_gradient(const orc_shell_t *A, const orc_shell_t *B, const orc_grpp_t *g, const double *C,
                            const double *D, double **ga, double **gx, double **gy, double **gz)
{
    int sz = A->cart_size * B->cart_size;
    for (int k = 0; k < 3; k++) {
        memset(ga[k], 0, sizeof(double) * sz);
        memset(gx[k], 0, sizeof(double) * sz);
        memset(gy[k], 0, sizeof(double) * sz);
        memset(gz[k], 0, sizeof(double) * sz);
    }
    const double *a = A->origin, *b = B->origin;
    int AC = orc_points_are_equal(a, C), BC = orc_points_are_equal(b, C), AB = orc_points_are_equal(a, b);
    int AD = orc_points_are_equal(a, D), BD = orc_points_are_equal(b, D), CD = orc_points_are_equal(C, D);
    if (AC && BC) return;
    if (!AD && !BD && !CD) return;
    if (!AC && !BC && !AB) {
        if (AD) contribution(A, B, g, C, ga, gx, gy, gz, 1, +1.0);
        if (BD) contribution(A, B, g, C, ga, gx, gy, gz, 0, +1.0);
        if (CD) {
            contribution(A, B, g, C, ga, gx, gy, gz, 1, -1.0);
            contribution(A, B, g, C, ga, gx, gy, gz, 0, -1.0);
        }
    }
    if (AB && !AC) {
        double f = AD ? +1.0 : -1.0;
        contribution(A, B, g, C, ga, gx, gy, gz, 1, f);
        contribution(A, B, g, C, ga, gx, gy, gz, 0, f);
    }
    if (!AC && BC) contribution(A, B, g, C, ga, gx, gy, gz, 1, AD ? +1.0 : -1.0);
    if (AC && !BC) contribution(A, B, g, C, ga, gx, gy, gz, 0, BD ? +1.0 : -1.0);
}
