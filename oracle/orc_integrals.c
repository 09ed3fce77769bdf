/*
 * Oracle (test infrastructure): the GRPP integrators --
 * overlap (Obara-Saika), type-1 (McMurchie-Davidson), type-2, spin-orbit, outercore, full GRPP.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

#define FOURPI2 (16.0 * M_PI * M_PI)

static double pw(double x, int n) { return pow(x, n); }

/* ---------------- overlap (overlap.c:39-138) ---------------- */
static void overlap_prim(const orc_shell_t *A, const orc_shell_t *B, double aA, double aB, double *out)
{
    int LA = A->L, LB = B->L;
    double NA = orc_gaussian_norm_factor(LA, 0, 0, aA), NB = orc_gaussian_norm_factor(LB, 0, 0, aB);
    double p = aA + aB, mu = aA * aB / (aA + aB);
    double S[3][16][16];
    for (int c = 0; c < 3; c++) {
        double P = (aA * A->origin[c] + aB * B->origin[c]) / p;
        double XAB = A->origin[c] - B->origin[c], XPA = P - A->origin[c], XPB = P - B->origin[c];
        double pf = 1.0 / (2.0 * p);
        for (int i = 0; i <= LA; i++)
            for (int j = 0; j <= LB; j++) {
                double s = 0.0;
                if (i + j == 0) { S[c][0][0] = sqrt(M_PI / p) * exp(-mu * XAB * XAB); continue; }
                if (i == 0) {
                    s += XPB * S[c][i][j - 1];
                    if (j - 1 > 0) s += (j - 1) * pf * S[c][i][j - 2];
                } else {
                    s += XPA * S[c][i - 1][j];
                    if (i - 1 > 0) s += (i - 1) * pf * S[c][i - 2][j];
                    if (j > 0) s += j * pf * S[c][i - 1][j - 1];
                }
                S[c][i][j] = s;
            }
    }
    for (int m = 0; m < A->cart_size; m++)
        for (int n = 0; n < B->cart_size; n++) {
            const int *a = A->cart_list + 3 * m, *b = B->cart_list + 3 * n;
            out[m * B->cart_size + n] = NA * NB * S[0][a[0]][b[0]] * S[1][a[1]][b[1]] * S[2][a[2]][b[2]];
        }
}

void orc_overlap(const orc_shell_t *A, const orc_shell_t *B, double *S)
{
    int sz = A->cart_size * B->cart_size;
    double *buf = calloc(sz, sizeof(double));
    memset(S, 0, sz * sizeof(double));
    for (int i = 0; i < A->num_primitives; i++)
        for (int j = 0; j < B->num_primitives; j++) {
            overlap_prim(A, B, A->alpha[i], B->alpha[j], buf);
            double f = A->coeffs[i] * B->coeffs[j];
            for (int k = 0; k < sz; k++) S[k] += f * buf[k];
        }
    free(buf);
}

/* ---------------- type-1, McMurchie-Davidson (type1_mcmurchie_davidson.c:129-644) ---------------- */
typedef struct {
    double A[3], B[3], C[3], Q[3], q, K[3];
    int LA, LB, N;
    double *E;       /* [3][LA+1][LB+1][N+1] */
    double *R;       /* [N+1]^4 : [t][u][v][n] */
    double aux[64];
} mmd_t;

#define EIX(d, c, i, j, t) ((d)->E[(((c) * ((d)->LA + 1) + (i)) * ((d)->LB + 1) + (j)) * ((d)->N + 1) + (t)])
#define RIX(d, t, u, v, n) ((d)->R[((((t) * ((d)->N + 1) + (u)) * ((d)->N + 1) + (v)) * ((d)->N + 1)) + (n)])

/* Hermite expansion coefficients E^{ij}_t (type1_mcmurchie_davidson.c:591-644) */
static void setup_E(mmd_t *d)
{
    double q = d->q;
    for (int c = 0; c < 3; c++)
        for (int i = 0; i <= d->LA; i++)
            for (int j = 0; j <= d->LB; j++)
                for (int t = 0; t <= i + j; t++) {
                    if (t == 0) {
                        if (i == 0 && j == 0) EIX(d, c, 0, 0, 0) = d->K[c];
                        else if (i == 1 && j == 0) EIX(d, c, 1, 0, 0) = (d->Q[c] - d->A[c]) * EIX(d, c, 0, 0, 0);
                        else if (i == 0 && j == 1) EIX(d, c, 0, 1, 0) = (d->Q[c] - d->B[c]) * EIX(d, c, 0, 0, 0);
                        else if (i == 0)
                            EIX(d, c, i, j, 0) = (d->Q[c] - d->B[c]) * EIX(d, c, i, j - 1, 0) + EIX(d, c, i, j - 1, 1);
                        else
                            EIX(d, c, i, j, 0) = (d->Q[c] - d->A[c]) * EIX(d, c, i - 1, j, 0) + EIX(d, c, i - 1, j, 1);
                    } else {
                        double e = 0.0, f = 1.0 / (2.0 * q * t);
                        if (i > 0) e += f * i * EIX(d, c, i - 1, j, t - 1);
                        if (j > 0) e += f * j * EIX(d, c, i, j - 1, t - 1);
                        EIX(d, c, i, j, t) = e;
                    }
                }
}

/* R^n_{tuv} for 1/r (sign -2q) or G^n_{tuv} for 1/r^2 (extra -2q R^n terms)
 * (type1_mcmurchie_davidson.c:479-585) */
static void setup_RG(mmd_t *d, int is_g)
{
    double q = d->q, X = d->Q[0] - d->C[0], Y = d->Q[1] - d->C[1], Z = d->Q[2] - d->C[2];
    int N = d->N;
    for (int t = 0; t <= N; t++)
        for (int u = 0; u <= N; u++)
            for (int v = 0; v <= N; v++) {
                if (t + u + v > N) continue;
                for (int n = 0; n <= N - (t + u + v); n++) {
                    double val = 0.0;
                    if (!is_g) {
                        if (t + u + v == 0) val = pow(-2.0 * q, n) * d->aux[n];
                        else if (t + u == 0) {
                            if (v > 1) val += (v - 1) * RIX(d, t, u, v - 2, n + 1);
                            val += Z * RIX(d, t, u, v - 1, n + 1);
                        } else if (t == 0) {
                            if (u > 1) val += (u - 1) * RIX(d, t, u - 2, v, n + 1);
                            val += Y * RIX(d, t, u - 1, v, n + 1);
                        } else {
                            if (t > 1) val += (t - 1) * RIX(d, t - 2, u, v, n + 1);
                            val += X * RIX(d, t - 1, u, v, n + 1);
                        }
                    } else {
                        if (t + u + v == 0) val = pow(2.0 * q, n) * d->aux[n];
                        else if (t + u == 0) {
                            if (v > 1) val += (v - 1) * (RIX(d, t, u, v - 2, n + 1) - 2.0 * q * RIX(d, t, u, v - 2, n));
                            val += Z * (RIX(d, t, u, v - 1, n + 1) - 2.0 * q * RIX(d, t, u, v - 1, n));
                        } else if (t == 0) {
                            if (u > 1) val += (u - 1) * (RIX(d, t, u - 2, v, n + 1) - 2.0 * q * RIX(d, t, u - 2, v, n));
                            val += Y * (RIX(d, t, u - 1, v, n + 1) - 2.0 * q * RIX(d, t, u - 1, v, n));
                        } else {
                            if (t > 1) val += (t - 1) * (RIX(d, t - 2, u, v, n + 1) - 2.0 * q * RIX(d, t - 2, u, v, n));
                            val += X * (RIX(d, t - 1, u, v, n + 1) - 2.0 * q * RIX(d, t - 1, u, v, n));
                        }
                    }
                    RIX(d, t, u, v, n) = val;
                }
            }
}

/* one primitive pair x one potential Gaussian r^{n-2} exp(-c r^2), n in {0,1,2}
 * (type1_mcmurchie_davidson.c:194-473) */
static void mmd_prim(const orc_shell_t *SA, double a, const orc_shell_t *SB, double b, const double *C, double c,
                     int npow, double *out)
{
    mmd_t d;
    d.LA = SA->L;
    d.LB = SB->L;
    d.N = d.LA + d.LB;
    double q = a + b + c, muAB = a * b / q, muAC = a * c / q, muBC = b * c / q;
    for (int i = 0; i < 3; i++) {
        d.A[i] = SA->origin[i];
        d.B[i] = SB->origin[i];
        d.C[i] = C[i];
        d.Q[i] = (a * SA->origin[i] + b * SB->origin[i] + c * C[i]) / q;
        double XAB = SA->origin[i] - SB->origin[i], XAC = SA->origin[i] - C[i], XBC = SB->origin[i] - C[i];
        d.K[i] = exp(-muAB * XAB * XAB) * exp(-muAC * XAC * XAC) * exp(-muBC * XBC * XBC);
    }
    d.q = q;
    double dx = d.Q[0] - d.C[0], dy = d.Q[1] - d.C[1], dz = d.Q[2] - d.C[2];
    double RQC2 = dx * dx + dy * dy + dz * dz;
    int N1 = d.N + 1;
    d.E = (double *) calloc((size_t) 3 * (d.LA + 1) * (d.LB + 1) * N1, sizeof(double));
    d.R = NULL;
    if (npow == 1) {
        for (int n = 0; n <= d.N; n++) d.aux[n] = orc_boys(n, q * RQC2);
    } else if (npow == 0) {
        orc_gfun_values(q * RQC2, d.N, d.aux);
    }
    setup_E(&d);
    if (npow != 2) {
        d.R = (double *) calloc((size_t) N1 * N1 * N1 * N1, sizeof(double));
        setup_RG(&d, npow == 0);
    }
    double NA = orc_gaussian_norm_factor(d.LA, 0, 0, a), NB = orc_gaussian_norm_factor(d.LB, 0, 0, b);
    for (int m = 0; m < SA->cart_size; m++)
        for (int n = 0; n < SB->cart_size; n++) {
            const int *ia = SA->cart_list + 3 * m, *ib = SB->cart_list + 3 * n;
            double s;
            if (npow == 2) {
                s = NA * NB * EIX(&d, 0, ia[0], ib[0], 0) * EIX(&d, 1, ia[1], ib[1], 0) * EIX(&d, 2, ia[2], ib[2], 0) *
                    pow(M_PI / q, 1.5);
            } else {
                s = 0.0;
                for (int t = 0; t <= ia[0] + ib[0]; t++) {
                    double Ex = EIX(&d, 0, ia[0], ib[0], t);
                    for (int u = 0; u <= ia[1] + ib[1]; u++) {
                        double Ey = EIX(&d, 1, ia[1], ib[1], u);
                        for (int v = 0; v <= ia[2] + ib[2]; v++) {
                            double Ez = EIX(&d, 2, ia[2], ib[2], v);
                            s += Ex * Ey * Ez * RIX(&d, t, u, v, 0);
                        }
                    }
                }
                if (npow == 1) s *= NA * NB * 2.0 * M_PI / q;
                else s *= NA * NB * 2.0 * pow(M_PI, 1.5) / sqrt(q);
            }
            out[m * SB->cart_size + n] = s;
        }
    free(d.E);
    free(d.R);
}

static void mmd_contracted(const orc_shell_t *A, const orc_shell_t *B, const double *C, double zeta, int npow,
                           double *out)
{
    if (npow < 0 || npow > 2) abort();   /* type1_mcmurchie_davidson.c:134 asserts n in {0,1,2} */
    int sz = A->cart_size * B->cart_size;
    double *buf = calloc(sz, sizeof(double));
    memset(out, 0, sz * sizeof(double));
    for (int i = 0; i < A->num_primitives; i++) {
        if (fabs(A->coeffs[i]) < ORC_ZERO_THRESH) continue;
        for (int j = 0; j < B->num_primitives; j++) {
            if (fabs(B->coeffs[j]) < ORC_ZERO_THRESH) continue;
            mmd_prim(A, A->alpha[i], B, B->alpha[j], C, zeta, npow, buf);
            double f = A->coeffs[i] * B->coeffs[j];
            for (int k = 0; k < sz; k++) out[k] += f * buf[k];
        }
    }
    free(buf);
}

/* one normalised primitive pair with an r^-1 exp(-zeta r_C^2) term (the point-charge nuclear
 * attraction with zeta = 0, nuclear_attraction.c:73-84) */
void orc_type1_prim_n1(const orc_shell_t *A, double a, const orc_shell_t *B, double b, const double *C, double zeta,
                       double *out)
{
    mmd_prim(A, a, B, b, C, zeta, 1, out);
}

/* sum_k d_k <A| r^{n_k-2} exp(-z_k r_C^2) |B> (type1_integrals.c:43-102) */
void orc_type1(const orc_shell_t *A, const orc_shell_t *B, const double *C, const orc_potential_t *pot, double *M)
{
    int sz = A->cart_size * B->cart_size;
    memset(M, 0, sz * sizeof(double));
    if (pot == NULL) return;
    double *buf = calloc(sz, sizeof(double));
    for (int k = 0; k < pot->num_primitives; k++) {
        mmd_contracted(A, B, C, pot->alpha[k], pot->powers[k], buf);
        for (int i = 0; i < sz; i++) M[i] += pot->coeffs[k] * buf[i];
    }
    free(buf);
}

/* ---------------- type-2 and spin-orbit (type2_integrals.c:34-243, spin_orbit_integrals.c:41-288) ---------------- */
#define RSH_W 64

typedef struct {
    double CA[3], CB[3], kA[3], kB[3];
    int L, LA, LB;
} semiloc_t;

static void semiloc_setup(semiloc_t *s, const orc_shell_t *A, const orc_shell_t *B, const double *C, int L)
{
    for (int i = 0; i < 3; i++) {
        s->CA[i] = C[i] - A->origin[i];
        s->CB[i] = C[i] - B->origin[i];
        s->kA[i] = -2.0 * (A->alpha[0] * s->CA[i]);
        s->kB[i] = -2.0 * (B->alpha[0] * s->CB[i]);
    }
    s->L = L;
    s->LA = A->cart_list[0] + A->cart_list[1] + A->cart_list[2];
    s->LB = B->cart_list[0] + B->cart_list[1] + B->cart_list[2];
}

static double q_at(const orc_radial_table_t *t, int l1, int l2, int n)
{
    return t->Q[(size_t) (t->lambda2_max + 1) * (t->n_max + 1) * l1 + (t->n_max + 1) * l2 + n];
}

void orc_type2(const orc_shell_t *A, const orc_shell_t *B, const double *C, const orc_potential_t *pot, double *M)
{
    orc_init();
    int nA = A->cart_size, nB = B->cart_size;
    memset(M, 0, sizeof(double) * nA * nB);
    semiloc_t s;
    semiloc_setup(&s, A, B, C, pot->L);
    int L = s.L, l1max = L + s.LA, l2max = L + s.LB, Nmax = s.LA + s.LB;
    int lmax = l1max > l2max ? l1max : l2max;
    if (L > lmax) lmax = L;
    double CA2 = s.CA[0] * s.CA[0] + s.CA[1] * s.CA[1] + s.CA[2] * s.CA[2];
    double CB2 = s.CB[0] * s.CB[0] + s.CB[1] * s.CB[1] + s.CB[2] * s.CB[2];
    orc_radial_table_t *tab = orc_tabulate_radial(l1max, l2max, Nmax, CA2, CB2, pot, A, B);
    double (*rA)[RSH_W] = malloc(sizeof(double) * RSH_W * (lmax + 1));
    double (*rB)[RSH_W] = malloc(sizeof(double) * RSH_W * (lmax + 1));
    for (int l = 0; l <= lmax; l++) {
        orc_rsh_values(l, s.kA, rA[l]);
        orc_rsh_values(l, s.kB, rB[l]);
    }
    for (int ic = 0; ic < nA; ic++)
        for (int jc = 0; jc < nB; jc++) {
            int na = A->cart_list[3 * ic], la = A->cart_list[3 * ic + 1], ma = A->cart_list[3 * ic + 2];
            int nb = B->cart_list[3 * jc], lb = B->cart_list[3 * jc + 1], mb = B->cart_list[3 * jc + 2];
            double gamma = 0.0;
            for (int a = 0; a <= na; a++) {
                double fa = (double) orc_binomial(na, a) * 1.0, pa = pw(s.CA[0], na - a);
                for (int b = 0; b <= la; b++) {
                    double fb = (double) orc_binomial(la, b), pb = pw(s.CA[1], la - b);
                    for (int c = 0; c <= ma; c++) {
                        double fc = (double) orc_binomial(ma, c), pc = pw(s.CA[2], ma - c);
                        for (int d = 0; d <= nb; d++) {
                            double fd = (double) orc_binomial(nb, d), pd = pw(s.CB[0], nb - d);
                            for (int e = 0; e <= lb; e++) {
                                double fe = (double) orc_binomial(lb, e), pe = pw(s.CB[1], lb - e);
                                for (int f = 0; f <= mb; f++) {
                                    double ff = (double) orc_binomial(mb, f), pf = pw(s.CB[2], mb - f);
                                    double factor = fa * fb * fc * fd * fe * ff * pa * pb * pc * pd * pe * pf;
                                    if (fabs(factor) < 1e-13) continue;
                                    int N = a + b + c + d + e + f;
                                    double som = 0.0;
                                    int l1lo = L - a - b - c > 0 ? L - a - b - c : 0;
                                    int l2lo = L - d - e - f > 0 ? L - d - e - f : 0;
                                    for (int l1 = l1lo; l1 <= L + a + b + c; l1++) {
                                        if ((L + a + b + c - l1) % 2 != 0) continue;
                                        for (int l2 = l2lo; l2 <= L + d + e + f; l2++) {
                                            if ((L + d + e + f - l2) % 2 != 0) continue;
                                            double QN = q_at(tab, l1, l2, N);
                                            if (fabs(QN) < 1e-16) continue;
                                            double sang = 0.0;
                                            for (int m = -L; m <= L; m++) {
                                                double o1 = orc_angular_type2(l1, L, m, a, b, c, rA[l1]);
                                                if (fabs(o1) < 1e-16) continue;
                                                double o2 = orc_angular_type2(l2, L, m, d, e, f, rB[l2]);
                                                sang += o1 * o2;
                                            }
                                            som += QN * sang;
                                        }
                                    }
                                    gamma += factor * som;
                                }
                            }
                        }
                    }
                }
            }
            gamma *= FOURPI2;
            M[ic * nB + jc] = gamma;
        }
    free(rA);
    free(rB);
    orc_free_radial(tab);
}

void orc_spin_orbit(const orc_shell_t *A, const orc_shell_t *B, const double *C, const orc_potential_t *pot,
                    double *X, double *Y, double *Z)
{
    orc_init();
    int nA = A->cart_size, nB = B->cart_size;
    memset(X, 0, sizeof(double) * nA * nB);
    memset(Y, 0, sizeof(double) * nA * nB);
    memset(Z, 0, sizeof(double) * nA * nB);
    semiloc_t s;
    semiloc_setup(&s, A, B, C, pot->L);
    int L = s.L, l1max = L + s.LA, l2max = L + s.LB, Nmax = s.LA + s.LB, dim = 2 * L + 1;
    double *Lx = calloc(dim * dim, sizeof(double)), *Ly = calloc(dim * dim, sizeof(double)),
           *Lz = calloc(dim * dim, sizeof(double));
    orc_lmatrix_rsh(L, Lx, Ly, Lz);
    int lmax = l1max > l2max ? l1max : l2max;
    if (L > lmax) lmax = L;
    double (*rA)[RSH_W] = malloc(sizeof(double) * RSH_W * (lmax + 1));
    double (*rB)[RSH_W] = malloc(sizeof(double) * RSH_W * (lmax + 1));
    for (int l = 0; l <= lmax; l++) {
        orc_rsh_values(l, s.kA, rA[l]);
        orc_rsh_values(l, s.kB, rB[l]);
    }
    double CA2 = s.CA[0] * s.CA[0] + s.CA[1] * s.CA[1] + s.CA[2] * s.CA[2];
    double CB2 = s.CB[0] * s.CB[0] + s.CB[1] * s.CB[1] + s.CB[2] * s.CB[2];
    orc_radial_table_t *tab = orc_tabulate_radial(l1max, l2max, Nmax, CA2, CB2, pot, A, B);
    for (int ic = 0; ic < nA; ic++)
        for (int jc = 0; jc < nB; jc++) {
            int na = A->cart_list[3 * ic], la = A->cart_list[3 * ic + 1], ma = A->cart_list[3 * ic + 2];
            int nb = B->cart_list[3 * jc], lb = B->cart_list[3 * jc + 1], mb = B->cart_list[3 * jc + 2];
            double sx = 0, sy = 0, sz = 0;
            for (int a = 0; a <= na; a++) {
                double fa = (double) orc_binomial(na, a), pa = pw(s.CA[0], na - a);
                for (int b = 0; b <= la; b++) {
                    double fb = (double) orc_binomial(la, b), pb = pw(s.CA[1], la - b);
                    for (int c = 0; c <= ma; c++) {
                        double fc = (double) orc_binomial(ma, c), pc = pw(s.CA[2], ma - c);
                        for (int d = 0; d <= nb; d++) {
                            double fd = (double) orc_binomial(nb, d), pd = pw(s.CB[0], nb - d);
                            for (int e = 0; e <= lb; e++) {
                                double fe = (double) orc_binomial(lb, e), pe = pw(s.CB[1], lb - e);
                                for (int f = 0; f <= mb; f++) {
                                    double ff = (double) orc_binomial(mb, f), pf = pw(s.CB[2], mb - f);
                                    int N = a + b + c + d + e + f;
                                    double factor = fa * fb * fc * fd * fe * ff * pa * pb * pc * pd * pe * pf;
                                    if (fabs(factor) < ORC_ZERO_THRESH) continue;
                                    double qx = 0, qy = 0, qz = 0;
                                    int l1lo = L - a - b - c > 0 ? L - a - b - c : 0;
                                    int l2lo = L - d - e - f > 0 ? L - d - e - f : 0;
                                    for (int l1 = l1lo; l1 <= L + a + b + c; l1++) {
                                        if ((L + a + b + c - l1) % 2 != 0) continue;
                                        for (int l2 = l2lo; l2 <= L + d + e + f; l2++) {
                                            if ((L + d + e + f - l2) % 2 != 0) continue;
                                            double QN = q_at(tab, l1, l2, N);
                                            if (fabs(QN) < ORC_ZERO_THRESH) continue;
                                            double ax = 0, ay = 0, az = 0;
                                            for (int m1 = -L; m1 <= L; m1++)
                                                for (int m2 = -L; m2 <= L; m2++) {
                                                    double lx = Lx[dim * (m1 + L) + m2 + L];
                                                    double ly = Ly[dim * (m1 + L) + m2 + L];
                                                    double lz = Lz[dim * (m1 + L) + m2 + L];
                                                    if (fabs(lx) < ORC_ZERO_THRESH && fabs(ly) < ORC_ZERO_THRESH &&
                                                        fabs(lz) < ORC_ZERO_THRESH)
                                                        continue;
                                                    double o1 = orc_angular_type2(l1, L, m1, a, b, c, rA[l1]);
                                                    if (fabs(o1) < ORC_ZERO_THRESH) continue;
                                                    double o2 = orc_angular_type2(l2, L, m2, d, e, f, rB[l2]);
                                                    ax += o1 * o2 * lx;
                                                    ay += o1 * o2 * ly;
                                                    az += o1 * o2 * lz;
                                                }
                                            qx += QN * ax;
                                            qy += QN * ay;
                                            qz += QN * az;
                                        }
                                    }
                                    sx += factor * qx;
                                    sy += factor * qy;
                                    sz += factor * qz;
                                }
                            }
                        }
                    }
                }
            }
            X[ic * nB + jc] = sx * FOURPI2;
            Y[ic * nB + jc] = sy * FOURPI2;
            Z[ic * nB + jc] = sz * FOURPI2;
        }
    orc_free_radial(tab);
    free(Lx); free(Ly); free(Lz);
    free(rA); free(rB);
}

/* ---------------- outercore (outercore_integrals.c:83-553) ---------------- */

/* normalisation of r^L exp(-a r^2): sqrt(2^{L+2} (2a)^{L+3/2} / ((2L+1)!! sqrt(pi)))
 * (outercore_integrals.c:472-500 tabulates the a-independent factor) */
static double radial_gto_norm(int L, double a)
{
    double f = sqrt(pow(2.0, L + 2) * pow(2.0, L + 1.5) / (orc_double_factorial(2 * L + 1) * sqrt(M_PI)));
    return f * pow(a, 0.75 + L / 2.0);
}

/* int r^{2L+n} exp(-(a1+a2+z) r^2) dr without normalisation (outercore_integrals.c:454-465) */
static double onecenter_prim(int L, double a1, double a2, int n, double z)
{
    double a = a1 + a2 + z;
    if (n % 2 == 0) {
        int k = L + n / 2;
        return orc_double_factorial(2 * k - 1) / (pow(2.0, k + 1) * pow(a, k)) * sqrt(M_PI / a);
    }
    int k = L + (n - 1) / 2;
    return orc_factorial(k) / (2.0 * pow(a, k + 1));
}

static double onecenter_contracted(const orc_shell_t *bra, const orc_shell_t *ket, const orc_potential_t *pot)
{
    double sum = 0.0;
    int L = pot->L;
    for (int i = 0; i < bra->num_primitives; i++) {
        double Ni = radial_gto_norm(L, bra->alpha[i]);
        for (int j = 0; j < ket->num_primitives; j++) {
            double Nj = radial_gto_norm(L, ket->alpha[j]);
            double f = Ni * Nj * bra->coeffs[i] * ket->coeffs[j];
            for (int k = 0; k < pot->num_primitives; k++)
                sum += f * pot->coeffs[k] * onecenter_prim(L, bra->alpha[i], ket->alpha[j], pot->powers[k], pot->alpha[k]);
        }
    }
    return sum;
}

static double ang_norm_factor(int L) { return 1.0 / (2.0 * sqrt(M_PI)) * sqrt(orc_double_factorial(2 * L + 1)); }

/* C += A(MxK) * B(KxN) (utils.c:64-75) */
static void gemm_acc(int M, int N, int K, const double *A, const double *B, double *C)
{
    for (int i = 0; i < M; i++)
        for (int j = 0; j < N; j++) {
            double s = 0.0;
            for (int k = 0; k < K; k++) s += A[i * K + k] * B[k * N + j];
            C[i * N + j] += s;
        }
}

/* S_lm coefficients of the OC shell: rows m, columns Cartesian components */
static double *slm_coefs(const orc_shell_t *oc)
{
    int L = oc->L, n = oc->cart_size;
    double *S = calloc((2 * L + 1) * n, sizeof(double));
    for (int m = -L; m <= L; m++)
        for (int c = 0; c < n; c++) {
            const int *t = oc->cart_list + 3 * c;
            S[n * (m + L) + c] = orc_rsh_coef(L, m, t[0], t[1], t[2]) / ang_norm_factor(L);
        }
    return S;
}

/* <A|X|oc_cart> (nA x ncart) -> (nA x (2L+1)) and <oc_cart|X|B> (ncart x nB) -> ((2L+1) x nB)
 * (outercore_integrals.c:508-542) */
static void to_sph_ket(int nA, int nc, int ns, const double *in, double *out, const double *S)
{
    for (int i = 0; i < nA; i++)
        for (int j = 0; j < ns; j++) {
            double s = 0.0;
            for (int c = 0; c < nc; c++) s += S[nc * j + c] * in[i * nc + c];
            out[i * ns + j] = s;
        }
}
static void to_sph_bra(int nc, int ns, int nB, const double *in, double *out, const double *S)
{
    for (int j = 0; j < nB; j++)
        for (int i = 0; i < ns; i++) {
            double s = 0.0;
            for (int c = 0; c < nc; c++) s += S[nc * i + c] * in[c * nB + j];
            out[i * nB + j] = s;
        }
}

static void oc_part1(const orc_shell_t *A, const orc_shell_t *B, const double *C, const orc_potential_t *pot,
                     const orc_shell_t *oc, double *arep, double *X, double *Y, double *Z)
{
    int L = oc->L, nA = A->cart_size, nB = B->cart_size, no = oc->cart_size, ns = 2 * L + 1;
    double J = pot->J / 2.0;
    double *S = slm_coefs(oc);
    double *t1 = calloc(no * nB, sizeof(double)), *t2 = calloc(nA * no, sizeof(double));
    double *Sb = calloc(ns * nB, sizeof(double)), *Sa = calloc(nA * ns, sizeof(double));
    double *Ub = calloc(ns * nB, sizeof(double)), *Ua = calloc(nA * ns, sizeof(double));
    orc_overlap(oc, B, t1);
    orc_overlap(A, oc, t2);
    to_sph_ket(nA, no, ns, t2, Sa, S);
    to_sph_bra(no, ns, nB, t1, Sb, S);
    orc_type1(oc, B, C, pot, t1);
    orc_type1(A, oc, C, pot, t2);
    to_sph_ket(nA, no, ns, t2, Ua, S);
    to_sph_bra(no, ns, nB, t1, Ub, S);
    double af = (J < L) ? (L / (2.0 * L + 1)) : ((L + 1) / (2.0 * L + 1));
    double *buf = calloc(nA * nB, sizeof(double));
    gemm_acc(nA, nB, ns, Sa, Ub, buf);
    gemm_acc(nA, nB, ns, Ua, Sb, buf);
    for (int i = 0; i < nA * nB; i++) arep[i] += af * buf[i];
    free(buf);
    double *Lm[3];
    for (int k = 0; k < 3; k++) Lm[k] = calloc(ns * ns, sizeof(double));
    orc_lmatrix_rsh(L, Lm[0], Lm[1], Lm[2]);
    double ef = (J < L) ? (-2.0 / (2 * L + 1)) : (+2.0 / (2 * L + 1));
    double *outs[3] = {X, Y, Z};
    int mx = nA > nB ? nA : nB;
    double *tb = calloc(ns * mx, sizeof(double)), *so = calloc(nA * nB, sizeof(double));
    for (int k = 0; k < 3; k++) {
        memset(so, 0, sizeof(double) * nA * nB);
        memset(tb, 0, sizeof(double) * ns * nB);
        gemm_acc(ns, nB, ns, Lm[k], Sb, tb);
        gemm_acc(nA, nB, ns, Ua, tb, so);
        memset(tb, 0, sizeof(double) * ns * nA);
        gemm_acc(nA, ns, ns, Sa, Lm[k], tb);
        gemm_acc(nA, nB, ns, tb, Ub, so);
        for (int i = 0; i < nA * nB; i++) outs[k][i] += ef * so[i];
    }
    for (int k = 0; k < 3; k++) free(Lm[k]);
    free(tb); free(so); free(S); free(t1); free(t2); free(Sa); free(Sb); free(Ua); free(Ub);
}

static void oc_part2(const orc_shell_t *A, const orc_shell_t *B, const orc_potential_t *p1, const orc_shell_t *o1,
                     const orc_potential_t *p2, const orc_shell_t *o2, double *arep, double *X, double *Y, double *Z)
{
    if (p1->L != p2->L || p1->J != p2->J) return;
    int L = p1->L, nA = A->cart_size, nB = B->cart_size, no = o1->cart_size, ns = 2 * L + 1;
    double J = p1->J / 2.0;
    double delta = 0.5 * (onecenter_contracted(o1, o2, p1) + onecenter_contracted(o1, o2, p2));
    double *S = slm_coefs(o1);
    double *t1 = calloc(no * nB, sizeof(double)), *t2 = calloc(nA * no, sizeof(double));
    double *Sb = calloc(ns * nB, sizeof(double)), *Sa = calloc(nA * ns, sizeof(double));
    orc_overlap(o2, B, t1);
    orc_overlap(A, o1, t2);
    to_sph_ket(nA, no, ns, t2, Sa, S);
    to_sph_bra(no, ns, nB, t1, Sb, S);
    double af = (J < L) ? (L / (2.0 * L + 1)) : ((L + 1) / (2.0 * L + 1));
    double *buf = calloc(nA * nB, sizeof(double));
    gemm_acc(nA, nB, ns, Sa, Sb, buf);
    for (int i = 0; i < nA * nB; i++) arep[i] += (-1.0) * delta * af * buf[i];
    free(buf);
    double *Lm[3];
    for (int k = 0; k < 3; k++) Lm[k] = calloc(ns * ns, sizeof(double));
    orc_lmatrix_rsh(L, Lm[0], Lm[1], Lm[2]);
    double ef = (J < L) ? (-2.0 / (2 * L + 1)) : (+2.0 / (2 * L + 1));
    double *outs[3] = {X, Y, Z};
    double *tb = calloc(ns * nB, sizeof(double)), *so = calloc(nA * nB, sizeof(double));
    for (int k = 0; k < 3; k++) {
        memset(tb, 0, sizeof(double) * ns * nB);
        memset(so, 0, sizeof(double) * nA * nB);
        gemm_acc(ns, nB, ns, Lm[k], Sb, tb);
        gemm_acc(nA, nB, ns, Sa, tb, so);
        for (int i = 0; i < nA * nB; i++) outs[k][i] += (-1.0) * delta * ef * so[i];
    }
    for (int k = 0; k < 3; k++) free(Lm[k]);
    free(tb); free(so); free(S); free(t1); free(t2); free(Sa); free(Sb);
}

/* The outercore shells are bound to the GRPP centre C.  The reference does so by writing
 * C into the caller's shells (outercore_integrals.c:107-111); the oracle works on copies. */
void orc_outercore(const orc_shell_t *A, const orc_shell_t *B, const double *C, int n_oc,
                   orc_potential_t *const *pots, orc_shell_t *const *shells, double *arep, double *X, double *Y,
                   double *Z)
{
    orc_init();
    int sz = A->cart_size * B->cart_size;
    memset(arep, 0, sizeof(double) * sz);
    memset(X, 0, sizeof(double) * sz);
    memset(Y, 0, sizeof(double) * sz);
    memset(Z, 0, sizeof(double) * sz);
    orc_shell_t **oc = malloc(sizeof(orc_shell_t *) * (n_oc > 0 ? n_oc : 1));
    for (int i = 0; i < n_oc; i++) {
        oc[i] = orc_shell_copy(shells[i]);
        for (int k = 0; k < 3; k++) oc[i]->origin[k] = C[k];
    }
    for (int i = 0; i < n_oc; i++) oc_part1(A, B, C, pots[i], oc[i], arep, X, Y, Z);
    for (int i = 0; i < n_oc; i++)
        for (int j = 0; j < n_oc; j++) oc_part2(A, B, pots[i], oc[i], pots[j], oc[j], arep, X, Y, Z);
    for (int i = 0; i < n_oc; i++) orc_delete_shell(oc[i]);
    free(oc);
}

/* arep = T1(U_L) + sum T2 + OC; so = sum 2/(2L+1) SO + OC (full_grpp_integrals.c:27-104) */
void orc_full_grpp(const orc_shell_t *A, const orc_shell_t *B, const orc_grpp_t *g, const double *C, double *arep,
                   double *X, double *Y, double *Z)
{
    int sz = A->cart_size * B->cart_size;
    double *b0 = calloc(sz, sizeof(double)), *bx = calloc(sz, sizeof(double)), *by = calloc(sz, sizeof(double)),
           *bz = calloc(sz, sizeof(double));
    memset(arep, 0, sizeof(double) * sz);
    memset(X, 0, sizeof(double) * sz);
    memset(Y, 0, sizeof(double) * sz);
    memset(Z, 0, sizeof(double) * sz);
    orc_type1(A, B, C, g->U_L, b0);
    for (int i = 0; i < sz; i++) arep[i] += 1.0 * b0[i];
    for (int l = 0; l < g->n_arep; l++) {
        orc_type2(A, B, C, g->U_arep[l], b0);
        for (int i = 0; i < sz; i++) arep[i] += 1.0 * b0[i];
    }
    for (int l = 0; l < g->n_esop; l++) {
        orc_spin_orbit(A, B, C, g->U_esop[l], bx, by, bz);
        int L = g->U_esop[l]->L;
        for (int i = 0; i < sz; i++) {
            X[i] += 2.0 / (2 * L + 1) * bx[i];
            Y[i] += 2.0 / (2 * L + 1) * by[i];
            Z[i] += 2.0 / (2 * L + 1) * bz[i];
        }
    }
    orc_outercore(A, B, C, g->n_oc_shells, g->U_oc, g->oc_shells, b0, bx, by, bz);
    for (int i = 0; i < sz; i++) {
        arep[i] += 1.0 * b0[i];
        X[i] += 1.0 * bx[i];
        Y[i] += 1.0 * by[i];
        Z[i] += 1.0 * bz[i];
    }
    free(b0); free(bx); free(by); free(bz);
}
