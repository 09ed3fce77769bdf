/*
 * Oracle (test infrastructure): angular algebra of the GRPP path --
 * real spherical harmonics, type-2 angular integrals, angular-momentum matrices.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

#define RSH_LMAX 40

typedef struct {
    int ncomb;
    int *comb;        /* (i,j,k), i ascending, then j, then k (spherical_harmonics.c:283-304) */
    double *coef;     /* [(m+L)*ncomb + icomb] */
} rsh_table_t;

static rsh_table_t rsh_tabs[RSH_LMAX + 1];
static int rsh_ready = 0;

static double binom_s(int n, int k)
{
    if (n < 0 || k < 0 || k > n) return 0.0;
    return (double) orc_binomial((uint64_t) n, (uint64_t) k);
}

static double sign_pow(int e) { return (e % 2 == 0) ? 1.0 : -1.0; }

/*
 * Coefficient of x^lx y^ly z^lz (on the unit sphere) in the real spherical harmonic S_lm,
 * Flores-Moreno et al., J. Comput. Chem. 27, 1009 (2006), Eq. 32, with libgrpp's
 * normalisation including the extra 1/sqrt(2) for m = 0 (spherical_harmonics.c:115-166).
 */
double orc_rsh_coef(int l, int m, int lx, int ly, int lz)
{
    int am = abs(m);
    int j = lx + ly - am;
    if (j % 2 != 0) return 0.0;
    j /= 2;
    if (!((m > 0 && (am - lx) % 2 == 0) || (m == 0 && lx % 2 == 0) || (m < 0 && (am - lx) % 2 != 0)))
        return 0.0;

    double pref = sqrt((2 * l + 1) / (2 * M_PI) * orc_factorial(l - am) / orc_factorial(l + am));
    pref /= pow(2, l) * orc_factorial(l);

    double u = 0.0;
    for (int i = j; i <= (l - am) / 2; i++) {
        if (2 * l - 2 * i < 0 || l - am - 2 * i < 0) { u = 0.0; break; }
        double f1 = binom_s(l, i) * binom_s(i, j) * sign_pow(i) * orc_factorial_ratio(2 * l - 2 * i, l - am - 2 * i);
        double s = 0.0;
        for (int k = 0; k <= j; k++) {
            /* C integer division truncating toward zero, as the reference's pow(-1, (|m|-lx+2k)/2) */
            s += binom_s(j, k) * binom_s(am, lx - 2 * k) * sign_pow((am - lx + 2 * k) / 2);
        }
        u += f1 * s;
    }
    u *= pref;
    if (m == 0 && (lx % 2 == 0)) u *= M_SQRT1_2;
    return u;
}

static void rsh_build(void)
{
#pragma omp critical(orc_rsh_lock)
    {
        if (!rsh_ready) {
            for (int L = 0; L <= RSH_LMAX; L++) {
                int nc = (L + 1) * (L + 2) / 2;
                rsh_tabs[L].ncomb = nc;
                rsh_tabs[L].comb = (int *) malloc(sizeof(int) * 3 * nc);
                rsh_tabs[L].coef = (double *) malloc(sizeof(double) * (2 * L + 1) * nc);
                int n = 0;
                for (int i = 0; i <= L; i++)
                    for (int jj = 0; jj <= L; jj++)
                        for (int k = 0; k <= L; k++)
                            if (i + jj + k == L) {
                                rsh_tabs[L].comb[3 * n] = i;
                                rsh_tabs[L].comb[3 * n + 1] = jj;
                                rsh_tabs[L].comb[3 * n + 2] = k;
                                n++;
                            }
                for (int m = -L; m <= L; m++)
                    for (int c = 0; c < nc; c++)
                        rsh_tabs[L].coef[(m + L) * nc + c] =
                            orc_rsh_coef(L, m, rsh_tabs[L].comb[3 * c], rsh_tabs[L].comb[3 * c + 1],
                                         rsh_tabs[L].comb[3 * c + 2]);
            }
            rsh_ready = 1;
        }
    }
}

/* S_lm(k/|k|) for m = -l..l; |k| <= 1e-14 gives the zero vector (spherical_harmonics.c:227-280). */
void orc_rsh_values(int l, const double *k, double *out)
{
    rsh_build();
    double ux = 0, uy = 0, uz = 0;
    double len = sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
    if (len > ORC_ZERO_THRESH) {
        double inv = 1.0 / len;
        ux = k[0] * inv;
        uy = k[1] * inv;
        uz = k[2] * inv;
    }
    double px[64], py[64], pz[64];
    px[0] = py[0] = pz[0] = 1.0;
    for (int i = 1; i <= l; i++) {
        px[i] = px[i - 1] * ux;
        py[i] = py[i - 1] * uy;
        pz[i] = pz[i - 1] * uz;
    }
    memset(out, 0, sizeof(double) * (2 * l + 1));
    rsh_table_t *t = &rsh_tabs[l];
    for (int c = 0; c < t->ncomb; c++) {
        double kk = px[t->comb[3 * c]] * py[t->comb[3 * c + 1]] * pz[t->comb[3 * c + 2]];
        for (int m = -l; m <= l; m++) out[m + l] += t->coef[(m + l) * t->ncomb + c] * kk;
    }
}

/* 4 pi (i-1)!!(j-1)!!(k-1)!!/(i+j+k+1)!! for all-even exponents, else 0 (angular_integrals.c:118-130) */
double orc_unit_sphere_integral(int i, int j, int k)
{
    if (i % 2 == 0 && j % 2 == 0 && k % 2 == 0)
        return 4 * M_PI * orc_double_factorial(i - 1) * orc_double_factorial(j - 1) * orc_double_factorial(k - 1)
               / orc_double_factorial(i + j + k + 1);
    return 0.0;
}

/* Omega^lambda_{Lm}(a,b,c; k) of McMurchie & Davidson (1981) Eq. 29 (angular_integrals.c:63-111). */
double orc_angular_type2(int lambda, int L, int m, int a, int b, int c, const double *rsh_values)
{
    rsh_build();
    rsh_table_t *tl = &rsh_tabs[lambda], *tL = &rsh_tabs[L];
    double sum = 0.0;
    for (int mu = -lambda; mu <= lambda; mu++) {
        double sk = rsh_values[mu + lambda];
        if (fabs(sk) < ORC_ZERO_THRESH) continue;
        double s2 = 0.0;
        for (int iu = 0; iu < tL->ncomb; iu++) {
            int u = tL->comb[3 * iu], v = tL->comb[3 * iu + 1], w = tL->comb[3 * iu + 2];
            double yL = tL->coef[(m + L) * tL->ncomb + iu];
            if (fabs(yL) < ORC_ZERO_THRESH) continue;
            for (int ir = 0; ir < tl->ncomb; ir++) {
                int r = tl->comb[3 * ir], s = tl->comb[3 * ir + 1], t = tl->comb[3 * ir + 2];
                double yl = tl->coef[(mu + lambda) * tl->ncomb + ir];
                if (fabs(yl) < ORC_ZERO_THRESH) continue;
                s2 += yl * yL * orc_unit_sphere_integral(a + r + u, b + s + v, c + t + w);
            }
        }
        sum += s2 * sk;
    }
    return sum;
}

/* Real-harmonic Lx, Ly, Lz matrices (lmatrix.c:28-141): the complex-harmonic ladder matrices
 * transformed with S_lm = a Y_{l,-m} + b Y_{l,m}; stored as Im, -Re, Im respectively. */
void orc_lmatrix_rsh(int L, double *lxm, double *lym, double *lzm)
{
    int dim = 2 * L + 1;
    double *cx = calloc(dim * dim, sizeof(double)), *cy = calloc(dim * dim, sizeof(double)),
           *cz = calloc(dim * dim, sizeof(double));
    for (int m1 = -L; m1 <= L; m1++)
        for (int m2 = -L; m2 <= L; m2++) {
            double lz = m2 * (m1 == m2);
            double lp = sqrt((L - m2) * (L + m2 + 1)) * (m1 == m2 + 1);
            double lm = sqrt((L + m2) * (L - m2 + 1)) * (m1 == m2 - 1);
            cx[(m1 + L) * dim + m2 + L] = 0.5 * (lp + lm);
            cy[(m1 + L) * dim + m2 + L] = 0.5 * (lp - lm);
            cz[(m1 + L) * dim + m2 + L] = lz;
        }
    for (int m1 = -L; m1 <= L; m1++)
        for (int m2 = -L; m2 <= L; m2++) {
            double complex ab[2][2];
            int ms[2] = {m1, m2};
            for (int q = 0; q < 2; q++) {
                int m = ms[q];
                if (m == 0) { ab[q][0] = 0.5; ab[q][1] = 0.5; }
                else if (m < 0) { ab[q][0] = I / sqrt(2); ab[q][1] = -I / sqrt(2) * sign_pow(abs(m)); }
                else { ab[q][0] = 1.0 / sqrt(2); ab[q][1] = 1.0 / sqrt(2) * sign_pow(m); }
            }
            int i1[2] = {-abs(m1) + L, abs(m1) + L}, i2[2] = {-abs(m2) + L, abs(m2) + L};
            double complex vx = 0, vy = 0, vz = 0;
            for (int p = 0; p < 2; p++)
                for (int q = 0; q < 2; q++) {
                    double complex f = conj(ab[0][p]) * ab[1][q];
                    vx += f * cx[dim * i1[p] + i2[q]];
                    vy += f * cy[dim * i1[p] + i2[q]];
                    vz += f * cz[dim * i1[p] + i2[q]];
                }
            lxm[(m1 + L) * dim + m2 + L] = cimag(vx);
            lym[(m1 + L) * dim + m2 + L] = -creal(vy);
            lzm[(m1 + L) * dim + m2 + L] = cimag(vz);
        }
    free(cx);
    free(cy);
    free(cz);
}
