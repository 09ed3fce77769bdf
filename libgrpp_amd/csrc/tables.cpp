// Constant tables of the GRPP engine, built on the host once per process.
//
//  * Log3 radial grid (C.-K. Skylaris et al., Chem. Phys. Lett. 296, 445 (1998)) with R = 5 in
//    nested-doubling order, as libgrpp generates it (radial_type2_integral.c:324-444), preceded by
//    one zero-weight padding point so every level starts on a multiple of 32 points;
//  * K_lambda(x) = exp(-x) i_lambda(x) values at x = 0..16 step 0.01 from the positive power
//    series in long double, and their first four derivatives by the order recursion
//    (specfunc_scaled_mod_sph_bessel.c:45-61,135-156); asymptotic coefficients (k+n)!/(k!(n-k)!2^k);
//  * Boys F_n(x) at x = 0..30 step 0.1 from the positive series (specfunc_boys.c:60 holds such a table);
//  * real spherical harmonic coefficients (Flores-Moreno et al. J. Comput. Chem. 27, 1009 (2006)
//    Eq. 32 with libgrpp's m = 0 convention, spherical_harmonics.c:115-166);
//  * real-harmonic angular-momentum matrices (lmatrix.c:28-141);
//  * angular coupling blocks W (see tables.h) from the coefficient tables and unit-sphere
//    monomial integrals (angular_integrals.c:63-130).
#include "tables.h"

#include <cmath>
#include <complex>
#include <cstdlib>
#include <mutex>

namespace grpp {

double fact(int n)
{
    if (n < 0) return 1.0;
    uint64_t e = 1;
    int k = 1;
    for (; k <= n && k <= 20; k++) e *= (uint64_t) k;
    double f = (double) e;
    for (; k <= n; k++) f *= k;
    return f;
}

static double fact_ratio(int n, int m)
{
    if (n == m) return 1.0;
    if (n < m) return 1.0 / fact_ratio(m, n);
    double p = 1.0;
    for (int i = m + 1; i <= n; i++) p *= i;
    return p;
}

double dfact(int n)
{
    if (n <= 0) return 1.0;
    uint64_t p = 1;
    for (int k = n; k > 1; k -= 2) p *= (uint64_t) k;
    return (double) p;
}

static double binom(int n, int k)
{
    if (n < 0 || k < 0 || k > n) return 0.0;
    if (k > n - k) k = n - k;
    uint64_t r = 1;
    for (int d = 1; d <= k; d++) r = r * (uint64_t) (n - k + d) / (uint64_t) d;
    return (double) r;
}

static double sgn(int e) { return (e % 2 == 0) ? 1.0 : -1.0; }

double gaussian_norm(int L, double a) { return std::pow(2 * a / M_PI, 0.75) * std::pow(4 * a, 0.5 * L); }

double rsh_coef(int l, int m, int lx, int ly, int lz)
{
    (void) lz;
    int am = std::abs(m);
    int j = lx + ly - am;
    if (j % 2 != 0) return 0.0;
    j /= 2;
    if (!((m > 0 && (am - lx) % 2 == 0) || (m == 0 && lx % 2 == 0) || (m < 0 && (am - lx) % 2 != 0))) return 0.0;
    double pref = std::sqrt((2 * l + 1) / (2 * M_PI) * fact(l - am) / fact(l + am));
    pref /= std::pow(2, l) * fact(l);
    double u = 0.0;
    for (int i = j; i <= (l - am) / 2; i++) {
        if (2 * l - 2 * i < 0 || l - am - 2 * i < 0) { u = 0.0; break; }
        double f1 = binom(l, i) * binom(i, j) * sgn(i) * fact_ratio(2 * l - 2 * i, l - am - 2 * i);
        double s = 0.0;
        for (int k = 0; k <= j; k++) s += binom(j, k) * binom(am, lx - 2 * k) * sgn((am - lx + 2 * k) / 2);
        u += f1 * s;
    }
    u *= pref;
    if (m == 0 && (lx % 2 == 0)) u *= M_SQRT1_2;
    return u;
}

static long double kseries(int l, long double x)
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

static double unit_sphere(int i, int j, int k)
{
    if (i % 2 == 0 && j % 2 == 0 && k % 2 == 0)
        return 4 * M_PI * dfact(i - 1) * dfact(j - 1) * dfact(k - 1) / dfact(i + j + k + 1);
    return 0.0;
}

static void build(HostTables &T)
{
    // ---- radial grid in doubling order, padding point first ----
    T.r.assign(NPTS, 1.0);
    T.w.assign(NPTS, 0.0);
    const double R = 5.0, R3 = R * R * R;
    int idx = 1;
    auto put = [&](int i, int nr) {
        double xi = i / (nr + 1.0), xi3 = xi * xi * xi, lx = std::log(1 - xi3);
        T.w[idx] = 3 * R3 * xi * xi * lx * lx / ((1 - xi3) * (nr + 1.0));
        T.r[idx] = -R * lx;
        idx++;
    };
    for (int i = 1; i <= 31; i++) put(i, 31);
    for (int nr = 63; nr <= 8191; nr = 2 * nr + 1)
        for (int i = 1; i <= nr; i += 2) put(i, nr);
    T.rpow.assign((MAXPOW + 1) * NPTS, 0.0);
    for (int n = 0; n <= MAXPOW; n++)
        for (int i = 0; i < NPTS; i++) T.rpow[n * NPTS + i] = std::pow(T.r[i], n);

    // ---- K_lambda table (orders 0..59 generated so the derivative recursion matches) ----
    const int NG = 60;
    std::vector<double> full((size_t) BES_NPT * NG * 5, 0.0);
    auto F = [&](int i, int l, int k) -> double & { return full[((size_t) i * NG + l) * 5 + k]; };
    for (int i = 0; i < BES_NPT; i++) {
        long double x = (long double) i / 100.0L;
        for (int l = 0; l < NG; l++) F(i, l, 0) = (double) kseries(l, x);
        for (int k = 1; k < 5; k++)
            for (int l = 0; l <= NG - 1 - 5; l++) {
                double d = 0.0;
                if (l > 0) d += l / (2.0 * l + 1.0) * F(i, l - 1, k - 1);
                d += (l + 1.0) / (2.0 * l + 1.0) * F(i, l + 1, k - 1);
                d -= F(i, l, k - 1);
                F(i, l, k) = d;
            }
    }
    T.bes.assign((size_t) BES_NPT * BES_NL * 5, 0.0);
    for (int i = 0; i < BES_NPT; i++)
        for (int l = 0; l < BES_NL; l++)
            for (int k = 0; k < 5; k++) T.bes[((size_t) i * BES_NL + l) * 5 + k] = F(i, l, k);
    // order-major copy: lanes at neighbouring grid points read neighbouring nodes of one order
    T.bes_t.assign((size_t) BES_NPT * BES_NL * 5, 0.0);
    for (int l = 0; l < BES_NL; l++)
        for (int i = 0; i < BES_NPT; i++)
            for (int k = 0; k < 5; k++) T.bes_t[((size_t) l * BES_NPT + i) * 5 + k] = F(i, l, k);
    // both orders of a Bessel ratio in one 80-byte entry (16-byte aligned): the steering
    // evaluations gather them with five 16-byte loads per lane
    T.besp.assign((size_t) BES_NPT * BES_NL * 10, 0.0);
    {
        const double inv_fact[5] = {1.0, 1.0, 0.5, 1.0 / 6.0, 1.0 / 24.0};
        for (int i = 0; i < BES_NPT; i++)
            for (int l = 0; l < BES_NL; l++) {
                const int lo = l > 0 ? l - 1 : 1;
                double *e = &T.besp[((size_t) i * BES_NL + l) * 10];
                for (int k = 0; k < 5; k++) {
                    e[k] = F(i, l, k) * inv_fact[k];
                    e[5 + k] = F(i, lo, k) * inv_fact[k];
                }
            }
    }
    T.rfun.assign(BES_NL * BES_NL, 0.0);
    for (int n = 0; n < BES_NL; n++)
        for (int k = 0; k <= n; k++) {
            long double v = 1.0L;
            for (int j = 1; j <= n + k; j++) v *= j;
            for (int j = 1; j <= k; j++) v /= j;
            for (int j = 1; j <= n - k; j++) v /= j;
            for (int j = 0; j < k; j++) v /= 2.0L;
            T.rfun[n * BES_NL + k] = (double) v;
        }

    // ---- Boys table ----
    T.boys.assign(BOYS_NPT * BOYS_NN, 0.0);
    for (int i = 0; i < BOYS_NPT; i++)
        for (int n = 0; n < BOYS_NN; n++) T.boys[i * BOYS_NN + n] = (double) boys_series(n, (long double) i / 10.0L);

    // ---- real spherical harmonics ----
    T.rsh_off.assign(RSH_LMAX + 2, 0);
    T.comb_off.assign(RSH_LMAX + 2, 0);
    for (int l = 0; l <= RSH_LMAX; l++) {
        int nc = ncomb(l);
        T.rsh_off[l + 1] = T.rsh_off[l] + (2 * l + 1) * nc;
        T.comb_off[l + 1] = T.comb_off[l] + 3 * nc;
    }
    T.rsh.assign(T.rsh_off[RSH_LMAX + 1], 0.0);
    T.comb.assign(T.comb_off[RSH_LMAX + 1], 0);
    for (int l = 0; l <= RSH_LMAX; l++) {
        int nc = ncomb(l), c = 0;
        int *cb = &T.comb[T.comb_off[l]];
        for (int i = 0; i <= l; i++)
            for (int j = 0; j <= l; j++)
                for (int k = 0; k <= l; k++)
                    if (i + j + k == l) {
                        cb[3 * c] = i;
                        cb[3 * c + 1] = j;
                        cb[3 * c + 2] = k;
                        c++;
                    }
        for (int m = -l; m <= l; m++)
            for (int c2 = 0; c2 < nc; c2++)
                T.rsh[T.rsh_off[l] + (m + l) * nc + c2] = rsh_coef(l, m, cb[3 * c2], cb[3 * c2 + 1], cb[3 * c2 + 2]);
    }

    // ---- angular momentum matrices ----
    T.lmat_off.assign(LMAX_POT + 2, 0);
    for (int L = 0; L <= LMAX_POT; L++) T.lmat_off[L + 1] = T.lmat_off[L] + 3 * (2 * L + 1) * (2 * L + 1);
    T.lmat.assign(T.lmat_off[LMAX_POT + 1], 0.0);
    for (int L = 0; L <= LMAX_POT; L++) {
        int dim = 2 * L + 1;
        std::vector<double> cx(dim * dim), cy(dim * dim), cz(dim * dim);
        for (int m1 = -L; m1 <= L; m1++)
            for (int m2 = -L; m2 <= L; m2++) {
                double lz = m2 * (m1 == m2);
                double lp = std::sqrt((double) (L - m2) * (L + m2 + 1)) * (m1 == m2 + 1);
                double lm = std::sqrt((double) (L + m2) * (L - m2 + 1)) * (m1 == m2 - 1);
                cx[(m1 + L) * dim + m2 + L] = 0.5 * (lp + lm);
                cy[(m1 + L) * dim + m2 + L] = 0.5 * (lp - lm);
                cz[(m1 + L) * dim + m2 + L] = lz;
            }
        typedef std::complex<double> cd;
        auto coefs = [](int m, cd &a, cd &b) {
            const cd I(0, 1);
            if (m == 0) { a = 0.5; b = 0.5; }
            else if (m < 0) { a = I / std::sqrt(2.0); b = -I / std::sqrt(2.0) * sgn(std::abs(m)); }
            else { a = 1.0 / std::sqrt(2.0); b = 1.0 / std::sqrt(2.0) * sgn(m); }
        };
        double *out = &T.lmat[T.lmat_off[L]];
        for (int m1 = -L; m1 <= L; m1++)
            for (int m2 = -L; m2 <= L; m2++) {
                cd ab1[2], ab2[2];
                coefs(m1, ab1[0], ab1[1]);
                coefs(m2, ab2[0], ab2[1]);
                int i1[2] = {-std::abs(m1) + L, std::abs(m1) + L}, i2[2] = {-std::abs(m2) + L, std::abs(m2) + L};
                cd vx = 0, vy = 0, vz = 0;
                for (int p = 0; p < 2; p++)
                    for (int q = 0; q < 2; q++) {
                        cd f = std::conj(ab1[p]) * ab2[q];
                        vx += f * cx[dim * i1[p] + i2[q]];
                        vy += f * cy[dim * i1[p] + i2[q]];
                        vz += f * cz[dim * i1[p] + i2[q]];
                    }
                int e = (m1 + L) * dim + m2 + L;
                out[e] = vx.imag();
                out[dim * dim + e] = -vy.real();
                out[2 * dim * dim + e] = vz.imag();
            }
    }

    // ---- type-1 angular couplings of the numeric nuclear-attraction path, summed over (r, s, t)
    //      in the reference's loop order so that each coupling rounds as there ----
    T.w1.assign((size_t) (NUC_LAB + 1) * (NUC_LAB + 1) * NUC_NIJK, 0.0);
    for (int lam = 0; lam <= NUC_LAB; lam++)
        for (int mu = -lam; mu <= lam; mu++)
            for (int s = 0; s <= NUC_LAB; s++)
                for (int I = s; I >= 0; I--)
                    for (int J = s - I; J >= 0; J--) {
                        const int K = s - I - J;
                        double sum2 = 0.0;
                        for (int r = 0; r <= lam; r++)
                            for (int q = 0; q <= lam; q++)
                                for (int t = 0; t <= lam; t++)
                                    if (r + q + t == lam) sum2 += rsh_coef(lam, mu, r, q, t) * unit_sphere(I + r, J + q, K + t);
                        // simplex index of (I, J, K): layer s, then I, then J (as the device's simplex_idx(t, u, v))
                        const int idx = s * (s + 1) * (s + 2) / 6 + I * (s + 1) - I * (I - 1) / 2 + J;
                        T.w1[(size_t) (lam * lam + mu + lam) * NUC_NIJK + idx] = sum2;
                    }

    // ---- OC-shell Cartesian -> real-harmonic coefficients (outercore_integrals.c:171-180,545-553) ----
    T.slm_off.assign(LMAX_SHELL + 2, 0);
    for (int L = 0; L <= LMAX_SHELL; L++) T.slm_off[L + 1] = T.slm_off[L] + (2 * L + 1) * ncart(L);
    T.slm.assign(T.slm_off[LMAX_SHELL + 1], 0.0);
    for (int L = 0; L <= LMAX_SHELL; L++) {
        int nc = ncart(L), c = 0;
        double an = 1.0 / (2.0 * std::sqrt(M_PI)) * std::sqrt(dfact(2 * L + 1));
        for (int r = L; r >= 0; r--)
            for (int s = L - r; s >= 0; s--, c++)
                for (int m = -L; m <= L; m++)
                    T.slm[T.slm_off[L] + (m + L) * nc + c] = rsh_coef(L, m, r, s, L - r - s) / an;
    }
}

const HostTables &host_tables()
{
    static HostTables T;
    static std::once_flag once;
    std::call_once(once, [] { build(T); });
    return T;
}

// ---- angular coupling blocks, built lazily per (L, n) ----
void angular_w_tables(int Lmax, int nmax, std::vector<double> &W, std::vector<WBlock> &dir)
{
    const HostTables &T = host_tables();
    W.clear();
    dir.clear();
    for (int L = 0; L <= Lmax; L++)
        for (int n = 0; n <= nmax; n++) {
            if (L + n > ANG_SUM_MAX) continue;
            WBlock b;
            b.L = L;
            b.n = n;
            b.off = (int) W.size();
            int lmx = L + n;
            b.mu_stride = 2 * lmx + 1;
            int nab = ncomb(n), cntl = lam_cnt(L, n), lo = lam_lo(L, n);
            W.resize(W.size() + (size_t) cntl * (2 * L + 1) * b.mu_stride * nab, 0.0);
            double *out = &W[b.off];
            // (a,b,c) with a+b+c = n in generator order (x power descending, then y)
            std::vector<int> abc;
            for (int a = n; a >= 0; a--)
                for (int bb = n - a; bb >= 0; bb--) {
                    abc.push_back(a);
                    abc.push_back(bb);
                    abc.push_back(n - a - bb);
                }
            const double *yL = &T.rsh[T.rsh_off[L]];
            const int *cL = &T.comb[T.comb_off[L]];
            int ncL = ncomb(L);
            for (int j = 0; j < cntl; j++) {
                int lam = lo + 2 * j;
                const double *yl = &T.rsh[T.rsh_off[lam]];
                const int *cl = &T.comb[T.comb_off[lam]];
                int ncl = ncomb(lam);
                for (int m = -L; m <= L; m++)
                    for (int mu = -lam; mu <= lam; mu++)
                        for (int q = 0; q < nab; q++) {
                            int a = abc[3 * q], bb = abc[3 * q + 1], c = abc[3 * q + 2];
                            double s2 = 0.0;
                            // same skips and summation order as angular_integrals.c:78-104
                            for (int iu = 0; iu < ncL; iu++) {
                                double y1 = yL[(m + L) * ncL + iu];
                                if (std::fabs(y1) < 1e-14) continue;
                                for (int ir = 0; ir < ncl; ir++) {
                                    double y2 = yl[(mu + lam) * ncl + ir];
                                    if (std::fabs(y2) < 1e-14) continue;
                                    s2 += y2 * y1 *
                                          unit_sphere(a + cl[3 * ir] + cL[3 * iu], bb + cl[3 * ir + 1] + cL[3 * iu + 1],
                                                      c + cl[3 * ir + 2] + cL[3 * iu + 2]);
                                }
                            }
                            out[((size_t) (j * (2 * L + 1) + (m + L)) * b.mu_stride + (mu + lam)) * nab + q] = s2;
                        }
            }
            dir.push_back(b);
        }
}

}  // namespace grpp
