// Host-side constant tables of the GRPP engine (built once per process, uploaded per device).
#pragma once
#include <cstdint>
#include <vector>

namespace grpp {

constexpr int NPTS = 8192;        // 1 padding point + 8191 Log3 points in doubling order
constexpr int NLEV = 9;           // levels 31, 63, ..., 8191 points
constexpr int BES_NPT = 1601;     // K_lambda table nodes x = 0, 0.01, ..., 16
constexpr int BES_NL = 32;        // orders kept on the device
constexpr int BOYS_NPT = 301;
constexpr int BOYS_NN = 40;
constexpr int RSH_LMAX = 30;      // real-harmonic coefficient tables 0..RSH_LMAX
constexpr int LMAX_SHELL = 10;    // basis L (incl. the L+1 shells of gradients)
constexpr int LMAX_POT = 10;
constexpr int MAXPOW = 20;        // r^n, n <= MAXPOW (numeric type-1: N <= la + lb)
constexpr int ANG_SUM_MAX = 14;   // L + nA bound: unit-sphere integrals need (2(L+nA)+1)!! <= 30!!

constexpr int NUC_LMAX = 6;        // nuclear attraction by the numeric radial quadrature: l <= 6 per shell
constexpr int NUC_LAB = 2 * NUC_LMAX;
constexpr int NUC_NIJK = (NUC_LAB + 1) * (NUC_LAB + 2) * (NUC_LAB + 3) / 6;   // (I, J, K), I + J + K <= NUC_LAB

struct HostTables {
    std::vector<double> r, w;                 // [NPTS]
    std::vector<double> rpow;                 // [(MAXPOW+1) * NPTS]
    // type-1 angular couplings sum_{r+s+t=lam} y_{lam mu}^{rst} int_sphere x^{I+r} y^{J+s} z^{K+t}
    // (angular_integrals.c:35-56 without the harmonic of k), [(lam^2 + mu + lam) * NUC_NIJK + simplex(I,J,K)]
    std::vector<double> w1;
    std::vector<double> bes;                  // [BES_NPT][BES_NL][5]
    std::vector<double> bes_t;                // the same, [BES_NL][BES_NPT][5] (order-major)
    std::vector<double> besp;                 // [BES_NPT][BES_NL][10]: Taylor coefficient k over k! (Horner form)
                                              // of K_l (0..4), then of K_{l-1} (K_1 for l = 0) (5..9)
    std::vector<double> rfun;                 // [BES_NL][BES_NL]
    std::vector<double> boys;                 // [BOYS_NPT][BOYS_NN]
    // real spherical harmonics y_{lm}^{ijk}: per l, ncomb = (l+1)(l+2)/2 combos (i ascending, j, k)
    std::vector<int> rsh_off;                 // [RSH_LMAX+2] offset of l's coefficients
    std::vector<int> comb_off;                // [RSH_LMAX+2] offset of l's combos (3 ints each)
    std::vector<double> rsh;                  // coefficients [(m+l)*ncomb + c]
    std::vector<int> comb;                    // (i,j,k) triples
    std::vector<int> lmat_off;                // [LMAX_POT+2]: offset of L's three (2L+1)^2 matrices
    std::vector<double> lmat;                 // Lx, Ly, Lz real-harmonic matrices
    std::vector<int> slm_off;                 // [LMAX_SHELL+2]: OC-shell Cartesian->RSH coefficients
    std::vector<double> slm;                  // [(m+L)*ncart + c] in the Cartesian generator order
};

const HostTables &host_tables();

// lambda range of the type-2 expansion: lambda = lo, lo+2, ..., L+n (type2_integrals.c:169-182)
inline int lam_lo(int L, int n) { return (n <= L) ? (L - n) : ((L + n) & 1); }
inline int lam_cnt(int L, int n) { return (L + n - lam_lo(L, n)) / 2 + 1; }
inline int ncomb(int n) { return (n + 1) * (n + 2) / 2; }
inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// W[L][n] block: sum_{uvw,rst} y_{Lm}^{uvw} y_{lam mu}^{rst} I(a+r+u, b+s+v, c+t+w) for all
// valid lambda (index j), m, mu, and (a,b,c) with a+b+c = n in generator order.
// layout [j][m+L][mu (0..2*lam_max)][abc]; mu runs over 2*lam_j+1 entries, row stride 2*(L+n)+1.
struct WBlock {
    int L, n, off, mu_stride;
};
// Returns the (cached) concatenated table and block directory for all (L, n) with L <= Lmax,
// n <= nmax, L + n <= ANG_SUM_MAX.
void angular_w_tables(int Lmax, int nmax, std::vector<double> &W, std::vector<WBlock> &dir);

double rsh_coef(int l, int m, int lx, int ly, int lz);
double gaussian_norm(int L, double alpha);
double dfact(int n);
double fact(int n);

}  // namespace grpp
