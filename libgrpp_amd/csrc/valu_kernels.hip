// FP64 VALU kernels of the GRPP engine that are compiled without machine loop-invariant code
// motion (-mllvm -disable-machine-licm, this file only): hoisting their Bessel, exp and erf
// constants out of the loops held them in registers and spilled them to scratch (k_screen: 42
// VGPRs of spill at 96 VGPRs, reloaded per entry); rematerialised in the loops they need none.
//
//   k_tab_F     F^{C,A}_lambda(r_i) on the level-ordered Log3 grid (radial_type2_integral.c:465-489);
//   k_screen    the Shaw-Hill envelope estimate of every radial entry (screening.c:193-378);
//   k_contract  sum over channels of 16 pi^2 sum_m P^A_m Q P^B_m^T, the spin-orbit analogue and the
//               outercore products (type2_integrals.c, spin_orbit_integrals.c, outercore_integrals.c).
#include <hip/hip_runtime.h>

#include <cmath>

#include "engine_dev.h"

// ================================================================== tabulation kernels
constexpr int MAXLAM = LMAX_POT + LMAX_SHELL;   // 20

__global__ __launch_bounds__(256) void k_tab_F(const TabItem *__restrict__ items, DevJob J, DevTables T,
                                               double *__restrict__ F)
{
    const TabItem it = items[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPTS) return;
    const DShell s = J.sh[it.A];
    const DCenter c = J.ctr[it.c];
    const double dx = c.x - s.x, dy = c.y - s.y, dz = c.z - s.z;
    const double CA = sqrt(dx * dx + dy * dy + dz * dz);
    const double r = T.r[i];
    double acc[MAXLAM + 1];
#pragma unroll
    for (int l = 0; l <= MAXLAM; l++) acc[l] = 0.0;
    const double d2 = (r - CA) * (r - CA);
    for (int k = 0; k < s.nprim; k++) {
        const double cf = J.cnorm[s.poff + k];
        if (cf == 0.0) continue;
        const double a = J.alpha[s.poff + k];
        const double e = cf * exp(-a * d2);
        if (e == 0.0) continue;   // the primitive vanishes at this point (K_l <= 1): adds exactly 0
        const double x = 2.0 * a * CA * r;
        if (x > 1e-7 && x <= 16.0) {
            const int i0 = (int) (rint(x * 100) + 1e-5);
            const double ddx = x - i0 * 0.01, dx2 = ddx * ddx, dx3 = dx2 * ddx, dx4 = dx3 * ddx;
            // order-major table: the lanes of a wave, at neighbouring grid points, read neighbouring
            // nodes of one order (fewer cache lines per load than node-major rows 1280 bytes apart)
            const double *row = T.bes_t + (size_t) i0 * 5;
#pragma unroll
            for (int l = 0; l <= MAXLAM; l++) {
                if (l <= it.lam) {
                    const double *d = row + (size_t) l * BES_NPT * 5;
                    const double K = d[0] + d[1] * ddx + d[2] * dx2 * 0.5 + d[3] * dx3 * (1.0 / 6.0) +
                                     d[4] * dx4 * (1.0 / 24.0);
                    acc[l] += e * K;
                }
            }
        } else {
#pragma unroll
            for (int l = 0; l <= MAXLAM; l++)
                if (l <= it.lam) acc[l] += e * bessel_k(T.bes, T.rfun, l, x);
        }
    }
#pragma unroll
    for (int l = 0; l <= MAXLAM; l++)
        if (l <= it.lam) F[(size_t) (it.frow + l) * NPTS + i] = acc[l];
}

// ================================================================== Shaw-Hill screening of the radial entries
// Envelope estimate of one radial integral from its primitive triples (screening.c:193-378):
// the integrand maximum r0 is found by <= 4 fixed-point steps of tolerance 1e-2 using Bessel ratios;
// any triple that does not settle makes the screening inert.  Inert screening is encoded as NaN,
// which the quadrature's tests treat exactly as the reference treats a failed screening.
//
// Steering-only evaluations of K_l for the envelope estimates: the reference's scheme (table
// Taylor quartic for x <= 16, asymptotic R_l(-x)/(2x) beyond) in Horner form with one reciprocal.
__device__ __forceinline__ double rcp_refined(double b)
{
    double r = __builtin_amdgcn_rcp(b);
    r = fma(fma(-b, r, 1.0), r, r);
    return fma(fma(-b, r, 1.0), r, r);
}

__device__ __forceinline__ double bessel_k_steer(const DevTables &T, int n, double x)
{
    if (x > 1e-7 && x <= 16.0) {
        const int i0 = (int) (rint(x * 100) + 1e-5);
        const double dx = x - i0 * 0.01;
        const double2 *d = (const double2 *) (T.besp + ((size_t) i0 * BES_NL + n) * 10);
        const double2 c01 = d[0], c23 = d[1];
        const double c4 = d[2].x;
        return c01.x + dx * (c01.y + dx * (c23.x + dx * (c23.y + dx * c4)));
    }
    if (x > 16.0) {
        const double inv = -rcp_refined(x);
        const double *c = T.rfun + n * BES_NL;
        double s = c[n];
        for (int k = n - 1; k >= 0; k--) s = fma(s, inv, c[k]);
        return s * (-0.5 * inv);
    }
    return bessel_k(T.bes, T.rfun, n, x);
}

// K_{l-1}(x)/K_l(x) (l > 0) or K_1(x)/K_0(x) (l = 0): both orders from one table node, or both
// asymptotic polynomials in one pass (their common 1/(2x) cancels)
__device__ __forceinline__ double bessel_ratio(const DevTables &T, int l, double x)
{
    const int a = (l == 0) ? 1 : l - 1, b = (l == 0) ? 0 : l;
    if (x > 1e-7 && x <= 16.0) {
        const int i0 = (int) (rint(x * 100) + 1e-5);
        const double dx = x - i0 * 0.01;
        // entry l holds K_b's coefficients, then K_a's: five 16-byte loads per lane
        const double2 *d = (const double2 *) (T.besp + ((size_t) i0 * BES_NL + l) * 10);
        const double2 b01 = d[0], b23 = d[1], b4a0 = d[2], a12 = d[3], a34 = d[4];
        const double ka = b4a0.y + dx * (a12.x + dx * (a12.y + dx * (a34.x + dx * a34.y)));
        const double kb = b01.x + dx * (b01.y + dx * (b23.x + dx * (b23.y + dx * b4a0.x)));
        return ka * rcp_refined(kb);
    }
    if (x > 16.0) {
        const double inv = -rcp_refined(x);
        const double *ca = T.rfun + a * BES_NL, *cb = T.rfun + b * BES_NL;
        const int nmax = a > b ? a : b;
        double sa = 0.0, sb = 0.0;
        for (int k = nmax; k >= 0; k--) {
            sa = fma(sa, inv, k <= a ? ca[k] : 0.0);
            sb = fma(sb, inv, k <= b ? cb[k] : 0.0);
        }
        return sa * rcp_refined(sb);
    }
    return bessel_k(T.bes, T.rfun, a, x) / bessel_k(T.bes, T.rfun, b, x);
}

// x^n for n >= 0 by binary powering
__device__ __forceinline__ double ipow(double x, int n)
{
    double r = 1.0;
    for (; n > 0; n >>= 1) {
        if (n & 1) r *= x;
        x *= x;
    }
    return r;
}

// One primitive triple of the envelope (screening.c:272-378).  Loop invariants of the triple are
// hoisted: 1/(2p), sqrt(pi/p), sqrt(p) and the integer shift of the fixed-point equation
//   r = sqrt((k1 r K_{l1-1}/K_{l1}(k1 r) + k2 r K_{l2-1}/K_{l2}(k2 r) + n - [l1>0](l1+1) - [l2>0](l2+1)) / 2p),
// iterated from r = (aA CA + aB CB)/p until |r - r_prev| <= 1e-2; a triple that has not settled
// after four evaluations fails the whole estimate (the reference's five-step rule).  The estimate
// only steers the quadrature (screened-out threshold, forced refinement), so few-ulp differences
// of these algebraically equal forms are immaterial.
__device__ __forceinline__ bool screen_fixed_point(const DevTables &T, int n, int l1, int l2, double aA, double aB,
                                                   double eta, double CA, double CB, double k1, double k2, double &r0,
                                                   unsigned &nstep)
{
    const double p = aA + aB + eta, invp = 1.0 / p, inv2p = 0.5 * invp;
    const double nsh = (double) (n - (l1 > 0 ? l1 + 1 : 0) - (l2 > 0 ? l2 + 1 : 0));
    double r0p;
    r0 = (aA * CA + aB * CB) * invp;
    int steps = 0;
    do {
        if (++steps == 5) return false;
        nstep++;
        r0p = r0;
        const double k1r = k1 * r0, k2r = k2 * r0;
        const double a = bessel_ratio(T, l1, k1r) * k1r + bessel_ratio(T, l2, k2r) * k2r + nsh;
        r0 = sqrt(a * inv2p);
    } while (fabs(r0 - r0p) > 1e-2);
    return true;
}

__device__ __forceinline__ double screen_value(const DevTables &T, int n, int l1, int l2, double aA, double aB,
                                              double eta, double CA, double CB, double k1, double k2, double r0)
{
    const double p = aA + aB + eta, dA = r0 - CA, dB = r0 - CB;
    return sqrt(M_PI / p) * ipow(r0, n) * bessel_k_steer(T, l1, k1 * r0) * bessel_k_steer(T, l2, k2 * r0) *
           exp(-eta * r0 * r0 - aA * dA * dA - aB * dB * dB) * 0.5 * (1.0 + erf(sqrt(p) * r0));
}

// triples per entry in the screening probe (64 / PROBE entries per wave): 1, 2 or 4, picked per
// plan by the autotuner (the estimates do not depend on it)

// Fixed point of triple q of an entry in the probe order (smallest exponents first: the triples most
// likely to fail); active is false for triples that vanish
__device__ __forceinline__ bool screen_probe_triple(const DevJob &J, const DevTables &T, const DShell &A,
                                                    const DShell &B, double CA, double CB, const DPot &P, int l1,
                                                    int l2, int N, int q, bool &active, unsigned &rsteps)
{
    const int nt = P.nterm, njt = B.nprim * nt;
    const int iq = q / njt, r = q - iq * njt, jq = r / nt, tq = r - jq * nt;
    const double2 pa = J.pprobe[A.poff + iq], pb = J.pprobe[B.poff + jq];
    const double aA = pa.x, aB = pb.x, k1 = 2 * aA * CA, k2 = 2 * aB * CB;
    active = pa.y != 0.0 && pb.y != 0.0 && !((l1 >= 1 && fabs(k1) <= 1e-14) || (l2 >= 1 && fabs(k2) <= 1e-14));
    if (!active || (l1 == 0 && l2 == 0 && fabs(k1) <= 1e-14 && fabs(k2) <= 1e-14)) return false;
    double r0;
    return !screen_fixed_point(T, N + J.tpow_probe[P.toff + tq], l1, l2, aA, aB, J.tprobe[P.toff + tq].x, CA, CB,
                               k1, k2, r0, rsteps);
}

// The envelope of one radial entry (screening.c:193-248), evaluated by a whole wave: lane q takes
// the primitive triples q, q+64, ... of the (i, j, t) space, so the lanes of a wave share the entry's
// (lambda1, lambda2, N) and run the same code.  A triple that fails the fixed point fails the
// entry (NaN, the reference's failed screening); the lane sums are combined by an xor butterfly,
// which leaves the bitwise same total in every lane.
__device__ double screen_entry_wave(const DevJob &J, const DevTables &T, const DShell &A, const DShell &B,
                                    double CA, double CB, const DPot &P, int l1, int l2, int N, unsigned &ntrip,
                                    unsigned &nstep)
{
    const int lane = threadIdx.x & 63;
    const int nt = P.nterm, njt = B.nprim * nt, ntot = A.nprim * njt;
    double est = 0.0;
#pragma unroll 1
    for (int q0 = 0; q0 < ntot; q0 += 64) {
        // fixed points of the round's 64 triples first: a failure ends the entry before any value
        // is computed (most entries of a typical workload fail)
        const int q = q0 + lane;
        bool active = false, analytic = false, fail = false;
        unsigned rsteps = 0;
        double ci = 0.0, cj = 0.0, aA = 0.0, aB = 0.0, k1 = 0.0, k2 = 0.0, eta = 0.0, coef = 0.0, r0 = 0.0;
        int n = 0;
        if (q < ntot) {
            // triples in the order of the screening probe (smallest exponents first)
            const int iq = q / njt, r = q - iq * njt, jq = r / nt, tq = r - jq * nt;
            const double2 pa = J.pprobe[A.poff + iq], pb = J.pprobe[B.poff + jq];
            aA = pa.x;
            ci = pa.y;
            aB = pb.x;
            cj = pb.y;
            k1 = 2 * aA * CA;
            k2 = 2 * aB * CB;
            // primitives dropped by libgrpp_shell_shrink; K_l(0) = 0 for l >= 1
            active = ci != 0.0 && cj != 0.0 && !((l1 >= 1 && fabs(k1) <= 1e-14) || (l2 >= 1 && fabs(k2) <= 1e-14));
            if (active) {
                const double2 tz = J.tprobe[P.toff + tq];
                eta = tz.x;
                coef = tz.y;
                n = N + J.tpow_probe[P.toff + tq];
                analytic = (l1 == 0 && l2 == 0 && fabs(k1) <= 1e-14 && fabs(k2) <= 1e-14);
                if (!analytic) fail = !screen_fixed_point(T, n, l1, l2, aA, aB, eta, CA, CB, k1, k2, r0, rsteps);
            }
        }
        // work accounting follows the reference's order: a failed entry counts the triples up to
        // and including its first failing one
        const unsigned long long fmask = __ballot(fail);
        const int qfail = fmask ? q0 + __ffsll((long long) fmask) - 1 : ntot;
        if (active && q <= qfail) {
            ntrip++;
            nstep += rsteps;
        }
        if (fmask) return __longlong_as_double(0x7ff8000000000000LL);
        if (active) {
            const double val = analytic ? exp(-aA * CA * CA - aB * CB * CB) * int_gauss_s(n, aA + aB + eta)
                                        : screen_value(T, n, l1, l2, aA, aB, eta, CA, CB, k1, k2, r0);
            est += ci * cj * coef * val;
        }
    }
    for (int o = 32; o > 0; o >>= 1) est += __shfl_xor(est, o);
    return est;
}

// SCR_SPLIT waves per 64 x 64 tile (4, 8 or 16, picked per plan): a wave takes the rows sub,
// sub + SCR_SPLIT, ... of its tile.  The estimates do not depend on it; finer splits shorten the tail
// of a launch with few tiles, whose waves differ widely in how many entries their probe leaves.
template <int SCR_PROBE, int SCR_SPLIT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_screen(
    const TileItem *__restrict__ tiles, const QJob *__restrict__ jobs, const RowDesc *__restrict__ rows, DevJob J,
    DevTables T, double *__restrict__ scr, unsigned long long *__restrict__ counters)
{
    // one wave per 1/SCR_SPLIT of a tile (rows sub, sub + SCR_SPLIT, ...).  The wave first lists the entries it must
    // estimate -- inside the table, upper triangle, canonical (lambda1, lambda2, N) split -- with the
    // full-cost ones first and the ones that vanish analytically (K_l(0) = 0 for l >= 1 on a shell
    // sitting on the centre) last, so lanes run entries of like cost side by side.
    __shared__ unsigned short list[RT * RT / SCR_SPLIT];   // positions < RT * RT / 4 <= 1024
    // per tile position e of the wave: 0 nothing to store (outside the table, lower triangle or not
    // a canonical split), 1 estimate pending, 2 vanishes analytically, 3 a triple failed (NaN),
    // 4 evaluated and stored
    __shared__ unsigned char st[RT * RT / SCR_SPLIT];
    __shared__ int n_full, n_zero;
    const TileItem t = tiles[blockIdx.x / SCR_SPLIT];
    const QJob jb = jobs[t.job];
    const RowDesc *R = rows + jb.rows;
    const DCenter c = J.ctr[jb.center];
    const DPot P = J.pot[jb.potential];
    const int sub = blockIdx.x % SCR_SPLIT;
    if (threadIdx.x == 0) {
        n_full = 0;
        n_zero = 0;
    }
    __syncthreads();
    constexpr int NE = RT * RT / SCR_SPLIT;
    // the quarter tiles cover the table in its logical row order (shell, n, lambda), whatever the
    // physical order of the rows, so the work of a wave does not depend on a row permutation
    auto phys = [&](int l) { return J.rowpos[jb.rows + l]; };
    for (int e = threadIdx.x; e < NE; e += blockDim.x) {
        const int lrow = t.ti * RT + (e / RT) * SCR_SPLIT + sub, lcol = t.tj * RT + e % RT;
        st[e] = 0;
        if (lrow >= jb.M || lcol >= jb.M) continue;
        if (t.ti == t.tj && lrow > lcol) continue;
        const RowDesc ra = R[phys(lrow)], rb = R[phys(lcol)];
        const DShell A = J.sh[ra.shell], B = J.sh[rb.shell];
        // the canonical split of (lambda1, lambda2, N) may sit on either side of the diagonal
        // (rows need not be in (shell, n) order), so both orientations are tested
        if (!is_canonical(jb.L, A.L, B.L, ra, rb) && !is_canonical(jb.L, B.L, A.L, rb, ra)) continue;
        const bool a_on = (c.x - A.x) * (c.x - A.x) + (c.y - A.y) * (c.y - A.y) + (c.z - A.z) * (c.z - A.z) == 0.0;
        const bool b_on = (c.x - B.x) * (c.x - B.x) + (c.y - B.y) * (c.y - B.y) + (c.z - B.z) * (c.z - B.z) == 0.0;
        const bool zero = (a_on && ra.lam >= 1) || (b_on && rb.lam >= 1);
        // analytically vanishing entries: every triple is zero, so the estimate is 0 (screened out)
        if (zero) {
            list[NE - 1 - atomicAdd(&n_zero, 1)] = e;
            st[e] = 2;
        } else {
            list[atomicAdd(&n_full, 1)] = e;
            st[e] = 1;
        }
    }
    __syncthreads();
    const int nf = n_full;
    unsigned ntrip = 0, nstep = 0;   // primitive triples and fixed-point evaluations (work accounting)
    // an entry with a triple that fails the fixed point has no estimate (NaN), whatever its other
    // triples give; the triples most likely to fail come first in the probe order, so a probe of
    // the first SCR_PROBE triples of 64 / SCR_PROBE entries at once settles most entries of a workload
    auto entry = [&](int q, int &row, int &col, RowDesc &ra, RowDesc &rb, DShell &A, DShell &B, double &CA,
                     double &CB) {
        const int e = list[q];
        row = phys(t.ti * RT + (e / RT) * SCR_SPLIT + sub);
        col = phys(t.tj * RT + e % RT);
        ra = R[row];
        rb = R[col];
        // a fixed orientation of the (A, B, lambda1, lambda2, N) tuple so that every split and both
        // triangles get bitwise the same estimate
        if (ra.shell > rb.shell || (ra.shell == rb.shell && ra.lam > rb.lam)) {
            const RowDesc tmp = ra;
            ra = rb;
            rb = tmp;
        }
        A = J.sh[ra.shell];
        B = J.sh[rb.shell];
        const double ax = c.x - A.x, ay = c.y - A.y, az = c.z - A.z, bx = c.x - B.x, by = c.y - B.y, bz = c.z - B.z;
        CA = sqrt(ax * ax + ay * ay + az * az);
        CB = sqrt(bx * bx + by * by + bz * bz);
    };
    const double NaN = __longlong_as_double(0x7ff8000000000000LL);
#pragma unroll 1
    for (int g0 = 0; g0 < nf; g0 += 64 / SCR_PROBE) {
        const int g = threadIdx.x / SCR_PROBE, k = threadIdx.x % SCR_PROBE, qi = g0 + g;
        bool fail = false, active = false;
        unsigned rsteps = 0;
        int row = 0, col = 0;
        if (qi < nf) {
            RowDesc ra, rb;
            DShell A, B;
            double CA, CB;
            entry(qi, row, col, ra, rb, A, B, CA, CB);
            if (k < A.nprim * B.nprim * P.nterm)
                fail = screen_probe_triple(J, T, A, B, CA, CB, P, ra.lam, rb.lam, ra.n + rb.n, k, active, rsteps);
        }
        const unsigned long long fmask = __ballot(fail);
        const unsigned gbits = (unsigned) (fmask >> (SCR_PROBE * g)) & ((1u << SCR_PROBE) - 1u);
        if (gbits) {
            if (active && k <= __ffs(gbits) - 1) {
                ntrip++;
                nstep += rsteps;
            }
            if (k == 0) st[list[qi]] = 3;
        }
    }
    __syncthreads();
    // the entries the probe did not settle: the whole wave per entry.  Their shells, (lambda1,
    // lambda2, N) and distances to the centre are gathered 64 at a time, one entry per lane (the
    // loads of all 64 in flight at once), into LDS, and each entry's evaluation reads them from there
    __shared__ int s_apoff[64], s_bpoff[64], s_np[64], s_lln[64], s_e[64];
    __shared__ double s_ca[64], s_cb[64];
    auto rfl = [](int v) { return __builtin_amdgcn_readfirstlane(v); };
    auto rfl_d = [](double v) {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_readfirstlane((int) (b & 0xffffffffLL));
        const int hi = __builtin_amdgcn_readfirstlane((int) (b >> 32));
        return __longlong_as_double(((long long) hi << 32) | (unsigned int) lo);
    };
#pragma unroll 1
    for (int c0 = 0; c0 < nf; c0 += 64) {
        const int qi = c0 + threadIdx.x;
        bool live = false;
        if (qi < nf) {
            const int e = list[qi];
            if (st[e] != 3) {
                int row, col;
                RowDesc ra, rb;
                DShell A, B;
                double CA, CB;
                entry(qi, row, col, ra, rb, A, B, CA, CB);
                s_apoff[threadIdx.x] = A.poff;
                s_bpoff[threadIdx.x] = B.poff;
                s_np[threadIdx.x] = A.nprim | (B.nprim << 16);
                s_lln[threadIdx.x] = ra.lam | (rb.lam << 8) | ((ra.n + rb.n) << 16);
                s_e[threadIdx.x] = e;
                s_ca[threadIdx.x] = CA;
                s_cb[threadIdx.x] = CB;
                live = true;
            }
        }
        unsigned long long m = __ballot(live);
        __syncthreads();
#pragma unroll 1
        while (m) {
            const int ln = __ffsll((long long) m) - 1;
            m &= m - 1;
            DShell A{}, B{};
            const int np = rfl(s_np[ln]), lln = rfl(s_lln[ln]), e = rfl(s_e[ln]);
            A.poff = rfl(s_apoff[ln]);
            A.nprim = np & 0xffff;
            B.poff = rfl(s_bpoff[ln]);
            B.nprim = np >> 16;
            const double v = screen_entry_wave(J, T, A, B, rfl_d(s_ca[ln]), rfl_d(s_cb[ln]), P, lln & 0xff,
                                               (lln >> 8) & 0xff, lln >> 16, ntrip, nstep);
            if (threadIdx.x == 0)
                scr[jb.qoff + (long long) (t.ti * RT + (e / RT) * SCR_SPLIT + sub) * jb.M + t.tj * RT + e % RT] = v;
        }
        __syncthreads();   // the next chunk rewrites the entry data
    }
    __syncthreads();
    // the failed (NaN) and vanishing (0) estimates in tile order: a row's 64 columns are 64
    // consecutive doubles of the logical upper triangle, one store per wave
    for (int e = threadIdx.x; e < NE; e += blockDim.x) {
        const unsigned char v = st[e];
        if (v == 2 || v == 3)
            scr[jb.qoff + (long long) (t.ti * RT + (e / RT) * SCR_SPLIT + sub) * jb.M + t.tj * RT + e % RT] =
                v == 2 ? 0.0 : NaN;
    }
    if (counters) {
        unsigned long long a = ntrip, b = nstep;
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_xor(a, o);
            b += __shfl_xor(b, o);
        }
        if (threadIdx.x == 0) {
            atomicAdd(&counters[3], a);
            atomicAdd(&counters[4], b);
        }
    }
}

// ================================================================== contraction
// k_contract workgroups are CT_THREADS = 64, 128 or 256 threads (picked per plan by the autotuner)
// with CT_EPT = 256 / CT_THREADS outputs each, so a row chunk holds up to 256 outputs in every case
// and every output is the same sum

__device__ __forceinline__ double int_gauss(int n, double a)
{
    if (n % 2 == 0) {
        const int k = n / 2;
        double df = 1.0;
        for (int z = 2 * k - 1; z > 1; z -= 2) df *= z;
        return df / (pow(2.0, k + 1) * pow(a, k)) * sqrt(M_PI / a);
    }
    const int k = (n - 1) / 2;
    double fk = 1.0;
    for (int z = 2; z <= k; z++) fk *= z;
    return fk / (2.0 * pow(a, k + 1));
}


// LDS visibility among the lanes of one wave (the wave-per-part contraction has no block barrier)
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The contraction of one (shell pair, centre) part by CT_THREADS threads (tid = 0 .. CT_THREADS-1)
// with up to CT_EPT outputs each per pass, the Q block and the intermediate in Qs / Ts; WAVE: the
// threads are one wave of a workgroup whose other waves work on other parts, so synchronisation is
// per wave
template <int CT_THREADS, int CT_EPT, bool WAVE>
__device__ __forceinline__ void contract_part(const PartItem pt, const WorkItem *__restrict__ items, const DevJob &J,
                                              const DevTables &T, double *__restrict__ partial, int parts, int nm_max,
                                              int ts_len, double *Qs, double *Ts, const int tid)
{
    auto sync = [] {
        if (WAVE) wave_sync();
        else __syncthreads();
    };
    auto sync_or = [](bool v) -> bool {
        if (WAVE) {
            wave_sync();
            return __ballot(v) != 0;
        }
        return __syncthreads_or(v);
    };
    WorkItem it = items[pt.item];
    it.cb = pt.ci;
    it.ce = pt.ci + 1;
    const DShell sa = J.sh[it.a], sb = J.sh[it.b];
    const int la = sa.L, lb = sb.L, ncA = dev_ncart(la), ncB = dev_ncart(lb);
    const double FOURPI2 = 16.0 * M_PI * M_PI;
    // rows of A per pass: the outputs fit CT_EPT per thread and T[i][m][k2] fits TS_MAX
    const int rows_per_chunk = max(1, min((CT_EPT * CT_THREADS) / ncB, ts_len / (nm_max * ncB)));
    for (int r0 = 0; r0 < ncA; r0 += rows_per_chunk) {
        const int r1 = min(ncA, r0 + rows_per_chunk), nr = r1 - r0, nel = nr * ncB;
        double acc[4][CT_EPT];
#pragma unroll
        for (int o = 0; o < 4; o++)
#pragma unroll
            for (int q = 0; q < CT_EPT; q++) acc[o][q] = 0.0;

        for (int ci = it.cb; ci < it.ce; ci++) {
            const int cidx = J.cidx[ci];
            const DCenter c = J.ctr[cidx];
            const double CAx = c.x - sa.x, CAy = c.y - sa.y, CAz = c.z - sa.z;
            const double CBx = c.x - sb.x, CBy = c.y - sb.y, CBz = c.z - sb.z;
            const double CA2 = CAx * CAx + CAy * CAy + CAz * CAz, CB2 = CBx * CBx + CBy * CBy + CBz * CBz;
            const bool onec = CA2 < 1e-14 && CB2 < 1e-14;
            // ---- semilocal and spin-orbit channels ----
            // one-wave parts: each channel's record and block offsets are loaded by one lane (the
            // dependent loads of up to 64 channels in flight at once) and read back with readlane
            for (int pb0 = c.pb; pb0 < c.pe; pb0 += 64) {
            const int nch = min(64, c.pe - pb0), lane = tid & 63;
            int m_L = 0, m_kind = 0, m_M = 0, m_ra = 0, m_rb = 0;
            long long m_qoff = 0, m_pa = 0, m_pb = 0, m_plb = 0;
            double m_scale = 0.0;
            if (WAVE && lane < nch) {
                const DPot pl = J.pot[pb0 + lane];
                const long long so = (long long) pl.slot * J.nshell;
                m_L = pl.L;
                m_kind = pl.kind;
                m_M = pl.M;
                m_qoff = pl.qoff;
                m_scale = pl.scale;
                m_ra = J.slot_rowoff[so + it.a];
                m_rb = J.slot_rowoff[so + it.b];
                m_pa = (pl.kind == K_SO ? J.slot_pb3 : J.slot_pb2)[so + it.a];
                m_pb = J.slot_poff[so + it.b];
                m_plb = pl.kind == K_SO ? J.slot_ploff[so + it.b] : 0;
            }
            auto rl = [](int v, int k) { return __builtin_amdgcn_readlane(v, k); };
            auto rl64 = [](long long v, int k) {
                const int lo = __builtin_amdgcn_readlane((int) (v & 0xffffffffLL), k);
                const int hi = __builtin_amdgcn_readlane((int) (v >> 32), k);
                return ((long long) hi << 32) | (unsigned int) lo;
            };
            for (int pi = pb0; pi < pb0 + nch; pi++) {
                const int ch = pi - pb0;
                DPot p;
                if (WAVE) {
                    p.L = rl(m_L, ch);
                    p.kind = rl(m_kind, ch);
                } else {
                    p = J.pot[pi];
                }
                const bool is_so = (p.kind == K_SO);
                if (is_so && !(parts & 4)) continue;
                if (!is_so && !(parts & 2)) continue;
                const int L = p.L, nm = 2 * L + 1;
                const int KA = dev_kcols(L, la), KB = dev_kcols(L, lb);
                // radial block Q[(lam1,nA)][(lam2,nB)]; nz: this thread loaded a nonzero entry
                bool nz = false;
                if (onec) {
                    // analytic one-centre branch (radial_type2_integral.c:126-155): only lambda1 = lambda2 = 0
                    if (WAVE) p = J.pot[pi];
                    for (int e = tid; e < KA * KB; e += CT_THREADS) Qs[e] = 0.0;
                    sync();
                    if (tid < (la + 1) * (lb + 1)) {
                        const int nA = tid / (lb + 1), nB = tid % (lb + 1);
                        if (dev_lam_lo(L, nA) == 0 && dev_lam_lo(L, nB) == 0) {
                            const int ka = dev_colbase(L, nA), kb = dev_colbase(L, nB);
                            double v = 0.0;
                            for (int i = 0; i < sa.nprim; i++) {
                                const double ci_ = J.craw[sa.poff + i] * gnorm(la, J.alpha[sa.poff + i]);
                                for (int j = 0; j < sb.nprim; j++) {
                                    const double cj = J.craw[sb.poff + j] * gnorm(lb, J.alpha[sb.poff + j]);
                                    for (int k = 0; k < p.nterm; k++) {
                                        const double pp = J.alpha[sa.poff + i] + J.alpha[sb.poff + j] + J.tzeta[p.toff + k];
                                        v += ci_ * cj * J.tcoef[p.toff + k] * int_gauss(J.tpow[p.toff + k] + nA + nB, pp);
                                    }
                                }
                            }
                            if (fabs(v) < p.thr) v = 0.0;
                            Qs[ka * KB + kb] = v;
                            nz = v != 0.0;
                        }
                    }
                } else {
                    const int ra = WAVE ? rl(m_ra, ch) : J.slot_rowoff[(long long) p.slot * J.nshell + it.a];
                    const int rb = WAVE ? rl(m_rb, ch) : J.slot_rowoff[(long long) p.slot * J.nshell + it.b];
                    const int M = WAVE ? rl(m_M, ch) : p.M;
                    const long long qoff = WAVE ? rl64(m_qoff, ch) : p.qoff;
                    for (int e = tid; e < KA * KB; e += CT_THREADS) {
                        const int k1 = e / KB, k2 = e % KB;
                        const double v = J.Q[qoff + (long long) (ra + k1) * M + rb + k2];   // logical order
                        Qs[e] = v;
                        nz = nz || v != 0.0;
                    }
                }
                // a block of zeros adds exactly zero to every output (acc is never -0): skip the channel
                if (!sync_or(nz)) continue;
                const double *PA, *PB, *PLB;
                double fac;
                if (WAVE) {
                    PA = J.P + rl64(m_pa, ch);
                    PB = J.P + rl64(m_pb, ch);
                    PLB = is_so ? J.P + rl64(m_plb, ch) : nullptr;
                    fac = FOURPI2 * __longlong_as_double(rl64(__double_as_longlong(m_scale), ch));
                } else {
                    PA = J.P + (is_so ? J.slot_pb3 : J.slot_pb2)[(long long) p.slot * J.nshell + it.a];
                    PB = J.P + J.slot_poff[(long long) p.slot * J.nshell + it.b];
                    PLB = is_so ? J.P + J.slot_ploff[(long long) p.slot * J.nshell + it.b] : nullptr;
                    fac = FOURPI2 * p.scale;
                }
                // T[i][m][k2] = sum_k1 P^A[i][m][k1] Q[k1][k2] for all m at once; row i of T is then
                // contiguous over (m, k2), as row j of the ket projectors is, so each output is one
                // dot product of length nm * KB (the same terms in the same order as a loop over m
                // and k2)
                const int nmk = nm * KB, KB2 = (KB + 1) >> 1;
                // two neighbouring k2 per thread share each P^A load (each sum still runs over k1 in order)
                for (int e = tid; e < nr * nm * KB2; e += CT_THREADS) {
                    const int im = e / KB2, k2 = 2 * (e - im * KB2);   // im = i * nm + m
                    const bool two = k2 + 1 < KB;
                    const int k2b = two ? k2 + 1 : k2;
                    const double *pa = PA + ((long long) r0 * nm + im) * KA;
                    double s0 = 0.0, s1 = 0.0;
#pragma unroll 8
                    for (int k1 = 0; k1 < KA; k1++) {
                        const double a = pa[k1];
                        s0 += a * Qs[k1 * KB + k2];
                        s1 += a * Qs[k1 * KB + k2b];
                    }
                    Ts[im * KB + k2] = s0;
                    if (two) Ts[im * KB + k2 + 1] = s1;
                }
                sync();
#pragma unroll
                for (int q = 0; q < CT_EPT; q++) {
                    const int e = tid + q * CT_THREADS;
                    if (e >= nel) break;
                    const int i = e / ncB, j = e - i * ncB;
                    const double *tr = Ts + i * nmk;
                    if (!is_so) {
                        const double *pb = PB + (long long) j * nmk;
                        double sacc = 0.0;
#pragma unroll 16
                        for (int x = 0; x < nmk; x++) sacc += tr[x] * pb[x];
                        acc[0][q] += fac * sacc;
                    } else {
                        // the three L-products side by side (each sum in its own order over x)
                        const double *pl0 = PLB + (long long) j * nmk, *pl1 = pl0 + (long long) ncB * nmk,
                                     *pl2 = pl1 + (long long) ncB * nmk;
                        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll 8
                        for (int x = 0; x < nmk; x++) {
                            const double t = tr[x];
                            s0 += t * pl0[x];
                            s1 += t * pl1[x];
                            s2 += t * pl2[x];
                        }
                        acc[1][q] += fac * s0;
                        acc[2][q] += fac * s1;
                        acc[3][q] += fac * s2;
                    }
                }
                sync();
            }
            }
            // ---- outercore projectors: part 1 (outercore_integrals.c:147-270) and part 2
            // (:277-390); a part-2-only job may restrict part 2 to one (oi, oj) term ----
            const bool oc1 = parts & (8 | 16), oc2 = parts & (8 | 32);
            const bool oc_pair = (parts & 32) && c.oc_i >= 0;
            if ((oc1 || oc2) && c.oe > c.ob) {
                const int no = c.oe - c.ob;
                for (int oi = 0; oi < no; oi++) {
                    const OcEntry e1 = J.oce[c.ob + oi];
                    const DPot p1 = J.pot[e1.pot];
                    const int L = J.sh[e1.shell].L, nm = 2 * L + 1;
                    const double Jv = p1.J / 2.0;
                    const double af = (Jv < L) ? (L / (2.0 * L + 1)) : ((L + 1) / (2.0 * L + 1));
                    const double ef = (Jv < L) ? (-2.0 / (2 * L + 1)) : (+2.0 / (2 * L + 1));
                    const double *SA = J.OCV + J.ocv_tab[(long long) (c.ob + oi) * J.nshell + it.a];
                    const double *UA = SA + ncA * nm;
                    const double *SB = J.OCV + J.ocv_tab[(long long) (c.ob + oi) * J.nshell + it.b];
                    const double *UB = SB + ncB * nm;
                    const double *Lm = T.lmat + T.lmat_off[L];
                    // The ket-side L-products sum_m2 L_eta[m1][m2] S_B[j][m2] (and of U_B) depend on j,
                    // eta and m1 only: formed once per (j, eta, m1), with the same sums, in the LDS
                    // the channels used (free by now), instead of once per output, when they fit
                    double *W = Qs;
                    const int avail = (int) (Ts - Qs) + ts_len, nlb = ncB * 3 * nm;
                    const bool pre1 = oc1 && 2 * nlb <= avail;
                    // the OC vectors of A and B (S then U, contiguous) copied beside them when they fit
                    const int nva = 2 * ncA * nm, nvb = 2 * ncB * nm, vbase = 2 * nlb;
                    const bool stage = vbase + nva + nvb <= avail;
                    if (stage) {
                        for (int e = tid; e < nva; e += CT_THREADS) W[vbase + e] = SA[e];
                        for (int e = tid; e < nvb; e += CT_THREADS) W[vbase + nva + e] = SB[e];
                        SA = W + vbase;
                        UA = SA + ncA * nm;
                        SB = W + vbase + nva;
                        UB = SB + ncB * nm;
                        sync();
                    }
                    if (pre1) {
                        for (int e = tid; e < nlb; e += CT_THREADS) {
                            const int m1 = e % nm, je = e / nm, eta = je % 3, j = je / 3;
                            const double *Le = Lm + eta * nm * nm;
                            double lsb = 0.0, lub = 0.0;
                            for (int m2 = 0; m2 < nm; m2++) {
                                lsb += Le[m1 * nm + m2] * SB[j * nm + m2];
                                lub += Le[m1 * nm + m2] * UB[j * nm + m2];
                            }
                            W[e] = lsb;
                            W[nlb + e] = lub;
                        }
                        sync();
                    }
#pragma unroll
                    for (int q = 0; q < CT_EPT; q++) {
                        const int e = tid + q * CT_THREADS;
                        if (e >= nel) break;
                        const int i = r0 + e / ncB, j = e % ncB;
                        if (oc1) {
                            double s = 0.0;
                            for (int m = 0; m < nm; m++) s += SA[i * nm + m] * UB[j * nm + m] + UA[i * nm + m] * SB[j * nm + m];
                            acc[0][q] += af * s;
#pragma unroll
                            for (int eta = 0; eta < 3; eta++) {
                                const double *Le = Lm + eta * nm * nm;
                                double so = 0.0;
                                for (int m1 = 0; m1 < nm; m1++) {
                                    double lsb = 0.0, lub = 0.0;
                                    if (pre1) {
                                        lsb = W[(j * 3 + eta) * nm + m1];
                                        lub = W[nlb + (j * 3 + eta) * nm + m1];
                                    } else {
                                        for (int m2 = 0; m2 < nm; m2++) {
                                            lsb += Le[m1 * nm + m2] * SB[j * nm + m2];
                                            lub += Le[m1 * nm + m2] * UB[j * nm + m2];
                                        }
                                    }
                                    so += UA[i * nm + m1] * lsb + SA[i * nm + m1] * lub;
                                }
                                acc[1 + eta][q] += ef * so;
                            }
                        }
                    }
                    // part 2: -P_nlj [(U + U')/2] P_n'lj for OC entries of equal (L, J).  Each output
                    // still takes part 1 and then the oj terms in order: the loops over outputs and
                    // oj are exchanged so that the L-products of an oj are formed once
                    for (int oj = 0; oc2 && oj < no; oj++) {
                        if (oc_pair && (oi != c.oc_i || oj != c.oc_j)) continue;
                        const double d = J.delta[c.delta_off + oi * no + oj];
                        if (d == 0.0) continue;
                        const double *SB2 = J.OCV + J.ocv_tab[(long long) (c.ob + oj) * J.nshell + it.b];
                        const bool pre2 = nlb <= avail;
                        if (pre2) {
                            sync();   // the previous products are read by every thread first
                            if (stage) {   // S_B' over the copy of U_B (part 2 reads S_A and S_B' only)
                                for (int e = tid; e < ncB * nm; e += CT_THREADS) W[vbase + nva + ncB * nm + e] = SB2[e];
                                SB2 = W + vbase + nva + ncB * nm;
                                sync();
                            }
                            for (int e = tid; e < nlb; e += CT_THREADS) {
                                const int m1 = e % nm, je = e / nm, eta = je % 3, j = je / 3;
                                const double *Le = Lm + eta * nm * nm;
                                double l2 = 0.0;
                                for (int m2 = 0; m2 < nm; m2++) l2 += Le[m1 * nm + m2] * SB2[j * nm + m2];
                                W[e] = l2;
                            }
                            sync();
                        }
#pragma unroll
                        for (int q = 0; q < CT_EPT; q++) {
                            const int e = tid + q * CT_THREADS;
                            if (e >= nel) break;
                            const int i = r0 + e / ncB, j = e % ncB;
                            double s2 = 0.0;
                            for (int m = 0; m < nm; m++) s2 += SA[i * nm + m] * SB2[j * nm + m];
                            acc[0][q] += (-1.0) * d * af * s2;
#pragma unroll
                            for (int eta = 0; eta < 3; eta++) {
                                const double *Le = Lm + eta * nm * nm;
                                double so = 0.0;
                                for (int m1 = 0; m1 < nm; m1++) {
                                    double l2 = 0.0;
                                    if (pre2) {
                                        l2 = W[(j * 3 + eta) * nm + m1];
                                    } else {
                                        for (int m2 = 0; m2 < nm; m2++) l2 += Le[m1 * nm + m2] * SB2[j * nm + m2];
                                    }
                                    so += SA[i * nm + m1] * l2;
                                }
                                acc[1 + eta][q] += (-1.0) * d * ef * so;
                            }
                        }
                    }
                    sync();   // W is rewritten by the next OC entry
                }
            }
        }
        // ---- this centre's slice of the partial blocks ----
#pragma unroll
        for (int q = 0; q < CT_EPT; q++) {
            const int e = tid + q * CT_THREADS;
            if (e >= nel) break;
            const int i = r0 + e / ncB, j = e % ncB;
#pragma unroll
            for (int op = 0; op < 4; op++) partial[pt.off + ((long long) op * ncA + i) * ncB + j] = acc[op][q];
        }
        sync();
    }
}

// One workgroup per (shell pair, centre) (PartItem): the centre's contribution to the four blocks
// goes to its own slice of the contraction's partial buffer; k_contract_reduce adds, in centre
// order, each centre's type-1 slice plus its contraction slice.
template <int CT_THREADS>
__global__ __launch_bounds__(CT_THREADS) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_contract(const PartItem *__restrict__ partlist, const WorkItem *__restrict__ items,
                                                  DevJob J, DevTables T, double *__restrict__ partial, int parts,
                                                  int nm_max, int qs_len, int ts_len)
{
    // dynamic LDS sized by the plan: Q block (qs_len doubles) + (i, m, k2) intermediate (ts_len)
    extern __shared__ double ct_lds[];
    // runs of 32 parts per XCD: neighbours in the launch order read the same bra projectors
    const int b = xcd_chunked(blockIdx.x, gridDim.x, 32);
    contract_part<CT_THREADS, 256 / CT_THREADS, false>(partlist[b], items, J, T, partial, parts, nm_max, ts_len, ct_lds,
                                     ct_lds + qs_len, threadIdx.x);
}

// Parts of at most 64 Cartesian pairs (the tail of the launch order): four waves per workgroup,
// one part per wave, each with its own LDS (qs_len + ts_len doubles); the sums are those of
// k_contract (same chunking of rows into one pass, same order of every sum)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_contract_w(const PartItem *__restrict__ partlist, int n_parts,
                                                  const WorkItem *__restrict__ items, DevJob J, DevTables T,
                                                  double *__restrict__ partial, int parts, int nm_max, int qs_len,
                                                  int ts_len)
{
    extern __shared__ double ct_lds[];
    // the part index is uniform over the wave: with readfirstlane its records are scalar loads kept
    // in scalar registers (118 VGPRs and no spill instead of 128 and 39 spilled)
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), p = xcd_chunked(blockIdx.x, gridDim.x, 32) * 4 + w;
    if (p >= n_parts) return;   // whole waves leave; nothing below synchronises the block
    double *base = ct_lds + (size_t) w * (qs_len + ts_len);
    contract_part<64, 1, true>(partlist[p], items, J, T, partial, parts, nm_max, ts_len, base, base + qs_len,
                            threadIdx.x & 63);
}

// ------------------------------------------------------------------ launchers
void launch_tab_F(int n_items, hipStream_t st, const TabItem *items, const DevJob &J, const DevTables &T, double *F)
{
    hipLaunchKernelGGL(k_tab_F, dim3(NPTS / 256, n_items), dim3(256), 0, st, items, J, T, F);
}

void launch_screen(int sp, int ss, int n_tiles, hipStream_t st, const TileItem *tiles, const QJob *jobs,
                   const RowDesc *rows, const DevJob &J, const DevTables &T, double *scr,
                   unsigned long long *counters)
{
    auto screen = ss == 16 ? (sp == 1 ? k_screen<1, 16> : sp == 2 ? k_screen<2, 16> : k_screen<4, 16>)
                : ss == 8  ? (sp == 1 ? k_screen<1, 8> : sp == 2 ? k_screen<2, 8> : k_screen<4, 8>)
                           : (sp == 1 ? k_screen<1, 4> : sp == 2 ? k_screen<2, 4> : k_screen<4, 4>);
    hipLaunchKernelGGL(screen, dim3(ss * n_tiles), dim3(64), 0, st, tiles, jobs, rows, J, T, scr, counters);
}

void launch_contract_w(int n_parts, size_t lds_bytes_per_wave, hipStream_t st, const PartItem *partlist,
                       const WorkItem *items, const DevJob &J, const DevTables &T, double *partial, int parts,
                       int nm_max, int qs_len, int ts_len)
{
    hipLaunchKernelGGL(k_contract_w, dim3((n_parts + 3) / 4), dim3(256), 4 * lds_bytes_per_wave, st, partlist, n_parts,
                       items, J, T, partial, parts, nm_max, qs_len, ts_len);
}

const void *contract_kernel_ptr(int threads)
{
    return threads == 64 ? (const void *) k_contract<64> : threads == 256 ? (const void *) k_contract<256>
                                                                        : (const void *) k_contract<128>;
}

void launch_contract(int threads, int n_parts, size_t lds_bytes, hipStream_t st, const PartItem *partlist,
                     const WorkItem *items, const DevJob &J, const DevTables &T, double *partial, int parts,
                     int nm_max, int qs_len, int ts_len)
{
    auto k = threads == 64 ? k_contract<64> : threads == 256 ? k_contract<256> : k_contract<128>;
    hipLaunchKernelGGL(k, dim3(n_parts), dim3(threads), lds_bytes, st, partlist, items, J, T, partial, parts,
                       nm_max, qs_len, ts_len);
}
