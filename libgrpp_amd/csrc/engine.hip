// MI355X (gfx950) engine for the GRPP integral path.
//
// A job is a list of shells, a list of GRPP centres and a list of output blocks.  It runs as a
// fixed sequence of kernels on one stream:
//
//   k_tab_F     F^{C,A}_lambda(r_i): contracted radial functions of every shell A about every
//               centre C on the level-ordered Log3 grid (radial_type2_integral.c:465-489);
//   (host)      w_i U_p(r_i) for every semilocal / spin-orbit channel p (potential.c:70-81) is
//               tabulated once per plan on the host with the C library's exp and pow, and kept as
//               sign * sqrt|.| so that both operands of the product carry half of it;
//   k_screen    the Shaw-Hill envelope estimate of every radial entry (screening.c:193-378), one
//               wave per entry; it steers the quadrature exactly as in the reference;
//   k_radial    the radial integrals of ALL shell pairs of a centre and channel as one FP64
//               product Q = G diag(wU) G^T over grid points (G sqrt|wU| times sign(wU) G^T sqrt|wU|), rows (shell, lambda, n_A) with
//               G = r^{n_A} F; the K loop runs level by level (31, 63, ..., 8191 points) and
//               applies libgrpp's per-entry nested-doubling stop rule at each level boundary
//               (radial_type2_integral.c:242-318); tiles stop early when every entry stopped;
//               k_radial_mfma is the same product on v_mfma_f64_16x16x4f64 (the default);
//   k_canon     every split of (lambda1, lambda2, N) reads one canonical split, as libgrpp keeps
//               one Q^N per N (type2_integrals.c:183-186);
//   k_proj      angular projectors P^{C,A,L}[i][m][(lambda,n_A)] = binomial expansion of the
//               Cartesian component about C times Omega^lambda_{Lm} (type2_integrals.c:117-211,
//               angular_integrals.c:63-111), and their L-matrix products for spin-orbit;
//   k_ocv       outercore overlap / type-1 vectors <A|phi_m>, <A|U|phi_m> (outercore_integrals.c:147-270);
//   k_delta     outercore one-centre radial constants (outercore_integrals.c:397-465);
//   k_type1     type-1 (local) McMurchie-Davidson blocks (type1_mcmurchie_davidson.c:129-644),
//               written first into the output blocks;
//   k_contract  sum over channels of 16 pi^2 sum_m P^A_m Q P^B_m^T (and the SO analogue with
//               L-matrices) plus the outercore products, added into the output blocks.
//
// k_tab_F, k_screen and k_contract are defined in valu_kernels.hip (compiled with other flags) and
// launched through the launch_* functions of engine_dev.h; the records they share are there too.
//
// All arithmetic is FP64.  Every shell-pair block is written by one workgroup, so no atomics
// are needed on the outputs; blocks of (B,A) are produced from (A,B) by (anti)symmetry.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "device_math.h"
#include "engine.h"
#include "tables.h"

using namespace grpp;

// ------------------------------------------------------------------ errors
static thread_local char g_err[512] = "";

extern "C" void eng_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

extern "C" const char *libgrpp_last_error(void) { return g_err; }

#define HIPCHK(x)                                                                               \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) {                                                                 \
            eng_set_error("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
            return 1;                                                                           \
        }                                                                                       \
    } while (0)

#include "engine_dev.h"


// libgrpp keeps ONE radial integral per (lambda1, lambda2, N = nA + nB) and uses it for every split
// (type2_integrals.c:183-186); the product above evaluates each split separately.  This pass makes
// every split read one canonical split (smallest n on the side of the smaller (shell, lambda) key,
// which keeps the table symmetric), so rounding differences between splits cannot leak into the
// cancellation of the binomial expansion.
__global__ __launch_bounds__(256) void k_canon(const TileItem *__restrict__ tiles, const QJob *__restrict__ jobs,
                                               const RowDesc *__restrict__ rows, DevJob J,
                                               const double *__restrict__ Q, double *__restrict__ Qc,
                                               int upper_only)
{
    // upper_only: every work item has shell a <= shell b, so the contraction reads only the entries
    // whose row shell is at most their column shell; the others are not written
    const TileItem t = tiles[blockIdx.x];
    const QJob jb = jobs[t.job];
    const RowDesc *R = rows + jb.rows;
    const int L = jb.L;
    for (int e = threadIdx.x; e < 2 * RT * RT; e += blockDim.x) {
        const int half = e / (RT * RT), q = e % (RT * RT);
        int row = t.ti * RT + q / RT, col = t.tj * RT + q % RT;
        if (half) {
            if (t.ti == t.tj) continue;
            const int tmp = row;
            row = col;
            col = tmp;
        }
        if (row >= jb.M || col >= jb.M) continue;
        const RowDesc ra = R[row], rb = R[col];
        if (upper_only && ra.shell > rb.shell) continue;
        const int la = J.sh[ra.shell].L, lb = J.sh[rb.shell].L, N = ra.n + rb.n;
        const bool a_first = (ra.shell < rb.shell) || (ra.shell == rb.shell && ra.lam <= rb.lam);
        const int cn = a_first ? dev_canon_nA(L, la, lb, ra.lam, rb.lam, N) : dev_canon_nA_max(L, la, lb, ra.lam, rb.lam, N);
        const int r2 = ra.lidx - dev_kcol(L, ra.n, ra.lam) + dev_kcol(L, cn, ra.lam);
        const int c2 = rb.lidx - dev_kcol(L, rb.n, rb.lam) + dev_kcol(L, N - cn, rb.lam);
        // written in the LOGICAL row order: the contraction reads a shell pair's block as KB
        // consecutive doubles per row, with no row permutation to look up
        Qc[jb.qoff + (long long) ra.lidx * jb.M + rb.lidx] =
            Q[jb.qoff + (long long) J.rowpos[jb.rows + r2] * jb.M + J.rowpos[jb.rows + c2]];
    }
}

// ================================================================== radial quadrature as a tiled product

__global__ __launch_bounds__(256) void k_radial(const TileItem *__restrict__ tiles, const QJob *__restrict__ jobs,
                                                const RowDesc *__restrict__ rows, const double *__restrict__ F,
                                                const double *__restrict__ rpow, const double *__restrict__ wU,
                                                double *__restrict__ Q, int min_level, double tol,
                                                unsigned long long *__restrict__ counters,
                                                const double *__restrict__ scr, const DShell *__restrict__ J_sh,
                                                const int *__restrict__ pos, signed char *__restrict__ lev_out)
{
    __shared__ double As[RK][RT + 2];
    __shared__ double Bs[RK][RT + 2];
    __shared__ double Rw[LMAX_SHELL + 1][RK], Rs[LMAX_SHELL + 1][RK];   // r^n sqrt|wU|, r^n sign(wU) sqrt|wU|
    __shared__ int sfa[RT], sna[RT], sfb[RT], snb[RT];
    const TileItem t = tiles[blockIdx.x];
    const QJob jb = jobs[t.job];
    const int tid = threadIdx.x;
    const RowDesc *R = rows + jb.rows;
    if (tid < RT) {
        int r = t.ti * RT + tid;
        sfa[tid] = (r < jb.M) ? R[r].frow : -1;
        sna[tid] = (r < jb.M) ? R[r].n : 0;
    } else if (tid < 2 * RT) {
        int r = t.tj * RT + tid - RT;
        sfb[tid - RT] = (r < jb.M) ? R[r].frow : -1;
        snb[tid - RT] = (r < jb.M) ? R[r].n : 0;
    }
    __syncthreads();
    const int tx = tid & 15, ty = tid >> 4;
    const double *wp = wU + (size_t) jb.pot * NPTS;

    double acc[4][4], S[4][4], res[4][4], est[4][4];
    unsigned done = 0;
    int stop[4][4];   // stopping level per entry: -1 screened out, 9 not converged by the last level
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            S[a][b] = 0.0;
            res[a][b] = 0.0;
            stop[a][b] = 9;
            est[a][b] = __longlong_as_double(0x7ff8000000000000LL);
            const int row = t.ti * RT + ty * 4 + a, col = t.tj * RT + tx * 4 + b;
            if (row >= jb.M || col >= jb.M) {
                done |= 1u << (a * 4 + b);
            } else if (scr) {
                // screened below the tolerance: the envelope estimate is the result (radial_type2_integral.c:277-280)
                est[a][b] = scr[jb.qoff + canon_index(R, J_sh, pos, jb, row, col)];
                if (!isnan(est[a][b]) && fabs(est[a][b]) < tol) {
                    done |= 1u << (a * 4 + b);
                    res[a][b] = est[a][b];
                    stop[a][b] = -1;
                }
            }
        }

    int lev_done = 0;
    for (int lev = 0; lev <= jb.maxlev; lev++) {
        const int kb = (lev == 0) ? 0 : (16 << lev);
        const int ke = (lev == 0) ? 32 : (32 << lev);
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) acc[a][b] = 0.0;
        for (int k0 = kb; k0 < ke; k0 += RK) {
            // the K step's point weights once per power n, then one product per staged element
            for (int e = tid; e < (LMAX_SHELL + 1) * RK; e += 256) {
                const int n = e / RK, kk = e % RK, k = k0 + kk;
                const double w = wp[k], rp = rpow[n * NPTS + k];
                Rw[n][kk] = rp * fabs(w);
                Rs[n][kk] = rp * w;
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int e = tid + 256 * q, r = e >> 5, kk = e & 31, k = k0 + kk;
                const int fa = sfa[r], fb = sfb[r];
                As[kk][r] = (fa >= 0) ? F[(size_t) fa * NPTS + k] * Rw[sna[r]][kk] : 0.0;
                Bs[kk][r] = (fb >= 0) ? F[(size_t) fb * NPTS + k] * Rs[snb[r]][kk] : 0.0;
            }
            __syncthreads();
#pragma unroll 8
            for (int kk = 0; kk < RK; kk++) {
                double av[4], bv[4];
#pragma unroll
                for (int a = 0; a < 4; a++) av[a] = As[kk][ty * 4 + a];
#pragma unroll
                for (int b = 0; b < 4; b++) bv[b] = Bs[kk][tx * 4 + b];
#pragma unroll
                for (int a = 0; a < 4; a++)
#pragma unroll
                    for (int b = 0; b < 4; b++) acc[a][b] = fma(av[a], bv[b], acc[a][b]);
            }
            __syncthreads();
        }
        // libgrpp's stop rule per entry: S_k = S_{k-1}/2 + (new points), stop once |S_k - S_{k-1}| <= tol
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const unsigned bit = 1u << (a * 4 + b);
                if (lev == 0) {
                    S[a][b] = acc[a][b];
                } else {
                    const double prev = S[a][b];
                    S[a][b] = 0.5 * S[a][b] + acc[a][b];
                    // no stop while |S| < 1e-3 |estimate| (radial_type2_integral.c:308-311)
                    const bool forced = !isnan(est[a][b]) && (fabs(S[a][b]) / fabs(est[a][b]) < 0.001);
                    if (!(done & bit) && !forced && lev >= min_level && fabs(S[a][b] - prev) <= tol) {
                        done |= bit;
                        res[a][b] = S[a][b];
                        stop[a][b] = lev;
                    }
                }
            }
        lev_done = lev;
        if (!__syncthreads_or(done != 0xFFFFu)) break;
    }
    if (tid == 0 && counters) {
        // algorithmic work: distinct table entries of the tile (upper triangle) x points integrated
        const int nr = min(RT, jb.M - t.ti * RT), nc = min(RT, jb.M - t.tj * RT);
        const unsigned long long ent = (t.ti == t.tj) ? (unsigned long long) nr * (nr + 1) / 2
                                                       : (unsigned long long) nr * nc;
        atomicAdd(&counters[0], (unsigned long long) (32ull << lev_done));   // grid points integrated
        atomicAdd(&counters[1], 1ull);                                       // tiles
        atomicAdd(&counters[2], ent * (32ull << lev_done));                  // entry x point products
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const unsigned bit = 1u << (a * 4 + b);
            const int row = t.ti * RT + ty * 4 + a, col = t.tj * RT + tx * 4 + b;
            if (row >= jb.M || col >= jb.M) continue;
            double v = (done & bit) ? res[a][b] : S[a][b];
            if (fabs(v) < jb.thr) v = 0.0;   // |Q| skips of type2_integrals.c:185 / spin_orbit_integrals.c:205
            if (t.ti < t.tj || row <= col) {
                Q[jb.qoff + (long long) row * jb.M + col] = v;
                Q[jb.qoff + (long long) col * jb.M + row] = v;
                if (lev_out) {
                    lev_out[jb.qoff + (long long) row * jb.M + col] = (signed char) stop[a][b];
                    lev_out[jb.qoff + (long long) col * jb.M + row] = (signed char) stop[a][b];
                }
            }
        }
}

// The same quadrature on the FP64 matrix cores: each wave owns a 32x32 quarter of the tile as
// 2x2 blocks of v_mfma_f64_16x16x4f64 (A[i][k] in lane i+16k, B[k][j] in lane j+16k, D[l/16+4r][l%16]
// in element r of lane l, as measured on gfx950); the per-entry stop rule is unchanged.
typedef double v4d __attribute__((ext_vector_type(4)));

template <bool REC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_radial_mfma(const TileItem *__restrict__ tiles, const QJob *__restrict__ jobs,
                                                     const RowDesc *__restrict__ rows, const double *__restrict__ F,
                                                     const double *__restrict__ rpow, const double *__restrict__ wU,
                                                     double *__restrict__ Q, int min_level, double tol,
                                                     unsigned long long *__restrict__ counters,
                                                     const double *__restrict__ scr, const DShell *__restrict__ J_sh,
                                                     const int *__restrict__ pos, signed char *__restrict__ lev_out,
                                                     int *__restrict__ tile_lev)
{
    __shared__ double As[RK][RT + 1];   // odd row length: the staging stores hit distinct banks
    __shared__ double Bs[RK][RT + 1];
    __shared__ double Rw[LMAX_SHELL + 1][RK], Rs[LMAX_SHELL + 1][RK];   // r^n sqrt|wU|, r^n sign(wU) sqrt|wU|
    __shared__ int sfa[RT], sna[RT], sfb[RT], snb[RT];
    const TileItem t = tiles[blockIdx.x];
    const QJob jb = jobs[t.job];
    const int tid = threadIdx.x;
    const RowDesc *R = rows + jb.rows;
    if (tid < RT) {
        int r = t.ti * RT + tid;
        sfa[tid] = (r < jb.M) ? R[r].frow : -1;
        sna[tid] = (r < jb.M) ? R[r].n : 0;
    } else if (tid < 2 * RT) {
        int r = t.tj * RT + tid - RT;
        sfb[tid - RT] = (r < jb.M) ? R[r].frow : -1;
        snb[tid - RT] = (r < jb.M) ? R[r].n : 0;
    }
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    const double *wp = wU + (size_t) jb.pot * NPTS;

    v4d acc[2][2];
    // S: the running sum of an entry, frozen once the entry is settled (converged or screened out,
    // then its estimate), so it is also the entry's result
    double S[2][2][4], est[2][2][4];
    unsigned done = 0;
    int stop[2][2][4];   // stopping level per entry: -1 screened out, 9 not converged by the last level
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                S[a][b][r] = 0.0;
                if (REC) stop[a][b][r] = 9;
                est[a][b][r] = __longlong_as_double(0x7ff8000000000000LL);
                const int row = t.ti * RT + wr + a * 16 + lk + 4 * r, col = t.tj * RT + wc + b * 16 + li;
                if (row >= jb.M || col >= jb.M) {
                    done |= 1u << (a * 8 + b * 4 + r);
                } else if (scr) {
                    est[a][b][r] = scr[jb.qoff + canon_index(R, J_sh, pos, jb, row, col)];
                    if (!isnan(est[a][b][r]) && fabs(est[a][b][r]) < tol) {
                        done |= 1u << (a * 8 + b * 4 + r);
                        S[a][b][r] = est[a][b][r];
                        if (REC) stop[a][b][r] = -1;
                    }
                }
            }

    // The K steps of the levels are contiguous in the point order (level l covers points
    // [16 << l, 32 << l)), so the global operands of the next step are loaded into registers while
    // the matrix cores work on the current one.
    constexpr int NW = ((LMAX_SHELL + 1) * RK + 255) / 256;
    double fa_n[8], fb_n[8], w_n[NW], rp_n[NW];
    // staging: thread tid owns row sr = tid / 4 of both operands and the 8 consecutive points
    // sk0 .. sk0 + 7 of a K step (four 16-byte loads per row, its F row and power n read once per tile)
    const int sr = tid >> 2, sk0 = (tid & 3) * 8;
    const int my_fa = sfa[sr], my_fb = sfb[sr], my_na = sna[sr], my_nb = snb[sr];
    auto prefetch = [&](int k0) {
        const double2 *pa = (const double2 *) (F + (size_t) (my_fa >= 0 ? my_fa : 0) * NPTS + k0 + sk0);
        const double2 *pb = (const double2 *) (F + (size_t) (my_fb >= 0 ? my_fb : 0) * NPTS + k0 + sk0);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const double2 va = pa[q], vb = pb[q];
            fa_n[2 * q] = (my_fa >= 0) ? va.x : 0.0;
            fa_n[2 * q + 1] = (my_fa >= 0) ? va.y : 0.0;
            fb_n[2 * q] = (my_fb >= 0) ? vb.x : 0.0;
            fb_n[2 * q + 1] = (my_fb >= 0) ? vb.y : 0.0;
        }
#pragma unroll
        for (int j = 0; j < NW; j++) {
            const int e = tid + 256 * j;
            if (e < (LMAX_SHELL + 1) * RK) {
                const int k = k0 + e % RK;
                w_n[j] = wp[k];
                rp_n[j] = rpow[(e / RK) * NPTS + k];
            }
        }
    };
    prefetch(0);
    int lev_done = 0;
    for (int lev = 0; lev <= jb.maxlev; lev++) {
        const int kb = (lev == 0) ? 0 : (16 << lev);
        const int ke = (lev == 0) ? 32 : (32 << lev);
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) acc[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};
        for (int k0 = kb; k0 < ke; k0 += RK) {
            // the K step's point weights once per power n, then one product per staged element
#pragma unroll
            for (int j = 0; j < NW; j++) {
                const int e = tid + 256 * j;
                if (e < (LMAX_SHELL + 1) * RK) {
                    Rw[e / RK][e % RK] = rp_n[j] * fabs(w_n[j]);
                    Rs[e / RK][e % RK] = rp_n[j] * w_n[j];
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 8; q++) {
                As[sk0 + q][sr] = (my_fa >= 0) ? fa_n[q] * Rw[my_na][sk0 + q] : 0.0;
                Bs[sk0 + q][sr] = (my_fb >= 0) ? fb_n[q] * Rs[my_nb][sk0 + q] : 0.0;
            }
            __syncthreads();
            if (k0 + RK < NPTS) prefetch(k0 + RK);   // in bounds: the last step ends at NPTS
#pragma unroll
            for (int kk = 0; kk < RK; kk += 4) {
                const double a0 = As[kk + lk][wr + li], a1 = As[kk + lk][wr + 16 + li];
                const double b0 = Bs[kk + lk][wc + li], b1 = Bs[kk + lk][wc + 16 + li];
                acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
                acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
                acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
                acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
            }
            __syncthreads();
        }
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const unsigned bit = 1u << (a * 8 + b * 4 + r);
                    const double v = acc[a][b][r];
                    if (done & bit) continue;
                    if (lev == 0) {
                        S[a][b][r] = v;
                    } else {
                        const double prev = S[a][b][r];
                        S[a][b][r] = 0.5 * S[a][b][r] + v;
                        const bool forced = !isnan(est[a][b][r]) && (fabs(S[a][b][r]) / fabs(est[a][b][r]) < 0.001);
                        if (!forced && lev >= min_level && fabs(S[a][b][r] - prev) <= tol) {
                            done |= bit;
                            if (REC) stop[a][b][r] = lev;
                        }
                    }
                }
        lev_done = lev;
        if (!__syncthreads_or(done != 0xFFFFu)) break;
    }
    if (tid == 0 && tile_lev) tile_lev[t.pad] = lev_done;   // the tile's last level, for the next launch order
    if (tid == 0 && counters) {
        const int nr = min(RT, jb.M - t.ti * RT), nc = min(RT, jb.M - t.tj * RT);
        const unsigned long long ent = (t.ti == t.tj) ? (unsigned long long) nr * (nr + 1) / 2
                                                       : (unsigned long long) nr * nc;
        atomicAdd(&counters[0], (unsigned long long) (32ull << lev_done));
        atomicAdd(&counters[1], 1ull);
        atomicAdd(&counters[2], ent * (32ull << lev_done));
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = t.ti * RT + wr + a * 16 + lk + 4 * r, col = t.tj * RT + wc + b * 16 + li;
                if (row >= jb.M || col >= jb.M) continue;
                double v = S[a][b][r];
                if (fabs(v) < jb.thr) v = 0.0;
                if (t.ti < t.tj || row <= col) {
                    Q[jb.qoff + (long long) row * jb.M + col] = v;
                    Q[jb.qoff + (long long) col * jb.M + row] = v;
                    if (REC) {
                        lev_out[jb.qoff + (long long) row * jb.M + col] = (signed char) stop[a][b][r];
                        lev_out[jb.qoff + (long long) col * jb.M + row] = (signed char) stop[a][b][r];
                    }
                }
            }
}

// ================================================================== angular projectors
constexpr int PROJ_LMAX = LMAX_POT + LMAX_SHELL;  // lambda <= L + l

// Om (the Omega block of one n) is dynamic LDS of the plan's largest size, om_len doubles
__global__ __launch_bounds__(256) void k_proj(const ProjItem *__restrict__ items, DevJob J, DevTables T,
                                              double *__restrict__ P)
{
    __shared__ double Sk[(PROJ_LMAX + 1) * (2 * PROJ_LMAX + 1)];
    __shared__ double CApow[3][LMAX_SHELL + 1];   // pow(CA_d, k), k <= l_A (type2_integrals.c:132-142)
    extern __shared__ double Om[];
    const ProjItem it = items[blockIdx.x];
    const DShell s = J.sh[it.A];
    const DCenter c = J.ctr[it.c];
    const int L = it.L, la = s.L, nm = 2 * L + 1, nc = dev_ncart(la), K = it.K;
    const int tid = threadIdx.x;
    const double CAx = c.x - s.x, CAy = c.y - s.y, CAz = c.z - s.z;
    // k_A = -2 alpha_0 CA; only its direction enters (type2_integrals.c:66-81)
    const double a0 = J.alpha[s.poff];
    const double kx = -2.0 * (a0 * CAx), ky = -2.0 * (a0 * CAy), kz = -2.0 * (a0 * CAz);
    const double len = sqrt(kx * kx + ky * ky + kz * kz);
    double ux = 0, uy = 0, uz = 0;
    if (len > 1e-14) {
        const double inv = 1.0 / len;
        ux = kx * inv;
        uy = ky * inv;
        uz = kz * inv;
    }
    const int lmx = L + la;
    if (tid < 3 * (la + 1)) {
        const int d = tid / (la + 1), k = tid - d * (la + 1);
        CApow[d][k] = pow(d == 0 ? CAx : d == 1 ? CAy : CAz, k);
    }
    // real harmonics S_{lambda mu}(k) with the |S| < 1e-14 skip of angular_integrals.c:77
    for (int e = tid; e < (lmx + 1) * (2 * PROJ_LMAX + 1); e += blockDim.x) {
        const int lam = e / (2 * PROJ_LMAX + 1), mu = e % (2 * PROJ_LMAX + 1);
        double v = 0.0;
        if (mu <= 2 * lam) {
            const int ncl = (lam + 1) * (lam + 2) / 2;
            const double *y = T.rsh + T.rsh_off[lam] + mu * ncl;
            const int *cb = T.comb + T.comb_off[lam];
            for (int q = 0; q < ncl; q++) {
                double p = 1.0;
                for (int z = 0; z < cb[3 * q]; z++) p *= ux;
                double py = 1.0;
                for (int z = 0; z < cb[3 * q + 1]; z++) py *= uy;
                double pz = 1.0;
                for (int z = 0; z < cb[3 * q + 2]; z++) pz *= uz;
                v += y[q] * (p * py * pz);
            }
            if (fabs(v) < 1e-14) v = 0.0;
        }
        Sk[e] = v;
    }
    __syncthreads();
    int colbase = 0;
    for (int n = 0; n <= la; n++) {
        const int cnt = dev_lam_cnt(L, n), lo = dev_lam_lo(L, n), nab = (n + 1) * (n + 2) / 2;
        const int wo = T.wdir[L * (T.w_nmax + 1) + n], ws = T.wstride[L * (T.w_nmax + 1) + n];
        // Omega[j][m][abc] = sum_mu S_{lam_j mu} W[j][m][mu][abc]
        const int nom = cnt * nm * nab;
        for (int e = tid; e < nom; e += blockDim.x) {
            const int q = e % nab, jm = e / nab, m = jm % nm, j = jm / nm;
            const int lam = lo + 2 * j;
            const double *w = T.W + wo + ((size_t) (j * nm + m) * ws) * nab + q;
            const double *sk = Sk + lam * (2 * PROJ_LMAX + 1);
            double o = 0.0;
            for (int mu = 0; mu <= 2 * lam; mu++) o += w[(size_t) mu * nab] * sk[mu];
            Om[e] = o;
        }
        __syncthreads();
        // three projectors: ket (all Omega), bra with the type-2 skip |Omega| < 1e-16
        // (type2_integrals.c:232-234) and bra with the SO skip |Omega| < 1e-14 (spin_orbit_integrals.c:275-277)
        // P[i][m][colbase + j] = sum_{abc <= cart_i, a+b+c = n} binomials * CA powers * Omega
        const int np = nc * nm * cnt;
        for (int e = tid; e < np; e += blockDim.x) {
            const int j = e % cnt, im = e / cnt, m = im % nm, i = im / nm;
            int cx, cy, cz;
            dev_cart(la, i, cx, cy, cz);
            double v = 0.0, v2 = 0.0, v3 = 0.0;
            for (int a = 0; a <= cx && a <= n; a++) {
                const double fa = dev_binom(cx, a) * CApow[0][cx - a];
                for (int b = 0; b <= cy && a + b <= n; b++) {
                    const int cc = n - a - b;
                    if (cc > cz) continue;
                    const double f = fa * dev_binom(cy, b) * CApow[1][cy - b] * dev_binom(cz, cc) * CApow[2][cz - cc];
                    const int q = (n - a) * (n - a + 1) / 2 + (n - a - b);
                    const double o = Om[(j * nm + m) * nab + q];
                    v += f * o;
                    if (fabs(o) >= 1e-16) v2 += f * o;
                    if (fabs(o) >= 1e-14) v3 += f * o;
                }
            }
            const long long idx = ((long long) i * nm + m) * K + colbase + j;
            P[it.poff + idx] = v;
            P[it.pb2 + idx] = v2;
            P[it.pb3 + idx] = v3;
        }
        colbase += cnt;
        __syncthreads();
    }
    if (it.ploff < 0) return;
    __threadfence();
    __syncthreads();
    // ket-side spin-orbit projectors PL_eta[i][m1][k] = sum_m2 L_eta[m1][m2] P[i][m2][k]
    const double *Lm = T.lmat + T.lmat_off[L];
    const int tot = 3 * nc * nm * K;
    for (int e = tid; e < tot; e += blockDim.x) {
        const int k = e % K, r1 = e / K, m1 = r1 % nm, r2 = r1 / nm, i = r2 % nc, eta = r2 / nc;
        const double *Le = Lm + eta * nm * nm + m1 * nm;
        double v = 0.0;
        for (int m2 = 0; m2 < nm; m2++) v += Le[m2] * P[it.poff + ((long long) i * nm + m2) * K + k];
        P[it.ploff + e] = v;
    }
}

// ================================================================== McMurchie-Davidson blocks (type-1 and outercore)
// E^{ij}_t for one Cartesian direction (type1_mcmurchie_davidson.c:591-644); layout [i][j][t], dims (la+1)(lb+1)(N+1)
__device__ void mmd_E_dir(double *E, int la, int lb, double q, double K, double XQA, double XQB)
{
    const int N1 = la + lb + 1;
#define EI(i, j, t) E[((i) * (lb + 1) + (j)) * N1 + (t)]
    for (int i = 0; i <= la; i++)
        for (int j = 0; j <= lb; j++)
            for (int t = 0; t <= i + j; t++) {
                if (t == 0) {
                    if (i == 0 && j == 0) EI(0, 0, 0) = K;
                    else if (i == 1 && j == 0) EI(1, 0, 0) = XQA * EI(0, 0, 0);
                    else if (i == 0 && j == 1) EI(0, 1, 0) = XQB * EI(0, 0, 0);
                    else if (i == 0) EI(i, j, 0) = XQB * EI(i, j - 1, 0) + EI(i, j - 1, 1);
                    else EI(i, j, 0) = XQA * EI(i - 1, j, 0) + EI(i - 1, j, 1);
                } else {
                    double e = 0.0, f = 1.0 / (2.0 * q * t);
                    if (i > 0) e += f * i * EI(i - 1, j, t - 1);
                    if (j > 0) e += f * j * EI(i, j - 1, t - 1);
                    EI(i, j, t) = e;
                }
            }
#undef EI
}

// The t = 0 column E^{ij}_0 alone, which is all an r^0 Gaussian term needs (its integral is
// (pi/q)^{3/2} E^x_0 E^y_0 E^z_0): E^{i-1,j}_1 is formed on the fly from the t = 0 entries by the
// same expression mmd_E_dir evaluates, so the values are those of the full table, at
// (la+1)(lb+1) entries per direction instead of (la+1)(lb+1)(la+lb+1).
__device__ void mmd_E0_dir(double *E, int la, int lb, double q, double K, double XQA, double XQB)
{
    const double f = 1.0 / (2.0 * q * 1);
#define E0(i, j) E[(i) * (lb + 1) + (j)]
    for (int i = 0; i <= la; i++)
        for (int j = 0; j <= lb; j++) {
            if (i == 0 && j == 0) {
                E0(0, 0) = K;
            } else if (i == 1 && j == 0) {
                E0(1, 0) = XQA * E0(0, 0);
            } else if (i == 0 && j == 1) {
                E0(0, 1) = XQB * E0(0, 0);
            } else if (i == 0) {
                double e1 = 0.0;   // E^{0,j-1}_1
                if (j - 1 > 0) e1 += f * (j - 1) * E0(0, j - 2);
                E0(0, j) = XQB * E0(0, j - 1) + e1;
            } else {
                double e1 = 0.0;   // E^{i-1,j}_1
                if (i - 1 > 0) e1 += f * (i - 1) * E0(i - 2, j);
                if (j > 0) e1 += f * j * E0(i - 1, j - 1);
                E0(i, j) = XQA * E0(i - 1, j) + e1;
            }
        }
#undef E0
}

__device__ __forceinline__ int simplex_idx(int t, int u, int v)
{
    const int s = t + u + v;
    return s * (s + 1) * (s + 2) / 6 + t * (s + 1) - t * (t - 1) / 2 + u;
}

struct MmdShell {
    int L, nprim, poff;
    double x, y, z;
};

// Accumulates sum_k d_k <A| r_C^{n_k-2} e^{-z_k r_C^2} |B> over the terms [t0, t1) into acc
// (LDS, element i*ncB + j owned by thread (i*ncB+j) % blockDim).  Coefficients are raw; the
// primitive factors (2a/pi)^{3/4}(4a)^{L/2} are applied as in type1_mcmurchie_davidson.c.
__device__ void mmd_accumulate(const MmdShell A, const MmdShell B, double Cx, double Cy, double Cz,
                               const int *tpow, const double *tcoef, const double *tzeta, int t0, int t1,
                               const double *alpha, const double *craw, double *acc, double *scratch,
                               int scratch_len, const double *boys)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    const int la = A.L, lb = B.L, N = la + lb, N1 = N + 1;
    const int ncA = dev_ncart(la), ncB = dev_ncart(lb), nel = ncA * ncB;
    const int esz = 3 * (la + 1) * (lb + 1) * N1;   // E for three directions
    const int e0sz = 3 * (la + 1) * (lb + 1);        // their t = 0 columns (r^0 terms)
    const int ssz = (N + 1) * (N + 2) * (N + 3) / 6;  // simplex layer
    __shared__ double w_s[64], q_s[64];
    __shared__ int n_s;
    const int npair = A.nprim * B.nprim;
    {
        // r^0 terms (n_k = 2, type1_mcmurchie_davidson.c:397-473) whose t = 0 E columns of every
        // (term, primitive pair) fit the scratch together: all tables in one pass and one barrier,
        // then every output sums them term by term and, within a term, in the same batches of
        // pairs as the loop below (bitwise the same sums, two barriers instead of two per term)
        bool all2 = true;
        for (int tk = t0; tk < t1; tk++) all2 = all2 && (tpow[tk] == 2);
        const int nunit = (t1 - t0) * npair, bmax = min(64, scratch_len / e0sz);
        if (all2 && nunit > 0 && nunit * (e0sz + 1) <= scratch_len) {
            double *wbuf = scratch + (size_t) nunit * e0sz;
            for (int e = tid; e < nunit * 3; e += nt) {
                const int u = e / 3, dir = e - 3 * u, tk = t0 + u / npair, pp = u - (u / npair) * npair;
                const int ia = pp / B.nprim, ib = pp % B.nprim;
                const double a = alpha[A.poff + ia], b = alpha[B.poff + ib], c = tzeta[tk];
                const double q = a + b + c;
                const double Ad = dir == 0 ? A.x : (dir == 1 ? A.y : A.z);
                const double Bd = dir == 0 ? B.x : (dir == 1 ? B.y : B.z);
                const double Cd = dir == 0 ? Cx : (dir == 1 ? Cy : Cz);
                const double Qd = (a * Ad + b * Bd + c * Cd) / q;
                const double XAB = Ad - Bd, XAC = Ad - Cd, XBC = Bd - Cd;
                const double K = exp(-(a * b / q) * XAB * XAB) * exp(-(a * c / q) * XAC * XAC) * exp(-(b * c / q) * XBC * XBC);
                mmd_E0_dir(scratch + (size_t) u * e0sz + dir * (la + 1) * (lb + 1), la, lb, q, K, Qd - Ad, Qd - Bd);
                if (dir == 0) {
                    const double ca = craw[A.poff + ia], cb = craw[B.poff + ib];
                    const double skip = (fabs(ca) < 1e-14 || fabs(cb) < 1e-14) ? 0.0 : 1.0;
                    const double NA = gnorm(la, a), NB = gnorm(lb, b);
                    wbuf[u] = skip * tcoef[tk] * ca * cb * (NA * NB * pow(M_PI / q, 1.5));
                }
            }
            __syncthreads();
            for (int e = tid; e < nel; e += nt) {
                const int i = e / ncB, j = e % ncB;
                int ax, ay, az, bx, by, bz;
                dev_cart(la, i, ax, ay, az);
                dev_cart(lb, j, bx, by, bz);
                const int ox = ax * (lb + 1) + bx, oy = (la + 1) * (lb + 1) + ay * (lb + 1) + by,
                          oz = 2 * (la + 1) * (lb + 1) + az * (lb + 1) + bz;
                double a_e = acc[e];
                for (int tk = 0; tk < t1 - t0; tk++)
                    for (int p0 = 0; p0 < npair; p0 += bmax) {
                        const int nb = min(bmax, npair - p0);
                        double sacc = 0.0;
                        for (int pb = 0; pb < nb; pb++) {
                            const int u = tk * npair + p0 + pb;
                            const double *E = scratch + (size_t) u * e0sz;
                            sacc += wbuf[u] * (E[ox] * E[oy] * E[oz]);
                        }
                        a_e += sacc;
                    }
                acc[e] = a_e;
            }
            __syncthreads();
            return;
        }
    }
    for (int tk = t0; tk < t1; tk++) {
        const int npow = tpow[tk];
        const double dk = tcoef[tk], zk = tzeta[tk];
        const int per = (npow == 2) ? e0sz : esz + 2 * ssz + 64;
        int bmax = scratch_len / per;
        if (bmax > 64) bmax = 64;
        if (npow != 2) bmax = 1;
        for (int p0 = 0; p0 < npair; p0 += bmax) {
            const int nb = min(bmax, npair - p0);
            // phase 1: E tables (one thread per (pair, direction))
            for (int e = tid; e < nb * 3; e += nt) {
                const int pb = e / 3, dir = e % 3, pp = p0 + pb;
                const int ia = pp / B.nprim, ib = pp % B.nprim;
                const double a = alpha[A.poff + ia], b = alpha[B.poff + ib], c = zk;
                const double q = a + b + c;
                const double Ad = dir == 0 ? A.x : (dir == 1 ? A.y : A.z);
                const double Bd = dir == 0 ? B.x : (dir == 1 ? B.y : B.z);
                const double Cd = dir == 0 ? Cx : (dir == 1 ? Cy : Cz);
                const double Qd = (a * Ad + b * Bd + c * Cd) / q;
                const double XAB = Ad - Bd, XAC = Ad - Cd, XBC = Bd - Cd;
                const double K = exp(-(a * b / q) * XAB * XAB) * exp(-(a * c / q) * XAC * XAC) * exp(-(b * c / q) * XBC * XBC);
                if (npow == 2)
                    mmd_E0_dir(scratch + pb * per + dir * (la + 1) * (lb + 1), la, lb, q, K, Qd - Ad, Qd - Bd);
                else
                    mmd_E_dir(scratch + pb * per + dir * (la + 1) * (lb + 1) * N1, la, lb, q, K, Qd - Ad, Qd - Bd);
                if (dir == 0) {
                    const double ca = craw[A.poff + ia], cb = craw[B.poff + ib];
                    const double skip = (fabs(ca) < 1e-14 || fabs(cb) < 1e-14) ? 0.0 : 1.0;
                    const double NA = gnorm(la, a), NB = gnorm(lb, b);
                    double w;
                    if (npow == 2) w = NA * NB * pow(M_PI / q, 1.5);
                    else if (npow == 1) w = NA * NB * 2.0 * M_PI / q;
                    else w = NA * NB * 2.0 * pow(M_PI, 1.5) / sqrt(q);
                    w_s[pb] = skip * dk * ca * cb * w;
                    q_s[pb] = q;
                }
            }
            __syncthreads();
            if (npow != 2) {
                // R^{(n)}_{tuv} (n = 1) or G^{(n)}_{tuv} (n = 0), layered from n = N down to 0
                // (type1_mcmurchie_davidson.c:479-585); layer n uses layer n+1 only
                double *Lcur = scratch + esz, *Lnxt = scratch + esz + ssz, *aux = scratch + esz + 2 * ssz;
                const int ia = p0 / B.nprim, ib = p0 % B.nprim;
                const double a = alpha[A.poff + ia], b = alpha[B.poff + ib], q = a + b + zk;
                const double Qx = (a * A.x + b * B.x + zk * Cx) / q, Qy = (a * A.y + b * B.y + zk * Cy) / q,
                             Qz = (a * A.z + b * B.z + zk * Cz) / q;
                const double X = Qx - Cx, Y = Qy - Cy, Z = Qz - Cz;
                if (tid == 0) {
                    const double arg = q * (X * X + Y * Y + Z * Z);
                    if (npow == 1)
                        for (int n = 0; n <= N; n++) aux[n] = boys_f(boys, n, arg);
                    else
                        gfun_values(arg, N, aux);
                }
                __syncthreads();
                for (int n = N; n >= 0; n--) {
                    const int lim = N - n;
                    const double g2 = (npow == 0) ? 2.0 * q : 0.0;
                    // level by level in t+u+v: the G recursion reads lower (t,u,v) of the same layer
                    for (int s = 0; s <= lim; s++) {
                        const int cnt_s = (s + 1) * (s + 2) / 2;
                        for (int e = tid; e < cnt_s; e += nt) {
                            int rem = e, t = 0;
                            while (rem >= s + 1 - t) {
                                rem -= s + 1 - t;
                                t++;
                            }
                            const int u = rem, v = s - t - u;
                            double val = 0.0;
                            if (s == 0) {
                                val = pow((npow == 1) ? -2.0 * q : 2.0 * q, n) * aux[n];
                            } else if (t + u == 0) {
                                if (v > 1) val += (v - 1) * (Lnxt[simplex_idx(t, u, v - 2)] - g2 * Lcur[simplex_idx(t, u, v - 2)]);
                                val += Z * (Lnxt[simplex_idx(t, u, v - 1)] - g2 * Lcur[simplex_idx(t, u, v - 1)]);
                            } else if (t == 0) {
                                if (u > 1) val += (u - 1) * (Lnxt[simplex_idx(t, u - 2, v)] - g2 * Lcur[simplex_idx(t, u - 2, v)]);
                                val += Y * (Lnxt[simplex_idx(t, u - 1, v)] - g2 * Lcur[simplex_idx(t, u - 1, v)]);
                            } else {
                                if (t > 1) val += (t - 1) * (Lnxt[simplex_idx(t - 2, u, v)] - g2 * Lcur[simplex_idx(t - 2, u, v)]);
                                val += X * (Lnxt[simplex_idx(t - 1, u, v)] - g2 * Lcur[simplex_idx(t - 1, u, v)]);
                            }
                            Lcur[simplex_idx(t, u, v)] = val;
                        }
                        __syncthreads();
                    }
                    // the layer just computed becomes the "n+1" layer of the next step
                    double *tmp = Lnxt;
                    Lnxt = Lcur;
                    Lcur = tmp;
                }
                // after the loop the n = 0 layer is in Lnxt
                for (int e = tid; e < nel; e += nt) {
                    const int i = e / ncB, j = e % ncB;
                    int ax, ay, az, bx, by, bz;
                    dev_cart(la, i, ax, ay, az);
                    dev_cart(lb, j, bx, by, bz);
                    const double *Ex = scratch, *Ey = scratch + (la + 1) * (lb + 1) * N1,
                                 *Ez = scratch + 2 * (la + 1) * (lb + 1) * N1;
                    double sacc = 0.0;
                    for (int t = 0; t <= ax + bx; t++) {
                        const double ex = Ex[(ax * (lb + 1) + bx) * N1 + t];
                        for (int u = 0; u <= ay + by; u++) {
                            const double ey = Ey[(ay * (lb + 1) + by) * N1 + u];
                            for (int v = 0; v <= az + bz; v++) {
                                const double ez = Ez[(az * (lb + 1) + bz) * N1 + v];
                                sacc += ex * ey * ez * Lnxt[simplex_idx(t, u, v)];
                            }
                        }
                    }
                    acc[e] += w_s[0] * sacc;
                }
                __syncthreads();
            } else {
                for (int e = tid; e < nel; e += nt) {
                    const int i = e / ncB, j = e % ncB;
                    int ax, ay, az, bx, by, bz;
                    dev_cart(la, i, ax, ay, az);
                    dev_cart(lb, j, bx, by, bz);
                    const int ox = ax * (lb + 1) + bx, oy = (la + 1) * (lb + 1) + ay * (lb + 1) + by,
                              oz = 2 * (la + 1) * (lb + 1) + az * (lb + 1) + bz;
                    double sacc = 0.0;
                    for (int pb = 0; pb < nb; pb++) {
                        const double *E = scratch + pb * per;
                        sacc += w_s[pb] * (E[ox] * E[oy] * E[oz]);
                    }
                    acc[e] += sacc;
                }
                __syncthreads();
            }
        }
    }
    (void) q_s;
    (void) n_s;
}

constexpr int MMD_SCRATCH = 12288;   // doubles of LDS for E / R tables (96 KiB), the upper bound
constexpr int MMD_SCRATCH_BATCH = 4096;   // doubles for a batch of primitive-pair E tables
constexpr int MAX_BLOCK = 66 * 66;
constexpr int MAX_OCB = 66 * 28;   // basis shell x outercore shell (L <= 6)

// type-1 (local) contribution of one (shell pair, centre): it opens the part's slice of the partial
// buffer (arep = type-1, spin-orbit = 0); k_contract adds the semilocal, SO and outercore terms
__global__ __launch_bounds__(256) void k_type1(const PartItem *__restrict__ partlist, const WorkItem *__restrict__ items,
                                               DevJob J, DevTables T, double *__restrict__ partial, int parts,
                                               int scratch_len)
{
    // dynamic LDS sized by the plan: E / R tables (scratch_len doubles), then the block accumulator
    extern __shared__ double t1_lds[];
    double *scratch = t1_lds, *acc = t1_lds + scratch_len;
    const PartItem pt = partlist[blockIdx.x];
    const WorkItem it = items[pt.item];
    const DShell sa = J.sh[it.a], sb = J.sh[it.b];
    const int ncA = dev_ncart(sa.L), ncB = dev_ncart(sb.L), nel = ncA * ncB;
    for (int e = threadIdx.x; e < nel; e += blockDim.x) acc[e] = 0.0;
    __syncthreads();
    const DCenter c = J.ctr[J.cidx[pt.ci]];
    if ((parts & 1) && c.local >= 0) {
        const MmdShell A = {sa.L, sa.nprim, sa.poff, sa.x, sa.y, sa.z};
        const MmdShell B = {sb.L, sb.nprim, sb.poff, sb.x, sb.y, sb.z};
        const DPot p = J.pot[c.local];
        mmd_accumulate(A, B, c.x, c.y, c.z, J.tpow, J.tcoef, J.tzeta, p.toff, p.toff + p.nterm, J.alpha, J.craw, acc,
                       scratch, scratch_len, T.boys);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 4 * nel; e += blockDim.x) partial[pt.off + e] = (e < nel) ? acc[e] : 0.0;
}

// outercore vectors: [S (ncA x (2L+1)), U (ncA x (2L+1))] per (centre, shell, OC entry)
__global__ __launch_bounds__(256) void k_ocv(const OcvItem *__restrict__ items, DevJob J, DevTables T,
                                             double *__restrict__ OCV, int scratch_len, int nel_max)
{
    // dynamic LDS sized by the plan (ocv_scratch, ocv_nel) instead of the upper bounds (96 KiB of
    // E / R tables, MAX_OCB elements), so that several workgroups fit on a CU beside the kernels of
    // the other stream: E / R tables, then the overlap and potential blocks
    extern __shared__ double ocv_lds[];
    double *scratch = ocv_lds, *bs = ocv_lds + scratch_len, *bu = bs + nel_max;
    __shared__ int one_pow;
    __shared__ double one_coef, zero_zeta;
    const OcvItem it = items[blockIdx.x];
    const DCenter c = J.ctr[it.c];
    const OcEntry oe = J.oce[it.oc];
    const DShell sa = J.sh[it.A], so = J.sh[oe.shell];
    const DPot p = J.pot[oe.pot];
    const int ncA = dev_ncart(sa.L), ncO = dev_ncart(so.L), nel = ncA * ncO, L = so.L, nm = 2 * L + 1;
    for (int e = threadIdx.x; e < nel; e += blockDim.x) {
        bs[e] = 0.0;
        bu[e] = 0.0;
    }
    if (threadIdx.x == 0) {
        one_pow = 2;
        one_coef = 1.0;
        zero_zeta = 0.0;
    }
    __syncthreads();
    const MmdShell A = {sa.L, sa.nprim, sa.poff, sa.x, sa.y, sa.z};
    const MmdShell O = {so.L, so.nprim, so.poff, c.x, c.y, c.z};
    // overlap <A|phi>: the n = 2 kernel with a zero exponent (overlap.c:39-138 gives the same integral)
    mmd_accumulate(A, O, c.x, c.y, c.z, &one_pow, &one_coef, &zero_zeta, 0, 1, J.alpha, J.craw, bs, scratch,
                   scratch_len, T.boys);
    mmd_accumulate(A, O, c.x, c.y, c.z, J.tpow, J.tcoef, J.tzeta, p.toff, p.toff + p.nterm, J.alpha, J.craw, bu,
                   scratch, scratch_len, T.boys);
    __syncthreads();
    const double *slm = T.slm + T.slm_off[L];
    for (int e = threadIdx.x; e < ncA * nm; e += blockDim.x) {
        const int i = e / nm, m = e % nm;
        double s = 0.0, u = 0.0;
        for (int q = 0; q < ncO; q++) {
            s += slm[m * ncO + q] * bs[i * ncO + q];
            u += slm[m * ncO + q] * bu[i * ncO + q];
        }
        OCV[it.off + e] = s;
        OCV[it.off + ncA * nm + e] = u;
    }
}

// one-centre radial constants delta(phi_i, phi_j) = (U1 + U2)/2 (outercore_integrals.c:397-500)
__device__ double rgto_norm(int L, double a)
{
    double df = 1.0;
    for (int k = 2 * L + 1; k > 1; k -= 2) df *= k;
    return sqrt(pow(2.0, L + 2) * pow(2.0, L + 1.5) / (df * sqrt(M_PI))) * pow(a, 0.75 + L / 2.0);
}

__device__ double onecenter_contracted(const DShell &b, const DShell &k, const DPot &p, const DevJob &J)
{
    double sum = 0.0;
    const int L = p.L;
    for (int i = 0; i < b.nprim; i++) {
        const double ai = J.alpha[b.poff + i], Ni = rgto_norm(L, ai);
        for (int j = 0; j < k.nprim; j++) {
            const double aj = J.alpha[k.poff + j], Nj = rgto_norm(L, aj);
            const double f = Ni * Nj * J.craw[b.poff + i] * J.craw[k.poff + j];
            for (int t = 0; t < p.nterm; t++) {
                const int n = J.tpow[p.toff + t];
                const double a = ai + aj + J.tzeta[p.toff + t];
                double u;
                if (n % 2 == 0) {
                    const int kk = L + n / 2;
                    double df = 1.0;
                    for (int z = 2 * kk - 1; z > 1; z -= 2) df *= z;
                    // 2^(kk+1) exactly, as pow(2, kk + 1) gives it
                    u = df / (ldexp(1.0, kk + 1) * pow(a, kk)) * sqrt(M_PI / a);
                } else {
                    const int kk = L + (n - 1) / 2;
                    double fk = 1.0;
                    for (int z = 2; z <= kk; z++) fk *= z;
                    u = fk / (2.0 * pow(a, kk + 1));
                }
                sum += f * J.tcoef[p.toff + t] * u;
            }
        }
    }
    return sum;
}

// term q = (i * k.nprim + j) * nterm + t of onecenter_contracted's sum, as the factor x and the
// integral u of its last product (the term is x * u), evaluated as the loop there evaluates it
__device__ __forceinline__ void onecenter_term(const DShell &b, const DShell &k, const DPot &p, const DevJob &J, int q,
                                               double &x, double &u)
{
    const int nt = p.nterm, i = q / (k.nprim * nt), r = q - i * (k.nprim * nt), j = r / nt, t = r - j * nt;
    const int L = p.L;
    const double ai = J.alpha[b.poff + i], Ni = rgto_norm(L, ai);
    const double aj = J.alpha[k.poff + j], Nj = rgto_norm(L, aj);
    const double f = Ni * Nj * J.craw[b.poff + i] * J.craw[k.poff + j];
    const int n = J.tpow[p.toff + t];
    const double a = ai + aj + J.tzeta[p.toff + t];
    if (n % 2 == 0) {
        const int kk = L + n / 2;
        double df = 1.0;
        for (int z = 2 * kk - 1; z > 1; z -= 2) df *= z;
        u = df / (ldexp(1.0, kk + 1) * pow(a, kk)) * sqrt(M_PI / a);
    } else {
        const int kk = L + (n - 1) / 2;
        double fk = 1.0;
        for (int z = 2; z <= kk; z++) fk *= z;
        u = fk / (2.0 * pow(a, kk + 1));
    }
    x = f * J.tcoef[p.toff + t];
}

// One wave per (centre, outercore entry pair) (grid: centres x the largest entry count squared):
// the lanes evaluate 64 terms of a sum at a time and lane 0 adds them in the sum's own order, so each
// constant is bitwise onecenter_contracted's one-thread loop
__global__ __launch_bounds__(64) void k_delta(const int *__restrict__ centers, int n, DevJob J,
                                              double *__restrict__ delta)
{
    __shared__ double xs[64], us[64];
    const DCenter c = J.ctr[centers[blockIdx.x]];
    const int no = c.oe - c.ob, e = blockIdx.y, lane = threadIdx.x;
    if (e >= no * no) return;   // the whole workgroup
    const int i = e / no, j = e % no;
    const OcEntry oi = J.oce[c.ob + i], oj = J.oce[c.ob + j];
    const DPot p1 = J.pot[oi.pot], p2 = J.pot[oj.pot];
    double d = 0.0;
    if (p1.L == p2.L && p1.J == p2.J) {
        const DShell s1 = J.sh[oi.shell], s2 = J.sh[oj.shell];
        double sum[2];
        for (int w = 0; w < 2; w++) {
            const DPot p = w ? p2 : p1;
            const int nq = s1.nprim * s2.nprim * p.nterm;
            double s = 0.0;
            for (int q0 = 0; q0 < nq; q0 += 64) {
                if (q0 + lane < nq) onecenter_term(s1, s2, p, J, q0 + lane, xs[lane], us[lane]);
                __syncthreads();
                if (lane == 0) {
                    const int m = nq - q0 < 64 ? nq - q0 : 64;
                    for (int k = 0; k < m; k++) s += xs[k] * us[k];
                }
                __syncthreads();
            }
            sum[w] = s;
        }
        d = 0.5 * (sum[0] + sum[1]);
    }
    if (lane == 0) delta[c.delta_off + e] = d;
    (void) n;
}


// out = sum over the item's centres, in centre order, of (type-1 slice + contraction slice);
// the (B, A) block of a mirrored item is the (anti)symmetric image
__global__ __launch_bounds__(256) void k_contract_reduce(const WorkItem *__restrict__ items,
                                                         const int *__restrict__ item_part0,
                                                         const PartItem *__restrict__ partlist, DevJob J,
                                                         const double *__restrict__ partial_t1,
                                                         const double *__restrict__ partial_ct, double *__restrict__ out)
{
    const WorkItem it = items[blockIdx.x];
    const int ncA = dev_ncart(J.sh[it.a].L), ncB = dev_ncart(J.sh[it.b].L), nel = ncA * ncB;
    const int p0 = item_part0[blockIdx.x], np = it.ce - it.cb;
    for (int e = threadIdx.x; e < 4 * nel; e += blockDim.x) {
        const int op = e / nel, ij = e % nel, i = ij / ncB, j = ij % ncB;
        double s = 0.0;
        for (int k = 0; k < np; k++) {
            const long long o = partlist[p0 + k].off + e;
            s += partial_t1[o] + partial_ct[o];   // the centre's type-1 slice plus its contraction slice
        }
        out[op * J.op_stride + it.off + (long long) i * it.stride + j] = s;
        if (it.mirror) out[op * J.op_stride + it.moff + (long long) j * it.stride + i] = (op == 0 ? 1.0 : -1.0) * s;
    }
}

// ================================================================== device tables (per process, current device)
struct DeviceTables {
    DevTables t{};
    std::vector<void *> bufs;
    int Lmax = -1, nmax = -1;
    bool ok = false;
};

static std::mutex g_tab_mutex;
static std::map<int, DeviceTables> g_dev_tables;

template <class T>
static int upload(const std::vector<T> &v, T **out, std::vector<void *> &keep)
{
    size_t n = std::max<size_t>(1, v.size()) * sizeof(T);
    HIPCHK(hipMalloc((void **) out, n));
    keep.push_back(*out);
    if (!v.empty()) HIPCHK(hipMemcpy(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

// w_i U_p(r_i) of every channel p that has a radial table, on the host, as
// copysign(sqrt|w U|, w U) (the product stages sqrt|w U| on both operands and the sign on one).
// U is a sum of terms c r^(n-2) exp(-zeta r^2) that cancel, so one ulp of one term's exp moves
// the result by one ulp of the largest term: on a g-g block of the reference's KCs case that alone
// moved arep elements by 4.6e-9.  The terms are therefore evaluated with the C library's exp and
// pow and accumulated as val = fma(c r^(n-2), exp(...), val), which is how the reference's
// libgrpp_potential_value (potential.c:70-81) compiles with contraction, so that the table is the
// reference's own potential on the same Log3 points.  Runs on up to 16 host threads.
static void host_wU_table(const std::vector<int> &wpots, const std::vector<DPot> &pots, const std::vector<int> &tpow,
                          const std::vector<double> &tcoef, const std::vector<double> &tzeta, std::vector<double> &out)
{
    const HostTables &H = host_tables();
    out.assign(wpots.size() * (size_t) NPTS, 0.0);
    auto work = [&](size_t b, size_t e) {
        for (size_t q = b; q < e; q++) {
            const DPot &p = pots[wpots[q]];
            double *o = out.data() + q * NPTS;
            for (int i = 0; i < NPTS; i++) {
                const double r = H.r[i], r2 = r * r;
                double v = 0.0;
                for (int k = p.toff; k < p.toff + p.nterm; k++)
                    v = std::fma(tcoef[k] * std::pow(r, tpow[k] - 2), std::exp(-tzeta[k] * r2), v);
                const double u = H.w[i] * v;
                o[i] = std::copysign(std::sqrt(std::fabs(u)), u);
            }
        }
    };
    const size_t n = wpots.size();
    const size_t nt = std::min<size_t>(std::max(1u, std::min(16u, std::thread::hardware_concurrency())), std::max<size_t>(1, n / 8));
    if (nt <= 1) {
        work(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (size_t t = 0; t < nt; t++) th.emplace_back(work, n * t / nt, n * (t + 1) / nt);
    for (auto &x : th) x.join();
}


// Dynamic LDS beyond the 64 KiB default has to be opted into per kernel, and that attribute belongs
// to the kernel for the whole process, not to one plan.  It is raised once per device, when the
// device tables are first set up (under g_tab_mutex), to everything a gfx950 CU offers beside the
// kernel's static LDS; each launch still asks only for its plan's size, which alone sets occupancy.
// Setting it per plan instead would let a thread lower it under another thread's larger launch.
static const void *nuc_kernel(void);

static int raise_lds_limits(void)
{
    const int cu_lds = 160 * 1024;
    const void *fns[] = {(const void *) k_ocv, (const void *) k_type1, (const void *) k_proj, contract_kernel_ptr(64),
                         contract_kernel_ptr(128), contract_kernel_ptr(256),
                         nuc_kernel()};
    for (const void *f : fns) {
        hipFuncAttributes a{};
        HIPCHK(hipFuncGetAttributes(&a, f));
        HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, cu_lds - (int) a.sharedSizeBytes));
    }
    return 0;
}

// Scratch memory of the allocation-free entry point (libgrpp_batch_full_grpp_integrals_noalloc):
// one static device array per GPU that loads this library, held until the process ends.  In host
// code its name is only a stand-in: the device address is looked up once per device, with that
// device current, and kept in g_scratch; kernels and copies are given that address.  Every call on
// one device shares the array, so calls that overlap on one device race (the entry point
// documents it); nothing in it is assumed to start at zero.
static __device__ __align__(256) unsigned char scratch[ENG_SCRATCH_BYTES];
static std::map<int, unsigned char *> g_scratch;   // device -> address of its scratch array

// under g_tab_mutex, with device dev current
static int lookup_scratch(int dev)
{
    if (g_scratch.count(dev)) return 0;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        eng_set_error("scratch array of device %d: hipGetSymbolAddress failed: %s", dev, hipGetErrorString(e));
        return 1;
    }
    g_scratch[dev] = (unsigned char *) ptr;
    return 0;
}

static int scratch_base(int dev, unsigned char **out)
{
    std::lock_guard<std::mutex> lk(g_tab_mutex);
    if (lookup_scratch(dev)) return 1;
    *out = g_scratch[dev];
    return 0;
}

// no_alloc: the tables must already be on the device at sufficient size (set up by
// eng_setup_device); refuses instead of uploading
static int device_tables(int Lmax, int nmax, DevTables *out, bool no_alloc = false)
{
    std::lock_guard<std::mutex> lk(g_tab_mutex);
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    DeviceTables &D = g_dev_tables[dev];
    if (no_alloc && (!D.ok || Lmax > D.Lmax || nmax > D.nmax)) {
        eng_set_error("the allocation-free entry point needs the device tables of device %d in place: call "
                      "libgrpp_init() with that device current before the first call", dev);
        return 1;
    }
    if (!D.ok) {
        if (raise_lds_limits()) return 1;
        if (lookup_scratch(dev)) return 1;
        const HostTables &H = host_tables();
        double *p;
        int *ip;
        if (upload(H.r, &p, D.bufs)) return 1;
        D.t.r = p;
        if (upload(H.w, &p, D.bufs)) return 1;
        D.t.w = p;
        if (upload(H.rpow, &p, D.bufs)) return 1;
        D.t.rpow = p;
        if (upload(H.bes, &p, D.bufs)) return 1;
        D.t.bes = p;
        if (upload(H.bes_t, &p, D.bufs)) return 1;
        D.t.bes_t = p;
        if (upload(H.besp, &p, D.bufs)) return 1;
        D.t.besp = p;
        if (upload(H.rfun, &p, D.bufs)) return 1;
        D.t.rfun = p;
        if (upload(H.boys, &p, D.bufs)) return 1;
        D.t.boys = p;
        if (upload(H.rsh, &p, D.bufs)) return 1;
        D.t.rsh = p;
        if (upload(H.rsh_off, &ip, D.bufs)) return 1;
        D.t.rsh_off = ip;
        if (upload(H.comb, &ip, D.bufs)) return 1;
        D.t.comb = ip;
        if (upload(H.comb_off, &ip, D.bufs)) return 1;
        D.t.comb_off = ip;
        if (upload(H.lmat, &p, D.bufs)) return 1;
        D.t.lmat = p;
        if (upload(H.lmat_off, &ip, D.bufs)) return 1;
        D.t.lmat_off = ip;
        if (upload(H.slm, &p, D.bufs)) return 1;
        D.t.slm = p;
        if (upload(H.slm_off, &ip, D.bufs)) return 1;
        D.t.slm_off = ip;
        if (upload(H.w1, &p, D.bufs)) return 1;
        D.t.w1 = p;
        D.ok = true;
    }
    if (Lmax > D.Lmax || nmax > D.nmax) {
        const int L2 = std::max(Lmax, D.Lmax), n2 = std::max(nmax, D.nmax);
        std::vector<double> W;
        std::vector<WBlock> dir;
        angular_w_tables(L2, n2, W, dir);
        std::vector<int> wd((L2 + 1) * (n2 + 1), -1), ws((L2 + 1) * (n2 + 1), 0);
        for (auto &b : dir) {
            wd[b.L * (n2 + 1) + b.n] = b.off;
            ws[b.L * (n2 + 1) + b.n] = b.mu_stride;
        }
        double *p;
        int *ip;
        if (upload(W, &p, D.bufs)) return 1;
        D.t.W = p;
        if (upload(wd, &ip, D.bufs)) return 1;
        D.t.wdir = ip;
        if (upload(ws, &ip, D.bufs)) return 1;
        D.t.wstride = ip;
        D.t.w_nmax = n2;
        D.Lmax = L2;
        D.nmax = n2;
    }
    *out = D.t;
    return 0;
}

// Library set-up on the current device (libgrpp_init): the constant tables at the largest
// angular momenta the engine supports, the dynamic-LDS limits and the scratch array's address, so
// that the allocation-free entry point finds everything in place.  Without a device: nothing.
extern "C" int eng_setup_device(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 0;
    DevTables t;
    return device_tables(LMAX_POT, LMAX_SHELL, &t);
}

extern "C" int eng_device_ok(void)
{
    static std::atomic<int> seen{0};   // a device, once found, stays for the process
    if (seen.load(std::memory_order_relaxed)) return 1;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        eng_set_error("no HIP device available: the GRPP integrals run only on the GPU");
        return 0;
    }
    seen.store(1, std::memory_order_relaxed);
    return 1;
}

// ================================================================== plan
// small parts (<= 64 Cartesian pairs) go to the wave-per-part contraction only when there are this many
constexpr int CTW_MIN_PARTS = 262144;

struct eng_plan {
    int nshell = 0, ncenter = 0, parts = 0, flags = 0;
    long long out_len = 0, op_stride = 0;
    DevTables T{};
    DevJob J{};
    int device = 0;
    std::vector<std::pair<void *, size_t>> bufs;   // device buffers and their pooled sizes
    unsigned char *arena = nullptr;               // ENG_F_ARENA: buffers carved from the scratch array
    size_t arena_used = 0;
    unsigned char *up_dev = nullptr, *up_host = nullptr;   // per-call plans: uploads staged in one block
    size_t up_used = 0;
    bool percall = false;                         // built by eng_run: one call, then destroyed
    bool up_pageable = false;                     // a per-call upload went straight from a host vector
    bool serial = false;                          // one stream (the caller's), no events
    double *d_out = nullptr, *d_F = nullptr, *d_wU = nullptr, *d_Q = nullptr, *d_P = nullptr, *d_OCV = nullptr,
           *d_delta = nullptr, *d_scr = nullptr, *d_Qc = nullptr;
    TabItem *d_tab = nullptr;
    int n_tab = 0;
    int n_wpots = 0;
    RowDesc *d_rows = nullptr;
    QJob *d_jobs = nullptr;
    TileItem *d_tiles = nullptr;
    TileItem *d_tiles_radial = nullptr;   // the radial product's launch order (TileItem.pad = tile index)
    int *d_tile_lev = nullptr;            // last level of every tile in the latest build
    bool radial_order_fresh = false;      // launch order taken from a build with the current rows
    int n_tiles = 0;
    std::vector<TileItem> tiles_host;
    ProjItem *d_proj = nullptr;
    int n_proj = 0;
    int proj_om = 1;                 // k_proj dynamic LDS doubles: the largest Omega block of one n
    OcvItem *d_ocvi = nullptr;
    int n_ocv = 0;
    int *d_delta_centers = nullptr;
    int n_delta_centers = 0;
    int delta_max_pairs = 0;   // the largest outercore entry count squared of a centre
    WorkItem *d_items = nullptr;
    int n_items = 0;
    PartItem *d_parts = nullptr;     // (item, centre) workgroups of the contraction, item order
    PartItem *d_parts_run = nullptr; // the same parts in launch order: largest blocks first
    int *d_item_part0 = nullptr;     // first part of each item
    double *d_partial = nullptr;     // per-part type-1 blocks (k_type1)
    double *d_partial_ct = nullptr;  // per-part semilocal, spin-orbit and outercore blocks (k_contract)
    int n_parts = 0;
    int n_parts_small = 0;           // parts of <= 64 Cartesian pairs: the tail of d_parts_run (k_contract_w)
    int canon_upper = 0;             // every item has a <= b: k_canon writes the upper shell triangle only
    unsigned long long *d_counters = nullptr;
    eng_tuning_t tune{1, 128, 256, 0, 4};   // launch shapes and stream layout; none of them changes a result
    hipEvent_t ev[8] = {};   // execution phases on the caller's stream: see eng_plan_execute   // execution phases on the caller's stream: see eng_plan_execute
    hipStream_t side = nullptr;   // second stream (default priority) for the radial-independent kernels
    hipStream_t cap = nullptr;    // capture stream of a graph plan (default priority)
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipEvent_t gev[2] = {};       // around a graph replay
    hipEvent_t evs[3] = {};       // phases on the side stream
    hipEvent_t evt1 = nullptr;    // start of type-1
    hipEvent_t evr = nullptr;     // start of the radial product (after its wait for the grid tables)
    // statistics
    long long n_tasks = 0, n_integrals = 0, n_entries = 0;
    double flops_fixed = 0.0;    // contraction + type-1 + projector work (grid independent)
    double flops_per_point = 0.0; // radial FLOPs per grid point summed over tiles (x points below)
    unsigned long long last_points = 0, last_tiles = 0, last_entry_points = 0;
    double last_seconds = 0.0, last_radial_seconds = 0.0;
    double phase_seconds[7] = {};
    unsigned long long last_screen_triples = 0, last_screen_steps = 0;
    int min_level = 0;
    double tol = 1e-16;
    int radial_mfma = 0;
    int nm_max = 1;
    int ct_qs = 1, ct_ts = 1;   // k_contract LDS doubles
    int ctw_qs = 64, ctw_ts = 64;   // k_contract_w LDS doubles per wave
    int t1_scratch = 64, t1_acc = 1;   // k_type1 LDS doubles
    int ocv_scratch = 64, ocv_nel = 1;  // k_ocv LDS doubles: E / R tables, each of its two blocks
    // radial rows: logical order per slot (shell, n, lambda), their physical order in the tables,
    // per-row features for the level predictor, and the recorded stopping levels
    std::vector<RowDesc> rows_logical;
    std::vector<int> slot_rows, slot_M;
    std::vector<int> rowpos;           // [slot row base + logical] -> physical
    std::vector<QJob> jobs;
    std::vector<float> features;       // ENG_NFEAT per logical row
    int *d_rowpos = nullptr;
    signed char *d_lev = nullptr;      // stopping level per radial entry (ENG_F_LEVELS)
    long long qlen = 0;
};

// Device buffers of small plans are recycled: a destroyed plan returns buffers of up to
// POOL_MAX_BUFFER bytes to a per-device free list (by size, rounded up to 256 bytes, at most
// POOL_MAX_TOTAL bytes held), so the per-call entry points, which build and destroy a one-task plan
// per call, do not pay hipMalloc / hipFree each time.  Contents are not cleared (hipMalloc does not
// clear either); every buffer a build reads is written before it is read.
constexpr size_t POOL_MAX_BUFFER = (size_t) 16 << 20, POOL_MAX_TOTAL = (size_t) 512 << 20;
static std::mutex g_pool_mutex;
static std::multimap<std::pair<int, size_t>, void *> g_pool;
static size_t g_pool_bytes = 0;

static int pool_alloc(int dev, size_t bytes, void **p)
{
    {
        std::lock_guard<std::mutex> lk(g_pool_mutex);
        auto it = g_pool.find({dev, bytes});
        if (it != g_pool.end()) {
            *p = it->second;
            g_pool.erase(it);
            g_pool_bytes -= bytes;
            return 0;
        }
    }
    HIPCHK(hipMalloc(p, bytes));
    return 0;
}

static void pool_free(int dev, void *p, size_t bytes)
{
    if (bytes <= POOL_MAX_BUFFER) {
        std::lock_guard<std::mutex> lk(g_pool_mutex);
        if (g_pool_bytes + bytes <= POOL_MAX_TOTAL) {
            g_pool.insert({{dev, bytes}, p});
            g_pool_bytes += bytes;
            return;
        }
    }
    (void) hipFree(p);
}

template <class T>
static int dalloc(eng_plan *P, T **p, size_t n)
{
    const size_t bytes = (std::max<size_t>(1, n) * sizeof(T) + 255) & ~(size_t) 255;
    void *q = nullptr;
    if (P->arena) {
        if (P->arena_used + bytes > ENG_SCRATCH_BYTES) {
            eng_set_error("input too large for the allocation-free entry point: its buffers exceed the %llu-byte "
                          "scratch array (already %zu bytes, next buffer %zu)",
                          (unsigned long long) ENG_SCRATCH_BYTES, P->arena_used, bytes);
            return 1;
        }
        *p = (T *) (P->arena + P->arena_used);
        P->arena_used += bytes;
        return 0;
    }
    if (pool_alloc(P->device, bytes, &q)) return 1;
    *p = (T *) q;
    P->bufs.push_back({q, bytes});
    return 0;
}

// the stream a plan's uploads are ordered on while it is created (per host thread; null: hipMemcpy).
// The per-call entry points set it to the calling thread's own stream, so that their uploads do not
// wait for other threads' work on the device
static thread_local hipStream_t t_upload_stream = nullptr;

// Per-call plans stage their uploads: each buffer is carved from one pooled device block of
// UPLOAD_BLOCK bytes and its data written at the same offset of a pinned host buffer of that size,
// and the whole block goes up in one copy at the end of plan creation (about 34 small pageable
// copies otherwise).  Pinned staging buffers are taken from a free list and returned after the
// call; each stays allocated until the process ends (one per host thread that calls at once).
constexpr size_t UPLOAD_BLOCK = (size_t) 4 << 20;
static std::vector<unsigned char *> g_stage_free;   // under g_pool_mutex

static unsigned char *stage_take()
{
    {
        std::lock_guard<std::mutex> lk(g_pool_mutex);
        if (!g_stage_free.empty()) {
            unsigned char *p = g_stage_free.back();
            g_stage_free.pop_back();
            return p;
        }
    }
    void *p = nullptr;
    return hipHostMalloc(&p, UPLOAD_BLOCK, hipHostMallocDefault) == hipSuccess ? (unsigned char *) p : nullptr;
}

static void stage_give(unsigned char *p)
{
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_pool_mutex);
    g_stage_free.push_back(p);
}

// the pinned staging buffer of the calling thread's per-call plan while it is created (null: none)
static thread_local unsigned char *t_stage = nullptr;

template <class T>
static int dput(eng_plan *P, T **p, const std::vector<T> &v)
{
    const size_t bytes = (std::max<size_t>(1, v.size()) * sizeof(T) + 255) & ~(size_t) 255;
    if (P->up_host && P->up_used + bytes <= UPLOAD_BLOCK) {
        *p = (T *) (P->up_dev + P->up_used);
        if (!v.empty()) memcpy(P->up_host + P->up_used, v.data(), v.size() * sizeof(T));
        P->up_used += bytes;
        return 0;
    }
    if (dalloc(P, p, v.size())) return 1;
    if (v.empty()) return 0;
    if (t_upload_stream) {
        P->up_pageable = true;
        HIPCHK(hipMemcpyAsync(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, t_upload_stream));
    }
    else
        HIPCHK(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

static int kcols_host(int L, int l)
{
    int k = 0;
    for (int n = 0; n <= l; n++) k += lam_cnt(L, n);
    return k;
}

extern "C" int eng_plan_create(int nshell, libgrpp_shell_t *const *shells, int ncenter, const eng_center_t *centers,
                               int nidx, const int *center_index, int nitem, const eng_item_t *items,
                               long long out_len, long long op_stride, int parts, int flags, eng_plan_t **plan_out)
{
    *plan_out = nullptr;
    if (!eng_device_ok()) return 1;
    eng_plan *P = new eng_plan();
    HIPCHK(hipGetDevice(&P->device));
    if (flags & ENG_F_ARENA) {
        if ((flags & ENG_F_GRAPH) || scratch_base(P->device, &P->arena)) {
            if (flags & ENG_F_GRAPH) eng_set_error("the allocation-free build does not capture a graph");
            delete P;
            return 1;
        }
        P->serial = true;
    }
    if (flags & ENG_F_SERIAL) P->serial = true;
    P->percall = t_upload_stream != nullptr;
    if (t_upload_stream && t_stage && !P->arena) {
        void *q = nullptr;
        if (pool_alloc(P->device, UPLOAD_BLOCK, &q)) {
            delete P;
            return 1;
        }
        P->bufs.push_back({q, UPLOAD_BLOCK});
        P->up_dev = (unsigned char *) q;
        P->up_host = t_stage;
    }
    P->nshell = nshell;
    P->ncenter = ncenter;
    P->parts = parts;
    P->flags = flags;
    P->out_len = out_len;
    P->op_stride = op_stride;

    // ---------------- pack shells (basis shells first, then outercore shells) ----------------
    std::vector<DShell> sh;
    std::vector<double> alpha, craw, cnorm;
    int lmax_shell = 0;
    auto add_shell = [&](const libgrpp_shell_t *s, const double *origin) {
        DShell d{};
        d.L = s->L;
        d.nprim = s->num_primitives;
        d.poff = (int) alpha.size();
        d.x = origin[0];
        d.y = origin[1];
        d.z = origin[2];
        for (int k = 0; k < s->num_primitives; k++) {
            alpha.push_back(s->alpha[k]);
            craw.push_back(s->coeffs[k]);
            // libgrpp_shell_shrink + libgrpp_shell_mult_normcoef (shell.c:67-92)
            cnorm.push_back(std::fabs(s->coeffs[k]) > 1e-14 ? s->coeffs[k] * gaussian_norm(s->L, s->alpha[k]) : 0.0);
        }
        sh.push_back(d);
        return (int) sh.size() - 1;
    };
    for (int i = 0; i < nshell; i++) {
        const libgrpp_shell_t *s = shells[i];
        if (s->L < 0 || s->L > LMAX_SHELL || s->num_primitives <= 0) {
            eng_set_error("shell %d: L=%d / %d primitives outside the supported range (L <= %d)", i, s->L,
                          s->num_primitives, LMAX_SHELL);
            delete P;
            return 2;
        }
        add_shell(s, s->origin);
        lmax_shell = std::max(lmax_shell, s->L);
    }

    // ---------------- potentials ----------------
    std::vector<DPot> pots;
    std::vector<int> tpow;
    std::vector<double> tcoef, tzeta;
    auto add_pot = [&](const libgrpp_potential_t *p, int kind, double scale) {
        DPot d{};
        d.L = p->L;
        d.J = p->J;
        d.kind = kind;
        d.toff = (int) tpow.size();
        d.slot = -1;
        d.qoff = -1;
        d.scale = scale;
        d.thr = (kind == K_SO) ? 1e-14 : 1e-16;
        for (int k = 0; k < p->num_primitives; k++) {
            if (std::fabs(p->coeffs[k]) <= 1e-14) continue;   // libgrpp_shrink_potential (potential.c:88-115)
            tpow.push_back(p->powers[k]);
            tcoef.push_back(p->coeffs[k]);
            tzeta.push_back(p->alpha[k]);
        }
        d.nterm = (int) tpow.size() - d.toff;
        pots.push_back(d);
        return (int) pots.size() - 1;
    };

    std::vector<DCenter> ctr(ncenter);
    std::vector<OcEntry> oce;
    int Lmax_pot = 0;
    for (int c = 0; c < ncenter; c++) {
        const eng_center_t &E = centers[c];
        DCenter &d = ctr[c];
        d.x = E.origin[0];
        d.y = E.origin[1];
        d.z = E.origin[2];
        d.sb = E.shell_begin < 0 ? 0 : E.shell_begin;
        d.se = E.shell_end < 0 ? nshell : E.shell_end;
        d.local = (E.local && (parts & ENG_TYPE1)) ? add_pot(E.local, K_LOCAL, 1.0) : -1;
        d.pb = (int) pots.size();
        if (parts & ENG_TYPE2)
            for (int k = 0; k < E.n_arep; k++) add_pot(E.arep[k], K_AREP, 1.0);
        if (parts & ENG_SO)
            for (int k = 0; k < E.n_so; k++) {
                const int L = E.so[k]->L;
                add_pot(E.so[k], K_SO, E.so_scale ? E.so_scale[k] : 2.0 / (2 * L + 1));
            }
        d.pe = (int) pots.size();
        for (int k = d.pb; k < d.pe; k++) {
            if (pots[k].L < 0 || pots[k].L > LMAX_POT || pots[k].L + lmax_shell > ANG_SUM_MAX) {
                eng_set_error("channel L=%d with basis L=%d exceeds the supported L + l <= %d", pots[k].L, lmax_shell,
                              ANG_SUM_MAX);
                delete P;
                return 2;
            }
            Lmax_pot = std::max(Lmax_pot, pots[k].L);
        }
        d.ob = (int) oce.size();
        d.oc_i = E.oc_i;
        d.oc_j = E.oc_j;
        if (parts & (ENG_OC | ENG_OC_P1 | ENG_OC_P2))
            for (int k = 0; k < E.n_oc; k++) {
                OcEntry oe;
                oe.pot = add_pot(E.oc_pot[k], K_OC, 1.0);
                oe.shell = add_shell(E.oc_shell[k], E.origin);
                if (E.oc_shell[k]->L > 6) {
                    eng_set_error("outercore shell L=%d too large", E.oc_shell[k]->L);
                    delete P;
                    return 2;
                }
                oce.push_back(oe);
            }
        d.oe = (int) oce.size();
    }

    // ---------------- radial machinery: slots (centre, L) ----------------
    struct Slot {
        int c, L, has_so;
    };
    std::vector<Slot> slots;
    std::vector<int> tabF_base;   // per (centre c, shell A): first F row, or -1
    std::vector<TabItem> tab;
    long long frows = 0;
    std::vector<int> slot_of_pot(pots.size(), -1);
    for (int c = 0; c < ncenter; c++) {
        DCenter &d = ctr[c];
        std::map<int, int> Ls;
        for (int k = d.pb; k < d.pe; k++) {
            int &flag = Ls[pots[k].L];
            if (pots[k].kind == K_SO) flag = 1;
        }
        std::map<int, int> slot_id;
        for (auto &kv : Ls) {
            slot_id[kv.first] = (int) slots.size();
            slots.push_back({c, kv.first, kv.second});
        }
        for (int k = d.pb; k < d.pe; k++) {
            pots[k].slot = slot_id[pots[k].L];
        }
    }
    const int nslot = (int) slots.size();
    // F rows per (centre, shell): lambda = 0 .. Lmax_c + l_A
    std::vector<int> cLmax(ncenter, -1);
    for (auto &s : slots) cLmax[s.c] = std::max(cLmax[s.c], s.L);
    tabF_base.assign((size_t) ncenter * nshell, -1);
    for (int c = 0; c < ncenter; c++) {
        if (cLmax[c] < 0) continue;
        for (int A = ctr[c].sb; A < ctr[c].se; A++) {
            TabItem t;
            t.c = c;
            t.A = A;
            t.frow = (int) frows;
            t.lam = cLmax[c] + sh[A].L;
            tabF_base[(size_t) c * nshell + A] = t.frow;
            frows += t.lam + 1;
            tab.push_back(t);
        }
    }
    // rows of each slot: shells in order, columns (n_A, lambda)
    std::vector<RowDesc> rows;
    std::vector<int> slot_rows(nslot), slot_M(nslot);
    std::vector<int> slot_rowoff((size_t) nslot * nshell, -1);
    std::vector<long long> slot_poff((size_t) nslot * nshell, -1), slot_ploff((size_t) nslot * nshell, -1);
    std::vector<long long> slot_pb2((size_t) nslot * nshell, -1), slot_pb3((size_t) nslot * nshell, -1);
    std::vector<ProjItem> proj;
    long long plen = 0;
    int proj_om = 1;
    for (int s = 0; s < nslot; s++) {
        const int c = slots[s].c, L = slots[s].L;
        slot_rows[s] = (int) rows.size();
        int M = 0;
        for (int A = ctr[c].sb; A < ctr[c].se; A++) {
            slot_rowoff[(size_t) s * nshell + A] = M;
            const int fb = tabF_base[(size_t) c * nshell + A];
            for (int n = 0; n <= sh[A].L; n++)
                for (int j = 0; j < lam_cnt(L, n); j++) {
                    RowDesc r;
                    r.frow = fb + lam_lo(L, n) + 2 * j;
                    r.n = n;
                    r.shell = A;
                    r.lam = lam_lo(L, n) + 2 * j;
                    r.lidx = M;
                    rows.push_back(r);
                    M++;
                }
            ProjItem pi;
            pi.c = c;
            pi.A = A;
            pi.L = L;
            pi.K = kcols_host(L, sh[A].L);
            const long long psz = (long long) ncart(sh[A].L) * (2 * L + 1) * pi.K;
            pi.poff = plen;
            plen += psz;
            pi.pb2 = plen;
            plen += psz;
            pi.pb3 = plen;
            plen += psz;
            pi.ploff = -1;
            if (slots[s].has_so) {
                pi.ploff = plen;
                plen += 3 * psz;
            }
            slot_poff[(size_t) s * nshell + A] = pi.poff;
            slot_ploff[(size_t) s * nshell + A] = pi.ploff;
            slot_pb2[(size_t) s * nshell + A] = pi.pb2;
            slot_pb3[(size_t) s * nshell + A] = pi.pb3;
            proj.push_back(pi);
            for (int n = 0; n <= sh[A].L; n++)
                proj_om = std::max(proj_om, lam_cnt(L, n) * (2 * L + 1) * (n + 1) * (n + 2) / 2);
        }
        slot_M[s] = M;
    }
    // radial jobs = semilocal / SO channels; tiles over the upper triangle
    std::vector<QJob> jobs;
    std::vector<TileItem> tiles;
    std::vector<int> wpots;
    long long qlen = 0;
    double flops_pp = 0.0;
    for (int k = 0; k < (int) pots.size(); k++) {
        DPot &p = pots[k];
        if (p.slot < 0) continue;
        QJob q;
        q.rows = slot_rows[p.slot];
        q.M = slot_M[p.slot];
        q.pot = (int) wpots.size();
        q.maxlev = NLEV - 1;
        q.qoff = qlen;
        q.thr = p.thr;
        q.center = slots[p.slot].c;
        q.potential = k;
        q.L = p.L;
        p.qoff = qlen;
        p.M = q.M;
        p.rbase = q.rows;
        qlen += (long long) q.M * q.M;
        wpots.push_back(k);
        const int nt = (q.M + RT - 1) / RT;
        for (int ti = 0; ti < nt; ti++)
            for (int tj = ti; tj < nt; tj++) tiles.push_back({(int) jobs.size(), ti, tj, 0});
        jobs.push_back(q);
        P->n_entries += (long long) q.M * (q.M + 1) / 2;
        flops_pp += 2.0 * RT * RT * nt * (nt + 1) / 2;
        (void) flops_pp;
    }
    // largest tiles first (entries of a diagonal tile count half): the long blocks start in the
    // first wave of the launch instead of forming its tail
    {
        auto tile_entries = [&](const TileItem &t) {
            const int M = jobs[t.job].M;
            const long long ri = std::min(RT, M - t.ti * RT), rj = std::min(RT, M - t.tj * RT);
            return (t.ti == t.tj) ? ri * (ri + 1) / 2 : ri * rj;
        };
        std::stable_sort(tiles.begin(), tiles.end(),
                         [&](const TileItem &a, const TileItem &b) { return tile_entries(a) > tile_entries(b); });
    }
    // outercore vectors and deltas
    std::vector<OcvItem> ocv;
    std::vector<long long> ocv_tab((size_t) std::max<size_t>(1, oce.size()) * nshell, -1);
    long long ocvlen = 0;
    std::vector<int> delta_centers;
    int delta_len = 0;
    for (int c = 0; c < ncenter; c++) {
        DCenter &d = ctr[c];
        const int no = d.oe - d.ob;
        d.delta_off = delta_len;
        if (no > 0) {
            delta_centers.push_back(c);
            delta_len += no * no;
        }
        for (int o = d.ob; o < d.oe; o++)
            for (int A = d.sb; A < d.se; A++) {
                OcvItem it;
                it.c = c;
                it.A = A;
                it.oc = o;
                it.off = ocvlen;
                ocv_tab[(size_t) o * nshell + A] = ocvlen;
                ocvlen += 2LL * ncart(sh[A].L) * (2 * sh[oce[o].shell].L + 1);
                ocv.push_back(it);
                // LDS of k_ocv that batches the primitive pairs exactly as the 96 KiB bound would
                // (mmd_accumulate takes min(64, scratch / e0sz) pairs per batch for r^0 terms, one
                // pair with two R/G layers otherwise), so the sums are the same
                const int la = sh[A].L, lo = sh[oce[o].shell].L, N = la + lo;
                const long long esz = 3LL * (la + 1) * (lo + 1) * (N + 1), e0sz = 3LL * (la + 1) * (lo + 1);
                const long long ssz = (long long) (N + 1) * (N + 2) * (N + 3) / 6;
                const long long npair = (long long) sh[A].nprim * sh[oce[o].shell].nprim;
                const long long bmax = std::min<long long>(64, MMD_SCRATCH / e0sz);
                long long need = std::max<long long>(1, std::min(npair, bmax)) * e0sz;
                const DPot &op = pots[oce[o].pot];
                for (int k = op.toff; k < op.toff + op.nterm; k++)
                    if (tpow[k] != 2) need = std::max(need, esz + 2 * ssz + 64);
                P->ocv_scratch = (int) std::max<long long>(P->ocv_scratch, std::min<long long>(need, MMD_SCRATCH));
                P->ocv_nel = std::max(P->ocv_nel, ncart(la) * ncart(lo));
            }
    }
    // work items
    std::vector<WorkItem> witems(nitem);
    for (int i = 0; i < nitem; i++) {
        const eng_item_t &e = items[i];
        if (ncart(sh[e.a].L) * ncart(sh[e.b].L) > MAX_BLOCK) {
            eng_set_error("shell pair block too large");
            delete P;
            return 2;
        }
        witems[i] = {e.a, e.b, e.cb, e.ce, e.off, e.moff, e.stride, e.mirror};
        const long long ntask = (long long) (e.ce - e.cb) * (e.mirror ? 2 : 1);
        P->n_tasks += ntask;
        P->n_integrals += 4LL * ncart(sh[e.a].L) * ncart(sh[e.b].L) * ntask;
    }
    std::vector<int> cidx(center_index, center_index + nidx);
    std::vector<PartItem> partlist;
    std::vector<int> item_part0(std::max(1, nitem), 0);
    long long partial_len = 0;
    for (int i = 0; i < nitem; i++) {
        item_part0[i] = (int) partlist.size();
        const long long blk = 4LL * ncart(sh[items[i].a].L) * ncart(sh[items[i].b].L);
        for (int ci = items[i].cb; ci < items[i].ce; ci++) {
            partlist.push_back({i, ci, partial_len});
            partial_len += blk;
        }
    }
    {
        // LDS of k_type1: t = 0 E columns of a batch of primitive pairs (r^0 terms), or E plus two
        // R/G layers of one pair (r^-1, r^-2 terms); the block accumulator
        bool other_pow = false;
        for (int v : tpow) other_pow |= (v != 2);
        long long sc = 64, acc = 1;
        for (int i = 0; i < nitem; i++) {
            const int la = sh[items[i].a].L, lb = sh[items[i].b].L, N = la + lb;
            const long long esz = 3LL * (la + 1) * (lb + 1) * (N + 1), e0sz = 3LL * (la + 1) * (lb + 1);
            const long long ssz = (long long) (N + 1) * (N + 2) * (N + 3) / 6;
            const long long npair = (long long) sh[items[i].a].nprim * sh[items[i].b].nprim;
            sc = std::max(sc, std::min<long long>(MMD_SCRATCH_BATCH, npair * e0sz));
            sc = std::max(sc, e0sz);
            if (other_pow) sc = std::max(sc, esz + 2 * ssz + 64);
            acc = std::max<long long>(acc, (long long) ncart(la) * ncart(lb));
        }
        P->t1_scratch = (int) std::min<long long>(sc, MMD_SCRATCH);
        P->t1_acc = (int) acc;
    }

    // ---------------- upload ----------------
    if (device_tables(Lmax_pot, lmax_shell, &P->T, P->arena != nullptr)) {
        delete P;
        return 1;
    }
    DShell *d_sh;
    double *d_alpha, *d_craw, *d_cnorm, *d_tcoef, *d_tzeta;
    DCenter *d_ctr;
    DPot *d_pot;
    int *d_tpow, *d_cidx, *d_slot_rowoff, *d_tpow_probe;
    double2 *d_pprobe, *d_tprobe;
    OcEntry *d_oce;
    long long *d_slot_poff, *d_slot_ploff, *d_ocv_tab, *d_slot_pb2, *d_slot_pb3;
    int rc = 0;
    rc |= dput(P, &d_sh, sh);
    rc |= dput(P, &d_alpha, alpha);
    {
        // screening probe order: each shell's primitives and each potential's terms by increasing
        // exponent (measured on heavy-atom clusters: the diffuse triples are the ones whose
        // envelope maximum does not settle, so they settle most failing entries in the probe)
        std::vector<int> pperm(alpha.size()), tperm(tzeta.size());
        for (const DShell &d : sh) {
            for (int i = 0; i < d.nprim; i++) pperm[d.poff + i] = d.poff + i;
            std::stable_sort(pperm.begin() + d.poff, pperm.begin() + d.poff + d.nprim,
                             [&](int a, int b) { return alpha[a] < alpha[b]; });
        }
        for (const DPot &d : pots) {
            for (int k = 0; k < d.nterm; k++) tperm[d.toff + k] = d.toff + k;
            std::stable_sort(tperm.begin() + d.toff, tperm.begin() + d.toff + d.nterm,
                             [&](int a, int b) { return tzeta[a] < tzeta[b]; });
        }
        // the screening reads the data of its probe order directly: {alpha, cnorm} per primitive,
        // {zeta, coef} and n per term, one 16-byte load per operand and no index indirection
        std::vector<double2> pprobe(alpha.size()), tprobe(tzeta.size());
        std::vector<int> tpow_probe(tzeta.size());
        for (size_t k = 0; k < pperm.size(); k++) pprobe[k] = make_double2(alpha[pperm[k]], cnorm[pperm[k]]);
        for (size_t k = 0; k < tperm.size(); k++) {
            tprobe[k] = make_double2(tzeta[tperm[k]], tcoef[tperm[k]]);
            tpow_probe[k] = tpow[tperm[k]];
        }
        rc |= dput(P, &d_pprobe, pprobe);
        rc |= dput(P, &d_tprobe, tprobe);
        rc |= dput(P, &d_tpow_probe, tpow_probe);
    }
    rc |= dput(P, &d_craw, craw);
    rc |= dput(P, &d_cnorm, cnorm);
    rc |= dput(P, &d_ctr, ctr);
    rc |= dput(P, &d_pot, pots);
    rc |= dput(P, &d_tpow, tpow);
    rc |= dput(P, &d_tcoef, tcoef);
    rc |= dput(P, &d_tzeta, tzeta);
    rc |= dput(P, &d_cidx, cidx);
    rc |= dput(P, &d_oce, oce);
    rc |= dput(P, &d_slot_rowoff, slot_rowoff);
    rc |= dput(P, &d_slot_poff, slot_poff);
    rc |= dput(P, &d_slot_ploff, slot_ploff);
    rc |= dput(P, &d_slot_pb2, slot_pb2);
    rc |= dput(P, &d_slot_pb3, slot_pb3);
    rc |= dput(P, &d_ocv_tab, ocv_tab);
    rc |= dput(P, &P->d_tab, tab);
    rc |= dput(P, &P->d_rows, rows);
    rc |= dput(P, &P->d_jobs, jobs);
    for (size_t i = 0; i < tiles.size(); i++) tiles[i].pad = (int) i;
    rc |= dput(P, &P->d_tiles, tiles);
    rc |= dput(P, &P->d_tiles_radial, tiles);
    rc |= dalloc(P, &P->d_tile_lev, tiles.size());
    P->tiles_host = tiles;
    rc |= dput(P, &P->d_proj, proj);
    rc |= dput(P, &P->d_ocvi, ocv);
    rc |= dput(P, &P->d_delta_centers, delta_centers);
    rc |= dput(P, &P->d_items, witems);
    P->canon_upper = 1;
    for (const WorkItem &w : witems)
        if (w.a > w.b) P->canon_upper = 0;
    rc |= dput(P, &P->d_parts, partlist);
    {
        // launch order of type-1 and the contraction: the costliest (largest Cartesian blocks)
        // first, so that they do not form the tail of the launch; results do not depend on it
        std::vector<PartItem> run(partlist);
        auto cost = [&](const PartItem &q) {
            const int la = sh[items[q.item].a].L, lb = sh[items[q.item].b].L;
            return (long long) ncart(la) * ncart(lb) * (la + lb + 1);
        };
        // parts of at most 64 Cartesian pairs go last, to the wave-per-part contraction (type-1 keeps
        // the same order)
        auto small = [&](const PartItem &q) {
            return ncart(sh[items[q.item].a].L) * ncart(sh[items[q.item].b].L) <= 64;
        };
        // within a cost, by (centre, A, B): the parts that read the same bra projectors run together
        auto loc = [&](const PartItem &x, const PartItem &y) {
            const int cx = cidx[x.ci], cy = cidx[y.ci];
            if (cx != cy) return cx < cy;
            const auto &ix = items[x.item];
            const auto &iy = items[y.item];
            return ix.a != iy.a ? ix.a < iy.a : ix.b < iy.b;
        };
        std::stable_sort(run.begin(), run.end(), [&](const PartItem &x, const PartItem &y) {
            const bool sx = small(x), sy = small(y);
            if (sx != sy) return sy;
            const long long cxv = cost(x), cyv = cost(y);
            return cxv != cyv ? cxv > cyv : loc(x, y);
        });
        P->n_parts_small = (int) std::count_if(run.begin(), run.end(), small);
        // the separate launch pays a second drain of the chip; below a few hundred thousand small
        // parts (heavy_cluster: 100,352 parts in all) the single launch is faster (measured)
        // GRPP_CTW_MIN_PARTS overrides the threshold (tests compare both paths on small inputs)
        const char *ctw = getenv("GRPP_CTW_MIN_PARTS");
        const long long ctw_min = ctw ? atoll(ctw) : CTW_MIN_PARTS;
        if (P->n_parts_small < ctw_min || P->n_parts_small == 0) {
            std::stable_sort(run.begin(), run.end(), [&](const PartItem &x, const PartItem &y) {
                const long long cxv = cost(x), cyv = cost(y);
                return cxv != cyv ? cxv > cyv : loc(x, y);
            });
            P->n_parts_small = 0;
        }
        rc |= dput(P, &P->d_parts_run, run);
    }
    rc |= dput(P, &P->d_item_part0, item_part0);
    rc |= dalloc(P, &P->d_partial, (size_t) partial_len);
    rc |= dalloc(P, &P->d_partial_ct, (size_t) partial_len);
    rc |= dalloc(P, &P->d_F, (size_t) frows * NPTS);
    {
        std::vector<double> wU;
        host_wU_table(wpots, pots, tpow, tcoef, tzeta, wU);
        rc |= dput(P, &P->d_wU, wU);
    }
    rc |= dalloc(P, &P->d_Q, (size_t) qlen);
    rc |= dalloc(P, &P->d_scr, (size_t) qlen);
    rc |= dalloc(P, &P->d_Qc, (size_t) qlen);
    rc |= dalloc(P, &P->d_P, (size_t) plen);
    rc |= dalloc(P, &P->d_OCV, (size_t) ocvlen);
    rc |= dalloc(P, &P->d_delta, (size_t) delta_len);
    rc |= dalloc(P, &P->d_out, (size_t) 4 * op_stride);
    rc |= dalloc(P, &P->d_counters, 8);
    std::vector<int> rowpos(rows.size());
    for (size_t i = 0; i < rows.size(); i++) rowpos[i] = rows[i].lidx;   // identity within each slot
    rc |= dput(P, &P->d_rowpos, rowpos);
    if (flags & ENG_F_LEVELS) rc |= dalloc(P, &P->d_lev, (size_t) qlen);
    if (rc) {
        eng_plan_destroy(P);
        return 1;
    }
    P->n_tab = (int) tab.size();
    P->n_wpots = (int) wpots.size();
    P->n_tiles = (int) tiles.size();
    // default screening split: 16 waves per tile when four per tile would not give the chip's
    // 5120 screening wave slots (256 CUs x 4 SIMDs x 5 waves) about six waves each, else 4
    P->tune.screen_split = 4LL * P->n_tiles < 8 * 4096 ? 16 : 4;
    P->n_proj = (int) proj.size();
    P->proj_om = proj_om;
    P->n_ocv = (int) ocv.size();
    P->n_delta_centers = (int) delta_centers.size();
    for (int c : delta_centers) P->delta_max_pairs = std::max(P->delta_max_pairs, (ctr[c].oe - ctr[c].ob) * (ctr[c].oe - ctr[c].ob));
    if (P->delta_max_pairs > 65535) {   // the launch's second grid dimension
        eng_set_error("a centre has more than 255 outercore entries");
        eng_plan_destroy(P);
        return 1;
    }
    P->n_items = nitem;
    P->n_parts = (int) partlist.size();
    P->nm_max = 2 * Lmax_pot + 1;
    P->rows_logical = rows;
    P->slot_rows = slot_rows;
    P->slot_M = slot_M;
    P->rowpos = rowpos;
    P->jobs = jobs;
    P->qlen = qlen;
    {
        // level-predictor features of every radial row (centre C, channel L, shell A, lambda, n)
        std::vector<double> zmin(nslot, 1e300), zmax(nslot, 0.0);
        for (const DPot &d : pots) {
            if (d.slot < 0) continue;
            for (int k = d.toff; k < d.toff + d.nterm; k++) {
                zmin[d.slot] = std::min(zmin[d.slot], tzeta[k]);
                zmax[d.slot] = std::max(zmax[d.slot], tzeta[k]);
            }
        }
        P->features.assign(rows.size() * ENG_NFEAT, 0.0f);
        for (int sl = 0; sl < nslot; sl++)
            for (int i = 0; i < slot_M[sl]; i++) {
                const RowDesc &r = rows[(size_t) slot_rows[sl] + i];
                const DShell &A = sh[r.shell];
                const DCenter &c = ctr[slots[sl].c];
                const double CA = std::sqrt((c.x - A.x) * (c.x - A.x) + (c.y - A.y) * (c.y - A.y) + (c.z - A.z) * (c.z - A.z));
                double amin = 1e300, amax = 0.0;
                for (int k = A.poff; k < A.poff + A.nprim; k++) {
                    amin = std::min(amin, alpha[k]);
                    amax = std::max(amax, alpha[k]);
                }
                float *f = &P->features[((size_t) slot_rows[sl] + i) * ENG_NFEAT];
                f[0] = (float) slots[sl].L;
                f[1] = (float) A.L;
                f[2] = (float) r.lam;
                f[3] = (float) r.n;
                f[4] = (float) CA;
                f[5] = (float) std::log(amin);
                f[6] = (float) std::log(amax);
                f[7] = (float) std::log(zmin[sl] < 1e299 ? zmin[sl] : 1.0);
                f[8] = (float) std::log(zmax[sl] > 0.0 ? zmax[sl] : 1.0);
                f[9] = (float) std::log1p(2.0 * amax * CA);
                f[10] = (float) A.nprim;
                f[11] = (CA < 1e-7) ? 1.0f : 0.0f;
            }
    }
    {
        // LDS of k_contract: the largest (lambda, n) block of one shell and the (m, i, k2) intermediate
        int kmax = 1, ncmax = 1;
        for (int A = 0; A < nshell; A++) {
            ncmax = std::max(ncmax, ncart(sh[A].L));
            for (int L = 0; L <= Lmax_pot; L++) kmax = std::max(kmax, kcols_host(L, sh[A].L));
        }
        P->ct_qs = kmax * kmax;
        P->ct_ts = std::min(TS_MAX, P->nm_max * ncmax * kmax);
        // small parts: KA <= ncart(la), KB <= ncart(lb), so a Q block has <= 64 doubles and the whole
        // intermediate nm x ncA x KB <= nm_max x 64 fits in one pass
        P->ctw_qs = 64;
        P->ctw_ts = P->nm_max * 64;
    }

    DevJob &J = P->J;
    J.sh = d_sh;
    J.alpha = d_alpha;
    J.craw = d_craw;
    J.cnorm = d_cnorm;
    J.ctr = d_ctr;
    J.pot = d_pot;
    J.tpow = d_tpow;
    J.rowpos = P->d_rowpos;
    J.pprobe = d_pprobe;
    J.tprobe = d_tprobe;
    J.tpow_probe = d_tpow_probe;
    J.tcoef = d_tcoef;
    J.tzeta = d_tzeta;
    J.cidx = d_cidx;
    J.oce = d_oce;
    J.slot_rowoff = d_slot_rowoff;
    J.slot_poff = d_slot_poff;
    J.slot_ploff = d_slot_ploff;
    J.slot_pb2 = d_slot_pb2;
    J.slot_pb3 = d_slot_pb3;
    J.ocv_tab = d_ocv_tab;
    J.delta = P->d_delta;
    J.Q = P->d_Qc;   // the contraction reads the canonicalised table
    J.P = P->d_P;
    J.OCV = P->d_OCV;
    J.nshell = nshell;   // lookup tables are indexed by basis shell; outercore shells follow them in sh
    J.op_stride = op_stride;
    if (!P->serial) {
        for (int i = 0; i < 8; i++) HIPCHK(hipEventCreate(&P->ev[i]));
        for (int i = 0; i < 3; i++) HIPCHK(hipEventCreate(&P->evs[i]));
        HIPCHK(hipEventCreate(&P->evt1));
        HIPCHK(hipEventCreate(&P->evr));
        HIPCHK(hipStreamCreateWithFlags(&P->side, hipStreamNonBlocking));
    }
    if (flags & ENG_F_GRAPH) {
        HIPCHK(hipStreamCreateWithFlags(&P->cap, hipStreamNonBlocking));
        for (int i = 0; i < 2; i++) HIPCHK(hipEventCreate(&P->gev[i]));
    }
    {
        // radial quadrature on the matrix cores unless GRPP_RADIAL_KERNEL=valu
        const char *k = getenv("GRPP_RADIAL_KERNEL");
        P->radial_mfma = !(k && strcmp(k, "valu") == 0);
    }
    // uploads ordered on the caller's stream: those from host vectors are done before the vectors go
    // away; the staging buffer outlives the call (eng_run returns it after its final synchronisation)
    if (P->up_used) HIPCHK(hipMemcpyAsync(P->up_dev, P->up_host, P->up_used, hipMemcpyHostToDevice, t_upload_stream));
    P->up_host = nullptr;
    if (t_upload_stream && P->up_pageable) HIPCHK(hipStreamSynchronize(t_upload_stream));
    *plan_out = P;
    return 0;
}

static int eager_phase_times(eng_plan_t *P);

__global__ void k_zero_u64(unsigned long long *p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

// Enqueues one whole build on st (and the plan's side stream, forked from and joined back to st).
// It makes no allocation, memset or synchronisation, so it can be captured into a graph.
static int enqueue_build(eng_plan_t *P, hipStream_t st)
{
    // Two streams of default priority.  The caller's stream runs the radial chain (screening,
    // radial product, canonical splits) and the contraction; the side stream runs what does not
    // depend on it (grid tables, projectors, outercore vectors, type-1), forked after the counters
    // are cleared.  The radial product waits for the grid tables (ev 5), the contraction for the
    // projectors (s1), the final reduction for the type-1 slices (s2).  Layout 1 (an autotuner
    // choice) runs the tables, projectors and outercore vectors on the caller's stream ahead of the
    // screening instead, and type-1 alone on the side stream from the start.
    // Events: caller 0 start = 1 | screening | 2 (wait for the tables) r | radial | 3 | canonical | 4 ...
    // 6 | contraction, reduction | 7; s0 | tables | 5 | projectors, outercore | s1 on the tables' stream;
    // side t1 | type-1 | s2.
    // serial plans (ENG_F_ARENA) run every kernel in this order on the caller's stream, which
    // orders them as the events below would; they own no stream or event
    const bool ser = P->serial;
    auto rec = [&](hipEvent_t e, hipStream_t s) { return ser ? hipSuccess : hipEventRecord(e, s); };
    auto wait = [&](hipStream_t s, hipEvent_t e) { return ser ? hipSuccess : hipStreamWaitEvent(s, e, 0); };
    // layout 2 runs every kernel on the caller's stream (nothing shares the chip)
    hipStream_t side = (ser || P->tune.layout == 2) ? st : P->side;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, P->d_counters, 8);
    HIPCHK(rec(P->ev[0], st));
    HIPCHK(wait(side, P->ev[0]));
    const bool front = P->tune.layout == 1;
    hipStream_t sd = front ? st : side;
    if (front) {   // type-1 needs no table: the side stream starts with it
        HIPCHK(rec(P->evt1, side));
        if (P->n_parts > 0)
            hipLaunchKernelGGL(k_type1, dim3(P->n_parts), dim3(P->tune.type1_threads),
                               (size_t) (P->t1_scratch + P->t1_acc) * sizeof(double), side, P->d_parts_run,
                               P->d_items, P->J, P->T, P->d_partial, P->parts, P->t1_scratch);
        HIPCHK(rec(P->evs[2], side));
    }
    HIPCHK(rec(P->evs[0], sd));
    // grid tables first: the radial product waits for them (event ev[5]); the screening does not
    if (P->n_tab > 0) {
        launch_tab_F(P->n_tab, sd, P->d_tab, P->J, P->T, P->d_F);
    }
    HIPCHK(rec(P->ev[5], sd));
    // one wave per projector item: most items have a few dozen outputs, so wider workgroups idled
    // (every output is one thread's sum: the width changes no result)
    if (P->n_proj > 0)
        hipLaunchKernelGGL(k_proj, dim3(P->n_proj), dim3(64), (size_t) P->proj_om * sizeof(double), sd, P->d_proj,
                           P->J, P->T, P->d_P);
    if (P->n_delta_centers > 0)
        hipLaunchKernelGGL(k_delta, dim3(P->n_delta_centers, P->delta_max_pairs), dim3(64), 0, sd, P->d_delta_centers,
                           P->n_delta_centers, P->J, P->d_delta);
    if (P->n_ocv > 0)   // one wave per item as well (the MMD sums are per thread)
        hipLaunchKernelGGL(k_ocv, dim3(P->n_ocv), dim3(64), (size_t) (P->ocv_scratch + 2 * P->ocv_nel) * sizeof(double),
                           sd, P->d_ocvi, P->J, P->T, P->d_OCV, P->ocv_scratch, P->ocv_nel);
    HIPCHK(rec(P->evs[1], sd));
    if (!front) {
        HIPCHK(rec(P->evt1, sd));
        if (P->n_parts > 0)
            hipLaunchKernelGGL(k_type1, dim3(P->n_parts), dim3(P->tune.type1_threads),
                               (size_t) (P->t1_scratch + P->t1_acc) * sizeof(double),
                               sd, P->d_parts_run, P->d_items, P->J, P->T, P->d_partial, P->parts, P->t1_scratch);
        HIPCHK(rec(P->evs[2], sd));
    }
    HIPCHK(rec(P->ev[1], st));
    if (P->n_tiles > 0) {
        const int sp = P->tune.screen_probe, ss = P->tune.screen_split;
        launch_screen(sp, ss, P->n_tiles, st, P->d_tiles, P->d_jobs, P->d_rows, P->J, P->T, P->d_scr, P->d_counters);
    }
    HIPCHK(rec(P->ev[2], st));    // end of the screening
    HIPCHK(wait(st, P->ev[5]));   // grid tables of the side stream
    HIPCHK(rec(P->evr, st));
    if (P->n_tiles > 0) {
        if (P->radial_mfma && P->d_lev)
            hipLaunchKernelGGL(k_radial_mfma<true>, dim3(P->n_tiles), dim3(256), 0, st, P->d_tiles_radial, P->d_jobs,
                               P->d_rows, P->d_F, P->T.rpow, P->d_wU, P->d_Q, P->min_level, P->tol, P->d_counters,
                               P->d_scr, P->J.sh, P->d_rowpos, P->d_lev, P->d_tile_lev);
        else if (P->radial_mfma)
            hipLaunchKernelGGL(k_radial_mfma<false>, dim3(P->n_tiles), dim3(256), 0, st, P->d_tiles_radial, P->d_jobs,
                               P->d_rows, P->d_F, P->T.rpow, P->d_wU, P->d_Q, P->min_level, P->tol, P->d_counters,
                               P->d_scr, P->J.sh, P->d_rowpos, P->d_lev, P->d_tile_lev);
        else
            hipLaunchKernelGGL(k_radial, dim3(P->n_tiles), dim3(256), 0, st, P->d_tiles, P->d_jobs, P->d_rows, P->d_F,
                               P->T.rpow, P->d_wU, P->d_Q, P->min_level, P->tol, P->d_counters, P->d_scr, P->J.sh,
                               P->d_rowpos, P->d_lev);
    }
    HIPCHK(rec(P->ev[3], st));
    if (P->n_tiles > 0)
        hipLaunchKernelGGL(k_canon, dim3(P->n_tiles), dim3(256), 0, st, P->d_tiles, P->d_jobs, P->d_rows, P->J, P->d_Q,
                           P->d_Qc, P->canon_upper);
    HIPCHK(rec(P->ev[4], st));
    HIPCHK(wait(st, P->evs[1]));   // the contraction reads projectors and outercore vectors
    HIPCHK(rec(P->ev[6], st));
    // the parts of more than 64 Cartesian pairs (largest first), then the small ones four per workgroup
    const int n_large = P->n_parts - P->n_parts_small;
    if (n_large > 0)
        launch_contract(P->tune.contract_threads, n_large, (size_t) (P->ct_qs + P->ct_ts) * sizeof(double), st,
                        P->d_parts_run, P->d_items, P->J, P->T, P->d_partial_ct, P->parts, P->nm_max, P->ct_qs,
                        P->ct_ts);
    if (P->n_parts_small > 0)
        launch_contract_w(P->n_parts_small, (size_t) (P->ctw_qs + P->ctw_ts) * sizeof(double), st,
                          P->d_parts_run + n_large, P->d_items, P->J, P->T, P->d_partial_ct, P->parts, P->nm_max,
                          P->ctw_qs, P->ctw_ts);
    HIPCHK(wait(st, P->evs[2]));   // join: the reduction reads the type-1 slices
    if (P->n_items > 0)   // also writes the zero blocks of items without centres
        hipLaunchKernelGGL(k_contract_reduce, dim3(P->n_items), dim3(256), 0, st, P->d_items, P->d_item_part0,
                           P->d_parts, P->J, P->d_partial, P->d_partial_ct, P->d_out);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int eng_plan_execute(eng_plan_t *P, void *stream_)
{
    hipStream_t st = (hipStream_t) stream_;
    if (P->flags & ENG_F_GRAPH) {
        // the build is captured once, on a plan-owned stream, and replayed on the caller's stream;
        // kernel arguments are device pointers and sizes fixed at plan creation, so a replay sees
        // the current contents of every buffer (e.g. a new row order)
        if (!P->gexec) {
            HIPCHK(hipStreamBeginCapture(P->cap, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue_build(P, P->cap);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(P->cap, &g);
            if (rc) {
                if (g) (void) hipGraphDestroy(g);
                return rc;
            }
            HIPCHK(e);
            P->graph = g;
            HIPCHK(hipGraphInstantiate(&P->gexec, P->graph, nullptr, nullptr, 0));
        }
        HIPCHK(hipEventRecord(P->gev[0], st));
        HIPCHK(hipGraphLaunch(P->gexec, st));
        HIPCHK(hipEventRecord(P->gev[1], st));
        HIPCHK(hipEventSynchronize(P->gev[1]));
        for (int i = 0; i < 7; i++) P->phase_seconds[i] = 0.0;   // phases are not timed inside a graph
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, P->gev[0], P->gev[1]));
        P->last_seconds = ms * 1e-3;
        P->last_radial_seconds = 0.0;
    } else if (P->serial) {
        const int rc = enqueue_build(P, st);
        if (rc) return rc;
        // a per-call plan's download follows on the same stream, and eng_run synchronises after it
        if (!P->percall) HIPCHK(hipStreamSynchronize(st));
        for (int i = 0; i < 7; i++) P->phase_seconds[i] = 0.0;   // no events: not timed
        P->last_seconds = 0.0;
        P->last_radial_seconds = 0.0;
    } else {
        const int rc = enqueue_build(P, st);
        if (rc) return rc;
        HIPCHK(hipEventRecord(P->ev[7], st));
        HIPCHK(hipEventSynchronize(P->ev[7]));
        const int rt = eager_phase_times(P);
        if (rt) return rt;
    }
    if (!P->serial && !P->radial_order_fresh && P->n_tiles > 0 && P->radial_mfma) {
        // the tiles that ran to the highest levels go first in the next builds (no result depends
        // on the order); refreshed after every change of the row order
        std::vector<int> lev(P->n_tiles);
        HIPCHK(hipMemcpy(lev.data(), P->d_tile_lev, lev.size() * sizeof(int), hipMemcpyDeviceToHost));
        std::vector<int> idx(P->n_tiles);
        for (int i = 0; i < P->n_tiles; i++) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return lev[a] > lev[b]; });
        std::vector<TileItem> order(P->n_tiles);
        for (int i = 0; i < P->n_tiles; i++) order[i] = P->tiles_host[(size_t) idx[i]];
        HIPCHK(hipMemcpy(P->d_tiles_radial, order.data(), order.size() * sizeof(TileItem), hipMemcpyHostToDevice));
        P->radial_order_fresh = true;
    }
    // work counters for eng_plan_stats; a per-call plan is destroyed unread, so it skips the
    // synchronous download (about 0.3 ms a call)
    if (P->percall) return 0;
    unsigned long long cnt[8];
    HIPCHK(hipMemcpy(cnt, P->d_counters, sizeof(cnt), hipMemcpyDeviceToHost));
    P->last_points = cnt[0];
    P->last_tiles = cnt[1];
    P->last_entry_points = cnt[2];
    P->last_screen_triples = cnt[3];
    P->last_screen_steps = cnt[4];
    return 0;
}

static int eager_phase_times(eng_plan_t *P)
{
    const hipEvent_t *pb[7] = {&P->evs[0], &P->ev[1], &P->evr, &P->ev[3], &P->ev[5], &P->evt1, &P->ev[6]};
    const hipEvent_t *pe[7] = {&P->ev[5], &P->ev[2], &P->ev[3], &P->ev[4], &P->evs[1], &P->evs[2], &P->ev[7]};
    for (int i = 0; i < 7; i++) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, *pb[i], *pe[i]));
        P->phase_seconds[i] = ms * 1e-3;
    }
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, P->ev[0], P->ev[7]));
    P->last_seconds = ms * 1e-3;
    P->last_radial_seconds = P->phase_seconds[2];
    return 0;
}

// ================================================================== nuclear attraction, numeric radial type-1
// Nuclear attraction of a finite (or point) nuclear charge distribution for one (shell A, shell B,
// nucleus) item and one primitive pair per workgroup (nuclear_attraction.c:54-133 with
// type1_integrals.c:144-270 and radial_type1_integral.c:100-325):
//   radial table Q[lam][N] = int w U(r) D exp(k r - (aA + aB) r^2 - aA CA^2 - aB CB^2) r^N K_lam(k r) dr,
//   lam, N <= la + lb, D = 4 pi N_A N_B, k = |k|, k = -2 (aA CA + aB CB),
// on the Log3 grid from 2047 points (levels 0..6 of the level-ordered grid, each point's weight
// scaled by the exact power of two 2^(level - 6)), doubled to 4095 and 8191: the entry keeps S7
// when |S7 - S6| <= tol, else S8 (the reference's loop stops at 10000 points).  U is the model's
// potential on the grid, tabulated on the host with the C library (as the reference evaluates it).
// Then chi_ij = sum over the binomial expansion about C of factor * sum_lam Q[lam][N] Omega_lam,
// Omega_lam(I, J, K) = sum_mu S_lam,mu(k) W1[lam][mu][(I, J, K)] (angular_integrals.c:35-56),
// with the reference's skips (|factor| < 1e-13, |Q| < 1e-16); the workgroup writes
// cA cB chi into its slice and k_nuc_sum adds the slices in primitive-pair order.
struct NucWork {
    int item, nucleus, ia, ib;   // (shell pair, nucleus, primitive pair)
    long long slice;             // first double of its ncart(A) x ncart(B) slice
};
struct NucItem {
    int a, b, w0, w1;            // shells; its work range [w0, w1) (nucleus-major, then primitive pairs)
    long long off, moff;         // output block start (row stride `stride`); start of the mirrored (B, A) block
    int stride, mirror;          // mirror: also write the transpose at moff
};

// x^n by repeated multiplication from 1 (the powers of spherical_harmonics.c:197-205)
__device__ __forceinline__ double rep_pow(double x, int n)
{
    double r = 1.0;
    for (int i = 0; i < n; i++) r *= x;
    return r;
}

__global__ __launch_bounds__(256) void k_nuc_prim(const NucWork *__restrict__ work, const NucItem *__restrict__ items,
                                                  DevJob J, DevTables T, const double *__restrict__ nuc_xyz,
                                                  const double *__restrict__ U, double tol, double *__restrict__ slices)
{
    extern __shared__ double nuc_lds[];
    const NucWork wk = work[blockIdx.x];
    const NucItem it = items[wk.item];
    const DShell A = J.sh[it.a], B = J.sh[it.b];
    const int la = A.L, lb = B.L, lab = la + lb, L1 = lab + 1, ncA = dev_ncart(la), ncB = dev_ncart(lb);
    const int nijk = L1 * (L1 + 1) * (L1 + 2) / 6;
    double *Kc = nuc_lds, *Rc = Kc + L1 * 64, *Gc = Rc + L1 * 64, *Q = Gc + 64, *SY = Q + L1 * L1, *Om = SY + L1 * L1;
    const int tid = threadIdx.x;
    const double cA = J.craw[A.poff + wk.ia], cB = J.craw[B.poff + wk.ib];
    double *out = slices + wk.slice;
    if (fabs(cA * cB) < 1e-13) {   // nuclear_attraction.c:66
        for (int e = tid; e < ncA * ncB; e += blockDim.x) out[e] = 0.0;
        return;
    }
    const double aA = J.alpha[A.poff + wk.ia], aB = J.alpha[B.poff + wk.ib];
    const double Cx = nuc_xyz[3 * wk.nucleus], Cy = nuc_xyz[3 * wk.nucleus + 1], Cz = nuc_xyz[3 * wk.nucleus + 2];
    const double CAx = Cx - A.x, CAy = Cy - A.y, CAz = Cz - A.z, CBx = Cx - B.x, CBy = Cy - B.y, CBz = Cz - B.z;
    const double CA2 = CAx * CAx + CAy * CAy + CAz * CAz, CB2 = CBx * CBx + CBy * CBy + CBz * CBz;
    const double kx = -2.0 * (aA * CAx + aB * CBx), ky = -2.0 * (aA * CAy + aB * CBy), kz = -2.0 * (aA * CAz + aB * CBz);
    const double k = sqrt(kx * kx + ky * ky + kz * kz);
    const double pref = 4 * M_PI * gnorm(la, aA) * gnorm(lb, aB);
    const double *Ug = U + (size_t) wk.nucleus * NPTS;
    // ---- radial table: thread e < L1^2 owns entry (lam = e / L1, N = e % L1) ----
    const int lam_e = tid / L1, n_e = tid % L1;
    const bool own = tid < L1 * L1;
    double s6 = 0.0, s7 = 0.0, s8 = 0.0;
    for (int p0 = 0; p0 < NPTS; p0 += 64) {
        if (tid < 64) {
            const int i = p0 + tid;
            const double r = T.r[i];
            const int lev = i < 32 ? 0 : 31 - __clz(i) - 4;   // level of the point (level l >= 1: [16 << l, 32 << l))
            const double w = i < 2048 ? T.w[i] * ldexp(1.0, lev - 6) : T.w[i];
            const double power = k * r - (aA + aB) * r * r - aA * CA2 - aB * CB2;
            Gc[tid] = (w * Ug[i]) * (pref * exp(power));
        }
        for (int e = tid; e < L1 * 64; e += blockDim.x) {
            const int l = e / 64, i = p0 + e % 64;
            Kc[e] = bessel_k(T.bes, T.rfun, l, T.r[i] * k);
            Rc[e] = T.rpow[(size_t) l * NPTS + i];
        }
        __syncthreads();
        if (own) {
            double acc = 0.0;
            for (int q = 0; q < 64; q++) acc += Gc[q] * Rc[n_e * 64 + q] * Kc[lam_e * 64 + q];
            if (p0 < 2048) s6 += acc;
            else if (p0 < 4096) s7 += acc;
            else s8 += acc;
        }
        __syncthreads();
    }
    if (own) {
        const double S6 = s6, S7 = 0.5 * S6 + s7, S8 = 0.5 * S7 + s8;
        Q[tid] = fabs(S7 - S6) <= tol ? S7 : S8;
    }
    // ---- real spherical harmonics of k (spherical_harmonics.c:172-230) ----
    if (tid < L1 * L1) {
        const int lam = (int) sqrt((double) tid), mu = tid - lam * lam - lam;
        double ux = 0.0, uy = 0.0, uz = 0.0;
        if (k > 1e-14) {
            ux = kx / k;
            uy = ky / k;
            uz = kz / k;
        }
        const double *cf = T.rsh + T.rsh_off[lam];
        const int *cm = T.comb + T.comb_off[lam];
        const int nc = (lam + 1) * (lam + 2) / 2;
        double v = 0.0;
        for (int c = 0; c < nc; c++)
            v += cf[(mu + lam) * nc + c] * (rep_pow(ux, cm[3 * c]) * rep_pow(uy, cm[3 * c + 1]) * rep_pow(uz, cm[3 * c + 2]));
        SY[tid] = v;
    }
    __syncthreads();
    // ---- Omega_lam(I, J, K) for every (I, J, K) with I + J + K <= la + lb ----
    for (int e = tid; e < L1 * nijk; e += blockDim.x) {
        const int lam = e / nijk, q = e % nijk;
        double sum = 0.0;
        for (int mu = -lam; mu <= lam; mu++)
            sum += T.w1[(size_t) (lam * lam + mu + lam) * NUC_NIJK + q] * SY[lam * lam + mu + lam];
        Om[e] = sum;
    }
    __syncthreads();
    // ---- Cartesian sum (type1_integrals.c:199-262) ----
    for (int e = tid; e < ncA * ncB; e += blockDim.x) {
        int ax, ay, az, bx, by, bz;
        dev_cart(la, e / ncB, ax, ay, az);
        dev_cart(lb, e % ncB, bx, by, bz);
        double chi = 0.0;
        for (int a = 0; a <= ax; a++)
            for (int b = 0; b <= ay; b++)
                for (int c = 0; c <= az; c++)
                    for (int d = 0; d <= bx; d++)
                        for (int f1 = 0; f1 <= by; f1++)
                            for (int f = 0; f <= bz; f++) {
                                const double factor = dev_binom(ax, a) * dev_binom(ay, b) * dev_binom(az, c) *
                                                      dev_binom(bx, d) * dev_binom(by, f1) * dev_binom(bz, f) *
                                                      pow(CAx, ax - a) * pow(CAy, ay - b) * pow(CAz, az - c) *
                                                      pow(CBx, bx - d) * pow(CBy, by - f1) * pow(CBz, bz - f);
                                if (fabs(factor) < 1e-13) continue;
                                const int N = a + b + c + d + f1 + f;
                                const int q = simplex_idx(a + d, b + f1, c + f);
                                double som = 0.0;
                                for (int lam = 0; lam <= lab; lam++) {
                                    const double qv = Q[lam * L1 + N];
                                    if (fabs(qv) < 1e-16) continue;
                                    som += Om[lam * nijk + q] * qv;
                                }
                                chi += factor * som;
                            }
        out[e] = cA * cB * chi;
    }
}

static const void *nuc_kernel(void) { return (const void *) k_nuc_prim; }

// out block of each item = sum of its slices in work order (nucleus, then primitive pair); mirrored
// (B, A) blocks are the transposes (the operator is real symmetric)
__global__ __launch_bounds__(256) void k_nuc_sum(const NucItem *__restrict__ items, const NucWork *__restrict__ work,
                                                 DevJob J, const double *__restrict__ slices, double *__restrict__ out)
{
    const NucItem it = items[blockIdx.x];
    const int ncA = dev_ncart(J.sh[it.a].L), ncB = dev_ncart(J.sh[it.b].L);
    for (int e = threadIdx.x; e < ncA * ncB; e += blockDim.x) {
        double v = 0.0;
        for (int w = it.w0; w < it.w1; w++) v += slices[work[w].slice + e];
        const int i = e / ncB, j = e % ncB;
        out[it.off + (long long) i * it.stride + j] = v;
        if (it.mirror) out[it.moff + (long long) j * it.stride + i] = v;
    }
}

// Host driver: the nuclear attraction blocks of `nitem` shell pairs summed over `nnuc` nuclei of
// model 1..5 (numeric radial quadrature), written to out_host (out_len doubles, row stride per item).
extern "C" int eng_nuclear(int nshell, libgrpp_shell_t *const *shells, int nnuc, const double *nuc_xyz,
                           const int *charge, int model, const double *params, double tol, int nitem,
                           const eng_item_t *items, long long out_len, double *out_host)
{
    if (!eng_device_ok()) return 1;
    for (int i = 0; i < nshell; i++)
        if (shells[i]->L < 0 || shells[i]->L > NUC_LMAX) {
            eng_set_error("nuclear attraction: shell %d has l = %d; the numeric nuclear models support l <= %d", i,
                          shells[i]->L, NUC_LMAX);
            return 2;
        }
    DevTables T;
    if (device_tables(0, NUC_LMAX, &T)) return 1;
    // shells: raw coefficients and exponents (nuclear_attraction.c uses shell->coeffs as given)
    std::vector<DShell> sh(nshell);
    std::vector<double> alpha, craw;
    for (int i = 0; i < nshell; i++) {
        DShell d{};
        d.L = shells[i]->L;
        d.nprim = shells[i]->num_primitives;
        d.poff = (int) alpha.size();
        d.x = shells[i]->origin[0];
        d.y = shells[i]->origin[1];
        d.z = shells[i]->origin[2];
        for (int k = 0; k < d.nprim; k++) {
            alpha.push_back(shells[i]->alpha[k]);
            craw.push_back(shells[i]->coeffs[k]);
        }
        sh[i] = d;
    }
    // the model's potential on the grid, per nucleus, with the C library (as the reference)
    const HostTables &H = host_tables();
    std::vector<double> U((size_t) nnuc * NPTS);
    for (int n = 0; n < nnuc; n++)
        if (nuc_model_potential(model, (double) charge[n], params + 3 * n, NPTS, H.r.data(), U.data() + (size_t) n * NPTS)) {
            eng_set_error("nuclear attraction: unknown nuclear model %d", model);
            return 2;
        }
    std::vector<NucItem> it(nitem);
    std::vector<NucWork> work;
    long long slen = 0;
    int lab_max = 0;
    for (int q = 0; q < nitem; q++) {
        const int a = items[q].a, b = items[q].b;
        it[q] = {a, b, (int) work.size(), 0, items[q].off, items[q].moff, items[q].stride, items[q].mirror};
        lab_max = std::max(lab_max, sh[a].L + sh[b].L);
        const long long blk = (long long) ncart(sh[a].L) * ncart(sh[b].L);
        for (int n = 0; n < nnuc; n++)
            for (int ia = 0; ia < sh[a].nprim; ia++)
                for (int ib = 0; ib < sh[b].nprim; ib++) {
                    work.push_back({q, n, ia, ib, slen});
                    slen += blk;
                }
        it[q].w1 = (int) work.size();
    }
    if (work.empty() || nitem == 0) {
        memset(out_host, 0, (size_t) out_len * sizeof(double));
        return 0;
    }
    const int L1 = lab_max + 1, nijk = L1 * (L1 + 1) * (L1 + 2) / 6;
    const size_t lds = (size_t) (2 * L1 * 64 + 64 + 2 * L1 * L1 + L1 * nijk) * sizeof(double);
    // one set of temporary device buffers for the call
    std::vector<void *> bufs;
    auto put = [&](const void *src, size_t bytes, void **dst) -> hipError_t {
        hipError_t e = hipMalloc(dst, std::max<size_t>(bytes, 8));
        if (e != hipSuccess) return e;
        bufs.push_back(*dst);
        return src ? hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };
    DShell *d_sh = nullptr;
    double *d_alpha = nullptr, *d_craw = nullptr, *d_xyz = nullptr, *d_U = nullptr, *d_slices = nullptr, *d_out = nullptr;
    NucItem *d_it = nullptr;
    NucWork *d_work = nullptr;
    hipError_t e = put(sh.data(), sh.size() * sizeof(DShell), (void **) &d_sh);
    if (e == hipSuccess) e = put(alpha.data(), alpha.size() * sizeof(double), (void **) &d_alpha);
    if (e == hipSuccess) e = put(craw.data(), craw.size() * sizeof(double), (void **) &d_craw);
    if (e == hipSuccess) e = put(nuc_xyz, (size_t) 3 * nnuc * sizeof(double), (void **) &d_xyz);
    if (e == hipSuccess) e = put(U.data(), U.size() * sizeof(double), (void **) &d_U);
    if (e == hipSuccess) e = put(it.data(), it.size() * sizeof(NucItem), (void **) &d_it);
    if (e == hipSuccess) e = put(work.data(), work.size() * sizeof(NucWork), (void **) &d_work);
    if (e == hipSuccess) e = put(nullptr, (size_t) slen * sizeof(double), (void **) &d_slices);
    if (e == hipSuccess) e = put(nullptr, (size_t) out_len * sizeof(double), (void **) &d_out);
    if (e == hipSuccess) e = hipMemset(d_out, 0, (size_t) out_len * sizeof(double));
    if (e == hipSuccess) {
        DevJob J{};
        J.sh = d_sh;
        J.alpha = d_alpha;
        J.craw = d_craw;
        hipLaunchKernelGGL(k_nuc_prim, dim3((unsigned) work.size()), dim3(256), lds, 0, d_work, d_it, J, T, d_xyz, d_U,
                           tol, d_slices);
        hipLaunchKernelGGL(k_nuc_sum, dim3((unsigned) nitem), dim3(256), 0, 0, d_it, d_work, J, d_slices, d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out_host, d_out, (size_t) out_len * sizeof(double), hipMemcpyDeviceToHost);
    for (void *b : bufs) (void) hipFree(b);
    if (e != hipSuccess) {
        eng_set_error("nuclear attraction: %s", hipGetErrorString(e));
        return 1;
    }
    return 0;
}

// ================================================================== spin-orbital Hamiltonian
// The step after the integrals (doc/manual.tex:594): libgrpp returns the real matrices arep and
// so_eta, where <A| U^SO P l_eta |B> = i so_eta[A][B] (its L matrices hold Im <S_m|l_eta|S_m'>),
// and leaves the spin s = sigma/2 to the host.  The one-electron GRPP Hamiltonian in the basis of
// spin-orbitals (all alpha functions, then all beta) is the complex Hermitian matrix
//   H = [ A + i Z/2        Y/2 + i X/2 ]
//       [ -Y/2 + i X/2     A - i Z/2   ],   A = T arep T^T, (X, Y, Z) = T so_(x,y,z) T^T,
// with T the identity (Cartesian functions) or, per shell, the Cartesian -> real-spherical-harmonic
// coefficients of libgrpp's outercore code (outercore_integrals.c:171-178, applied on both sides as
// in transform_to_sph_basis_ket/_bra, :505-542).  One workgroup per ordered shell pair.
struct HamItem {
    int la, lb, sph_row, sph_col, nsph;   // block position in the molecule's spherical (or Cartesian) dim
    int cart_stride;                      // row stride of the Cartesian matrices
    long long cart_off, h_off;            // first Cartesian element of the block; first element of H
};

constexpr int HAM_NC = (LMAX_SHELL + 1) * (LMAX_SHELL + 2) / 2, HAM_NS = 2 * LMAX_SHELL + 1;

__global__ __launch_bounds__(64) void k_spin_ham(const HamItem *__restrict__ items, const double *__restrict__ out4,
                                                 long long op_stride, DevTables T, int spherical, double2 *__restrict__ H)
{
    __shared__ double Y[HAM_NC * HAM_NS];
    __shared__ double Z[4][HAM_NS * HAM_NS];
    const HamItem it = items[blockIdx.x];
    const int ncA = dev_ncart(it.la), ncB = dev_ncart(it.lb);
    const int nsA = spherical ? 2 * it.la + 1 : ncA, nsB = spherical ? 2 * it.lb + 1 : ncB;
    const double *TA = T.slm + T.slm_off[it.la], *TB = T.slm + T.slm_off[it.lb];
    for (int op = 0; spherical && op < 4; op++) {
        const double *X = out4 + (long long) op * op_stride + it.cart_off;
        // Y = X T_B^T, then Z = T_A Y
        for (int e = threadIdx.x; e < ncA * nsB; e += blockDim.x) {
            const int a = e / nsB, q = e % nsB;
            double acc = 0.0;
            for (int b = 0; b < ncB; b++) acc += X[(long long) a * it.cart_stride + b] * TB[q * ncB + b];
            Y[e] = acc;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < nsA * nsB; e += blockDim.x) {
            const int pp = e / nsB, q = e % nsB;
            double acc = 0.0;
            for (int a = 0; a < ncA; a++) acc += TA[pp * ncA + a] * Y[a * nsB + q];
            Z[op][e] = acc;
        }
        __syncthreads();
    }
    __syncthreads();
    const long long n = it.nsph, n2 = 2LL * it.nsph;
    for (int e = threadIdx.x; e < nsA * nsB; e += blockDim.x) {
        const long long r = it.sph_row + e / nsB, c = it.sph_col + e % nsB;
        double v[4];
        for (int op = 0; op < 4; op++)
            v[op] = spherical ? Z[op][e]
                              : out4[(long long) op * op_stride + it.cart_off + (long long) (e / nsB) * it.cart_stride + e % nsB];
        const double a = v[0], x = 0.5 * v[1], y = 0.5 * v[2], z = 0.5 * v[3];
        H[it.h_off + r * n2 + c] = make_double2(a, z);
        H[it.h_off + (n + r) * n2 + n + c] = make_double2(a, -z);
        H[it.h_off + r * n2 + n + c] = make_double2(y, x);
        H[it.h_off + (n + r) * n2 + c] = make_double2(-y, x);
    }
}

// molecules m = 0..nmol-1 own shells [mol_shell_off[m], mol_shell_off[m+1]) of shell_L and their
// Cartesian matrices start at mol_cart_off[m] of each operator; H (host, interleaved complex)
// receives, molecule after molecule, the (2 n_m) x (2 n_m) spin-orbital matrices
extern "C" int eng_plan_spin_hamiltonian(eng_plan_t *P, const int *shell_L, const int *mol_shell_off, int nmol,
                                         const long long *mol_cart_off, int spherical, double *H_host,
                                         long long *h_len)
{
    std::vector<HamItem> items;
    long long hlen = 0;
    for (int m = 0; m < nmol; m++) {
        const int s0 = mol_shell_off[m], s1 = mol_shell_off[m + 1];
        int ncart_m = 0, nsph_m = 0;
        std::vector<int> coff, soff;
        for (int a = s0; a < s1; a++) {
            if (shell_L[a] < 0 || shell_L[a] > LMAX_SHELL) {
                eng_set_error("spin_hamiltonian: shell %d has l = %d outside 0..%d", a, shell_L[a], LMAX_SHELL);
                return 2;
            }
            coff.push_back(ncart_m);
            soff.push_back(nsph_m);
            ncart_m += ncart(shell_L[a]);
            nsph_m += spherical ? 2 * shell_L[a] + 1 : ncart(shell_L[a]);
        }
        for (int a = s0; a < s1; a++)
            for (int b = s0; b < s1; b++) {
                HamItem it;
                it.la = shell_L[a];
                it.lb = shell_L[b];
                it.sph_row = soff[a - s0];
                it.sph_col = soff[b - s0];
                it.nsph = nsph_m;
                it.cart_stride = ncart_m;
                it.cart_off = mol_cart_off[m] + (long long) coff[a - s0] * ncart_m + coff[b - s0];
                it.h_off = hlen;
                items.push_back(it);
            }
        hlen += 4LL * nsph_m * nsph_m;
    }
    *h_len = hlen;
    if (!H_host) return 0;
    if (items.empty()) return 0;
    HamItem *d_items = nullptr;
    double2 *d_H = nullptr;
    HIPCHK(hipMalloc((void **) &d_items, items.size() * sizeof(HamItem)));
    if (hipMalloc((void **) &d_H, (size_t) hlen * sizeof(double2)) != hipSuccess) {
        (void) hipFree(d_items);
        eng_set_error("spin_hamiltonian: %lld complex elements do not fit on the device", hlen);
        return 1;
    }
    hipError_t e = hipMemcpy(d_items, items.data(), items.size() * sizeof(HamItem), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_spin_ham, dim3((unsigned) items.size()), dim3(64), 0, 0, d_items, P->d_out, P->op_stride,
                           P->T, spherical, d_H);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(H_host, d_H, (size_t) hlen * sizeof(double2), hipMemcpyDeviceToHost);
    (void) hipFree(d_items);
    (void) hipFree(d_H);
    if (e != hipSuccess) {
        eng_set_error("spin_hamiltonian: %s", hipGetErrorString(e));
        return 1;
    }
    return 0;
}

extern "C" int eng_plan_download(eng_plan_t *P, double *out)
{
    HIPCHK(hipMemcpy(out, P->d_out, (size_t) 4 * P->op_stride * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int eng_plan_download_ops(eng_plan_t *P, double *o0, double *o1, double *o2, double *o3)
{
    double *o[4] = {o0, o1, o2, o3};
    for (int k = 0; k < 4; k++)
        HIPCHK(hipMemcpy(o[k], P->d_out + (size_t) k * P->op_stride, (size_t) P->out_len * sizeof(double),
                         hipMemcpyDeviceToHost));
    return 0;
}

extern "C" double *eng_plan_device_output(eng_plan_t *P) { return P->d_out; }

extern "C" void eng_plan_destroy(eng_plan_t *P)
{
    if (!P) return;
    if (!P->arena)
        for (auto &b : P->bufs) pool_free(P->device, b.first, b.second);
    for (int i = 0; i < 8; i++)
        if (P->ev[i]) (void) hipEventDestroy(P->ev[i]);
    for (int i = 0; i < 3; i++)
        if (P->evs[i]) (void) hipEventDestroy(P->evs[i]);
    if (P->evt1) (void) hipEventDestroy(P->evt1);
    if (P->evr) (void) hipEventDestroy(P->evr);
    if (P->side) (void) hipStreamDestroy(P->side);
    if (P->gexec) (void) hipGraphExecDestroy(P->gexec);
    if (P->graph) (void) hipGraphDestroy(P->graph);
    if (P->cap) (void) hipStreamDestroy(P->cap);
    for (int i = 0; i < 2; i++)
        if (P->gev[i]) (void) hipEventDestroy(P->gev[i]);
    delete P;
}

extern "C" int eng_plan_stats(const eng_plan_t *P, libgrpp_batch_stats_t *s)
{
    memset(s, 0, sizeof(*s));
    s->n_tasks = P->n_tasks;
    s->n_integrals = P->n_integrals;
    s->n_radial_entries = P->n_entries;
    s->n_radial_points = (long long) P->last_points;
    s->n_tiles = (long long) P->last_tiles;
    // radial quadrature: 2 FLOP (one FMA) per distinct (entry, grid point) integrated
    s->fp64_flops = 2.0 * (double) P->last_entry_points;
    s->device_seconds = P->last_seconds;
    s->radial_seconds = P->last_radial_seconds;
    s->tables_seconds = P->phase_seconds[0];
    s->screen_seconds = P->phase_seconds[1];
    s->canon_seconds = P->phase_seconds[3];
    s->proj_seconds = P->phase_seconds[4];
    s->type1_seconds = P->phase_seconds[5];
    s->contract_seconds = P->phase_seconds[6];
    s->screen_triples = (long long) P->last_screen_triples;
    s->screen_steps = (long long) P->last_screen_steps;
    // envelope estimates in FLOP equivalents (K_l evaluation 10, exp and erf 20 each):
    // per fixed-point evaluation 2 Bessel ratios (2 K + 1 division each) + 5 = 47;
    // per triple 2 K + exp + erf + 20 (r^n, prefactors, Gaussian exponent, products) = 80
    s->screen_flops = 80.0 * (double) P->last_screen_triples + 47.0 * (double) P->last_screen_steps;
    return 0;
}

extern "C" int eng_plan_features(eng_plan_t *P, int *n, float *features)
{
    *n = (int) P->rows_logical.size();
    if (features) memcpy(features, P->features.data(), P->features.size() * sizeof(float));
    return 0;
}

// Per logical row: [max, mean, fraction >= 7] of the stopping levels of its entries over every
// job of its slot (screened entries count as level 0, not converged as 8), from the last execute.
extern "C" int eng_plan_get_tuning(const eng_plan_t *P, eng_tuning_t *t)
{
    *t = P->tune;
    return 0;
}

extern "C" int eng_plan_set_tuning(eng_plan_t *P, const eng_tuning_t *t)
{
    auto in3 = [](int v, int a, int b, int c) { return v == a || v == b || v == c; };
    if (!in3(t->screen_probe, 1, 2, 4) || !in3(t->screen_split, 4, 8, 16) || !in3(t->contract_threads, 64, 128, 256) ||
        !in3(t->type1_threads, 64, 128, 256) || !in3(t->layout, 0, 1, 2)) {
        eng_set_error("tuning: screen_probe must be 1, 2 or 4, screen_split 4, 8 or 16, the thread counts 64, 128 or "
                      "256, the layout 0, 1 or 2");
        return 2;
    }
    P->tune = *t;
    if (P->gexec) {   // the captured graph holds the old launch shapes
        (void) hipGraphExecDestroy(P->gexec);
        (void) hipGraphDestroy(P->graph);
        P->gexec = nullptr;
        P->graph = nullptr;
    }
    return 0;
}

extern "C" int eng_plan_execute_eager(eng_plan_t *P, void *stream)
{
    const int f = P->flags;
    P->flags &= ~ENG_F_GRAPH;
    const int rc = eng_plan_execute(P, stream);
    P->flags = f;
    return rc;
}

extern "C" int eng_plan_row_levels(eng_plan_t *P, int *n, float *out)
{
    *n = (int) P->rows_logical.size();
    if (!out) return 0;
    if (!P->d_lev) {
        eng_set_error("plan was created without level recording");
        return 2;
    }
    std::vector<signed char> lev((size_t) P->qlen);
    HIPCHK(hipMemcpy(lev.data(), P->d_lev, lev.size(), hipMemcpyDeviceToHost));
    const size_t nr = P->rows_logical.size();
    std::vector<int> logical_of(nr);   // physical -> logical, per slot
    for (size_t sl = 0; sl < P->slot_rows.size(); sl++)
        for (int i = 0; i < P->slot_M[sl]; i++)
            logical_of[(size_t) P->slot_rows[sl] + P->rowpos[(size_t) P->slot_rows[sl] + i]] = i;
    std::vector<double> mx(nr, 0.0), sum(nr, 0.0), hi(nr, 0.0), cnt(nr, 0.0);
    for (const QJob &q : P->jobs)
        for (int pr = 0; pr < q.M; pr++) {
            const size_t gi = (size_t) q.rows + logical_of[(size_t) q.rows + pr];
            for (int pc = 0; pc < q.M; pc++) {
                int v = lev[(size_t) q.qoff + (size_t) pr * q.M + pc];
                v = v < 0 ? 0 : (v > 8 ? 8 : v);
                mx[gi] = std::max(mx[gi], (double) v);
                sum[gi] += v;
                hi[gi] += (v >= 7);
                cnt[gi] += 1.0;
            }
        }
    for (size_t i = 0; i < nr; i++) {
        out[3 * i] = (float) mx[i];
        out[3 * i + 1] = (float) (cnt[i] > 0 ? sum[i] / cnt[i] : 0.0);
        out[3 * i + 2] = (float) (cnt[i] > 0 ? hi[i] / cnt[i] : 0.0);
    }
    return 0;
}

// Orders the rows of every radial table by predicted stopping level (pred[3*i] for logical row
// i, highest first), so that 64x64 tiles hold entries of like level and stop together.  Results
// do not depend on the order: every entry keeps the reference's stopping rule.
extern "C" int eng_plan_set_predictions(eng_plan_t *P, int n, const float *pred)
{
    if (n != (int) P->rows_logical.size()) {
        eng_set_error("set_predictions: %d rows given, plan has %zu", n, P->rows_logical.size());
        return 2;
    }
    std::vector<RowDesc> phys(P->rows_logical.size());
    for (size_t sl = 0; sl < P->slot_rows.size(); sl++) {
        const int base = P->slot_rows[sl], M = P->slot_M[sl];
        std::vector<int> order(M);
        for (int i = 0; i < M; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return pred[3 * (base + a)] > pred[3 * (base + b)]; });
        for (int pos = 0; pos < M; pos++) {
            phys[(size_t) base + pos] = P->rows_logical[(size_t) base + order[pos]];
            P->rowpos[(size_t) base + order[pos]] = pos;
        }
    }
    HIPCHK(hipMemcpy(P->d_rows, phys.data(), phys.size() * sizeof(RowDesc), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(P->d_rowpos, P->rowpos.data(), P->rowpos.size() * sizeof(int), hipMemcpyHostToDevice));
    P->radial_order_fresh = false;
    return 0;
}

// diagnostics: copies a radial buffer of the last execute (0 = Q in the physical row order, 1 =
// canonical Q -- only the entries the contraction reads: the upper shell triangle when every item
// has a <= b -- and 2 = screening estimates in the logical order) and the row permutation
// (logical -> physical) to the host
extern "C" int eng_plan_debug_table(eng_plan_t *P, int which, double *out, int *rowpos, long long *qoff, int *M,
                                    int *rows)
{
    const double *src = which == 0 ? P->d_Q : (which == 1 ? P->d_Qc : P->d_scr);
    HIPCHK(hipMemcpy(out, src, (size_t) P->qlen * sizeof(double), hipMemcpyDeviceToHost));
    memcpy(rowpos, P->rowpos.data(), P->rowpos.size() * sizeof(int));
    for (size_t j = 0; j < P->jobs.size(); j++) {
        qoff[j] = P->jobs[j].qoff;
        M[j] = P->jobs[j].M;
        rows[j] = P->jobs[j].rows;
    }
    return (int) P->jobs.size();
}

extern "C" int eng_run(int nshell, libgrpp_shell_t *const *shells, int ncenter, const eng_center_t *centers, int nidx,
                       const int *center_index, int nitem, const eng_item_t *items, long long out_len,
                       long long op_stride, int parts, double *out)
{
    // a one-task plan on the calling thread's own stream, in order (no stream or events of its own):
    // calls from several host threads overlap on the device
    eng_plan_t *P = nullptr;
    const hipStream_t st = hipStreamPerThread;
    unsigned char *stage = stage_take();   // null (no pinned memory): the plain uploads
    t_upload_stream = st;
    t_stage = stage;
    int rc = eng_plan_create(nshell, shells, ncenter, centers, nidx, center_index, nitem, items, out_len, op_stride,
                             parts, ENG_F_SERIAL, &P);
    t_upload_stream = nullptr;
    t_stage = nullptr;
    if (rc) {
        stage_give(stage);
        return rc;
    }
    rc = eng_plan_execute(P, st);   // enqueues only: the download below is ordered after it on st
    if (!rc) {
        // the download through the pinned buffer when it fits
        const size_t bytes = (size_t) 4 * P->op_stride * sizeof(double);
        const bool staged = stage && bytes <= UPLOAD_BLOCK;
        const hipError_t e = hipMemcpyAsync(staged ? (void *) stage : (void *) out, P->d_out, bytes,
                                            hipMemcpyDeviceToHost, st);
        const hipError_t e2 = e == hipSuccess ? hipStreamSynchronize(st) : e;
        if (e2 != hipSuccess) {
            eng_set_error("HIP error %s in eng_run's download", hipGetErrorString(e2));
            rc = 1;
        } else if (staged) {
            memcpy(out, stage, bytes);
        }
    } else {
        (void) hipStreamSynchronize(st);   // what was enqueued ends before its buffers are pooled again
    }
    eng_plan_destroy(P);
    stage_give(stage);
    return rc;
}
