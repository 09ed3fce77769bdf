// Device-side records and small device helpers of the GRPP engine, shared by its translation units
// (engine.hip: plans, radial product, projectors, type-1, outercore; valu_kernels.hip: grid
// tabulation, screening, contraction).
#pragma once

#include <hip/hip_runtime.h>

#include "device_math.h"
#include "tables.h"

using namespace grpp;

// ------------------------------------------------------------------ device-side records
struct DShell {
    int L, nprim, poff, pad;
    double x, y, z;
};

struct DCenter {
    double x, y, z;
    int local;            // potential index of U_L or -1
    int pb, pe;           // semilocal + SO channels
    int ob, oe;           // outercore entries (global OC index)
    int sb, se;           // shells acted on
    int delta_off;        // offset of the n_oc x n_oc delta block
    int oc_i, oc_j;       // part 2 restricted to the (oc_i, oc_j) term; oc_i < 0: every term
};

struct DPot {
    int L, J, kind, nterm, toff, slot, M, rbase;   // rbase: first row of the slot
    long long qoff;
    double scale, thr;
};

enum { K_LOCAL = 0, K_AREP = 1, K_SO = 2, K_OC = 3 };

struct TabItem {
    int c, A, frow, lam;
};
struct RowDesc {
    int frow, n, shell, lam, lidx;   // lidx: logical index in the slot (shell, n, lambda order)
};
struct QJob {
    int rows, M, pot, maxlev;
    long long qoff;
    double thr;
    int center, potential;   // centre and potential record (screening)
    int L, pad;
};
struct TileItem {
    int job, ti, tj, pad;
};
struct ProjItem {
    int c, A, L, K;
    long long poff, ploff;   // ket projector, ket L-products (SO)
    long long pb2, pb3;      // bra projectors with |Omega| < 1e-16 (type-2) / < 1e-14 (SO) dropped
};
struct OcvItem {
    int c, A, oc, pad;      // oc: global OC index (potential + shell)
    long long off;
};
struct OcEntry {
    int pot, shell;         // potential index, OC-shell index (in the shell array)
};
// one (shell pair, centre) of a work item and its slice of the partial buffer
struct PartItem {
    int item, ci;
    long long off;   // first double of the slice: 4 operators x ncart(A) x ncart(B)
};
struct WorkItem {
    int a, b, cb, ce;
    long long off, moff;
    int stride, mirror;
};

struct DevJob {
    const DShell *sh;
    const double *alpha, *craw, *cnorm;
    const DCenter *ctr;
    const DPot *pot;
    const int *tpow;
    const double *tcoef, *tzeta;
    const int *cidx;
    const OcEntry *oce;
    // lookup tables, slot = (centre, L) pair
    const int *slot_rowoff;        // [slot * nshell + A]
    const int *rowpos;             // [slot row base + logical row] -> physical row of the radial table
    const double2 *pprobe;         // [poff + i] {alpha, cnorm} of the shell's primitive with the i-th smallest exponent
    const double2 *tprobe;         // [toff + t] {zeta, coef} of the potential's term with the t-th smallest exponent
    const int *tpow_probe;         // [toff + t] its power n
    const long long *slot_poff;    // [slot * nshell + A]
    const long long *slot_ploff;   // [slot * nshell + A] (-1 if the slot has no SO channel)
    const long long *slot_pb2;     // bra projector, type-2 thresholds
    const long long *slot_pb3;     // bra projector, spin-orbit thresholds
    const long long *ocv_tab;      // [oc * nshell + A]
    const double *delta;
    const double *Q, *P, *OCV;
    int nshell;
    long long op_stride;
};

// ------------------------------------------------------------------ small helpers
__device__ __forceinline__ int dev_ncart(int l) { return (l + 1) * (l + 2) / 2; }
__device__ __forceinline__ int dev_lam_lo(int L, int n) { return (n <= L) ? (L - n) : ((L + n) & 1); }
__device__ __forceinline__ int dev_lam_cnt(int L, int n) { return (L + n - dev_lam_lo(L, n)) / 2 + 1; }
__device__ __forceinline__ int dev_colbase(int L, int n)
{
    int k = 0;
    for (int q = 0; q < n; q++) k += dev_lam_cnt(L, q);
    return k;
}
__device__ __forceinline__ bool dev_valid(int L, int n, int lam)
{
    return n >= 0 && lam >= dev_lam_lo(L, n) && lam <= L + n && ((L + n - lam) & 1) == 0;
}
// column of (n, lambda) in a shell's projector block
__device__ __forceinline__ int dev_kcol(int L, int n, int lam) { return dev_colbase(L, n) + (lam - dev_lam_lo(L, n)) / 2; }
// libgrpp keeps one radial integral per (lambda1, lambda2, N = nA + nB); the canonical split is the
// smallest valid nA, so every split of the same integral reads the same number
__device__ __forceinline__ int dev_canon_nA(int L, int la, int lb, int lam1, int lam2, int N)
{
    for (int n = max(0, N - lb); n <= min(la, N); n++)
        if (dev_valid(L, n, lam1) && dev_valid(L, N - n, lam2)) return n;
    return -1;
}
__device__ __forceinline__ int dev_canon_nA_max(int L, int la, int lb, int lam1, int lam2, int N)
{
    for (int n = min(la, N); n >= max(0, N - lb); n--)
        if (dev_valid(L, n, lam1) && dev_valid(L, N - n, lam2)) return n;
    return -1;
}
__device__ __forceinline__ int dev_kcols(int L, int l)
{
    int k = 0;
    for (int n = 0; n <= l; n++) k += dev_lam_cnt(L, n);
    return k;
}
// Cartesian component c of angular momentum l in generator order (x desc, then y desc)
__device__ __forceinline__ void dev_cart(int l, int c, int &nx, int &ny, int &nz)
{
    int r = l, base = 0;
    while (c >= base + (l - r + 1)) {
        base += l - r + 1;
        r--;
    }
    nx = r;
    ny = l - r - (c - base);
    nz = l - nx - ny;
}
__device__ __forceinline__ int dev_cart_index(int nx, int ny, int nz)
{
    int l = nx + ny + nz;
    return (l - nx) * (l - nx + 1) / 2 + (l - nx - ny);
}
__device__ __forceinline__ double dev_binom(int n, int k)
{
    if (k < 0 || k > n) return 0.0;
    double r = 1.0;
    if (k > n - k) k = n - k;
    for (int d = 1; d <= k; d++) r = r * (n - k + d) / d;
    return r;
}
__device__ __forceinline__ double gnorm(int L, double a) { return pow(2 * a / M_PI, 0.75) * pow(4 * a, 0.5 * L); }

// int_0^inf r^n exp(-a r^2) dr (screening.c:255-265)
__device__ __forceinline__ double int_gauss_s(int n, double a)
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

// Workgroup b of a launch runs on XCD b % 8 (the dispatcher deals workgroups round-robin over the
// eight XCDs, each with its own L2).  This maps b to a work index so that, in every full group of
// 8 * C workgroups, each XCD takes C consecutive work indices (neighbours in the launch order share
// inputs); the indices of a last, partial group are kept.  A bijection of [0, n).
__device__ __forceinline__ int xcd_chunked(int b, int n, int C)
{
    const int G = 8 * C, g0 = (b / G) * G;
    return g0 + G <= n ? g0 + (b % 8) * C + (b / 8) % C : b;
}

constexpr int RT = 64;   // radial tile edge (entries)
constexpr int RK = 32;   // grid points per K step
constexpr int TS_MAX = 2048;   // upper bound of the LDS doubles for the (m, i, k2) intermediate

// flat index of the canonical split of entry (row, col) of a radial table
// entry (row a, column b) is the canonical split of its (lambda1, lambda2, N): the row carries
// the smallest valid n_A when its (shell, lambda) key is the smaller one, else the largest
__device__ __forceinline__ bool is_canonical(int L, int la, int lb, const RowDesc &ra, const RowDesc &rb)
{
    const int N = ra.n + rb.n;
    const bool a_first = (ra.shell < rb.shell) || (ra.shell == rb.shell && ra.lam <= rb.lam);
    const int cn = a_first ? dev_canon_nA(L, la, lb, ra.lam, rb.lam, N) : dev_canon_nA_max(L, la, lb, ra.lam, rb.lam, N);
    return cn == ra.n;
}

__device__ __forceinline__ long long canon_index(const RowDesc *R, const DShell *sh, const int *pos, const QJob &jb,
                                                int row, int col)
{
    const RowDesc ra = R[row], rb = R[col];
    const int L = jb.L, la = sh[ra.shell].L, lb = sh[rb.shell].L, N = ra.n + rb.n;
    const bool a_first = (ra.shell < rb.shell) || (ra.shell == rb.shell && ra.lam <= rb.lam);
    const int cn = a_first ? dev_canon_nA(L, la, lb, ra.lam, rb.lam, N) : dev_canon_nA_max(L, la, lb, ra.lam, rb.lam, N);
    const int r2 = ra.lidx - dev_kcol(L, ra.n, ra.lam) + dev_kcol(L, cn, ra.lam);
    const int c2 = rb.lidx - dev_kcol(L, rb.n, rb.lam) + dev_kcol(L, N - cn, rb.lam);
    (void) pos;
    // the estimate table is kept once, in the upper triangle of the LOGICAL row order (the order
    // in which the screening walks its tiles), whatever the physical order of the rows
    return (long long) min(r2, c2) * jb.M + max(r2, c2);
}

// ------------------------------------------------------------------ launchers (valu_kernels.hip)
// k_tab_F over items [blockIdx.y], NPTS / 256 workgroups of 256 threads per item
void launch_tab_F(int n_items, hipStream_t st, const TabItem *items, const DevJob &J, const DevTables &T, double *F);
// k_screen<probe, split>: split workgroups of one wave per radial tile
void launch_screen(int probe, int split, int n_tiles, hipStream_t st, const TileItem *tiles, const QJob *jobs,
                   const RowDesc *rows, const DevJob &J, const DevTables &T, double *scr,
                   unsigned long long *counters);
// k_contract<threads>: one workgroup per (shell pair, centre), lds_bytes of dynamic LDS
void launch_contract(int threads, int n_parts, size_t lds_bytes, hipStream_t st, const PartItem *partlist,
                     const WorkItem *items, const DevJob &J, const DevTables &T, double *partial, int parts,
                     int nm_max, int qs_len, int ts_len);
// k_contract_w: parts of at most 64 Cartesian pairs, four per workgroup (one per wave), each wave
// with lds_bytes_per_wave of dynamic LDS (qs_len + ts_len doubles)
void launch_contract_w(int n_parts, size_t lds_bytes_per_wave, hipStream_t st, const PartItem *partlist,
                       const WorkItem *items, const DevJob &J, const DevTables &T, double *partial, int parts,
                       int nm_max, int qs_len, int ts_len);
// the k_contract instance of a workgroup size (kernel attributes)
const void *contract_kernel_ptr(int threads);
