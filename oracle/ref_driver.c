/*
 * Matrix-assembly and timing loop linked against the genuine libgrpp sources
 * (test infrastructure: built only into oracle/_ref/, never shipped).
 *
 * It assembles what the reference C driver assembles (test_libgrpp_c/eval_integrals.c:42-130):
 * for every ordered shell pair (A,B) and every GRPP centre C, the full-GRPP blocks
 * libgrpp_full_grpp_integrals(A,B,grpp_C,C) (full_grpp_integrals.c:27-104) summed into
 * dense row-major dim x dim matrices.  Pairs are spread over OpenMP threads with a
 * dynamic schedule; blocks of different pairs are disjoint, so no reduction is needed.
 *
 * Every centre must own its libgrpp_grpp_t: libgrpp_outercore_potential_integrals()
 * writes the centre into the outercore shells (outercore_integrals.c:107-111).
 */
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include "libgrpp.h"

static int *offsets_of(int nshell, libgrpp_shell_t **shells, int *dim)
{
    int *off = (int *) malloc(sizeof(int) * (nshell + 1));
    off[0] = 0;
    for (int i = 0; i < nshell; i++) off[i + 1] = off[i] + shells[i]->cart_size;
    *dim = off[nshell];
    return off;
}

static void add_block(int dim, double *m, int r0, int c0, int nr, int nc, const double *blk, double f)
{
    for (int i = 0; i < nr; i++)
        for (int j = 0; j < nc; j++)
            m[(size_t) (r0 + i) * dim + c0 + j] += f * blk[i * nc + j];
}

/* Returns wall seconds spent in the pair loop (steady clock via omp_get_wtime). */
double ref_build_grpp_matrices(int nshell, libgrpp_shell_t **shells, int ncenter, const double *xyz,
                               libgrpp_grpp_t **grpps, double *arep, double *so_x, double *so_y, double *so_z,
                               int nthreads, int npsel, const int *pair_list)
{
    int dim;
    int *off = offsets_of(nshell, shells, &dim);
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
        libgrpp_shell_t *a = shells[i], *b = shells[j];
        int sz = a->cart_size * b->cart_size;
        double *buf = (double *) malloc(sizeof(double) * 4 * sz);
        for (int c = 0; c < ncenter; c++) {
            if (grpps[c] == NULL) continue;
            double origin[3] = {xyz[3 * c], xyz[3 * c + 1], xyz[3 * c + 2]};
            libgrpp_full_grpp_integrals(a, b, grpps[c], origin, buf, buf + sz, buf + 2 * sz, buf + 3 * sz);
            add_block(dim, arep, off[i], off[j], a->cart_size, b->cart_size, buf, 1.0);
            add_block(dim, so_x, off[i], off[j], a->cart_size, b->cart_size, buf + sz, 1.0);
            add_block(dim, so_y, off[i], off[j], a->cart_size, b->cart_size, buf + 2 * sz, 1.0);
            add_block(dim, so_z, off[i], off[j], a->cart_size, b->cart_size, buf + 3 * sz, 1.0);
        }
        free(buf);
    }
    double t1 = omp_get_wtime();
    free(off);
    return t1 - t0;
}

/*
 * Gradient matrices as the reference C driver assembles them
 * (test_libgrpp_c/eval_integrals_grad.c:37-163): grad[op][3*atom+k] is a dim x dim matrix,
 * op = 0..3 for arep, so_x, so_y, so_z; atoms are the centre list itself.
 * out must hold 4 * 3 * natom * dim * dim doubles.
 */
double ref_build_grpp_gradients(int nshell, libgrpp_shell_t **shells, int natom, const double *xyz,
                                libgrpp_grpp_t **grpps, double *out, int nthreads)
{
    int dim;
    int *off = offsets_of(nshell, shells, &dim);
    size_t n2 = (size_t) dim * dim;
    memset(out, 0, sizeof(double) * 12 * natom * n2);
    if (nthreads > 0) omp_set_num_threads(nthreads);
    double t0 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic)
    for (int p = 0; p < nshell * nshell; p++) {
        int i = p / nshell, j = p % nshell;
        libgrpp_shell_t *a = shells[i], *b = shells[j];
        int sz = a->cart_size * b->cart_size;
        double *g[4][3];
        for (int op = 0; op < 4; op++)
            for (int k = 0; k < 3; k++) g[op][k] = (double *) calloc(sz, sizeof(double));
        for (int d = 0; d < natom; d++) {
            double pt[3] = {xyz[3 * d], xyz[3 * d + 1], xyz[3 * d + 2]};
            for (int c = 0; c < natom; c++) {
                if (grpps[c] == NULL) continue;
                double origin[3] = {xyz[3 * c], xyz[3 * c + 1], xyz[3 * c + 2]};
                libgrpp_full_grpp_integrals_gradient(a, b, grpps[c], origin, pt, g[0], g[1], g[2], g[3]);
                for (int op = 0; op < 4; op++)
                    for (int k = 0; k < 3; k++)
                        add_block(dim, out + ((size_t) (op * 3 * natom + 3 * d + k)) * n2, off[i], off[j],
                                  a->cart_size, b->cart_size, g[op][k], 1.0);
            }
        }
        for (int op = 0; op < 4; op++)
            for (int k = 0; k < 3; k++) free(g[op][k]);
    }
    double t1 = omp_get_wtime();
    free(off);
    return t1 - t0;
}
