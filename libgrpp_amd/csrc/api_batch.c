/*
 * libgrpp_amd C API, part 2: gradients (grpp_integrals_gradient.c) and the batched entry points.
 *
 * Gradients follow the reference's construction: d/dA of a Cartesian Gaussian is
 * -n <L-1| + 2a <L+1| (diff_gaussian.c:35-63), so the engine evaluates full-GRPP blocks of the
 * L-1 / L+1 shells and the host scatters them with the reference's geometry cases
 * (grpp_integrals_gradient.c:280-393):
 *     A != C :  D = A gets +<dA|U_C|B>,  D = C gets -<dA|U_C|B>
 *     B != C :  D = B gets +<A|U_C|dB>,  D = C gets -<A|U_C|dB>
 * which reproduces the ACB / ACA / ACC / CCB branches (translational invariance).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "engine.h"
#include "libgrpp.h"

static void fail(const char *who)
{
    fprintf(stderr, "libgrpp_amd: %s: %s\n", who, libgrpp_last_error());
    abort();
}

static int points_equal(const double *a, const double *b)
{
    const double t = 1e-12;   /* utils.c:96-107 */
    return fabs(a[0] - b[0]) < t && fabs(a[1] - b[1]) < t && fabs(a[2] - b[2]) < t;
}

static int cidx_of(int n, int l, int m)
{
    int L = n + l + m;
    return (L - n) * (L - n + 1) / 2 + (L - n - l);
}

static double nfac(double a, int L) { return pow(2 * a / M_PI, 0.75) * pow(4 * a, 0.5 * L); }

/* L-1 and L+1 shells with the norm-factor ratios absorbed (diff_gaussian.c:35-63) */
static void diff_shell(const libgrpp_shell_t *s, libgrpp_shell_t **down, libgrpp_shell_t **up)
{
    double *c = (double *) malloc(sizeof(double) * (s->num_primitives + 1));
    double origin[3] = {s->origin[0], s->origin[1], s->origin[2]};
    *down = NULL;
    if (s->L > 0) {
        for (int i = 0; i < s->num_primitives; i++) c[i] = s->coeffs[i] * nfac(s->alpha[i], s->L) / nfac(s->alpha[i], s->L - 1);
        *down = libgrpp_new_shell(origin, s->L - 1, s->num_primitives, c, s->alpha);
    }
    for (int i = 0; i < s->num_primitives; i++)
        c[i] = s->coeffs[i] * 2.0 * s->alpha[i] * nfac(s->alpha[i], s->L) / nfac(s->alpha[i], s->L + 1);
    *up = libgrpp_new_shell(origin, s->L + 1, s->num_primitives, c, s->alpha);
    free(c);
}

/*
 * bra-derivative blocks: out[op][k][a][b] = -n_k(a) Dn[op][a-e_k][b] + Up[op][a+e_k][b]
 * (canonical Cartesian order; Dn may be NULL for L = 0).  Dn/Up are op-major blocks
 * with op stride sdn / sup and row length nB.
 */
static void bra_derivative(int LA, int nB, const double *Dn, long long sdn, const double *Up, long long sup,
                           double *out /* [4][3][ncA*nB] */)
{
    int ncA = (LA + 1) * (LA + 2) / 2;
    for (int a = 0; a < ncA; a++) {
        int nl[3];
        /* canonical component a */
        int r = LA, base = 0;
        while (a >= base + (LA - r + 1)) {
            base += LA - r + 1;
            r--;
        }
        nl[0] = r;
        nl[1] = LA - r - (a - base);
        nl[2] = LA - nl[0] - nl[1];
        for (int k = 0; k < 3; k++) {
            int up[3] = {nl[0], nl[1], nl[2]};
            up[k] += 1;
            int iu = cidx_of(up[0], up[1], up[2]);
            int id = -1;
            if (nl[k] > 0 && Dn) {
                int dn[3] = {nl[0], nl[1], nl[2]};
                dn[k] -= 1;
                id = cidx_of(dn[0], dn[1], dn[2]);
            }
            for (int op = 0; op < 4; op++)
                for (int b = 0; b < nB; b++) {
                    double v = Up[op * sup + (long long) iu * nB + b];
                    if (id >= 0) v -= nl[k] * Dn[op * sdn + (long long) id * nB + b];
                    out[((op * 3 + k) * ncA + a) * nB + b] = v;
                }
        }
    }
}

/* ---------------------------------------------------------------- per-call gradients */
static void gradient_task(const char *who, libgrpp_shell_t *A, libgrpp_shell_t *B, const eng_center_t *cen,
                          const double *C, const double *D, int parts, double **ga, double **gx, double **gy,
                          double **gz)
{
    if (!libgrpp_is_initialized()) {
        eng_set_error("libgrpp_init() has not been called");
        fail(who);
    }
    /* the engine evaluates every Cartesian component of both shells (canonical order, nA x nB);
     * the caller's cart lists, possibly shorter, select the outputs at the end */
    int nA = (A->L + 1) * (A->L + 2) / 2, nB = (B->L + 1) * (B->L + 2) / 2, sz = nA * nB;
    int nAc = A->cart_size, nBc = B->cart_size;
    double **outs[4] = {ga, gx, gy, gz};
    for (int op = 0; op < 4; op++)
        for (int k = 0; k < 3; k++)
            if (outs[op]) memset(outs[op][k], 0, sizeof(double) * nAc * nBc);
    int AC = points_equal(A->origin, C), BC = points_equal(B->origin, C);
    if (AC && BC) return;
    if (!points_equal(A->origin, D) && !points_equal(B->origin, D) && !points_equal(C, D)) return;

    libgrpp_shell_t *Ad, *Au, *Bd, *Bu;
    diff_shell(A, &Ad, &Au);
    diff_shell(B, &Bd, &Bu);
    libgrpp_shell_t *sh[6] = {A, B, Au, Bu, Ad, Bd};
    int ns = 4 + (Ad != NULL) + (Bd != NULL);
    int iAd = Ad ? 4 : -1, iBd = Bd ? (Ad ? 5 : 4) : -1;
    if (!Ad && Bd) sh[4] = Bd;
    /* blocks: (Au,B) (Ad,B) (A,Bu) (A,Bd), op-major inside one buffer each */
    int ncAu = Au->cart_size, ncAd = Ad ? Ad->cart_size : 0, ncBu = Bu->cart_size, ncBd = Bd ? Bd->cart_size : 0;
    long long o0 = 0, o1 = (long long) ncAu * nB, o2 = o1 + (long long) ncAd * nB, o3 = o2 + (long long) nA * ncBu;
    long long tot = o3 + (long long) nA * ncBd;
    eng_item_t it[4];
    int nit = 0;
    it[nit++] = (eng_item_t){2, 1, 0, 1, o0, 0, nB, 0};
    if (Ad) it[nit++] = (eng_item_t){iAd, 1, 0, 1, o1, 0, nB, 0};
    it[nit++] = (eng_item_t){0, 3, 0, 1, o2, 0, ncBu, 0};
    if (Bd) it[nit++] = (eng_item_t){0, iBd, 0, 1, o3, 0, ncBd, 0};
    int cidx = 0;
    double *buf = (double *) malloc(sizeof(double) * 4 * (tot > 0 ? tot : 1));
    if (eng_run(ns, sh, 1, cen, 1, &cidx, nit, it, tot, tot, parts, buf)) fail(who);

    double *bra = (double *) calloc(12 * sz, sizeof(double));
    double *ket = (double *) calloc(12 * sz, sizeof(double));
    bra_derivative(A->L, nB, Ad ? buf + o1 : NULL, tot, buf + o0, tot, bra);
    /* ket derivative: transpose B's blocks into (B', A) form, differentiate as a bra, transpose back */
    {
        double *tu = (double *) malloc(sizeof(double) * 4 * ncBu * nA);
        double *td = Bd ? (double *) malloc(sizeof(double) * 4 * ncBd * nA) : NULL;
        for (int op = 0; op < 4; op++) {
            for (int i = 0; i < nA; i++)
                for (int j = 0; j < ncBu; j++) tu[op * ncBu * nA + j * nA + i] = buf[op * tot + o2 + i * ncBu + j];
            if (Bd)
                for (int i = 0; i < nA; i++)
                    for (int j = 0; j < ncBd; j++) td[op * ncBd * nA + j * nA + i] = buf[op * tot + o3 + i * ncBd + j];
        }
        double *kt = (double *) malloc(sizeof(double) * 12 * sz);
        bra_derivative(B->L, nA, td, (long long) ncBd * nA, tu, (long long) ncBu * nA, kt);
        for (int op = 0; op < 4; op++)
            for (int k = 0; k < 3; k++)
                for (int j = 0; j < nB; j++)
                    for (int i = 0; i < nA; i++) ket[((op * 3 + k) * nA + i) * nB + j] = kt[((op * 3 + k) * nB + j) * nA + i];
        free(tu);
        free(td);
        free(kt);
    }
    int AD = points_equal(A->origin, D), BD = points_equal(B->origin, D), CD = points_equal(C, D);
    double *gfull = (double *) calloc(12 * sz, sizeof(double));
    for (int op = 0; op < 4; op++) {
        if (!outs[op]) continue;
        for (int k = 0; k < 3; k++) {
            double *g = gfull + (op * 3 + k) * sz;
            const double *br = bra + (op * 3 + k) * sz, *kt = ket + (op * 3 + k) * sz;
            if (!AC) {
                if (AD) for (int e = 0; e < sz; e++) g[e] += br[e];
                if (CD) for (int e = 0; e < sz; e++) g[e] -= br[e];
            }
            if (!BC) {
                if (BD) for (int e = 0; e < sz; e++) g[e] += kt[e];
                if (CD) for (int e = 0; e < sz; e++) g[e] -= kt[e];
            }
        }
    }
    /* caller's Cartesian components and order */
    for (int op = 0; op < 4; op++) {
        if (!outs[op]) continue;
        for (int k = 0; k < 3; k++) {
            double *g = outs[op][k];
            const double *full = gfull + (op * 3 + k) * sz;
            for (int i = 0; i < nAc; i++) {
                int ci = cidx_of(A->cart_list[3 * i], A->cart_list[3 * i + 1], A->cart_list[3 * i + 2]);
                for (int j = 0; j < nBc; j++) {
                    int cj = cidx_of(B->cart_list[3 * j], B->cart_list[3 * j + 1], B->cart_list[3 * j + 2]);
                    g[i * nBc + j] = full[ci * nB + cj];
                }
            }
        }
    }
    free(gfull);
    free(bra);
    free(ket);
    free(buf);
    libgrpp_delete_shell(Au);
    libgrpp_delete_shell(Bu);
    if (Ad) libgrpp_delete_shell(Ad);
    if (Bd) libgrpp_delete_shell(Bd);
}

void libgrpp_full_grpp_integrals_gradient(libgrpp_shell_t *A, libgrpp_shell_t *B, libgrpp_grpp_t *g, double *C,
                                          double *D, double **ga, double **gx, double **gy, double **gz)
{
    eng_center_t c;
    memset(&c, 0, sizeof(c));
    c.origin = C;
    c.local = g->U_L;
    c.n_arep = g->n_arep;
    c.arep = g->U_arep;
    c.n_so = g->n_esop;
    c.so = g->U_esop;
    c.n_oc = g->n_oc_shells;
    c.oc_pot = g->U_oc;
    c.oc_shell = g->oc_shells;
    c.shell_end = -1;
    gradient_task("libgrpp_full_grpp_integrals_gradient", A, B, &c, C, D, ENG_ALL, ga, gx, gy, gz);
}

void libgrpp_type1_integrals_gradient(libgrpp_shell_t *A, libgrpp_shell_t *B, double *C, libgrpp_potential_t *pot,
                                      double *D, double **ga)
{
    eng_center_t c;
    memset(&c, 0, sizeof(c));
    c.origin = C;
    c.local = pot;
    c.shell_end = -1;
    gradient_task("libgrpp_type1_integrals_gradient", A, B, &c, C, D, ENG_TYPE1, ga, NULL, NULL, NULL);
}

void libgrpp_type2_integrals_gradient(libgrpp_shell_t *A, libgrpp_shell_t *B, double *C, libgrpp_potential_t *pot,
                                      double *D, double **ga)
{
    eng_center_t c;
    memset(&c, 0, sizeof(c));
    c.origin = C;
    c.n_arep = 1;
    c.arep = &pot;
    c.shell_end = -1;
    gradient_task("libgrpp_type2_integrals_gradient", A, B, &c, C, D, ENG_TYPE2, ga, NULL, NULL, NULL);
}

void libgrpp_spin_orbit_integrals_gradient(libgrpp_shell_t *A, libgrpp_shell_t *B, double *C, libgrpp_potential_t *pot,
                                           double *D, double **gx, double **gy, double **gz)
{
    /* without the 2/(2L+1) factor (grpp_integrals_gradient.c:229-237) */
    const double one = 1.0;
    eng_center_t c;
    memset(&c, 0, sizeof(c));
    c.origin = C;
    c.n_so = 1;
    c.so = &pot;
    c.so_scale = &one;
    c.shell_end = -1;
    gradient_task("libgrpp_spin_orbit_integrals_gradient", A, B, &c, C, D, ENG_SO, NULL, gx, gy, gz);
}

void libgrpp_outercore_potential_integrals_gradient(libgrpp_shell_t *A, libgrpp_shell_t *B, double *C, int n_oc,
                                                    libgrpp_potential_t **pots, libgrpp_shell_t **shells, double *D,
                                                    double **ga, double **gx, double **gy, double **gz)
{
    eng_center_t c;
    memset(&c, 0, sizeof(c));
    c.origin = C;
    c.n_oc = n_oc;
    c.oc_pot = pots;
    c.oc_shell = shells;
    c.shell_end = -1;
    gradient_task("libgrpp_outercore_potential_integrals_gradient", A, B, &c, C, D, ENG_OC, ga, gx, gy, gz);
}

/* ---------------------------------------------------------------- batched entry points */
/* host side of a gradient plan: what the scatter of the engine's L-1 / L+1 blocks needs */
typedef struct {
    int nshells, ncenters, nidx;
    int *L, *nc, *off;        /* per basis shell: angular momentum, Cartesian size, offset */
    int *ncu, *ncd;           /* Cartesian sizes of the L+1 and L-1 shells (0: none) */
    int *atom;                /* centre at the shell's origin, or -1 */
    int *cidx;                /* centres with a GRPP */
    double *xyz, *origin;     /* centre and shell coordinates */
    long long *bup, *bdn;     /* block offsets of <i+|U_C|j>, <i-|U_C|j> (-1: none) per (i, j, centre) */
    long long tot;            /* doubles per operator of the engine's output */
} grad_host_t;

struct libgrpp_batch_plan {
    eng_plan_t *eng;
    int dim;
    int flags;
    long long out_len;   /* doubles per operator in the engine's output */
    grad_host_t *grad;   /* gradient plans only */
    /* matrix plans: shell angular momenta and the molecules' shell ranges and output offsets
       (one molecule for a single-system plan), for the spin-orbital Hamiltonian */
    int nmol;
    int *shell_L;
    int *mol_shell_off;
    long long *mol_cart_off;
};

/* records the layout of a matrix plan's output (nmol molecules, shells [off[m], off[m+1])) */
static void keep_layout(libgrpp_batch_plan_t *p, int nmol, const int *shell_off, libgrpp_shell_t *const *shells)
{
    const int ns = shell_off[nmol];
    p->nmol = nmol;
    p->shell_L = (int *) malloc(sizeof(int) * (ns > 0 ? ns : 1));
    p->mol_shell_off = (int *) malloc(sizeof(int) * (nmol + 1));
    p->mol_cart_off = (long long *) malloc(sizeof(long long) * (nmol > 0 ? nmol : 1));
    long long off = 0;
    for (int m = 0; m <= nmol; m++) p->mol_shell_off[m] = shell_off[m];
    for (int i = 0; i < ns; i++) p->shell_L[i] = shells[i]->L;
    for (int m = 0; m < nmol; m++) {
        long long d = 0;
        for (int i = shell_off[m]; i < shell_off[m + 1]; i++) d += shells[i]->cart_size;
        p->mol_cart_off[m] = off;
        off += d * d;
    }
}

static int check_canonical(int n, libgrpp_shell_t *const *shells)
{
    for (int i = 0; i < n; i++)
        for (int c = 0; c < shells[i]->cart_size; c++) {
            const int *t = shells[i]->cart_list + 3 * c;
            if (cidx_of(t[0], t[1], t[2]) != c) {
                eng_set_error("shell %d: batched entry points need the standard Cartesian order", i);
                return 0;
            }
        }
    return 1;
}

/* Flags a plan accepts.  LIBGRPP_BATCH_LEARNED_GRID and LIBGRPP_BATCH_GRADIENT are refused with
 * an error rather than ignored: the level predictor reaches a plan through
 * libgrpp_batch_plan_set_predictions, and gradients through libgrpp_batch_gradients_plan_create. */
static int check_flags(const char *who, int flags)
{
    const int known = LIBGRPP_BATCH_NO_SYMMETRY | LIBGRPP_BATCH_RECORD_LEVELS | LIBGRPP_BATCH_GRAPH;
    if (flags & ~known) {
        eng_set_error("%s: unsupported flags 0x%x (LEARNED_GRID: use libgrpp_batch_plan_set_predictions; "
                      "GRADIENT: use libgrpp_batch_gradients_plan_create)", who, flags & ~known);
        return 0;
    }
    return 1;
}

static eng_center_t *make_centers(int ncenters, const double *xyz, libgrpp_grpp_t *const *g, int *cidx, int *nidx)
{
    eng_center_t *cs = (eng_center_t *) calloc(ncenters > 0 ? ncenters : 1, sizeof(eng_center_t));
    *nidx = 0;
    for (int c = 0; c < ncenters; c++) {
        cs[c].origin = xyz + 3 * c;
        cs[c].shell_end = -1;
        if (!g[c]) continue;
        cs[c].local = g[c]->U_L;
        cs[c].n_arep = g[c]->n_arep;
        cs[c].arep = g[c]->U_arep;
        cs[c].n_so = g[c]->n_esop;
        cs[c].so = g[c]->U_esop;
        cs[c].n_oc = g[c]->n_oc_shells;
        cs[c].oc_pot = g[c]->U_oc;
        cs[c].oc_shell = g[c]->oc_shells;
        cidx[(*nidx)++] = c;
    }
    return cs;
}

static int batch_plan_create(int nshells, libgrpp_shell_t *const *shells, int ncenters, const double *xyz,
                             libgrpp_grpp_t *const *grpps, int flags, int eng_extra, libgrpp_batch_plan_t **plan);

int libgrpp_batch_plan_create(int nshells, libgrpp_shell_t *const *shells, int ncenters, const double *xyz,
                              libgrpp_grpp_t *const *grpps, int flags, libgrpp_batch_plan_t **plan)
{
    return batch_plan_create(nshells, shells, ncenters, xyz, grpps, flags, 0, plan);
}

/* eng_extra: engine flags beyond those the public flags map to (ENG_F_ARENA) */
static int batch_plan_create(int nshells, libgrpp_shell_t *const *shells, int ncenters, const double *xyz,
                             libgrpp_grpp_t *const *grpps, int flags, int eng_extra, libgrpp_batch_plan_t **plan)
{
    *plan = NULL;
    if (!check_flags("batch_plan_create", flags)) return 2;
    if (!check_canonical(nshells, shells)) return 2;
    int *off = (int *) malloc(sizeof(int) * (nshells + 1));
    off[0] = 0;
    for (int i = 0; i < nshells; i++) off[i + 1] = off[i] + shells[i]->cart_size;
    int dim = off[nshells];
    int *cidx = (int *) malloc(sizeof(int) * (ncenters > 0 ? ncenters : 1));
    int nidx = 0;
    eng_center_t *cs = make_centers(ncenters, xyz, grpps, cidx, &nidx);
    int sym = !(flags & LIBGRPP_BATCH_NO_SYMMETRY);
    int nit = sym ? nshells * (nshells + 1) / 2 : nshells * nshells;
    eng_item_t *it = (eng_item_t *) malloc(sizeof(eng_item_t) * (nit > 0 ? nit : 1));
    int k = 0;
    for (int i = 0; i < nshells; i++)
        for (int j = sym ? i : 0; j < nshells; j++) {
            eng_item_t e;
            e.a = i;
            e.b = j;
            e.cb = 0;
            e.ce = nidx;
            e.off = (long long) off[i] * dim + off[j];
            e.moff = (long long) off[j] * dim + off[i];
            e.stride = dim;
            e.mirror = sym && (i != j);
            it[k++] = e;
        }
    eng_plan_t *eng = NULL;
    long long n2 = (long long) dim * dim;
    int eflags = ((flags & LIBGRPP_BATCH_RECORD_LEVELS) ? ENG_F_LEVELS : 0) |
                 ((flags & LIBGRPP_BATCH_GRAPH) ? ENG_F_GRAPH : 0) | eng_extra;
    int rc = eng_plan_create(nshells, shells, ncenters, cs, nidx, cidx, nit, it, n2, n2, ENG_ALL, eflags, &eng);
    free(off);
    free(cidx);
    free(cs);
    free(it);
    if (rc) return rc;
    libgrpp_batch_plan_t *p = (libgrpp_batch_plan_t *) calloc(1, sizeof(*p));
    p->eng = eng;
    p->dim = dim;
    p->flags = flags;
    p->out_len = n2;
    {
        const int so[2] = {0, nshells};
        keep_layout(p, 1, so, shells);
    }
    *plan = p;
    return 0;
}

int libgrpp_batch_molecules_plan_create(int nmol, const int *shell_off, const int *center_off,
                                        libgrpp_shell_t *const *shells, const double *xyz,
                                        libgrpp_grpp_t *const *grpps, int flags, libgrpp_batch_plan_t **plan)
{
    *plan = NULL;
    if (!check_flags("batch_molecules_plan_create", flags)) return 2;
    if (nmol <= 0) {
        eng_set_error("batch_molecules: no molecules");
        return 2;
    }
    const int nshells = shell_off[nmol], ncenters = center_off[nmol];
    for (int m = 0; m < nmol; m++)
        if (shell_off[m + 1] < shell_off[m] || center_off[m + 1] < center_off[m]) {
            eng_set_error("batch_molecules: molecule %d has decreasing shell or centre offsets", m);
            return 2;
        }
    if (!check_canonical(nshells, shells)) return 2;
    int *cidx = (int *) malloc(sizeof(int) * (ncenters > 0 ? ncenters : 1));
    int nidx = 0;
    eng_center_t *cs = make_centers(ncenters, xyz, grpps, cidx, &nidx);
    /* each centre acts on its own molecule's shells; the centre-index list stays in molecule order */
    int *cb = (int *) malloc(sizeof(int) * (nmol + 1));
    int k = 0;
    for (int m = 0; m < nmol; m++) {
        cb[m] = k;
        for (int c = center_off[m]; c < center_off[m + 1]; c++) {
            cs[c].shell_begin = shell_off[m];
            cs[c].shell_end = shell_off[m + 1];
        }
        while (k < nidx && cidx[k] < center_off[m + 1]) k++;
    }
    cb[nmol] = k;
    int sym = !(flags & LIBGRPP_BATCH_NO_SYMMETRY);
    long long nit = 0, total = 0;
    for (int m = 0; m < nmol; m++) {
        const long long n = shell_off[m + 1] - shell_off[m];
        nit += sym ? n * (n + 1) / 2 : n * n;
    }
    eng_item_t *it = (eng_item_t *) malloc(sizeof(eng_item_t) * (nit > 0 ? nit : 1));
    int *off = (int *) malloc(sizeof(int) * (nshells + 1));
    long long q = 0;
    for (int m = 0; m < nmol; m++) {
        int dim = 0;
        for (int i = shell_off[m]; i < shell_off[m + 1]; i++) {
            off[i] = dim;
            dim += shells[i]->cart_size;
        }
        for (int i = shell_off[m]; i < shell_off[m + 1]; i++)
            for (int j = sym ? i : shell_off[m]; j < shell_off[m + 1]; j++) {
                eng_item_t e;
                e.a = i;
                e.b = j;
                e.cb = cb[m];
                e.ce = cb[m + 1];
                e.off = total + (long long) off[i] * dim + off[j];
                e.moff = total + (long long) off[j] * dim + off[i];
                e.stride = dim;
                e.mirror = sym && (i != j);
                it[q++] = e;
            }
        total += (long long) dim * dim;
    }
    eng_plan_t *eng = NULL;
    int eflags = ((flags & LIBGRPP_BATCH_RECORD_LEVELS) ? ENG_F_LEVELS : 0) |
                 ((flags & LIBGRPP_BATCH_GRAPH) ? ENG_F_GRAPH : 0);
    int rc = nit > 0x7fffffffLL ? 2
                                : eng_plan_create(nshells, shells, ncenters, cs, nidx, cidx, (int) nit, it, total, total,
                                                  ENG_ALL, eflags, &eng);
    if (nit > 0x7fffffffLL) eng_set_error("batch_molecules: too many shell pairs");
    free(off);
    free(cidx);
    free(cs);
    free(cb);
    free(it);
    if (rc) return rc;
    libgrpp_batch_plan_t *p = (libgrpp_batch_plan_t *) calloc(1, sizeof(*p));
    p->eng = eng;
    p->dim = 0;
    p->flags = flags;
    p->out_len = total;
    keep_layout(p, nmol, shell_off, shells);
    *plan = p;
    return 0;
}

long long libgrpp_batch_plan_output_len(const libgrpp_batch_plan_t *p) { return p->out_len; }

int libgrpp_batch_plan_download_flat(libgrpp_batch_plan_t *p, double *out) { return eng_plan_download(p->eng, out); }

int libgrpp_batch_plan_execute(libgrpp_batch_plan_t *p, void *stream) { return eng_plan_execute(p->eng, stream); }

static void grad_host_free(grad_host_t *g);

int libgrpp_batch_plan_download(libgrpp_batch_plan_t *p, double *arep, double *so_x, double *so_y, double *so_z)
{
    if (p->grad) {
        eng_set_error("download: a gradient plan's output is read by libgrpp_batch_plan_download_gradients");
        return 2;
    }
    long long n2 = (long long) p->dim * p->dim;
    double *buf = (double *) malloc(sizeof(double) * 4 * (n2 > 0 ? n2 : 1));
    int rc = eng_plan_download(p->eng, buf);
    if (!rc) {
        memcpy(arep, buf, sizeof(double) * n2);
        memcpy(so_x, buf + n2, sizeof(double) * n2);
        memcpy(so_y, buf + 2 * n2, sizeof(double) * n2);
        memcpy(so_z, buf + 3 * n2, sizeof(double) * n2);
    }
    free(buf);
    return rc;
}

int libgrpp_batch_plan_dim(const libgrpp_batch_plan_t *p) { return p->dim; }

int libgrpp_batch_plan_set_tuning(libgrpp_batch_plan_t *p, const libgrpp_batch_tuning_t *t)
{
    eng_tuning_t e = {t->screen_probe, t->contract_threads, t->type1_threads, t->stream_layout, t->screen_split};
    if (e.screen_split == 0) {   /* 0 keeps the plan's current split */
        eng_tuning_t cur;
        eng_plan_get_tuning(p->eng, &cur);
        e.screen_split = cur.screen_split;
    }
    return eng_plan_set_tuning(p->eng, &e);
}

/* fastest of `repeats` eager builds: the whole build and the phase named by `which`
 * (0 whole, 1 screening, 2 contraction) */
static int time_builds(libgrpp_batch_plan_t *p, int repeats, int which, double *best)
{
    *best = 1e300;
    for (int r = 0; r < repeats; r++) {
        int rc = eng_plan_execute_eager(p->eng, NULL);
        if (rc) return rc;
        libgrpp_batch_stats_t st;
        eng_plan_stats(p->eng, &st);
        double v = which == 1 ? st.screen_seconds : which == 2 ? st.contract_seconds : st.device_seconds;
        if (v < *best) *best = v;
    }
    return 0;
}

int libgrpp_batch_plan_autotune(libgrpp_batch_plan_t *p, int repeats, libgrpp_batch_tuning_t *chosen)
{
    if (repeats < 1) repeats = 1;
    eng_tuning_t t;
    eng_plan_get_tuning(p->eng, &t);
    double before = 0.0;
    int rc = time_builds(p, repeats, 0, &before);
    if (rc) return rc;
    /* one parameter at a time, each timed by the phase it acts on */
    static const int probes[3] = {1, 2, 4}, splits[3] = {4, 8, 16}, threads[3] = {64, 128, 256}, layouts[3] = {0, 1, 2};
    /* par: 0 probe width, 1 contraction threads, 2 type-1 threads, 3 stream layout, 4 screening split */
    for (int par = 0; par < 5; par++) {
        int *field = par == 0 ? &t.screen_probe : par == 1 ? &t.contract_threads : par == 2 ? &t.type1_threads
                   : par == 3 ? &t.layout : &t.screen_split;
        const int *cand = par == 0 ? probes : par == 3 ? layouts : par == 4 ? splits : threads;
        const int ncand = 3;
        int best_v = *field;
        double best_t = 1e300;
        for (int c = 0; c < ncand; c++) {
            *field = cand[c];
            if ((rc = eng_plan_set_tuning(p->eng, &t))) return rc;
            double v;
            if ((rc = time_builds(p, repeats, par == 0 || par == 4 ? 1 : par == 1 ? 2 : 0, &v))) return rc;
            if (v < best_t) {
                best_t = v;
                best_v = cand[c];
            }
        }
        *field = best_v;
    }
    if ((rc = eng_plan_set_tuning(p->eng, &t))) return rc;
    double after = 0.0;
    if ((rc = time_builds(p, repeats, 0, &after))) return rc;
    if (chosen) {
        chosen->screen_probe = t.screen_probe;
        chosen->contract_threads = t.contract_threads;
        chosen->type1_threads = t.type1_threads;
        chosen->stream_layout = t.layout;
        chosen->screen_split = t.screen_split;
        chosen->seconds_before = before;
        chosen->seconds_after = after;
    }
    return 0;
}
int libgrpp_batch_plan_stats(const libgrpp_batch_plan_t *p, libgrpp_batch_stats_t *s) { return eng_plan_stats(p->eng, s); }
int libgrpp_batch_plan_features(libgrpp_batch_plan_t *p, int *n, float *f) { return eng_plan_features(p->eng, n, f); }
int libgrpp_batch_plan_row_levels(libgrpp_batch_plan_t *p, int *n, float *v) { return eng_plan_row_levels(p->eng, n, v); }
int libgrpp_batch_plan_set_predictions(libgrpp_batch_plan_t *p, int n, const float *v)
{
    return eng_plan_set_predictions(p->eng, n, v);
}

void libgrpp_batch_plan_destroy(libgrpp_batch_plan_t *p)
{
    if (!p) return;
    eng_plan_destroy(p->eng);
    grad_host_free(p->grad);
    free(p->shell_L);
    free(p->mol_shell_off);
    free(p->mol_cart_off);
    free(p);
}

long long libgrpp_batch_plan_spin_hamiltonian_len(const libgrpp_batch_plan_t *p, int spherical)
{
    if (p->grad || !p->shell_L) return -1;
    long long n = 0;
    for (int m = 0; m < p->nmol; m++) {
        long long d = 0;
        for (int i = p->mol_shell_off[m]; i < p->mol_shell_off[m + 1]; i++) {
            const int l = p->shell_L[i];
            d += spherical ? 2 * l + 1 : (l + 1) * (l + 2) / 2;
        }
        n += 4 * d * d;
    }
    return n;
}

int libgrpp_batch_plan_spin_hamiltonian(libgrpp_batch_plan_t *p, int spherical, double *H)
{
    if (p->grad || !p->shell_L) {
        eng_set_error("spin_hamiltonian: a gradient plan has no matrices to assemble");
        return 2;
    }
    long long len = 0;
    return eng_plan_spin_hamiltonian(p->eng, p->shell_L, p->mol_shell_off, p->nmol, p->mol_cart_off, spherical != 0,
                                     H, &len);
}

int libgrpp_batch_full_grpp_integrals(int nshells, libgrpp_shell_t *const *shells, int ncenters, const double *xyz,
                                      libgrpp_grpp_t *const *grpps, int flags, double *arep, double *so_x,
                                      double *so_y, double *so_z, libgrpp_batch_stats_t *stats)
{
    libgrpp_batch_plan_t *p = NULL;
    int rc = libgrpp_batch_plan_create(nshells, shells, ncenters, xyz, grpps, flags, &p);
    if (rc) return rc;
    rc = libgrpp_batch_plan_execute(p, NULL);
    if (!rc) rc = libgrpp_batch_plan_download(p, arep, so_x, so_y, so_z);
    if (!rc && stats) libgrpp_batch_plan_stats(p, stats);
    libgrpp_batch_plan_destroy(p);
    return rc;
}

/*
 * The same matrices with no device allocation: every device buffer of the build is carved from
 * the engine's static scratch array of this device (ENG_SCRATCH_BYTES), and the kernels run in
 * order on the caller's stream with no stream or event of their own; the four matrices are then
 * copied into the caller's host arrays.  Inputs that need more scratch than that are refused.
 */
int libgrpp_batch_full_grpp_integrals_noalloc(int nshells, libgrpp_shell_t *const *shells, int ncenters,
                                              const double *xyz, libgrpp_grpp_t *const *grpps, int flags,
                                              double *arep, double *so_x, double *so_y, double *so_z, void *stream,
                                              libgrpp_batch_stats_t *stats)
{
    if (flags & (LIBGRPP_BATCH_GRAPH | LIBGRPP_BATCH_RECORD_LEVELS)) {
        eng_set_error("batch_full_grpp_integrals_noalloc: GRAPH and RECORD_LEVELS are not available without allocation");
        return 2;
    }
    libgrpp_batch_plan_t *p = NULL;
    int rc = batch_plan_create(nshells, shells, ncenters, xyz, grpps, flags, ENG_F_ARENA, &p);
    if (rc) return rc;
    rc = libgrpp_batch_plan_execute(p, stream);
    if (!rc) rc = eng_plan_download_ops(p->eng, arep, so_x, so_y, so_z);
    if (!rc && stats) libgrpp_batch_plan_stats(p, stats);
    libgrpp_batch_plan_destroy(p);
    return rc;
}

/*
 * Nuclear attraction matrix of a shell list: every shell pair i <= j once, mirrored (the operator is
 * real symmetric), summed over the nuclei.  The point-charge model runs the analytic type-1 kernels
 * with one -Z r^-1 term of exponent 0 per nucleus (nuclear_attraction.c:73-84); the finite models
 * and the numeric point charge run the numeric radial quadrature (nuclear_attraction.c:85-160).
 */
int libgrpp_batch_nuclear_attraction_integrals(int nshells, libgrpp_shell_t *const *shells, int nnuc,
                                               const double *nuc_xyz, const int *charges, int nuclear_model,
                                               const double *model_params, double *V)
{
    if (!check_canonical(nshells, shells)) return 2;
    if (nuclear_model < LIBGRPP_NUCLEAR_MODEL_POINT_CHARGE || nuclear_model > LIBGRPP_NUCLEAR_MODEL_POINT_CHARGE_NUMERICAL) {
        eng_set_error("batch_nuclear_attraction_integrals: unknown nuclear model %d", nuclear_model);
        return 2;
    }
    if (nuclear_model != LIBGRPP_NUCLEAR_MODEL_POINT_CHARGE && nuclear_model != LIBGRPP_NUCLEAR_MODEL_POINT_CHARGE_NUMERICAL &&
        model_params == NULL) {
        eng_set_error("batch_nuclear_attraction_integrals: model %d needs model_params", nuclear_model);
        return 2;
    }
    int *off = (int *) malloc(sizeof(int) * (nshells + 1));
    off[0] = 0;
    for (int i = 0; i < nshells; i++) off[i + 1] = off[i] + shells[i]->cart_size;
    const int dim = off[nshells];
    const int nit = nshells * (nshells + 1) / 2;
    eng_item_t *it = (eng_item_t *) malloc(sizeof(eng_item_t) * (nit > 0 ? nit : 1));
    int k = 0;
    for (int i = 0; i < nshells; i++)
        for (int j = i; j < nshells; j++) {
            eng_item_t e;
            e.a = i;
            e.b = j;
            e.cb = 0;
            e.ce = nnuc;
            e.off = (long long) off[i] * dim + off[j];
            e.moff = (long long) off[j] * dim + off[i];
            e.stride = dim;
            e.mirror = i != j;
            it[k++] = e;
        }
    free(off);
    const long long n2 = (long long) dim * dim;
    int rc;
    if (nuclear_model == LIBGRPP_NUCLEAR_MODEL_POINT_CHARGE) {
        libgrpp_potential_t **pot = (libgrpp_potential_t **) calloc(nnuc > 0 ? nnuc : 1, sizeof(*pot));
        eng_center_t *cs = (eng_center_t *) calloc(nnuc > 0 ? nnuc : 1, sizeof(eng_center_t));
        int *cidx = (int *) malloc(sizeof(int) * (nnuc > 0 ? nnuc : 1));
        for (int n = 0; n < nnuc; n++) {
            int pw = 1;
            double cf = -(double) charges[n], ze = 0.0;
            pot[n] = libgrpp_new_potential(0, 0, 1, &pw, &cf, &ze);
            cs[n].origin = nuc_xyz + 3 * n;
            cs[n].local = pot[n];
            cs[n].shell_end = -1;
            cidx[n] = n;
        }
        double *buf = (double *) malloc(sizeof(double) * 4 * (n2 > 0 ? n2 : 1));
        rc = eng_run(nshells, shells, nnuc, cs, nnuc, cidx, nit, it, n2, n2, ENG_TYPE1, buf);
        if (!rc) memcpy(V, buf, sizeof(double) * n2);
        free(buf);
        for (int n = 0; n < nnuc; n++) libgrpp_delete_potential(pot[n]);
        free(pot);
        free(cs);
        free(cidx);
    } else {
        double *p3 = (double *) calloc(3 * (nnuc > 0 ? nnuc : 1), sizeof(double));
        for (int i = 0; model_params && i < 3 * nnuc; i++) p3[i] = model_params[i];
        rc = eng_nuclear(nshells, shells, nnuc, nuc_xyz, charges, nuclear_model, p3, libgrpp_params.radial_tolerance,
                         nit, it, n2, V);
        free(p3);
    }
    free(it);
    return rc;
}

/*
 * Batched gradients: the engine evaluates, per centre, the blocks <i-|U_C|j> and <i+|U_C|j> for
 * every ordered pair; ket derivatives reuse them through <i|U|j+-> = s <j+-|U|i>^T
 * (s = +1 arep, -1 spin-orbit).  A gradient plan holds the engine plan of those blocks and the
 * host data of the scatter, so the inputs stay resident between executions.
 */
static void grad_host_free(grad_host_t *g)
{
    if (!g) return;
    free(g->L);
    free(g->nc);
    free(g->off);
    free(g->ncu);
    free(g->ncd);
    free(g->atom);
    free(g->cidx);
    free(g->xyz);
    free(g->origin);
    free(g->bup);
    free(g->bdn);
    free(g);
}

int libgrpp_batch_gradients_plan_create(int nshells, libgrpp_shell_t *const *shells, int ncenters, const double *xyz,
                                        libgrpp_grpp_t *const *grpps, int flags, libgrpp_batch_plan_t **plan)
{
    *plan = NULL;
    if (!check_flags("batch_gradients_plan_create", flags)) return 2;
    if (!check_canonical(nshells, shells)) return 2;
    grad_host_t *g = (grad_host_t *) calloc(1, sizeof(*g));
    const int ns1 = nshells + 1;
    g->nshells = nshells;
    g->ncenters = ncenters;
    g->L = (int *) malloc(sizeof(int) * ns1);
    g->nc = (int *) malloc(sizeof(int) * ns1);
    g->off = (int *) malloc(sizeof(int) * ns1);
    g->ncu = (int *) malloc(sizeof(int) * ns1);
    g->ncd = (int *) malloc(sizeof(int) * ns1);
    g->atom = (int *) malloc(sizeof(int) * ns1);
    g->origin = (double *) malloc(sizeof(double) * 3 * ns1);
    g->xyz = (double *) malloc(sizeof(double) * 3 * (ncenters > 0 ? ncenters : 1));
    if (ncenters > 0) memcpy(g->xyz, xyz, sizeof(double) * 3 * ncenters);
    g->off[0] = 0;
    for (int i = 0; i < nshells; i++) {
        g->L[i] = shells[i]->L;
        g->nc[i] = shells[i]->cart_size;
        g->off[i + 1] = g->off[i] + g->nc[i];
        memcpy(g->origin + 3 * i, shells[i]->origin, sizeof(double) * 3);
        g->atom[i] = -1;
        for (int c = 0; c < ncenters; c++)
            if (points_equal(shells[i]->origin, xyz + 3 * c)) {
                g->atom[i] = c;
                break;
            }
    }
    const int dim = g->off[nshells];

    /* extended shell list: basis, then (down, up) per basis shell */
    int next = nshells;
    libgrpp_shell_t **sh = (libgrpp_shell_t **) malloc(sizeof(void *) * 3 * ns1);
    int *idn = (int *) malloc(sizeof(int) * ns1), *iup = (int *) malloc(sizeof(int) * ns1);
    for (int i = 0; i < nshells; i++) sh[i] = shells[i];
    for (int i = 0; i < nshells; i++) {
        libgrpp_shell_t *d, *u;
        diff_shell(shells[i], &d, &u);
        idn[i] = d ? next : -1;
        g->ncd[i] = d ? d->cart_size : 0;
        if (d) sh[next++] = d;
        iup[i] = next;
        g->ncu[i] = u->cart_size;
        sh[next++] = u;
    }
    g->cidx = (int *) malloc(sizeof(int) * (ncenters > 0 ? ncenters : 1));
    eng_center_t *cs = make_centers(ncenters, xyz, grpps, g->cidx, &g->nidx);
    const int nidx = g->nidx;
    /* items: per (i, j, centre) the down and up blocks, op-major in one buffer */
    const long long nkey = (long long) nshells * nshells * (nidx > 0 ? nidx : 1);
    eng_item_t *it = (eng_item_t *) malloc(sizeof(eng_item_t) * 2 * nkey);
    g->bup = (long long *) malloc(sizeof(long long) * nkey);
    g->bdn = (long long *) malloc(sizeof(long long) * nkey);
    long long tot = 0;
    int nit = 0;
    for (int i = 0; i < nshells; i++)
        for (int j = 0; j < nshells; j++)
            for (int ci = 0; ci < nidx; ci++) {
                const long long key = ((long long) i * nshells + j) * nidx + ci;
                const int ncb = g->nc[j];
                g->bdn[key] = -1;
                if (idn[i] >= 0) {
                    it[nit++] = (eng_item_t){idn[i], j, ci, ci + 1, tot, 0, ncb, 0};
                    g->bdn[key] = tot;
                    tot += (long long) g->ncd[i] * ncb;
                }
                it[nit++] = (eng_item_t){iup[i], j, ci, ci + 1, tot, 0, ncb, 0};
                g->bup[key] = tot;
                tot += (long long) g->ncu[i] * ncb;
            }
    g->tot = tot;
    const int eflags = ((flags & LIBGRPP_BATCH_RECORD_LEVELS) ? ENG_F_LEVELS : 0) |
                       ((flags & LIBGRPP_BATCH_GRAPH) ? ENG_F_GRAPH : 0);
    eng_plan_t *eng = NULL;
    int rc = eng_plan_create(next, sh, ncenters, cs, nidx, g->cidx, nit, it, tot, tot, ENG_ALL, eflags, &eng);
    /* the engine holds its own copies of the shells */
    for (int i = nshells; i < next; i++) libgrpp_delete_shell(sh[i]);
    free(sh);
    free(idn);
    free(iup);
    free(cs);
    free(it);
    if (rc) {
        grad_host_free(g);
        return rc;
    }
    libgrpp_batch_plan_t *p = (libgrpp_batch_plan_t *) calloc(1, sizeof(*p));
    p->eng = eng;
    p->dim = dim;
    p->flags = flags;
    p->out_len = tot;
    p->grad = g;
    *plan = p;
    return 0;
}

long long libgrpp_batch_plan_gradient_len(const libgrpp_batch_plan_t *p)
{
    if (!p->grad) return 0;
    return 12LL * p->grad->ncenters * p->dim * p->dim;
}

/* the scatter of the engine's blocks into out: slot[atom] is the position of atom's gradients in
 * out (nslot atoms), -1 for atoms that are not downloaded */
static int download_gradients_slots(libgrpp_batch_plan_t *p, int nslot, const int *slot, double *out)
{
    const grad_host_t *g = p->grad;
    const int nshells = g->nshells, nidx = g->nidx, dim = p->dim;
    const long long n2 = (long long) dim * dim, tot = g->tot;
    double *buf = (double *) malloc(sizeof(double) * 4 * (tot > 0 ? tot : 1));
    if (!buf) {
        eng_set_error("download_gradients: out of host memory for %lld doubles", 4 * tot);
        return 1;
    }
    int rc = eng_plan_download(p->eng, buf);
    if (rc) {
        free(buf);
        return rc;
    }
    memset(out, 0, sizeof(double) * 12 * (size_t) nslot * n2);
    int maxnc = 1, maxnu = 1;
    for (int i = 0; i < nshells; i++) {
        if (g->nc[i] > maxnc) maxnc = g->nc[i];
        if (g->ncu[i] > maxnu) maxnu = g->ncu[i];
    }
    double *bra = (double *) malloc(sizeof(double) * 12 * maxnc * maxnc);
    double *ket = (double *) malloc(sizeof(double) * 12 * maxnc * maxnc);
    double *kt = (double *) malloc(sizeof(double) * 12 * maxnc * maxnc);
    double *tu = (double *) malloc(sizeof(double) * 4 * maxnu * maxnc);
    double *td = (double *) malloc(sizeof(double) * 4 * maxnu * maxnc);
    for (int i = 0; i < nshells; i++)
        for (int j = 0; j < nshells; j++)
            for (int ci = 0; ci < nidx; ci++) {
                const int c = g->cidx[ci];
                const double *C = g->xyz + 3 * c;
                const int AC = points_equal(g->origin + 3 * i, C), BC = points_equal(g->origin + 3 * j, C);
                if (AC && BC) continue;
                if (slot[c] < 0 && (g->atom[i] < 0 || slot[g->atom[i]] < 0) && (g->atom[j] < 0 || slot[g->atom[j]] < 0))
                    continue;   /* touches no downloaded atom */
                const int nA = g->nc[i], nB = g->nc[j];
                const long long kij = ((long long) i * nshells + j) * nidx + ci;
                const long long kji = ((long long) j * nshells + i) * nidx + ci;
                if (!AC)
                    bra_derivative(g->L[i], nB, g->bdn[kij] >= 0 ? buf + g->bdn[kij] : NULL, tot, buf + g->bup[kij], tot,
                                   bra);
                if (!BC) {
                    /* <i|U|j+-> = s <j+-|U|i>^T, then differentiate as a bra over j */
                    const int ncu = g->ncu[j], ncd = g->ncd[j];
                    for (int op = 0; op < 4; op++) {
                        for (int a = 0; a < ncu; a++)
                            for (int b = 0; b < nA; b++) tu[op * ncu * nA + a * nA + b] = buf[op * tot + g->bup[kji] + a * nA + b];
                        for (int a = 0; a < ncd; a++)
                            for (int b = 0; b < nA; b++) td[op * ncd * nA + a * nA + b] = buf[op * tot + g->bdn[kji] + a * nA + b];
                    }
                    bra_derivative(g->L[j], nA, ncd ? td : NULL, (long long) ncd * nA, tu, (long long) ncu * nA, kt);
                    for (int op = 0; op < 4; op++) {
                        const double s = op == 0 ? 1.0 : -1.0;
                        for (int k = 0; k < 3; k++)
                            for (int b = 0; b < nB; b++)
                                for (int a = 0; a < nA; a++)
                                    ket[((op * 3 + k) * nA + a) * nB + b] = s * kt[((op * 3 + k) * nB + b) * nA + a];
                    }
                }
                for (int op = 0; op < 4; op++)
                    for (int k = 0; k < 3; k++) {
                        const double *br = bra + (op * 3 + k) * nA * nB, *kk = ket + (op * 3 + k) * nA * nB;
                        const int sa = g->atom[i] >= 0 ? slot[g->atom[i]] : -1;
                        const int sb = g->atom[j] >= 0 ? slot[g->atom[j]] : -1;
                        double *gA = sa >= 0 ? out + ((long long) (op * 3 * nslot + 3 * sa + k)) * n2 : NULL;
                        double *gB = sb >= 0 ? out + ((long long) (op * 3 * nslot + 3 * sb + k)) * n2 : NULL;
                        double *gC = slot[c] >= 0 ? out + ((long long) (op * 3 * nslot + 3 * slot[c] + k)) * n2 : NULL;
                        for (int a = 0; a < nA; a++)
                            for (int b = 0; b < nB; b++) {
                                const long long e = (long long) (g->off[i] + a) * dim + g->off[j] + b;
                                if (!AC) {
                                    if (gA) gA[e] += br[a * nB + b];
                                    if (gC) gC[e] -= br[a * nB + b];
                                }
                                if (!BC) {
                                    if (gB) gB[e] += kk[a * nB + b];
                                    if (gC) gC[e] -= kk[a * nB + b];
                                }
                            }
                    }
            }
    free(bra);
    free(ket);
    free(kt);
    free(tu);
    free(td);
    free(buf);
    return 0;
}

int libgrpp_batch_plan_download_gradients(libgrpp_batch_plan_t *p, double *out)
{
    if (!p->grad) {
        eng_set_error("download_gradients: the plan was not created by libgrpp_batch_gradients_plan_create");
        return 2;
    }
    const int n = p->grad->ncenters;
    int *slot = (int *) malloc(sizeof(int) * (n > 0 ? n : 1));
    for (int c = 0; c < n; c++) slot[c] = c;
    const int rc = download_gradients_slots(p, n, slot, out);
    free(slot);
    return rc;
}

int libgrpp_batch_plan_download_gradients_atoms(libgrpp_batch_plan_t *p, int natoms, const int *atoms, double *out)
{
    if (!p->grad) {
        eng_set_error("download_gradients_atoms: the plan was not created by libgrpp_batch_gradients_plan_create");
        return 2;
    }
    const int n = p->grad->ncenters;
    if (natoms < 0 || (natoms > 0 && !atoms)) {
        eng_set_error("download_gradients_atoms: invalid atom list");
        return 2;
    }
    int *slot = (int *) malloc(sizeof(int) * (n > 0 ? n : 1));
    for (int c = 0; c < n; c++) slot[c] = -1;
    for (int a = 0; a < natoms; a++) {
        if (atoms[a] < 0 || atoms[a] >= n || slot[atoms[a]] >= 0) {
            eng_set_error("download_gradients_atoms: atom %d out of range [0, %d) or listed twice", atoms[a], n);
            free(slot);
            return 2;
        }
        slot[atoms[a]] = a;
    }
    const int rc = download_gradients_slots(p, natoms, slot, out);
    free(slot);
    return rc;
}

int libgrpp_batch_full_grpp_gradients(int nshells, libgrpp_shell_t *const *shells, int ncenters, const double *xyz,
                                      libgrpp_grpp_t *const *grpps, int flags, double *out,
                                      libgrpp_batch_stats_t *stats)
{
    libgrpp_batch_plan_t *p = NULL;
    int rc = libgrpp_batch_gradients_plan_create(nshells, shells, ncenters, xyz, grpps, flags, &p);
    if (rc) return rc;
    rc = libgrpp_batch_plan_execute(p, NULL);
    if (!rc) rc = libgrpp_batch_plan_download_gradients(p, out);
    if (!rc && stats) libgrpp_batch_plan_stats(p, stats);
    libgrpp_batch_plan_destroy(p);
    return rc;
}
