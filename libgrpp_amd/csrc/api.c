/*
 * libgrpp_amd C API: the libgrpp object model and integrators (libgrpp.h), implemented as host
 * code over the HIP engine (engine.h).  The per-call integrators evaluate one (A, B, C) task as
 * a one-item engine job; the batched entry points evaluate all pairs x centres at once.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "engine.h"
#include "libgrpp.h"

#define MAX_POTENTIALS 100   /* grpp.c:12 */

/* ---------------------------------------------------------------- parameters (parameters.c) */
static int generator_canonical(int L, int *cart_list)
{
    int c = 0;
    for (int r = L; r >= 0; r--)
        for (int s = L - r; s >= 0; s--) {
            cart_list[3 * c] = r;
            cart_list[3 * c + 1] = s;
            cart_list[3 * c + 2] = L - r - s;
            c++;
        }
    return c;
}

/* tolerances are kept for API compatibility; the reference never reads the last two
 * (parameters.c:17-27) and its type-2 quadrature hard-codes 1e-16 (radial_type2_integral.c:188) */
libgrpp_parameters_t libgrpp_params = {1e-16, 1e-16, 1e-16, generator_canonical};

void libgrpp_set_default_parameters(void)
{
    libgrpp_params.radial_tolerance = 1e-16;
    libgrpp_params.angular_screening_tolerance = 1e-16;
    libgrpp_params.modified_bessel_tolerance = 1e-16;
    libgrpp_params.cartesian_generator = generator_canonical;
}
void libgrpp_set_radial_tolerance(double t) { libgrpp_params.radial_tolerance = t; }
void libgrpp_set_angular_screening_tolerance(double t) { libgrpp_params.angular_screening_tolerance = t; }
void libgrpp_set_modified_bessel_tolerance(double t) { libgrpp_params.modified_bessel_tolerance = t; }
/* DIRAC and TURBOMOLE orders are the same list (parameters.c:94-134) */
void libgrpp_set_cartesian_order(int order)
{
    if (order != LIBGRPP_CART_ORDER_DIRAC && order != LIBGRPP_CART_ORDER_TURBOMOLE) {
        eng_set_error("libgrpp_set_cartesian_order: unknown order %d", order);
        abort();
    }
    libgrpp_params.cartesian_generator = generator_canonical;
}
void libgrpp_set_cartesian_generator(int (*g)(int L, int *cart_list)) { libgrpp_params.cartesian_generator = g; }

/* ---------------------------------------------------------------- init (init_finalize.c) */
static int g_initialized = 0;
void libgrpp_init(void)
{
    /* on a GPU host the current device is set up here (constant tables, scratch address), so that
       later calls, the allocation-free one included, find them in place; without a GPU nothing */
    if (eng_setup_device()) fprintf(stderr, "libgrpp_amd: libgrpp_init: %s\n", libgrpp_last_error());
    g_initialized = 1;
}
void libgrpp_finalize(void) { g_initialized = 0; }
int libgrpp_is_initialized(void) { return g_initialized; }

static void require_init(const char *who)
{
    if (!g_initialized) {
        eng_set_error("%s: libgrpp_init() has not been called", who);
        fprintf(stderr, "libgrpp_amd: %s\n", libgrpp_last_error());
        abort();   /* the reference asserts (type2_integrals.c:42) */
    }
}

static void fail(const char *who)
{
    fprintf(stderr, "libgrpp_amd: %s: %s\n", who, libgrpp_last_error());
    abort();
}

/* ---------------------------------------------------------------- potentials (potential.c) */
libgrpp_potential_t *libgrpp_new_potential(int L, int J, int n, int *powers, double *coeffs, double *alpha)
{
    libgrpp_potential_t *p = (libgrpp_potential_t *) calloc(1, sizeof(*p));
    p->L = L;
    p->J = J;
    p->powers = (int *) calloc(n > 0 ? n : 1, sizeof(int));
    p->coeffs = (double *) calloc(n > 0 ? n : 1, sizeof(double));
    p->alpha = (double *) calloc(n > 0 ? n : 1, sizeof(double));
    for (int i = 0; i < n; i++) {
        if (fabs(coeffs[i]) < LIBGRPP_ZERO_THRESH) continue;   /* potential.c:34 */
        p->powers[p->num_primitives] = powers[i];
        p->coeffs[p->num_primitives] = coeffs[i];
        p->alpha[p->num_primitives] = alpha[i];
        p->num_primitives++;
    }
    return p;
}

libgrpp_potential_t *libgrpp_create_potential(int L, int J, int n, int *powers, double *coeffs, double *alpha)
{
    return libgrpp_new_potential(L, J, n, powers, coeffs, alpha);
}

void libgrpp_delete_potential(libgrpp_potential_t *p)
{
    if (!p) return;
    free(p->powers);
    free(p->coeffs);
    free(p->alpha);
    free(p);
}

double libgrpp_potential_value(libgrpp_potential_t *p, double r)
{
    double v = 0.0, r2 = r * r;
    for (int i = 0; i < p->num_primitives; i++) v += p->coeffs[i] * pow(r, p->powers[i] - 2) * exp(-p->alpha[i] * r2);
    return v;
}

libgrpp_potential_t *libgrpp_shrink_potential(libgrpp_potential_t *src)
{
    /* a copy without |c| <= 1e-14 terms (potential.c:88-115) */
    int n = 0;
    int *pw = (int *) calloc(src->num_primitives + 1, sizeof(int));
    double *c = (double *) calloc(src->num_primitives + 1, sizeof(double));
    double *a = (double *) calloc(src->num_primitives + 1, sizeof(double));
    for (int i = 0; i < src->num_primitives; i++)
        if (fabs(src->coeffs[i]) > LIBGRPP_ZERO_THRESH) {
            pw[n] = src->powers[i];
            c[n] = src->coeffs[i];
            a[n] = src->alpha[i];
            n++;
        }
    libgrpp_potential_t *p = libgrpp_new_potential(src->L, src->J, n, pw, c, a);
    free(pw);
    free(c);
    free(a);
    return p;
}

/* ---------------------------------------------------------------- shells (shell.c) */
int *libgrpp_generate_shell_cartesians(int L)
{
    int n = (L + 1) * (L + 2) / 2;
    int *cl = (int *) calloc(3 * n, sizeof(int));
    libgrpp_params.cartesian_generator(L, cl);
    return cl;
}

libgrpp_shell_t *libgrpp_new_shell(double *origin, int L, int n, double *coeffs, double *alpha)
{
    libgrpp_shell_t *s = (libgrpp_shell_t *) calloc(1, sizeof(*s));
    s->L = L;
    s->origin[0] = origin[0];
    s->origin[1] = origin[1];
    s->origin[2] = origin[2];
    s->cart_size = (L + 1) * (L + 2) / 2;
    s->cart_list = libgrpp_generate_shell_cartesians(L);
    s->num_primitives = n;
    s->coeffs = (double *) calloc(n > 0 ? n : 1, sizeof(double));
    s->alpha = (double *) calloc(n > 0 ? n : 1, sizeof(double));
    for (int i = 0; i < n; i++) {
        s->coeffs[i] = coeffs[i];
        s->alpha[i] = alpha[i];
    }
    return s;
}

libgrpp_shell_t *libgrpp_create_shell(double *origin, int L, int n, double *coeffs, double *alpha)
{
    return libgrpp_new_shell(origin, L, n, coeffs, alpha);
}

libgrpp_shell_t *libgrpp_shell_deep_copy(libgrpp_shell_t *s)
{
    return libgrpp_new_shell(s->origin, s->L, s->num_primitives, s->coeffs, s->alpha);
}

void libgrpp_shell_shrink(libgrpp_shell_t *s)
{
    int n = 0;
    for (int i = 0; i < s->num_primitives; i++)
        if (fabs(s->coeffs[i]) > LIBGRPP_ZERO_THRESH) {
            s->coeffs[n] = s->coeffs[i];
            s->alpha[n] = s->alpha[i];
            n++;
        }
    s->num_primitives = n;
}

void libgrpp_shell_mult_normcoef(libgrpp_shell_t *s)
{
    for (int i = 0; i < s->num_primitives; i++)
        s->coeffs[i] *= pow(2 * s->alpha[i] / M_PI, 0.75) * pow(4 * s->alpha[i], 0.5 * s->L);
}

int libgrpp_get_shell_size(libgrpp_shell_t *s) { return s->cart_size; }

void libgrpp_delete_shell(libgrpp_shell_t *s)
{
    if (!s) return;
    free(s->cart_list);
    free(s->coeffs);
    free(s->alpha);
    free(s);
}

/* ---------------------------------------------------------------- GRPP operator (grpp.c) */
libgrpp_grpp_t *libgrpp_new_grpp(void) { return (libgrpp_grpp_t *) calloc(1, sizeof(libgrpp_grpp_t)); }

void libgrpp_grpp_set_local_potential(libgrpp_grpp_t *g, libgrpp_potential_t *p)
{
    if (g->U_L) libgrpp_delete_potential(g->U_L);
    g->U_L = p;
}

void libgrpp_grpp_add_averaged_potential(libgrpp_grpp_t *g, libgrpp_potential_t *p)
{
    if (g->n_arep == 0) g->U_arep = (libgrpp_potential_t **) calloc(MAX_POTENTIALS, sizeof(void *));
    if (g->n_arep >= MAX_POTENTIALS) { eng_set_error("too many AREP channels"); fail("libgrpp_grpp_add_averaged_potential"); }
    g->U_arep[g->n_arep++] = p;
}

void libgrpp_grpp_add_spin_orbit_potential(libgrpp_grpp_t *g, libgrpp_potential_t *p)
{
    if (g->n_esop == 0) g->U_esop = (libgrpp_potential_t **) calloc(MAX_POTENTIALS, sizeof(void *));
    if (g->n_esop >= MAX_POTENTIALS) { eng_set_error("too many SO channels"); fail("libgrpp_grpp_add_spin_orbit_potential"); }
    g->U_esop[g->n_esop++] = p;
}

void libgrpp_grpp_add_outercore_potential(libgrpp_grpp_t *g, libgrpp_potential_t *p, libgrpp_shell_t *s)
{
    if (g->n_oc_shells == 0) {
        g->U_oc = (libgrpp_potential_t **) calloc(MAX_POTENTIALS, sizeof(void *));
        g->oc_shells = (libgrpp_shell_t **) calloc(MAX_POTENTIALS, sizeof(void *));
    }
    if (g->n_oc_shells >= MAX_POTENTIALS) { eng_set_error("too many outercore shells"); fail("libgrpp_grpp_add_outercore_potential"); }
    g->U_oc[g->n_oc_shells] = p;
    g->oc_shells[g->n_oc_shells] = s;
    g->n_oc_shells++;
}

void libgrpp_delete_grpp(libgrpp_grpp_t *g)
{
    if (!g) return;
    libgrpp_delete_potential(g->U_L);
    for (int i = 0; i < g->n_arep; i++) libgrpp_delete_potential(g->U_arep[i]);
    for (int i = 0; i < g->n_esop; i++) libgrpp_delete_potential(g->U_esop[i]);
    for (int i = 0; i < g->n_oc_shells; i++) {
        libgrpp_delete_potential(g->U_oc[i]);
        libgrpp_delete_shell(g->oc_shells[i]);
    }
    free(g->U_arep);
    free(g->U_esop);
    free(g->U_oc);
    free(g->oc_shells);
    free(g);
}

/* ---------------------------------------------------------------- helpers over the engine */
/* canonical position of (n,l,m) in the x-descending / y-descending order */
static int canon_index(int n, int l, int m)
{
    int L = n + l + m;
    return (L - n) * (L - n + 1) / 2 + (L - n - l);
}

/* copies engine block (canonical Cartesian order, row-major ncA x ncB) into the caller's order */
/* the caller's components (cart_list, possibly fewer than the shell's (L+1)(L+2)/2) of a block the
 * engine computed over all components of both shells in canonical order */
static void to_shell_order(const libgrpp_shell_t *A, const libgrpp_shell_t *B, const double *src, double *dst)
{
    int nA = A->cart_size, nB = B->cart_size, nBf = (B->L + 1) * (B->L + 2) / 2;
    for (int i = 0; i < nA; i++) {
        int ci = canon_index(A->cart_list[3 * i], A->cart_list[3 * i + 1], A->cart_list[3 * i + 2]);
        for (int j = 0; j < nB; j++) {
            int cj = canon_index(B->cart_list[3 * j], B->cart_list[3 * j + 1], B->cart_list[3 * j + 2]);
            dst[i * nB + j] = src[ci * nBf + cj];
        }
    }
}

static void center_from_grpp(eng_center_t *c, const double *origin, const libgrpp_grpp_t *g)
{
    memset(c, 0, sizeof(*c));
    c->origin = origin;
    c->local = g->U_L;
    c->n_arep = g->n_arep;
    c->arep = g->U_arep;
    c->n_so = g->n_esop;
    c->so = g->U_esop;
    c->so_scale = NULL;
    c->n_oc = g->n_oc_shells;
    c->oc_pot = g->U_oc;
    c->oc_shell = g->oc_shells;
    c->shell_begin = 0;
    c->shell_end = -1;
}

/* one (A, B, C) task: 4 blocks [arep, so_x, so_y, so_z] in the caller's Cartesian order */
static void run_task(const char *who, libgrpp_shell_t *A, libgrpp_shell_t *B, const eng_center_t *c, int parts,
                     double *out4)
{
    libgrpp_shell_t *sh[2] = {A, B};
    /* the engine evaluates every Cartesian component of both shells, whatever their cart lists */
    int nAf = (A->L + 1) * (A->L + 2) / 2, nBf = (B->L + 1) * (B->L + 2) / 2, szf = nAf * nBf;
    int sz = A->cart_size * B->cart_size;
    eng_item_t it = {0, 1, 0, 1, 0, 0, nBf, 0};
    int cidx = 0;
    double *buf = (double *) malloc(sizeof(double) * 4 * szf);
    if (eng_run(2, sh, 1, c, 1, &cidx, 1, &it, szf, szf, parts, buf)) {
        free(buf);
        fail(who);
    }
    for (int op = 0; op < 4; op++) to_shell_order(A, B, buf + op * szf, out4 + op * sz);
    free(buf);
}

/* ---------------------------------------------------------------- integrators */
void libgrpp_type1_integrals(libgrpp_shell_t *A, libgrpp_shell_t *B, double *C, libgrpp_potential_t *pot,
                             double *matrix)
{
    require_init("libgrpp_type1_integrals");
    int sz = A->cart_size * B->cart_size;
    memset(matrix, 0, sizeof(double) * sz);
    if (pot == NULL) return;   /* type1_integrals.c:57-59 */
    eng_center_t c;
    memset(&c, 0, sizeof(c));
    c.origin = C;
    c.local = pot;
    c.shell_end = -1;
    double *buf = (double *) malloc(sizeof(double) * 4 * sz);
    run_task("libgrpp_type1_integrals", A, B, &c, ENG_TYPE1, buf);
    memcpy(matrix, buf, sizeof(double) * sz);
    free(buf);
}

/* ---------------------------------------------------------------- nuclear attraction */
/* nuclear_attraction.c:54-133: the point charge through the type-1 kernels (a -Z r^-1 term of
 * exponent 0), the finite models (and the numeric point charge) through the numeric radial
 * quadrature on the device */
void libgrpp_nuclear_attraction_integrals(libgrpp_shell_t *A, libgrpp_shell_t *B, double *C, int charge, int model,
                                          double *params, double *matrix)
{
    require_init("libgrpp_nuclear_attraction_integrals");
    const int sz = A->cart_size * B->cart_size;
    memset(matrix, 0, sizeof(double) * sz);
    if (model == LIBGRPP_NUCLEAR_MODEL_POINT_CHARGE) {
        int pw = 1;
        double cf = -(double) charge, ze = 0.0;
        libgrpp_potential_t *u = libgrpp_new_potential(0, 0, 1, &pw, &cf, &ze);
        eng_center_t c;
        memset(&c, 0, sizeof(c));
        c.origin = C;
        c.local = u;
        c.shell_end = -1;
        double *buf = (double *) malloc(sizeof(double) * 4 * sz);
        run_task("libgrpp_nuclear_attraction_integrals", A, B, &c, ENG_TYPE1, buf);
        memcpy(matrix, buf, sizeof(double) * sz);
        free(buf);
        libgrpp_delete_potential(u);
        return;
    }
    if (model < LIBGRPP_NUCLEAR_MODEL_CHARGED_BALL || model > LIBGRPP_NUCLEAR_MODEL_POINT_CHARGE_NUMERICAL) {
        eng_set_error("libgrpp_nuclear_attraction_integrals: unknown nuclear model %d", model);
        fail("libgrpp_nuclear_attraction_integrals");
    }
    libgrpp_shell_t *sh[2] = {A, B};
    const int nAf = (A->L + 1) * (A->L + 2) / 2, nBf = (B->L + 1) * (B->L + 2) / 2;
    eng_item_t it = {0, 1, 0, 1, 0, 0, nBf, 0};
    double p3[3] = {0.0, 0.0, 0.0};
    const int np = model == LIBGRPP_NUCLEAR_MODEL_CHARGED_BALL || model == LIBGRPP_NUCLEAR_MODEL_GAUSSIAN ? 1
                   : model == LIBGRPP_NUCLEAR_MODEL_FERMI ? 2 : model == LIBGRPP_NUCLEAR_MODEL_FERMI_BUBBLE ? 3 : 0;
    for (int i = 0; i < np; i++) p3[i] = params[i];
    double *buf = (double *) malloc(sizeof(double) * nAf * nBf);
    if (eng_nuclear(2, sh, 1, C, &charge, model, p3, libgrpp_params.radial_tolerance, 1, &it, (long long) nAf * nBf,
                    buf)) {
        free(buf);
        fail("libgrpp_nuclear_attraction_integrals");
    }
    to_shell_order(A, B, buf, matrix);
    free(buf);
}

void libgrpp_nuclear_attraction_integrals_point_charge(libgrpp_shell_t *A, libgrpp_shell_t *B, double *C, int charge,
                                                       double *matrix)
{
    libgrpp_nuclear_attraction_integrals(A, B, C, charge, LIBGRPP_NUCLEAR_MODEL_POINT_CHARGE, NULL, matrix);
}

void libgrpp_nuclear_attraction_integrals_charged_ball(libgrpp_shell_t *A, libgrpp_shell_t *B, double *C, int charge,
                                                       double r_rms, double *matrix)
{
    libgrpp_nuclear_attraction_integrals(A, B, C, charge, LIBGRPP_NUCLEAR_MODEL_CHARGED_BALL, &r_rms, matrix);
}

void libgrpp_nuclear_attraction_integrals_gaussian_model(libgrpp_shell_t *A, libgrpp_shell_t *B, double *C, int charge,
                                                         double r_rms, double *matrix)
{
    libgrpp_nuclear_attraction_integrals(A, B, C, charge, LIBGRPP_NUCLEAR_MODEL_GAUSSIAN, &r_rms, matrix);
}

void libgrpp_nuclear_attraction_integrals_fermi_model(libgrpp_shell_t *A, libgrpp_shell_t *B, double *C, int charge,
                                                      double c, double a, double *matrix)
{
    double p[2] = {c, a};
    libgrpp_nuclear_attraction_integrals(A, B, C, charge, LIBGRPP_NUCLEAR_MODEL_FERMI, p, matrix);
}

void libgrpp_nuclear_attraction_integrals_fermi_bubble_model(libgrpp_shell_t *A, libgrpp_shell_t *B, double *C,
                                                             int charge, double c, double a, double k, double *matrix)
{
    double p[3] = {c, a, k};
    libgrpp_nuclear_attraction_integrals(A, B, C, charge, LIBGRPP_NUCLEAR_MODEL_FERMI_BUBBLE, p, matrix);
}

void libgrpp_type2_integrals(libgrpp_shell_t *A, libgrpp_shell_t *B, double *C, libgrpp_potential_t *pot,
                             double *matrix)
{
    require_init("libgrpp_type2_integrals");
    int sz = A->cart_size * B->cart_size;
    eng_center_t c;
    memset(&c, 0, sizeof(c));
    c.origin = C;
    c.n_arep = 1;
    c.arep = &pot;
    c.shell_end = -1;
    double *buf = (double *) malloc(sizeof(double) * 4 * sz);
    run_task("libgrpp_type2_integrals", A, B, &c, ENG_TYPE2, buf);
    memcpy(matrix, buf, sizeof(double) * sz);
    free(buf);
}

void libgrpp_spin_orbit_integrals(libgrpp_shell_t *A, libgrpp_shell_t *B, double *C, libgrpp_potential_t *pot,
                                  double *so_x, double *so_y, double *so_z)
{
    require_init("libgrpp_spin_orbit_integrals");
    int sz = A->cart_size * B->cart_size;
    const double one = 1.0;   /* this entry point excludes 2/(2L+1) (spin_orbit_integrals.c:233-235) */
    eng_center_t c;
    memset(&c, 0, sizeof(c));
    c.origin = C;
    c.n_so = 1;
    c.so = &pot;
    c.so_scale = &one;
    c.shell_end = -1;
    double *buf = (double *) malloc(sizeof(double) * 4 * sz);
    run_task("libgrpp_spin_orbit_integrals", A, B, &c, ENG_SO, buf);
    memcpy(so_x, buf + sz, sizeof(double) * sz);
    memcpy(so_y, buf + 2 * sz, sizeof(double) * sz);
    memcpy(so_z, buf + 3 * sz, sizeof(double) * sz);
    free(buf);
}

void libgrpp_outercore_potential_integrals(libgrpp_shell_t *A, libgrpp_shell_t *B, double *C, int n_oc,
                                           libgrpp_potential_t **pots, libgrpp_shell_t **shells, double *arep,
                                           double *esop_x, double *esop_y, double *esop_z)
{
    require_init("libgrpp_outercore_potential_integrals");
    int sz = A->cart_size * B->cart_size;
    eng_center_t c;
    memset(&c, 0, sizeof(c));
    c.origin = C;
    c.n_oc = n_oc;
    c.oc_pot = pots;
    c.oc_shell = shells;
    c.shell_end = -1;
    double *buf = (double *) malloc(sizeof(double) * 4 * sz);
    run_task("libgrpp_outercore_potential_integrals", A, B, &c, ENG_OC, buf);
    memcpy(arep, buf, sizeof(double) * sz);
    memcpy(esop_x, buf + sz, sizeof(double) * sz);
    memcpy(esop_y, buf + 2 * sz, sizeof(double) * sz);
    memcpy(esop_z, buf + 3 * sz, sizeof(double) * sz);
    free(buf);
}

/* one outercore part of one task, ADDED to the four outputs as the reference's parts accumulate */
static void oc_part(const char *who, libgrpp_shell_t *A, libgrpp_shell_t *B, double *C, int n_oc,
                    libgrpp_potential_t **pots, libgrpp_shell_t **shells, int parts, double *arep, double *so_x,
                    double *so_y, double *so_z)
{
    require_init(who);
    int sz = A->cart_size * B->cart_size;
    eng_center_t c;
    memset(&c, 0, sizeof(c));
    c.origin = C;
    c.n_oc = n_oc;
    c.oc_pot = pots;
    c.oc_shell = shells;
    c.shell_end = -1;
    c.oc_i = 0;   /* part 2: the (first, second) term only */
    c.oc_j = 1;
    double *buf = (double *) malloc(sizeof(double) * 4 * sz);
    run_task(who, A, B, &c, parts, buf);
    double *outs[4] = {arep, so_x, so_y, so_z};
    for (int op = 0; op < 4; op++)
        for (int e = 0; e < sz; e++) outs[op][e] += buf[op * sz + e];
    free(buf);
}

/* U|nlj><nlj| + |nlj><nlj|U of one outercore entry (outercore_integrals.c:147-270) */
void libgrpp_outercore_potential_integrals_part_1(libgrpp_shell_t *A, libgrpp_shell_t *B, double *C,
                                                  libgrpp_potential_t *oc_potential, libgrpp_shell_t *oc_shell,
                                                  double *arep, double *so_x, double *so_y, double *so_z)
{
    oc_part("libgrpp_outercore_potential_integrals_part_1", A, B, C, 1, &oc_potential, &oc_shell, ENG_OC_P1, arep,
            so_x, so_y, so_z);
}

/* -|nlj><nlj| (U + U')/2 |n'lj><n'lj| of one ordered pair of outercore entries, zero unless their
 * (L, J) agree (outercore_integrals.c:277-390) */
void libgrpp_outercore_potential_integrals_part_2(libgrpp_shell_t *A, libgrpp_shell_t *B, double *C,
                                                  libgrpp_potential_t *oc_potential_1, libgrpp_shell_t *oc_shell_1,
                                                  libgrpp_potential_t *oc_potential_2, libgrpp_shell_t *oc_shell_2,
                                                  double *arep, double *so_x, double *so_y, double *so_z)
{
    libgrpp_potential_t *pots[2] = {oc_potential_1, oc_potential_2};
    libgrpp_shell_t *shells[2] = {oc_shell_1, oc_shell_2};
    oc_part("libgrpp_outercore_potential_integrals_part_2", A, B, C, 2, pots, shells, ENG_OC_P2, arep, so_x, so_y,
            so_z);
}

void libgrpp_full_grpp_integrals(libgrpp_shell_t *A, libgrpp_shell_t *B, libgrpp_grpp_t *g, double *C,
                                 double *arep, double *so_x, double *so_y, double *so_z)
{
    require_init("libgrpp_full_grpp_integrals");
    int sz = A->cart_size * B->cart_size;
    eng_center_t c;
    center_from_grpp(&c, C, g);
    double *buf = (double *) malloc(sizeof(double) * 4 * sz);
    run_task("libgrpp_full_grpp_integrals", A, B, &c, ENG_ALL, buf);
    memcpy(arep, buf, sizeof(double) * sz);
    memcpy(so_x, buf + sz, sizeof(double) * sz);
    memcpy(so_y, buf + 2 * sz, sizeof(double) * sz);
    memcpy(so_z, buf + 3 * sz, sizeof(double) * sz);
    free(buf);
}

void libgrpp_overlap_integrals(libgrpp_shell_t *A, libgrpp_shell_t *B, double *S)
{
    /* <A|B> as the type-1 integral of exp(0 r^2) r^0 at any centre (the n = 2 McMurchie-Davidson
     * kernel with a zero exponent reduces to the Obara-Saika overlap of overlap.c:39-138) */
    int one = 2;
    double c1 = 1.0, z0 = 0.0;
    libgrpp_potential_t pot = {0, 0, 1, &one, &c1, &z0};
    int was = g_initialized;
    g_initialized = 1;
    libgrpp_type1_integrals(A, B, A->origin, &pot, S);
    g_initialized = was;
}
