/*
 * CPU oracle for the GRPP integral path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of libgrpp's algorithms for the GRPP path (type-1, type-2,
 * spin-orbit, outercore, full GRPP and their nuclear gradients).  It exists to check the
 * HIP engine; nothing in the product links, loads or calls it.  Only tests/, the smoke
 * check and the CPU-baseline leg of bench.py may use it.
 *
 * Every routine cites the reference file:line it restates (paths relative to the
 * libgrpp source tree).  Object layouts deliberately equal libgrpp.h:42-103 so the same
 * host-side description can be handed to the oracle, the genuine reference and the
 * product library.
 *
 * Pinning (how each part is checked, see tests/test_oracle_*.py):
 *   - the whole path against the reference-held golden matrices (H2O incl. gradients,
 *     Cs one-centre, ThO/UO2 subsets) and against the genuine reference compiled from its
 *     own sources (oracle/_ref) on seeded inputs;
 *   - special functions against closed forms / scipy / mpmath and the reference tables;
 *   - the quadrature against closed-form Gaussian moments.
 */
#ifndef ORC_H
#define ORC_H

#include <stddef.h>
#include <stdint.h>

#define ORC_ZERO_THRESH 1e-14      /* libgrpp.h:26 */

typedef struct {
    int L;
    int J;
    int num_primitives;
    int *powers;
    double *coeffs;
    double *alpha;
} orc_potential_t;                 /* libgrpp.h:42-49 */

typedef struct {
    int L;
    int cart_size;
    int *cart_list;
    int num_primitives;
    double *coeffs;
    double *alpha;
    double origin[3];
} orc_shell_t;                     /* libgrpp.h:66-74 */

typedef struct {
    int n_arep;
    int n_esop;
    int n_oc_shells;
    orc_potential_t *U_L;
    orc_potential_t **U_arep;
    orc_potential_t **U_esop;
    orc_potential_t **U_oc;
    orc_shell_t **oc_shells;
} orc_grpp_t;                      /* libgrpp.h:94-103 */

/* ---- radial quadrature knobs (the reference hard-codes tol=1e-16, screening on) ---- */
typedef struct {
    double tol;          /* absolute stop |S_k - S_{k-1}| <= tol (radial_type2_integral.c:188) */
    int screening;       /* 1: Shaw-Hill screening as in screening.c:193-248; 0: none          */
    int min_level;       /* never stop before this doubling level (0 = reference behaviour)   */
    int fixed_level;     /* >=0: evaluate exactly this level, no adaptivity (diagnostics)      */
    int engine_order;    /* diagnostics: bit 0 each level's new points in a fresh sum added to
                            S/2, bit 1 the point products as F1 (r^nA s) * F2 (r^nB sign s) with
                            s = sqrt|w U|, bit 2 new points summed in groups of four (0 = reference) */
} orc_radial_opts_t;

extern orc_radial_opts_t orc_radial_opts;

/* ---- math utilities (factorial.c, binomial.c, norm_gaussian.c, parameters.c) ---- */
double orc_factorial(int n);
double orc_factorial_ratio(int n, int m);
double orc_double_factorial(int n);
uint64_t orc_binomial(uint64_t n, uint64_t k);
double orc_gaussian_norm_factor(int n, int l, int m, double alpha);
int orc_cartesian_generator(int L, int *cart_list);
int orc_points_are_equal(const double *a, const double *b);
double orc_gaussian_integral(int n, double a);

/* ---- special functions ---- */
void orc_init(void);
double orc_bessel_scaled(int n, double x);          /* specfunc_scaled_mod_sph_bessel.c:67-122 */
double orc_bessel_table_value(int i, int lambda, int k);
double orc_boys(int n, double x);                   /* specfunc_boys.c:21-46 */
double orc_boys_table_value(int i, int n);
void orc_gfun_values(double x, int nmax, double *g);/* specfunc_gfun.c:40-65 */
double orc_dawson(double x);

/* ---- angular algebra ---- */
double orc_rsh_coef(int l, int m, int lx, int ly, int lz);     /* spherical_harmonics.c:115-166 */
void orc_rsh_values(int l, const double *k, double *out);      /* spherical_harmonics.c:227-280 */
double orc_angular_type2(int lambda, int L, int m, int a, int b, int c, const double *rsh_values);
double orc_unit_sphere_integral(int i, int j, int k);
void orc_lmatrix_rsh(int L, double *lx, double *ly, double *lz); /* lmatrix.c:28-84 */

/* ---- radial type-2 table ---- */
typedef struct {
    int lambda1_max, lambda2_max, n_max;
    double *Q;              /* index ((L2+1)(N+1))*l1 + (N+1)*l2 + n */
    int *level;             /* doubling level at which each entry stopped (-1 screened) */
} orc_radial_table_t;

orc_radial_table_t *orc_tabulate_radial(int l1max, int l2max, int nmax, double CA2, double CB2,
                                        const orc_potential_t *pot, const orc_shell_t *bra,
                                        const orc_shell_t *ket);
void orc_free_radial(orc_radial_table_t *t);

/* single contracted radial integral with explicit options (label generation / tests) */
double orc_radial_integral(int lambda1, int lambda2, int n, double CA, double CB,
                           const orc_potential_t *pot, const orc_shell_t *bra, const orc_shell_t *ket,
                           int *level_out);

/* ---- objects (shell.c, potential.c, grpp.c) ---- */
orc_shell_t *orc_new_shell(const double *origin, int L, int nprim, const double *coeffs, const double *alpha);
void orc_delete_shell(orc_shell_t *s);
orc_shell_t *orc_shell_copy(const orc_shell_t *s);
orc_potential_t *orc_new_potential(int L, int J, int nprim, const int *powers, const double *coeffs,
                                   const double *alpha);
void orc_delete_potential(orc_potential_t *p);
double orc_potential_value(const orc_potential_t *p, double r);
orc_grpp_t *orc_new_grpp(void);
void orc_grpp_set_local(orc_grpp_t *g, orc_potential_t *p);
void orc_grpp_add_arep(orc_grpp_t *g, orc_potential_t *p);
void orc_grpp_add_esop(orc_grpp_t *g, orc_potential_t *p);
void orc_grpp_add_oc(orc_grpp_t *g, orc_potential_t *p, orc_shell_t *s);
void orc_delete_grpp(orc_grpp_t *g);

/* ---- integrators (outputs zeroed then overwritten, row-major [i*ncB + j]) ---- */
void orc_overlap(const orc_shell_t *A, const orc_shell_t *B, double *S);
void orc_type1(const orc_shell_t *A, const orc_shell_t *B, const double *C, const orc_potential_t *pot, double *M);
void orc_type2(const orc_shell_t *A, const orc_shell_t *B, const double *C, const orc_potential_t *pot, double *M);
void orc_spin_orbit(const orc_shell_t *A, const orc_shell_t *B, const double *C, const orc_potential_t *pot,
                    double *X, double *Y, double *Z);
void orc_outercore(const orc_shell_t *A, const orc_shell_t *B, const double *C, int n_oc,
                   orc_potential_t *const *pots, orc_shell_t *const *shells,
                   double *arep, double *X, double *Y, double *Z);
void orc_full_grpp(const orc_shell_t *A, const orc_shell_t *B, const orc_grpp_t *g, const double *C,
                   double *arep, double *X, double *Y, double *Z);
void orc_full_grpp_gradient(const orc_shell_t *A, const orc_shell_t *B, const orc_grpp_t *g, const double *C,
                            const double *D, double **g_arep, double **g_x, double **g_y, double **g_z);

/* ---- whole-matrix builders (OpenMP over ordered shell pairs) ---- */
double orc_build_matrices(int nshell, orc_shell_t **shells, int ncenter, const double *xyz, orc_grpp_t **grpps,
                          double *arep, double *so_x, double *so_y, double *so_z, int nthreads,
                          int npsel, const int *pair_list);
double orc_build_gradients(int nshell, orc_shell_t **shells, int natom, const double *xyz, orc_grpp_t **grpps,
                           double *out, int nthreads);


/* nuclear attraction (orc_nuclear.c) */
void orc_type1_prim_n1(const orc_shell_t *A, double a, const orc_shell_t *B, double b, const double *C, double zeta,
                       double *out);
void orc_nuclear_attraction(const orc_shell_t *A, const orc_shell_t *B, const double *C, int charge, int model,
                            const double *params, double tol, double *M);
double orc_fermi_sk(int k, double x);
double orc_fermi_sk_table_value(int i, int col);
double orc_potential_point(double r, double Z);
double orc_potential_ball(double r, double Z, double R_rms);
double orc_potential_gaussian(double r, double Z, double R_rms);
double orc_potential_fermi(double r, double Z, double c, double a);
double orc_potential_fermi_bubble(double r, double Z, double c, double a, double k);
double orc_rms_johnson_1985(int A);
double orc_rms_golovko_2008(int A);
int orc_fermi_params(double R_rms, double *c, double *a);

#endif
