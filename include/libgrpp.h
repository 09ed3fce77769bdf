/*
 * libgrpp_amd -- MI355X-native GRPP integral engine, drop-in for libgrpp's C API.
 *
 * Every declaration below keeps the name, signature, object layout and semantics of the
 * reference header (libgrpp/libgrpp.h, libgrpp/parameters.h); the comment on each names the
 * reference interface it replaces.  Integrals are evaluated by HIP kernels on the current
 * device (gfx950); if no device or kernel image is usable, the integrators report the error
 * through libgrpp_last_error() and abort, as the reference does on failed assertions --
 * there is no CPU fallback.
 *
 * Output conventions (as the reference): caller-allocated row-major blocks
 * [i_cart_A * ncart_B + j_cart_B], zeroed then overwritten; Cartesian order x-power
 * descending, then y-power descending.
 */
#ifndef LIBGRPP_AMD_LIBGRPP_H
#define LIBGRPP_AMD_LIBGRPP_H

#ifdef __cplusplus
extern "C" {
#endif

#define LIBGRPP_MAX_BASIS_L 10      /* libgrpp.h:16 */
#define LIBGRPP_MAX_RPP_L 10        /* libgrpp.h:21 */
#define LIBGRPP_ZERO_THRESH 1e-14   /* libgrpp.h:26 */
#define LIBGRPP_RADIAL_TOL 1e-14    /* libgrpp.h:31 */

/* ---------------- parameters (parameters.h:11-35) ---------------- */
enum { LIBGRPP_CART_ORDER_DIRAC, LIBGRPP_CART_ORDER_TURBOMOLE };

typedef struct {
    double radial_tolerance;
    double angular_screening_tolerance;
    double modified_bessel_tolerance;
    int (*cartesian_generator)(int L, int *cart_list);
} libgrpp_parameters_t;

extern libgrpp_parameters_t libgrpp_params;                               /* parameters.h:23 */
void libgrpp_set_default_parameters(void);                                 /* parameters.h:25 */
void libgrpp_set_radial_tolerance(double tolerance);                       /* parameters.h:27 */
void libgrpp_set_angular_screening_tolerance(double tolerance);            /* parameters.h:29 */
void libgrpp_set_modified_bessel_tolerance(double tolerance);              /* parameters.h:31 */
void libgrpp_set_cartesian_order(int order);                               /* parameters.h:33 */
void libgrpp_set_cartesian_generator(int (*generator)(int L, int *cart_list)); /* parameters.h:35 */

/* ---------------- library state (libgrpp.h:33-37, init_finalize.c:19-50) ---------------- */
void libgrpp_init(void);
void libgrpp_finalize(void);
int libgrpp_is_initialized(void);

/* ---------------- potentials (libgrpp.h:42-58, potential.c) ---------------- */
typedef struct {
    int L;
    int J;
    int num_primitives;
    int *powers;
    double *coeffs;
    double *alpha;
} libgrpp_potential_t;

libgrpp_potential_t *libgrpp_new_potential(int L, int J, int num_primitives, int *powers, double *coeffs,
                                           double *alpha);                 /* libgrpp.h:51-52 */
void libgrpp_delete_potential(libgrpp_potential_t *potential);            /* libgrpp.h:54 */
double libgrpp_potential_value(libgrpp_potential_t *potential, double r); /* libgrpp.h:56 */
libgrpp_potential_t *libgrpp_shrink_potential(libgrpp_potential_t *src);  /* libgrpp.h:58 */

/* ---------------- shells (libgrpp.h:66-88, shell.c) ---------------- */
typedef struct {
    int L;
    int cart_size;
    int *cart_list;
    int num_primitives;
    double *coeffs;
    double *alpha;
    double origin[3];
} libgrpp_shell_t;

libgrpp_shell_t *libgrpp_new_shell(double *origin, int L, int num_primitives, double *coeffs,
                                   double *alpha);                         /* libgrpp.h:76 */
int libgrpp_get_shell_size(libgrpp_shell_t *shell);                        /* libgrpp.h:78 */
void libgrpp_delete_shell(libgrpp_shell_t *shell);                         /* libgrpp.h:80 */
int *libgrpp_generate_shell_cartesians(int L);                             /* libgrpp.h:82 */
libgrpp_shell_t *libgrpp_shell_deep_copy(libgrpp_shell_t *src_shell);      /* libgrpp.h:84 */
void libgrpp_shell_shrink(libgrpp_shell_t *shell);                         /* libgrpp.h:86 */
void libgrpp_shell_mult_normcoef(libgrpp_shell_t *shell);                  /* libgrpp.h:88 */

/* aliases under the names the project brief uses for the two constructors */
libgrpp_shell_t *libgrpp_create_shell(double *origin, int L, int num_primitives, double *coeffs, double *alpha);
libgrpp_potential_t *libgrpp_create_potential(int L, int J, int num_primitives, int *powers, double *coeffs,
                                              double *alpha);

/* ---------------- whole GRPP operator (libgrpp.h:94-116, grpp.c) ---------------- */
typedef struct {
    int n_arep;
    int n_esop;
    int n_oc_shells;
    libgrpp_potential_t *U_L;
    libgrpp_potential_t **U_arep;
    libgrpp_potential_t **U_esop;
    libgrpp_potential_t **U_oc;
    libgrpp_shell_t **oc_shells;
} libgrpp_grpp_t;

libgrpp_grpp_t *libgrpp_new_grpp(void);                                                           /* :106 */
void libgrpp_grpp_set_local_potential(libgrpp_grpp_t *grpp, libgrpp_potential_t *pot);            /* :108 */
void libgrpp_grpp_add_averaged_potential(libgrpp_grpp_t *grpp, libgrpp_potential_t *pot);         /* :110 */
void libgrpp_grpp_add_spin_orbit_potential(libgrpp_grpp_t *grpp, libgrpp_potential_t *pot);       /* :112 */
void libgrpp_grpp_add_outercore_potential(libgrpp_grpp_t *grpp, libgrpp_potential_t *pot,
                                          libgrpp_shell_t *oc_shell);                             /* :114 */
void libgrpp_delete_grpp(libgrpp_grpp_t *grpp);                                                   /* :116 */

/* ---------------- integrators (libgrpp.h:123-171) ---------------- */
void libgrpp_type1_integrals(libgrpp_shell_t *shell_A, libgrpp_shell_t *shell_B, double *rpp_origin,
                             libgrpp_potential_t *potential, double *matrix);           /* type1_integrals.c:43 */
void libgrpp_type2_integrals(libgrpp_shell_t *shell_A, libgrpp_shell_t *shell_B, double *rpp_origin,
                             libgrpp_potential_t *potential, double *matrix);           /* type2_integrals.c:34 */
void libgrpp_spin_orbit_integrals(libgrpp_shell_t *shell_A, libgrpp_shell_t *shell_B, double *rpp_origin,
                                  libgrpp_potential_t *potential, double *so_x_matrix, double *so_y_matrix,
                                  double *so_z_matrix);                                 /* spin_orbit_integrals.c:41 */
void libgrpp_outercore_potential_integrals(libgrpp_shell_t *shell_A, libgrpp_shell_t *shell_B,
                                           double *rpp_origin, int num_oc_shells,
                                           libgrpp_potential_t **oc_potentials, libgrpp_shell_t **oc_shells,
                                           double *arep, double *esop_x, double *esop_y,
                                           double *esop_z);                             /* outercore_integrals.c:83 */
/* the two parts of the outercore operator, each ADDED to the outputs (extern in the reference library,
 * declared by its Fortran layer, fortran.c:263-267 and :367-372) */
void libgrpp_outercore_potential_integrals_part_1(libgrpp_shell_t *shell_A, libgrpp_shell_t *shell_B, double *C,
                                                  libgrpp_potential_t *oc_potential, libgrpp_shell_t *oc_shell,
                                                  double *arep_matrix, double *so_x_matrix, double *so_y_matrix,
                                                  double *so_z_matrix);              /* outercore_integrals.c:147 */
void libgrpp_outercore_potential_integrals_part_2(libgrpp_shell_t *shell_A, libgrpp_shell_t *shell_B, double *C,
                                                  libgrpp_potential_t *oc_potential_1, libgrpp_shell_t *oc_shell_1,
                                                  libgrpp_potential_t *oc_potential_2, libgrpp_shell_t *oc_shell_2,
                                                  double *arep_matrix, double *so_x_matrix, double *so_y_matrix,
                                                  double *so_z_matrix);              /* outercore_integrals.c:277 */
void libgrpp_full_grpp_integrals(libgrpp_shell_t *shell_A, libgrpp_shell_t *shell_B,
                                 libgrpp_grpp_t *grpp_operator, double *grpp_origin, double *arep_matrix,
                                 double *so_x_matrix, double *so_y_matrix,
                                 double *so_z_matrix);                                  /* full_grpp_integrals.c:27 */

/* ---------------- gradients (libgrpp.h:178-231) ---------------- */
void libgrpp_type1_integrals_gradient(libgrpp_shell_t *shell_A, libgrpp_shell_t *shell_B, double *grpp_origin,
                                      libgrpp_potential_t *potential, double *point_3d,
                                      double **grad_arep);                     /* grpp_integrals_gradient.c:123 */
void libgrpp_type2_integrals_gradient(libgrpp_shell_t *shell_A, libgrpp_shell_t *shell_B, double *grpp_origin,
                                      libgrpp_potential_t *potential, double *point_3d,
                                      double **grad_arep);                     /* grpp_integrals_gradient.c:163 */
void libgrpp_spin_orbit_integrals_gradient(libgrpp_shell_t *shell_A, libgrpp_shell_t *shell_B,
                                           double *grpp_origin, libgrpp_potential_t *potential,
                                           double *point_3d, double **grad_so_x, double **grad_so_y,
                                           double **grad_so_z);                /* grpp_integrals_gradient.c:203 */
void libgrpp_outercore_potential_integrals_gradient(libgrpp_shell_t *shell_A, libgrpp_shell_t *shell_B,
                                                    double *rpp_origin, int num_oc_shells,
                                                    libgrpp_potential_t **oc_potentials,
                                                    libgrpp_shell_t **oc_shells, double *point_3d,
                                                    double **grad_arep, double **grad_so_x, double **grad_so_y,
                                                    double **grad_so_z);       /* grpp_integrals_gradient.c:248 */
void libgrpp_full_grpp_integrals_gradient(libgrpp_shell_t *shell_A, libgrpp_shell_t *shell_B,
                                          libgrpp_grpp_t *grpp_operator, double *grpp_origin, double *point_3d,
                                          double **grad_arep, double **grad_so_x, double **grad_so_y,
                                          double **grad_so_z);                 /* grpp_integrals_gradient.c:280 */

/* overlap of two contracted shells (overlap.h / overlap.c:39), used by hosts for normalisation */
void libgrpp_overlap_integrals(libgrpp_shell_t *shell_A, libgrpp_shell_t *shell_B, double *overlap_matrix);

/*
 * Nuclear attraction of a shell pair for a nucleus of charge `charge` at charge_origin
 * (libgrpp.h:244-316, nuclear_attraction.c:54-256).  Models: the point charge, analytic
 * (type-1 McMurchie-Davidson with an r^-1 term); a charged ball (model_params[0] = R_rms), a
 * Gaussian distribution (R_rms), a Fermi distribution (c, a), Fermi with a quadratic "bubble" term
 * (c, a, k) and the point charge again by quadrature: numeric radial type-1 integrals on the Log3
 * grid (radial_type1_integral.c), shells with l <= 6.  Lengths in bohr.  Output as the other
 * integrators (overwritten, row-major).
 */
enum {
    LIBGRPP_NUCLEAR_MODEL_POINT_CHARGE = 0,
    LIBGRPP_NUCLEAR_MODEL_CHARGED_BALL,
    LIBGRPP_NUCLEAR_MODEL_GAUSSIAN,
    LIBGRPP_NUCLEAR_MODEL_FERMI,
    LIBGRPP_NUCLEAR_MODEL_FERMI_BUBBLE,
    LIBGRPP_NUCLEAR_MODEL_POINT_CHARGE_NUMERICAL
};
void libgrpp_nuclear_attraction_integrals(libgrpp_shell_t *shell_A, libgrpp_shell_t *shell_B, double *charge_origin,
                                          int charge, int nuclear_model, double *model_params,
                                          double *coulomb_matrix);                        /* libgrpp.h:256 */
void libgrpp_nuclear_attraction_integrals_point_charge(libgrpp_shell_t *shell_A, libgrpp_shell_t *shell_B,
                                                       double *charge_origin, int charge,
                                                       double *coulomb_matrix);           /* libgrpp.h:267 */
void libgrpp_nuclear_attraction_integrals_charged_ball(libgrpp_shell_t *shell_A, libgrpp_shell_t *shell_B,
                                                       double *charge_origin, int charge, double r_rms,
                                                       double *coulomb_matrix);           /* libgrpp.h:276 */
void libgrpp_nuclear_attraction_integrals_gaussian_model(libgrpp_shell_t *shell_A, libgrpp_shell_t *shell_B,
                                                         double *charge_origin, int charge, double r_rms,
                                                         double *coulomb_matrix);         /* libgrpp.h:286 */
void libgrpp_nuclear_attraction_integrals_fermi_model(libgrpp_shell_t *shell_A, libgrpp_shell_t *shell_B,
                                                      double *charge_origin, int charge, double fermi_param_c,
                                                      double fermi_param_a, double *coulomb_matrix); /* libgrpp.h:296 */
void libgrpp_nuclear_attraction_integrals_fermi_bubble_model(libgrpp_shell_t *shell_A, libgrpp_shell_t *shell_B,
                                                             double *charge_origin, int charge, double param_c,
                                                             double param_a, double param_k,
                                                             double *coulomb_matrix);     /* libgrpp.h:307 */

/* nuclear charge distribution models (nuclear_models.h:13-39): radii in fm for the estimates */
double libgrpp_estimate_nuclear_rms_radius_johnson_1985(int A);
double libgrpp_estimate_nuclear_rms_radius_golovko_2008(int A);
int libgrpp_estimate_fermi_model_parameters(double R_rms, double *c, double *a);
double libgrpp_charge_density_ball(double r, double Z, double R_rms);
double libgrpp_charge_density_gaussian(double r, double Z, double R_rms);
double libgrpp_charge_density_fermi(double r, double Z, double c, double a);
double libgrpp_charge_density_fermi_bubble(double r, double Z, double c, double a, double k);
double libgrpp_rms_radius_fermi(int Z, double c, double a);
double libgrpp_rms_radius_fermi_bubble(int Z, double c, double a, double k);
double libgrpp_coulomb_potential_point(double r, double Z);
double libgrpp_coulomb_potential_ball(double r, double Z, double R_rms);
double libgrpp_coulomb_potential_gaussian(double r, double Z, double R_rms);
double libgrpp_coulomb_potential_fermi(double r, double Z, double c, double a);
double libgrpp_coulomb_potential_fermi_bubble(double r, double Z, double c, double a, double k);

#ifdef __cplusplus
}
#endif

#include "libgrpp_amd.h"

#endif
