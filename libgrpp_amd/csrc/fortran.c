/*
 * Fortran-77 ABI of the GRPP path (the reference's fortran.c): every argument by reference,
 * integers as int32_t (integer(4)), reals as double (real(8)), one trailing underscore.  Each
 * routine builds the C objects from the flat arrays, calls the C entry point of this library and
 * frees them.  Routine names, argument order and meaning follow fortran.c line by line (cited per
 * routine); the nuclear-attraction, kinetic and momentum routines of that file are not on the GRPP
 * path and are not provided.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "libgrpp.h"

static libgrpp_shell_t *f_shell(double *origin, const int32_t *L, const int32_t *nprim, double *coeffs, double *alpha)
{
    return libgrpp_new_shell(origin, (int) *L, (int) *nprim, coeffs, alpha);
}

/* a radial potential of the flat Fortran arrays (powers converted from int32_t) */
static libgrpp_potential_t *f_potential(int L, int J, const int32_t *nprim, const int32_t *powers, double *coeffs,
                                        double *alpha)
{
    const int n = (int) *nprim;
    int *pw = (int *) malloc(sizeof(int) * (n > 0 ? n : 1));
    for (int k = 0; k < n; k++) pw[k] = (int) powers[k];
    libgrpp_potential_t *p = libgrpp_new_potential(L, J, n, pw, coeffs, alpha);
    free(pw);
    return p;
}

/* ---------------------------------------------------------------- parameters (fortran.c:29-57) */
void libgrpp_set_default_parameters_(void) { libgrpp_set_default_parameters(); }
void libgrpp_set_radial_tolerance_(const double *t) { libgrpp_set_radial_tolerance(*t); }
void libgrpp_set_angular_screening_tolerance_(const double *t) { libgrpp_set_angular_screening_tolerance(*t); }
void libgrpp_set_modified_bessel_tolerance_(const double *t) { libgrpp_set_modified_bessel_tolerance(*t); }
void libgrpp_set_cartesian_order_(const int32_t *order) { libgrpp_set_cartesian_order((int) *order); }

/* ---------------------------------------------------------------- init / finalize (fortran.c:62-71) */
void libgrpp_init_(void) { libgrpp_init(); }
void libgrpp_finalize_(void) { libgrpp_finalize(); }

/* ---------------------------------------------------------------- integrals */
/* fortran.c:78-117: local term; the potential is U_L (L = J = 0 in the object) */
void libgrpp_type1_integrals_(double *origin_A, int32_t *L_A, int32_t *np_A, double *c_A, double *a_A,
                              double *origin_B, int32_t *L_B, int32_t *np_B, double *c_B, double *a_B,
                              double *rpp_origin, int32_t *rpp_np, int32_t *rpp_powers, double *rpp_c,
                              double *rpp_a, double *matrix)
{
    libgrpp_shell_t *A = f_shell(origin_A, L_A, np_A, c_A, a_A), *B = f_shell(origin_B, L_B, np_B, c_B, a_B);
    libgrpp_potential_t *U = f_potential(0, 0, rpp_np, rpp_powers, rpp_c, rpp_a);
    libgrpp_type1_integrals(A, B, rpp_origin, U, matrix);
    libgrpp_delete_potential(U);
    libgrpp_delete_shell(A);
    libgrpp_delete_shell(B);
}

/* fortran.c:124-165: one semilocal channel L */
void libgrpp_type2_integrals_(double *origin_A, int32_t *L_A, int32_t *np_A, double *c_A, double *a_A,
                              double *origin_B, int32_t *L_B, int32_t *np_B, double *c_B, double *a_B,
                              double *pot_origin, int32_t *pot_L, int32_t *pot_np, int32_t *pot_powers,
                              double *pot_c, double *pot_a, double *matrix)
{
    libgrpp_shell_t *A = f_shell(origin_A, L_A, np_A, c_A, a_A), *B = f_shell(origin_B, L_B, np_B, c_B, a_B);
    libgrpp_potential_t *U = f_potential((int) *pot_L, 0, pot_np, pot_powers, pot_c, pot_a);
    libgrpp_type2_integrals(A, B, pot_origin, U, matrix);
    libgrpp_delete_potential(U);
    libgrpp_delete_shell(A);
    libgrpp_delete_shell(B);
}

/* fortran.c:172-224: one spin-orbit channel L, three matrices (no 2/(2L+1) factor) */
void libgrpp_spin_orbit_integrals_(double *origin_A, int32_t *L_A, int32_t *np_A, double *c_A, double *a_A,
                                   double *origin_B, int32_t *L_B, int32_t *np_B, double *c_B, double *a_B,
                                   double *pot_origin, int32_t *pot_L, int32_t *pot_np, int32_t *pot_powers,
                                   double *pot_c, double *pot_a, double *so_x, double *so_y, double *so_z)
{
    libgrpp_shell_t *A = f_shell(origin_A, L_A, np_A, c_A, a_A), *B = f_shell(origin_B, L_B, np_B, c_B, a_B);
    libgrpp_potential_t *U = f_potential((int) *pot_L, 0, pot_np, pot_powers, pot_c, pot_a);
    libgrpp_spin_orbit_integrals(A, B, pot_origin, U, so_x, so_y, so_z);
    libgrpp_delete_potential(U);
    libgrpp_delete_shell(A);
    libgrpp_delete_shell(B);
}

/* fortran.c:231-318: outercore part 1 for the (L, J) potential and its outercore shell, added to
 * the four outputs; the outercore shell sits at the potential's origin */
void libgrpp_outercore_potential_integrals_part_1_(double *origin_A, int32_t *L_A, int32_t *np_A, double *c_A,
                                                   double *a_A, double *origin_B, int32_t *L_B, int32_t *np_B,
                                                   double *c_B, double *a_B, double *pot_origin, int32_t *pot_L,
                                                   int32_t *pot_J, int32_t *pot_np, int32_t *pot_powers,
                                                   double *pot_c, double *pot_a, int32_t *oc_np, double *oc_c,
                                                   double *oc_a, double *arep, double *so_x, double *so_y,
                                                   double *so_z)
{
    libgrpp_shell_t *A = f_shell(origin_A, L_A, np_A, c_A, a_A), *B = f_shell(origin_B, L_B, np_B, c_B, a_B);
    libgrpp_potential_t *U = f_potential((int) *pot_L, (int) *pot_J, pot_np, pot_powers, pot_c, pot_a);
    libgrpp_shell_t *nlj = f_shell(pot_origin, pot_L, oc_np, oc_c, oc_a);
    libgrpp_outercore_potential_integrals_part_1(A, B, pot_origin, U, nlj, arep, so_x, so_y, so_z);
    libgrpp_delete_shell(nlj);
    libgrpp_delete_potential(U);
    libgrpp_delete_shell(A);
    libgrpp_delete_shell(B);
}

/* fortran.c:325-436: outercore part 2 for one ordered pair of (potential, shell) entries, added
 * to the four outputs */
void libgrpp_outercore_potential_integrals_part_2_(
    double *origin_A, int32_t *L_A, int32_t *np_A, double *c_A, double *a_A, double *origin_B, int32_t *L_B,
    int32_t *np_B, double *c_B, double *a_B, double *pot_origin, int32_t *L_1, int32_t *J_1, int32_t *pot1_np,
    int32_t *pot1_powers, double *pot1_c, double *pot1_a, int32_t *oc1_np, double *oc1_c, double *oc1_a,
    int32_t *L_2, int32_t *J_2, int32_t *pot2_np, int32_t *pot2_powers, double *pot2_c, double *pot2_a,
    int32_t *oc2_np, double *oc2_c, double *oc2_a, double *arep, double *so_x, double *so_y, double *so_z)
{
    libgrpp_shell_t *A = f_shell(origin_A, L_A, np_A, c_A, a_A), *B = f_shell(origin_B, L_B, np_B, c_B, a_B);
    libgrpp_potential_t *U1 = f_potential((int) *L_1, (int) *J_1, pot1_np, pot1_powers, pot1_c, pot1_a);
    libgrpp_potential_t *U2 = f_potential((int) *L_2, (int) *J_2, pot2_np, pot2_powers, pot2_c, pot2_a);
    libgrpp_shell_t *s1 = f_shell(pot_origin, L_1, oc1_np, oc1_c, oc1_a);
    libgrpp_shell_t *s2 = f_shell(pot_origin, L_2, oc2_np, oc2_c, oc2_a);
    libgrpp_outercore_potential_integrals_part_2(A, B, pot_origin, U1, s1, U2, s2, arep, so_x, so_y, so_z);
    libgrpp_delete_shell(s1);
    libgrpp_delete_shell(s2);
    libgrpp_delete_potential(U1);
    libgrpp_delete_potential(U2);
    libgrpp_delete_shell(A);
    libgrpp_delete_shell(B);
}

/* ---------------------------------------------------------------- gradients */
/* fortran.c:443-490: d/d(point) of the local-term matrix, three matrices */
void libgrpp_type1_integrals_gradient_(double *origin_A, int32_t *L_A, int32_t *np_A, double *c_A, double *a_A,
                                       double *origin_B, int32_t *L_B, int32_t *np_B, double *c_B, double *a_B,
                                       double *rpp_origin, int32_t *rpp_np, int32_t *rpp_powers, double *rpp_c,
                                       double *rpp_a, double *point_3d, double *gx, double *gy, double *gz)
{
    libgrpp_shell_t *A = f_shell(origin_A, L_A, np_A, c_A, a_A), *B = f_shell(origin_B, L_B, np_B, c_B, a_B);
    libgrpp_potential_t *U = f_potential(0, 0, rpp_np, rpp_powers, rpp_c, rpp_a);
    double *g[3] = {gx, gy, gz};
    libgrpp_type1_integrals_gradient(A, B, rpp_origin, U, point_3d, g);
    libgrpp_delete_potential(U);
    libgrpp_delete_shell(A);
    libgrpp_delete_shell(B);
}

/* fortran.c:497-546: d/d(point) of one semilocal channel */
void libgrpp_type2_integrals_gradient_(double *origin_A, int32_t *L_A, int32_t *np_A, double *c_A, double *a_A,
                                       double *origin_B, int32_t *L_B, int32_t *np_B, double *c_B, double *a_B,
                                       double *pot_origin, int32_t *pot_L, int32_t *pot_np, int32_t *pot_powers,
                                       double *pot_c, double *pot_a, double *point_3d, double *gx, double *gy,
                                       double *gz)
{
    libgrpp_shell_t *A = f_shell(origin_A, L_A, np_A, c_A, a_A), *B = f_shell(origin_B, L_B, np_B, c_B, a_B);
    libgrpp_potential_t *U = f_potential((int) *pot_L, 0, pot_np, pot_powers, pot_c, pot_a);
    double *g[3] = {gx, gy, gz};
    libgrpp_type2_integrals_gradient(A, B, pot_origin, U, point_3d, g);
    libgrpp_delete_potential(U);
    libgrpp_delete_shell(A);
    libgrpp_delete_shell(B);
}

/* fortran.c:553-634: d/d(point) of one spin-orbit channel, nine matrices (so_x by x, y, z; so_y ...) */
void libgrpp_spin_orbit_integrals_gradient_(double *origin_A, int32_t *L_A, int32_t *np_A, double *c_A,
                                            double *a_A, double *origin_B, int32_t *L_B, int32_t *np_B, double *c_B,
                                            double *a_B, double *pot_origin, int32_t *pot_L, int32_t *pot_np,
                                            int32_t *pot_powers, double *pot_c, double *pot_a, double *point_3d,
                                            double *gxx, double *gxy, double *gxz, double *gyx, double *gyy,
                                            double *gyz, double *gzx, double *gzy, double *gzz)
{
    libgrpp_shell_t *A = f_shell(origin_A, L_A, np_A, c_A, a_A), *B = f_shell(origin_B, L_B, np_B, c_B, a_B);
    libgrpp_potential_t *U = f_potential((int) *pot_L, 0, pot_np, pot_powers, pot_c, pot_a);
    double *sx[3] = {gxx, gxy, gxz}, *sy[3] = {gyx, gyy, gyz}, *sz[3] = {gzx, gzy, gzz};
    libgrpp_spin_orbit_integrals_gradient(A, B, pot_origin, U, point_3d, sx, sy, sz);
    libgrpp_delete_potential(U);
    libgrpp_delete_shell(A);
    libgrpp_delete_shell(B);
}

/* ---------------------------------------------------------------- overlap and normalisation */
/* fortran.c:641-682: <x^nA y^lA z^mA ...|x^nB y^lB z^mB ...> of two contractions: one element of the
 * overlap block of the two whole shells */
void evaluate_overlap_integral_contracted_(double *origin_A, int32_t *n_A, int32_t *l_A, int32_t *m_A,
                                           int32_t *np_A, double *c_A, double *a_A, double *origin_B, int32_t *n_B,
                                           int32_t *l_B, int32_t *m_B, int32_t *np_B, double *c_B, double *a_B,
                                           double *overlap)
{
    const int LA = (int) (*n_A + *l_A + *m_A), LB = (int) (*n_B + *l_B + *m_B);
    libgrpp_shell_t *A = libgrpp_new_shell(origin_A, LA, (int) *np_A, c_A, a_A);
    libgrpp_shell_t *B = libgrpp_new_shell(origin_B, LB, (int) *np_B, c_B, a_B);
    /* the one requested component of each shell */
    A->cart_size = 1;
    A->cart_list[0] = (int) *n_A;
    A->cart_list[1] = (int) *l_A;
    A->cart_list[2] = (int) *m_A;
    B->cart_size = 1;
    B->cart_list[0] = (int) *n_B;
    B->cart_list[1] = (int) *l_B;
    B->cart_list[2] = (int) *m_B;
    libgrpp_overlap_integrals(A, B, overlap);
    libgrpp_delete_shell(A);
    libgrpp_delete_shell(B);
}

/* fortran.c:689-702: sqrt((2L-1)!!) / sqrt(<x^L|x^L>) of a contraction about the origin */
void radial_gto_norm_factor_(int32_t *L, int32_t *num_primitives, double *coeffs, double *alpha, double *norm)
{
    double origin[3] = {0.0, 0.0, 0.0}, S = 0.0;
    int32_t zero = 0;
    evaluate_overlap_integral_contracted_(origin, L, &zero, &zero, num_primitives, coeffs, alpha, origin, L, &zero,
                                          &zero, num_primitives, coeffs, alpha, &S);
    double df = 1.0;
    for (int k = 2 * (int) *L - 1; k > 1; k -= 2) df *= k;
    *norm = sqrt(df) / sqrt(S);
}

/* fortran.c:709-733: overlap block of two shells */
void libgrpp_overlap_integrals_(double *origin_A, int32_t *L_A, int32_t *np_A, double *c_A, double *a_A,
                                double *origin_B, int32_t *L_B, int32_t *np_B, double *c_B, double *a_B,
                                double *matrix)
{
    libgrpp_shell_t *A = f_shell(origin_A, L_A, np_A, c_A, a_A), *B = f_shell(origin_B, L_B, np_B, c_B, a_B);
    libgrpp_overlap_integrals(A, B, matrix);
    libgrpp_delete_shell(A);
    libgrpp_delete_shell(B);
}
