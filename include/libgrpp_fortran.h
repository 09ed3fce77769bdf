/*
 * libgrpp_amd: the Fortran-77 ABI of the GRPP path (libgrpp's fortran.c).  Every argument is passed
 * by reference; integer(4) is int32_t and real(8) is double.  Fortran callers use these through
 * their own interface blocks (as test_libgrpp_f90/libgrpp.f90 does for the reference); C callers
 * may include this header.  Outputs of the outercore parts are added to, all others overwritten.
 */
#ifndef LIBGRPP_AMD_FORTRAN_H
#define LIBGRPP_AMD_FORTRAN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* fortran.c:29 */
void libgrpp_set_default_parameters_(void);
/* fortran.c:35 */
void libgrpp_set_radial_tolerance_(const double *t);
/* fortran.c:41 */
void libgrpp_set_angular_screening_tolerance_(const double *t);
/* fortran.c:47 */
void libgrpp_set_modified_bessel_tolerance_(const double *t);
/* fortran.c:53 */
void libgrpp_set_cartesian_order_(const int32_t *order);
/* fortran.c:62 */
void libgrpp_init_(void);
/* fortran.c:68 */
void libgrpp_finalize_(void);
/* fortran.c:78 */
void libgrpp_type1_integrals_(double *origin_A, int32_t *L_A, int32_t *np_A, double *c_A, double *a_A, double *origin_B, int32_t *L_B, int32_t *np_B, double *c_B, double *a_B, double *rpp_origin, int32_t *rpp_np, int32_t *rpp_powers, double *rpp_c, double *rpp_a, double *matrix);
/* fortran.c:124 */
void libgrpp_type2_integrals_(double *origin_A, int32_t *L_A, int32_t *np_A, double *c_A, double *a_A, double *origin_B, int32_t *L_B, int32_t *np_B, double *c_B, double *a_B, double *pot_origin, int32_t *pot_L, int32_t *pot_np, int32_t *pot_powers, double *pot_c, double *pot_a, double *matrix);
/* fortran.c:172 */
void libgrpp_spin_orbit_integrals_(double *origin_A, int32_t *L_A, int32_t *np_A, double *c_A, double *a_A, double *origin_B, int32_t *L_B, int32_t *np_B, double *c_B, double *a_B, double *pot_origin, int32_t *pot_L, int32_t *pot_np, int32_t *pot_powers, double *pot_c, double *pot_a, double *so_x, double *so_y, double *so_z);
/* fortran.c:231 */
void libgrpp_outercore_potential_integrals_part_1_(double *origin_A, int32_t *L_A, int32_t *np_A, double *c_A, double *a_A, double *origin_B, int32_t *L_B, int32_t *np_B, double *c_B, double *a_B, double *pot_origin, int32_t *pot_L, int32_t *pot_J, int32_t *pot_np, int32_t *pot_powers, double *pot_c, double *pot_a, int32_t *oc_np, double *oc_c, double *oc_a, double *arep, double *so_x, double *so_y, double *so_z);
/* fortran.c:325 */
void libgrpp_outercore_potential_integrals_part_2_( double *origin_A, int32_t *L_A, int32_t *np_A, double *c_A, double *a_A, double *origin_B, int32_t *L_B, int32_t *np_B, double *c_B, double *a_B, double *pot_origin, int32_t *L_1, int32_t *J_1, int32_t *pot1_np, int32_t *pot1_powers, double *pot1_c, double *pot1_a, int32_t *oc1_np, double *oc1_c, double *oc1_a, int32_t *L_2, int32_t *J_2, int32_t *pot2_np, int32_t *pot2_powers, double *pot2_c, double *pot2_a, int32_t *oc2_np, double *oc2_c, double *oc2_a, double *arep, double *so_x, double *so_y, double *so_z);
/* fortran.c:443 */
void libgrpp_type1_integrals_gradient_(double *origin_A, int32_t *L_A, int32_t *np_A, double *c_A, double *a_A, double *origin_B, int32_t *L_B, int32_t *np_B, double *c_B, double *a_B, double *rpp_origin, int32_t *rpp_np, int32_t *rpp_powers, double *rpp_c, double *rpp_a, double *point_3d, double *gx, double *gy, double *gz);
/* fortran.c:497 */
void libgrpp_type2_integrals_gradient_(double *origin_A, int32_t *L_A, int32_t *np_A, double *c_A, double *a_A, double *origin_B, int32_t *L_B, int32_t *np_B, double *c_B, double *a_B, double *pot_origin, int32_t *pot_L, int32_t *pot_np, int32_t *pot_powers, double *pot_c, double *pot_a, double *point_3d, double *gx, double *gy, double *gz);
/* fortran.c:553 */
void libgrpp_spin_orbit_integrals_gradient_(double *origin_A, int32_t *L_A, int32_t *np_A, double *c_A, double *a_A, double *origin_B, int32_t *L_B, int32_t *np_B, double *c_B, double *a_B, double *pot_origin, int32_t *pot_L, int32_t *pot_np, int32_t *pot_powers, double *pot_c, double *pot_a, double *point_3d, double *gxx, double *gxy, double *gxz, double *gyx, double *gyy, double *gyz, double *gzx, double *gzy, double *gzz);
/* fortran.c:641 */
void evaluate_overlap_integral_contracted_(double *origin_A, int32_t *n_A, int32_t *l_A, int32_t *m_A, int32_t *np_A, double *c_A, double *a_A, double *origin_B, int32_t *n_B, int32_t *l_B, int32_t *m_B, int32_t *np_B, double *c_B, double *a_B, double *overlap);
/* fortran.c:689 */
void radial_gto_norm_factor_(int32_t *L, int32_t *num_primitives, double *coeffs, double *alpha, double *norm);
/* fortran.c:709 */
void libgrpp_overlap_integrals_(double *origin_A, int32_t *L_A, int32_t *np_A, double *c_A, double *a_A, double *origin_B, int32_t *L_B, int32_t *np_B, double *c_B, double *a_B, double *matrix);

#ifdef __cplusplus
}
#endif

#endif
