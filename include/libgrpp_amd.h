/*
 * libgrpp_amd -- batched entry points (no counterpart in the reference API).
 *
 * The reference evaluates one (shell A, shell B, GRPP centre C) task per call and its C driver
 * loops over all ordered shell pairs and centres (test_libgrpp_c/eval_integrals.c:42-130,
 * eval_integrals_grad.c:37-163).  These entry points take the whole shell list and centre list
 * at once and return the same dense matrices that loop assembles, computed by one sequence of
 * HIP kernels on the current device.
 *
 * Return value: 0 on success, non-zero error code otherwise (message: libgrpp_last_error()).
 */
#ifndef LIBGRPP_AMD_BATCH_H
#define LIBGRPP_AMD_BATCH_H

#ifdef __cplusplus
extern "C" {
#endif

struct libgrpp_batch_plan;
typedef struct libgrpp_batch_plan libgrpp_batch_plan_t;

/* flags */
enum {
    LIBGRPP_BATCH_DEFAULT = 0,
    LIBGRPP_BATCH_LEARNED_GRID = 1,   /* predicted fixed radial grids + embedded-rule check + adaptive fallback */
    LIBGRPP_BATCH_GRADIENT = 2,       /* plan also evaluates nuclear gradients (see ..._gradients) */
    LIBGRPP_BATCH_NO_SYMMETRY = 4     /* evaluate every ordered pair instead of mirroring B,A from A,B */
};

typedef struct {
    long long n_tasks;            /* ordered shell pairs x GRPP centres */
    long long n_integrals;        /* 4 * sum over tasks of ncart(A) * ncart(B) */
    long long n_radial_entries;   /* radial-table entries evaluated (per centre and potential) */
    long long n_radial_points;    /* sum over radial tiles of grid points integrated */
    long long n_fallback_tiles;   /* learned-grid tiles that failed the check and were refined */
    long long n_tiles;            /* radial tiles evaluated */
    double fp64_flops;            /* algorithmic FP64 operations of the last execution */
    double device_seconds;        /* kernel time of the last execution (HIP events) */
    double radial_seconds;        /* time of the radial-quadrature kernel alone */
    /* per-phase kernel times of the last execution (HIP events on the plan's stream) */
    double tables_seconds;        /* grid tabulation of the shell factors and potentials */
    double screen_seconds;        /* Shaw-Hill envelope estimates of the radial entries */
    double canon_seconds;         /* canonical-split gather of the radial tables */
    double proj_seconds;          /* angular projectors and outercore vectors */
    double type1_seconds;         /* type-1 (local) McMurchie-Davidson blocks */
    double contract_seconds;      /* angular x radial contraction and outercore terms */
    long long screen_triples;     /* primitive triples of the envelope estimates */
    long long screen_steps;       /* fixed-point evaluations of the envelope maxima */
    double screen_flops;          /* envelope work in FLOP equivalents (see the engine) */
} libgrpp_batch_stats_t;

/*
 * Full GRPP matrices for all ordered shell pairs and all centres:
 *   arep = sum_C [ T1(U_L) + sum_L T2(U_L) + OC_arep ],
 *   so_k = sum_C [ sum_L 2/(2L+1) SO_k(U_L) + OC_so_k ],
 * each dim x dim row-major, dim = sum of ncart over the shell list, blocks at the cumulative
 * Cartesian offsets (same result as libgrpp_full_grpp_integrals summed by the reference driver).
 * grpp_per_center[c] may be NULL (centre skipped).  Inputs are read, never modified.
 */
int libgrpp_batch_full_grpp_integrals(int nshells, libgrpp_shell_t *const *shells, int ncenters,
                                      const double *center_xyz, libgrpp_grpp_t *const *grpp_per_center, int flags,
                                      double *arep, double *so_x, double *so_y, double *so_z,
                                      libgrpp_batch_stats_t *stats);

/*
 * Nuclear gradients of the same matrices with respect to every centre's coordinates, treating
 * the centre list as the atom list (eval_integrals_grad.c:37-163 with the reference's
 * libgrpp_full_grpp_integrals_gradient per task).  out holds 4 * 3*ncenters * dim * dim doubles:
 * out[((op*3*ncenters) + 3*atom + k) * dim*dim + i*dim + j], op = arep, so_x, so_y, so_z.
 * Basis shells must sit on centres of the list (as in the reference driver).
 */
int libgrpp_batch_full_grpp_gradients(int nshells, libgrpp_shell_t *const *shells, int ncenters,
                                      const double *center_xyz, libgrpp_grpp_t *const *grpp_per_center, int flags,
                                      double *out, libgrpp_batch_stats_t *stats);

/*
 * Plan interface: pack and upload once, then execute repeatedly with inputs and outputs
 * resident in device memory.  stream is a hipStream_t (NULL = default stream).
 */
int libgrpp_batch_plan_create(int nshells, libgrpp_shell_t *const *shells, int ncenters, const double *center_xyz,
                              libgrpp_grpp_t *const *grpp_per_center, int flags, libgrpp_batch_plan_t **plan);
int libgrpp_batch_plan_execute(libgrpp_batch_plan_t *plan, void *stream);
int libgrpp_batch_plan_download(libgrpp_batch_plan_t *plan, double *arep, double *so_x, double *so_y, double *so_z);
int libgrpp_batch_plan_dim(const libgrpp_batch_plan_t *plan);
int libgrpp_batch_plan_stats(const libgrpp_batch_plan_t *plan, libgrpp_batch_stats_t *stats);
int libgrpp_batch_plan_set_predictions(libgrpp_batch_plan_t *plan, int n, const float *level_and_interval);
int libgrpp_batch_plan_features(libgrpp_batch_plan_t *plan, int *n, float *features);
void libgrpp_batch_plan_destroy(libgrpp_batch_plan_t *plan);

/* thread-local description of the last error of any libgrpp_amd call */
const char *libgrpp_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
