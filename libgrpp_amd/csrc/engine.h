// Internal interface between the C API layer (api.c) and the HIP engine (engine.hip).
#pragma once

#include "libgrpp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* which operator parts a job evaluates */
enum {
    ENG_TYPE1 = 1,      /* local U_L (type-1) */
    ENG_TYPE2 = 2,      /* semilocal AREP channels */
    ENG_SO = 4,         /* spin-orbit channels, each scaled by so_scale */
    ENG_OC = 8,         /* outercore projectors (both parts) */
    ENG_ALL = 15,
    ENG_OC_P1 = 16,     /* outercore part 1 only: U|nlj><nlj| + |nlj><nlj|U */
    ENG_OC_P2 = 32      /* outercore part 2 only: -|nlj><nlj| (U + U')/2 |n'lj><n'lj| */
};

typedef struct {
    const double *origin;               /* [3] */
    const libgrpp_potential_t *local;   /* may be NULL */
    int n_arep;
    libgrpp_potential_t *const *arep;
    int n_so;
    libgrpp_potential_t *const *so;
    const double *so_scale;             /* per SO channel factor (NULL: 2/(2L+1)) */
    int n_oc;
    libgrpp_potential_t *const *oc_pot;
    libgrpp_shell_t *const *oc_shell;   /* centred at origin regardless of their own origin field */
    int shell_begin, shell_end;         /* shells this centre acts on (end < 0: all) */
    int oc_i, oc_j;                     /* part 2 restricted to the (oc_i, oc_j) term (oc_i < 0: all) */
} eng_center_t;

/* one output block: shells (a, b), centres [cb, ce) of the centre-index list, destination
 * out[op*op_stride + off + i*stride + j]; mirror != 0 also writes the transposed block at moff
 * (arep symmetric, SO antisymmetric) */
typedef struct {
    int a, b, cb, ce;
    long long off, moff;
    int stride, mirror;
} eng_item_t;

typedef struct eng_plan eng_plan_t;

/* launch shapes of a plan that do not change any result (the autotuner's choices) */
typedef struct {
    int screen_probe;       /* triples per entry in the screening probe: 1, 2 or 4 */
    int contract_threads;   /* k_contract workgroup: 64, 128 or 256 threads */
    int type1_threads;      /* k_type1 workgroup: 64, 128 or 256 threads */
    int layout;             /* 0: tables, projectors, type-1 on the side stream; 1: tables and projectors
                               ahead of the screening, type-1 alone on the side stream; 2: every kernel
                               in order on the caller's stream */
    int screen_split;       /* k_screen waves per 64 x 64 radial tile: 4, 8 or 16 */
} eng_tuning_t;

/* flags */
enum { ENG_F_STATS = 2, ENG_F_LEVELS = 4, ENG_F_GRAPH = 8,
       ENG_F_ARENA = 16 /* every device buffer carved from the device's static scratch array, no
                           stream or event of the plan's own: the build runs in order on the
                           caller's stream (the allocation-free entry point) */,
       ENG_F_SERIAL = 32 /* no stream or event of the plan's own (buffers from the pool): the build
                            runs in order on the caller's stream (one-task plans of the per-call
                            entry points) */ };
#define ENG_SCRATCH_BYTES (64ull << 20)   /* size of that static scratch array per device */
#define ENG_NFEAT 12   /* level-predictor features per radial row */

int eng_plan_create(int nshell, libgrpp_shell_t *const *shells, int ncenter, const eng_center_t *centers,
                    int nidx, const int *center_index, int nitem, const eng_item_t *items, long long out_len,
                    long long op_stride, int parts, int flags, eng_plan_t **plan);
int eng_plan_execute(eng_plan_t *plan, void *stream);
int eng_plan_download(eng_plan_t *plan, double *out);   /* out_len * 4 doubles */
/* spin-orbital one-electron Hamiltonian of the last execute (see engine.hip, k_spin_ham); with
   H_host NULL only *h_len (complex elements) is set */
int eng_plan_spin_hamiltonian(eng_plan_t *plan, const int *shell_L, const int *mol_shell_off, int nmol,
                              const long long *mol_cart_off, int spherical, double *H_host, long long *h_len);
/* the four operators' first out_len doubles straight into four host arrays */
int eng_plan_download_ops(eng_plan_t *plan, double *o0, double *o1, double *o2, double *o3);
double *eng_plan_device_output(eng_plan_t *plan);
void eng_plan_destroy(eng_plan_t *plan);
int eng_plan_stats(const eng_plan_t *plan, libgrpp_batch_stats_t *stats);
int eng_plan_features(eng_plan_t *plan, int *n, float *features);
int eng_plan_set_predictions(eng_plan_t *plan, int n, const float *pred);
int eng_plan_row_levels(eng_plan_t *plan, int *n, float *levels);
int eng_plan_get_tuning(const eng_plan_t *plan, eng_tuning_t *t);
int eng_plan_set_tuning(eng_plan_t *plan, const eng_tuning_t *t);   /* invalidates a captured graph */
int eng_plan_execute_eager(eng_plan_t *plan, void *stream);   /* never the graph: phases are timed */

/* convenience: create + execute + download + destroy */
int eng_run(int nshell, libgrpp_shell_t *const *shells, int ncenter, const eng_center_t *centers, int nidx,
            const int *center_index, int nitem, const eng_item_t *items, long long out_len, long long op_stride,
            int parts, double *out);

void eng_set_error(const char *fmt, ...);
int eng_device_ok(void);
int eng_setup_device(void);
/* nuclear attraction by the numeric radial quadrature (models 1..5, l <= 6), summed over nuclei */
int eng_nuclear(int nshell, libgrpp_shell_t *const *shells, int nnuc, const double *nuc_xyz, const int *charge,
                int model, const double *params /* [nnuc][3] */, double tol, int nitem, const eng_item_t *items,
                long long out_len, double *out);
/* host: a nuclear model's potential at n radii (nuclear_models.c) */
int nuc_model_potential(int model, double Z, const double *p, int n, const double *r, double *U);   /* tables, LDS limits, scratch address on the current device (if any) */

#ifdef __cplusplus
}
#endif
