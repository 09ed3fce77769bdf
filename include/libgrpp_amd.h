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
    LIBGRPP_BATCH_LEARNED_GRID = 1,   /* reserved, refused: a plan takes the level predictor's output through
                                         libgrpp_batch_plan_set_predictions */
    LIBGRPP_BATCH_GRADIENT = 2,       /* reserved, refused: gradients are built by
                                         libgrpp_batch_gradients_plan_create / libgrpp_batch_full_grpp_gradients */
    LIBGRPP_BATCH_NO_SYMMETRY = 4,    /* evaluate every ordered pair instead of mirroring B,A from A,B */
    LIBGRPP_BATCH_RECORD_LEVELS = 8,  /* plan records each radial entry's stopping level (training labels) */
    LIBGRPP_BATCH_GRAPH = 16          /* execute replays one hipGraph of the whole build, captured by the
                                         first execute; per-phase times are then not measured (zero) */
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
 * Allocation-free variant of libgrpp_batch_full_grpp_integrals (no counterpart in the reference):
 * for callers that cannot allocate device memory -- tight serving loops, harnesses that forbid
 * hipMalloc.  Same inputs and host outputs; flags may hold LIBGRPP_BATCH_NO_SYMMETRY only.  Every
 * device buffer of the build lives in one static 64 MiB scratch array per device, and the kernels
 * run in order on `stream` (a hipStream_t, NULL = default stream); the call returns when the
 * matrices are in the caller's arrays.  No hipMalloc / hipFree or other allocating library routine
 * is called, so libgrpp_init() must have run with this device current (it puts the constant tables
 * in place).  Inputs whose build needs more than the scratch array are refused with an error
 * (libgrpp_last_error()).  All calls on one device share that array: calls that overlap on one
 * device race, and must be serialised by the caller.  stats: as for the batched call, without
 * kernel times (zero).
 */
int libgrpp_batch_full_grpp_integrals_noalloc(int nshells, libgrpp_shell_t *const *shells, int ncenters,
                                              const double *center_xyz, libgrpp_grpp_t *const *grpp_per_center,
                                              int flags, double *arep, double *so_x, double *so_y, double *so_z,
                                              void *stream, libgrpp_batch_stats_t *stats);

/*
 * Nuclear attraction matrix of a whole shell list (no counterpart in the reference, whose
 * nuclear-attraction call is per shell pair): V = sum over the nnuc nuclei of the attraction blocks
 * of every ordered shell pair, dim x dim row-major (dim = sum of cart_size).  nuclear_model as for
 * libgrpp_nuclear_attraction_integrals, the same for every nucleus; model_params holds 3 doubles per
 * nucleus (R_rms, or c, a[, k]).  The point-charge model runs the analytic type-1 kernels, the
 * others the numeric radial quadrature (l <= 6).
 */
int libgrpp_batch_nuclear_attraction_integrals(int nshells, libgrpp_shell_t *const *shells, int nnuc,
                                               const double *nuc_xyz, const int *charges, int nuclear_model,
                                               const double *model_params, double *V);

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

/*
 * Many independent molecules in one plan (uneven batches of shell pairs x centres).  Molecule m
 * owns shells [shell_off[m], shell_off[m+1]) and centres [center_off[m], center_off[m+1]) of the
 * concatenated lists; its centres act on its shells only.  Its four dim_m x dim_m matrices
 * (dim_m = sum of cart_size over its shells) follow those of molecule m-1: download_flat writes
 * out[op * output_len + sum_{k<m} dim_k^2 + i * dim_m + j], op = arep, so_x, so_y, so_z.
 */
int libgrpp_batch_molecules_plan_create(int nmol, const int *shell_off, const int *center_off,
                                        libgrpp_shell_t *const *shells, const double *center_xyz,
                                        libgrpp_grpp_t *const *grpp_per_center, int flags,
                                        libgrpp_batch_plan_t **plan);
long long libgrpp_batch_plan_output_len(const libgrpp_batch_plan_t *plan);   /* doubles per operator */
int libgrpp_batch_plan_download_flat(libgrpp_batch_plan_t *plan, double *out);   /* 4 * output_len doubles */
int libgrpp_batch_plan_dim(const libgrpp_batch_plan_t *plan);
/*
 * Spin-orbital GRPP Hamiltonian of a matrix plan's last execute (no counterpart in the reference,
 * which leaves this step to the host: doc/manual.tex:594).  libgrpp's so_eta are real, with
 * <A| U^SO P l_eta |B> = i so_eta[A][B], and the spin s = sigma/2 is not applied; this assembles
 *   H = [ A + i Z/2        Y/2 + i X/2 ]     A = T arep T^T, (X, Y, Z) = T so_(x,y,z) T^T,
 *       [ -Y/2 + i X/2     A - i Z/2   ]
 * over spin-orbitals (all alpha functions, then all beta), a complex Hermitian matrix of order
 * 2n, row-major with interleaved (re, im) doubles.  spherical = 0: T is the identity (n = number of
 * Cartesian functions); spherical = 1: T applies, per shell, the Cartesian -> real spherical
 * harmonic coefficients of libgrpp's outercore code (outercore_integrals.c:171-178), n = sum of
 * 2l+1.  A molecules plan gives one such matrix per molecule, one after the other.
 * _len: complex elements H must hold (-1 for a gradient plan).
 */
long long libgrpp_batch_plan_spin_hamiltonian_len(const libgrpp_batch_plan_t *plan, int spherical);
int libgrpp_batch_plan_spin_hamiltonian(libgrpp_batch_plan_t *plan, int spherical, double *H);

/*
 * Autotuner: launch shapes of a plan that change no result (bitwise), picked by timing the plan's
 * own build: the screening probe width and waves per tile (per-phase time of the screening), the contraction's and
 * the type-1 workgroup sizes (contraction phase, whole build) and the stream layout (whole build).  Each candidate runs `repeats`
 * eager builds and keeps its fastest; the choice stays with the plan (a captured graph is
 * recaptured at the next execute).
 */
typedef struct {
    int screen_probe;         /* 1, 2 or 4 triples per entry in the screening probe */
    int contract_threads;     /* 64, 128 or 256 */
    int type1_threads;        /* 64, 128 or 256 */
    int stream_layout;        /* 0: grid tables, projectors and type-1 beside the screening on a second
                                 stream; 1: tables and projectors ahead of it, type-1 beside it;
                                 2: every kernel in order on the caller's stream */
    int screen_split;         /* 4, 8 or 16 screening waves per 64 x 64 radial tile; 0 in
                                 libgrpp_batch_plan_set_tuning keeps the plan's current split (the
                                 plan's default: 16 when it has fewer than 8192 radial tiles, else 4).
                                 The struct is 40 bytes (this field at offset 16); callers compiled
                                 against an older 32-byte layout must be rebuilt. */
    double seconds_before;    /* fastest build with the shapes in place before tuning */
    double seconds_after;     /* fastest build with the chosen shapes */
} libgrpp_batch_tuning_t;
int libgrpp_batch_plan_autotune(libgrpp_batch_plan_t *plan, int repeats, libgrpp_batch_tuning_t *chosen);
int libgrpp_batch_plan_set_tuning(libgrpp_batch_plan_t *plan, const libgrpp_batch_tuning_t *tuning);
int libgrpp_batch_plan_stats(const libgrpp_batch_plan_t *plan, libgrpp_batch_stats_t *stats);
/*
 * Gradient plan: the inputs of libgrpp_batch_full_grpp_gradients packed once.  execute evaluates
 * the engine's L-1 / L+1 blocks; download_gradients scatters them into gradient_len doubles laid
 * out as in libgrpp_batch_full_grpp_gradients.  The level-predictor hooks below apply as well;
 * libgrpp_batch_plan_download is refused for a gradient plan.
 */
int libgrpp_batch_gradients_plan_create(int nshells, libgrpp_shell_t *const *shells, int ncenters,
                                        const double *center_xyz, libgrpp_grpp_t *const *grpp_per_center,
                                        int flags, libgrpp_batch_plan_t **plan);
long long libgrpp_batch_plan_gradient_len(const libgrpp_batch_plan_t *plan);   /* 12 * ncenters * dim^2 */
int libgrpp_batch_plan_download_gradients(libgrpp_batch_plan_t *plan, double *out);
/*
 * The gradients of the listed atoms only (no counterpart in the reference): out holds
 * 12 * natoms * dim^2 doubles, out[((op * 3 * natoms) + 3 * a + k) * dim^2 + i * dim + j] for the
 * a-th atom of the list (each atom index in [0, ncenters), at most once), the same values as the
 * full download.  A caller streams the gradients of a large system atom by atom instead of holding
 * all 12 * ncenters * dim^2 doubles.
 */
int libgrpp_batch_plan_download_gradients_atoms(libgrpp_batch_plan_t *plan, int natoms, const int *atoms,
                                                double *out);
/*
 * Radial-grid level predictor hooks.  The plan's radial tables have one row per (centre, channel
 * L, shell, lambda, n_A).  features: *n rows x 12 floats (L, l, lambda, n, |CA|, ln alpha_min,
 * ln alpha_max, ln zeta_min, ln zeta_max, ln(1 + 2 alpha_max |CA|), primitives, on-centre flag);
 * features == NULL returns *n only.  row_levels (plans created with LIBGRPP_BATCH_RECORD_LEVELS,
 * after an execute): *n rows x 3 floats = max, mean and fraction >= 7 of the stopping levels of
 * the row's entries (level k = 32 * 2^k - 1 Log3 points).  set_predictions: n rows x 3 floats in
 * the same layout; the plan orders each radial table's rows by the first value so that tiles hold
 * entries of like level.  Every entry still stops by the reference's rule, so results do not
 * depend on the predictions.
 */
int libgrpp_batch_plan_set_predictions(libgrpp_batch_plan_t *plan, int n, const float *level_and_interval);
int libgrpp_batch_plan_features(libgrpp_batch_plan_t *plan, int *n, float *features);
int libgrpp_batch_plan_row_levels(libgrpp_batch_plan_t *plan, int *n, float *levels);
void libgrpp_batch_plan_destroy(libgrpp_batch_plan_t *plan);

/*
 * Radial-grid level predictor: a 12 -> 128 -> 128 -> 128 -> 3 SiLU multilayer perceptron trained
 * and evaluated by HIP kernels (FP32 training with Adam; inference in FP32 or with BF16 weights).
 * Inputs are the 12 row features of libgrpp_batch_plan_features, outputs the 3 level statistics
 * of libgrpp_batch_plan_row_levels.  Host arrays: X n x 12, T and Y n x 3 (row-major floats).
 */
typedef struct libgrpp_mlp libgrpp_mlp_t;
int libgrpp_mlp_nparams(void);
int libgrpp_mlp_create(unsigned seed, libgrpp_mlp_t **mlp);
int libgrpp_mlp_set_normalisation(libgrpp_mlp_t *mlp, const float *mean, const float *scale);
int libgrpp_mlp_params(libgrpp_mlp_t *mlp, float *get, const float *set);
int libgrpp_mlp_gradient(libgrpp_mlp_t *mlp, int n, const float *X, const float *T, float *grad, float *loss);
int libgrpp_mlp_train(libgrpp_mlp_t *mlp, int n, const float *X, const float *T, int epochs, int batch, float lr,
                      unsigned seed, float *loss);
int libgrpp_mlp_predict(libgrpp_mlp_t *mlp, int n, const float *X, float *Y, int bf16);
void libgrpp_mlp_destroy(libgrpp_mlp_t *mlp);

/* thread-local description of the last error of any libgrpp_amd call */
const char *libgrpp_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
