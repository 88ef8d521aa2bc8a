/*
 * vjx -- MI355X-native transient core for vajax-style circuit simulation.
 *
 * Plain C ABI (extern "C", plain pointers and sizes).  Every function returns
 * an int status: 0 = ok, > 0 = warning (result usable), < 0 = error; the text
 * of the last error or warning on the calling thread is vjx_last_error().
 *
 * Pointer conventions: arguments named *_host are host memory borrowed for the
 * duration of the call; arguments named *_dev are device (HBM) memory on the
 * current HIP device, also borrowed.  The library owns everything behind its
 * handles and frees it in the matching *_destroy call.
 */
#ifndef VJX_H
#define VJX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VJX_OK 0
#define VJX_WARN_PIVOT 1      /* numeric pivot degraded; result still returned */
#define VJX_WARN_NOCONV 2     /* Newton iteration did not converge */
#define VJX_ERR_ARG (-1)      /* invalid argument / malformed CSR */
#define VJX_ERR_HIP (-2)      /* HIP runtime error */
#define VJX_ERR_SINGULAR (-3) /* structurally singular matrix */
#define VJX_ERR_STATE (-4)    /* call out of sequence (e.g. solve before refactor) */
#define VJX_ERR_FILL (-5)     /* analysis abandoned: LU fill beyond the fill limit */

#define VJX_PREC_F64 0 /* factorise and solve in FP64 */
#define VJX_PREC_F32 1 /* factorise in FP32, FP64 residual + refinement */

typedef struct vjx_pattern vjx_pattern;
typedef struct vjx_circuit vjx_circuit;

/* Library version (major*10000 + minor*100 + patch). */
int vjx_version(void);
/* Text of the last error/warning on this thread ("" if none). */
const char* vjx_last_error(void);
/* Number of visible HIP devices (0 when none; never fails). */
int vjx_device_count(void);

/* ---------------------------------------------------------------------------
 * Sparse LU of an n x n CSR matrix with a fixed pattern.
 * Replaces the reference's XLA FFI linear solver
 * (vajax/sparse/umfpack_ffi.cpp:189-338 solve, :345-370 dot) and its
 * pattern-keyed symbolic cache (:140-179): one handle per pattern, symbolic
 * analysis and pivot order chosen once, numeric refactorisation per call.
 * ------------------------------------------------------------------------- */

/* Validate the CSR pattern (indptr[0]=0, monotonic, indptr[n]=nnz, columns
 * strictly increasing in [0,n)) and create a handle.  umfpack_ffi.cpp:234-251. */
int vjx_pattern_create(int32_t n, int32_t nnz, const int32_t* indptr_host,
                       const int32_t* indices_host, vjx_pattern** out);
void vjx_pattern_destroy(vjx_pattern* pat);

/* Ordering + threshold-pivoted symbolic/numeric analysis from host values.
 * Called implicitly by the first vjx_lu_refactor; call again to re-pivot.
 * (umfpack_ffi.cpp:266-290, umfpack_di_symbolic + first numeric.) */
int vjx_lu_analyze(vjx_pattern* pat, const double* data_host);

/* Numeric refactorisation on the GPU with the pivot order fixed by the last
 * analysis (umfpack_ffi.cpp:292-313, umfpack_di_numeric).  precision is
 * VJX_PREC_F64 or VJX_PREC_F32.  Returns VJX_WARN_PIVOT when pivot growth
 * suggests calling vjx_lu_analyze again. */
int vjx_lu_refactor(vjx_pattern* pat, const double* data_dev, int32_t precision);

/* Solve A x = b with the last factorisation.  For VJX_PREC_F32, refine_steps
 * FP64 residual corrections r = b - A x use data_dev (the FP64 matrix), as in
 * vajax/analysis/solver_factories.py:895-922 (refine_steps = 1 there). */
int vjx_lu_solve(vjx_pattern* pat, const double* data_dev, const double* b_dev, double* x_dev,
                 int32_t refine_steps);

/* vjx_lu_solve with no allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops).  Same arguments, kernels and results as vjx_lu_solve;
 * the intermediate vectors and hub partials live in a fixed device array of
 * about 16.4 MiB per GPU (held until the process ends), so systems of more
 * than 524288 unknowns or hub plans of more than 16384 chunks are rejected
 * with VJX_ERR_ARG.  The pattern must have been factorised by
 * vjx_lu_refactor (VJX_ERR_STATE otherwise).  Calls that overlap on one GPU
 * share that array and race. */
int vjx_lu_solve_noalloc(vjx_pattern* pat, const double* data_dev, const double* b_dev,
                         double* x_dev, int32_t refine_steps);

/* y = A x (umfpack_ffi.cpp:345-370, UmfpackDotF64Impl). */
int vjx_spmv(vjx_pattern* pat, const double* data_dev, const double* x_dev, double* y_dev);

/* Solve A^T x = b (umfpack_ffi.cpp:388-481, UmfpackSolveTransposeF64Impl,
 * UMFPACK_At; the reverse-mode rule of vajax/sparse/umfpack_jax.py:148-177).
 * A^T is gathered from data_dev on the device into a second handle that the
 * pattern keeps, with its own ordering and pivots (analysed on the first
 * call, re-pivoted on a pivot warning), so each call costs one numeric
 * refactorisation and one solve, as the reference's numeric factor per call
 * does.  precision and refine_steps as for vjx_lu_refactor / vjx_lu_solve. */
int vjx_lu_solve_transpose(vjx_pattern* pat, const double* data_dev, const double* b_dev,
                           double* x_dev, int32_t precision, int32_t refine_steps);

/* One-shot host-buffer solve with the reference FFI's semantics:
 * solve(indptr, indices, data, b) -> x (umfpack_ffi.cpp:329-338).  Keeps a
 * handle cached on the (n, nnz, pattern) key like the reference's global
 * symbolic cache; vjx_clear_cache() drops it. */
int vjx_solve_host(int32_t n, int32_t nnz, const int32_t* indptr_host, const int32_t* indices_host,
                   const double* data_host, const double* b_host, double* x_host);
/* Host-buffer forms of the FFI's transposed solve and product
 * (umfpack_ffi.cpp:388-481 solve_transpose, :345-370 dot), on the same cache. */
int vjx_solve_transpose_host(int32_t n, int32_t nnz, const int32_t* indptr_host,
                             const int32_t* indices_host, const double* data_host,
                             const double* b_host, double* x_host);
int vjx_dot_host(int32_t n, int32_t nnz, const int32_t* indptr_host, const int32_t* indices_host,
                 const double* data_host, const double* x_host, double* y_host);
void vjx_clear_cache(void);

typedef struct {
  int32_t n;
  int32_t nnz_a;
  int64_t nnz_l;       /* strictly lower */
  int64_t nnz_u;       /* upper incl. diagonal */
  int64_t n_updates;   /* multiply-add contributions per refactorisation */
  double flops;        /* sum_k (|L_k| + 2 |L_k| |U_k|) */
  int32_t fact_levels, fwd_levels, bwd_levels;
  int32_t fact_launches, fwd_launches, bwd_launches;
  int32_t n_dense_rows;
  int32_t n_offdiag_pivots;
  double min_pivot_ratio;
  /* block-diagonal (batched) systems: diagonal blocks and the number of
   * distinct block patterns, each analysed once and shared */
  int32_t n_blocks;
  int32_t n_block_patterns;
  /* triangular-solve band programs: bands (launches) and hub rows */
  int32_t fwd_bands, bwd_bands, fwd_hub_rows, bwd_hub_rows;
  /* refactorisation band program: update contributions and targets run by
   * the band tasks and by the Schur steps (all in the tasks' counts when the
   * level program is used) */
  int64_t fact_task_updates, fact_schur_updates, fact_task_targets, fact_schur_targets;
  /* analysis time in ms: ordering, pivoting, LU pattern, refactorisation
   * program, solve programs (single-block systems; zero for packed batches) */
  double analyze_ms[5];
  int32_t band_width, solve_band_width; /* launch plan in use (solve 0: band_width) */
  double tuned_solve_ms;                /* solve time the autotuner measured (0: not run) */
  double tuned_factor_ms;               /* refactorisation time it measured (0: not run) */
  int32_t lu_pipeline;                  /* LU band tasks: 0 by task count, 1 shallow, 2 deep */
  /* dense trailing block factorised with MFMA: its size (0: none), the
   * earlier pivots contributing to it and its FP64 operations per refactor */
  int32_t dense_k, dense_m;
  double dense_flops;
  /* nested-dissection leaf of the ordering in use: ordering 3 reports the
   * leaf taken (with leaf 0, automatic, systems of up to 4 096 rows take
   * 256- or 4-row leaves, whichever program has fewer pivot levels for its
   * updates, larger ones 256); ordering 2 the configured leaf (0: max(48,
   * n / 128) per diagonal block); 0 for the other orderings */
  int32_t nd_leaf;
} vjx_lu_info;

int vjx_lu_get_info(const vjx_pattern* pat, vjx_lu_info* out);

/* Host-side copies of the analysis (for tests): permutations, LU pattern. */
int vjx_lu_get_perm(const vjx_pattern* pat, int32_t* p_host, int32_t* q_host);
int vjx_lu_get_pattern(const vjx_pattern* pat, int32_t* rowptr_host, int32_t* col_host);
/* Dependency check of the analysed solve programs: every row is scheduled
 * once and after every row it reads.  VJX_OK or VJX_ERR_STATE with a message. */
int vjx_lu_check_schedules(const vjx_pattern* pat);
/* Triangular solves in bands of `width` levels, one workgroup per group of
 * connected rows (default 16; 0 = one launch per level).  A new value drops
 * the analysis; the next refactorisation analyses again.  No counterpart in
 * the reference (a launch-plan knob). */
int vjx_pattern_set_band_width(vjx_pattern* pat, int32_t width);
/* Fill-reducing column order of the analysis: 0 = minimum degree on A+A^T,
 * 1 = natural, 2 or 3 = nested dissection with minimum-degree leaves of at
 * most nd_leaf rows, its vertex separators from BFS level structures (2) or
 * from multilevel bisection with node refinement (3); nd_leaf 0 = max(48,
 * n / 128) per diagonal block (2) or automatic (3: 256, or 4 on systems of
 * up to 4 096 rows when that program is estimated faster, see
 * vjx_lu_info.nd_leaf); default 3, 0.  The
 * role of UMFPACK's ordering choice (vajax/sparse/umfpack_ffi.cpp:140-179
 * caches one symbolic analysis per pattern); a new value drops the analysis
 * like a new band width. */
int vjx_pattern_set_ordering(vjx_pattern* pat, int32_t ordering, int32_t nd_leaf);
/* Fill limit of later analyses: an analysis whose L + U would hold more than
 * max_lu_nnz entries stops early with VJX_ERR_FILL and leaves the previous
 * analysis in effect (0 = no limit, the default).  The transient loop's pivot
 * monitor re-analyses under a limit of four times the current factor.  No
 * counterpart in the reference (UMFPACK's numeric factorisation re-pivots
 * within its fixed symbolic analysis, vajax/sparse/umfpack_ffi.cpp:287-316). */
int vjx_pattern_set_fill_limit(vjx_pattern* pat, int64_t max_lu_nnz);
/* Dense trailing block of later analyses: the largest trailing block of the
 * LU pattern with at least min_rows rows whose density reaches `density`
 * (0 < density <= 1) is factorised as a dense matrix on the FP64 matrix
 * cores; min_rows 0 = never (default 96, 0.4).  A change drops the analysis.
 * No counterpart in the reference (a supernode-shape choice). */
int vjx_pattern_set_dense(vjx_pattern* pat, int32_t min_rows, double density);
/* Autotuner of the launch plan: with the matrix data_dev (device), time the
 * numeric refactorisation for refactorisation bands of 12, 16 and 24 pivot
 * levels (a fresh analysis each: same ordering and pivots) and each LU
 * band-task pipeline, keep the fastest, then time the triangular solves for
 * band widths 8..64 and keep the fastest (solve results do not depend on
 * the width).  No counterpart in the reference (a launch-plan choice). */
int vjx_lu_autotune(vjx_pattern* pat, const double* data_dev);
/* Factorise / solve on the CPU with the same numeric program as the GPU path. */
int vjx_lu_refactor_cpu(vjx_pattern* pat, const double* data_host, int32_t precision,
                        double* lu_host);
int vjx_lu_solve_cpu(vjx_pattern* pat, const double* data_host, int32_t precision,
                     const double* b_host, double* x_host);
/* The dense-block kernels of the refactorisation on their own (device
 * arrays, row-major), for tests: dense LU without pivoting of a K x K matrix
 * in place (unit L below the diagonal, U on and above; panels of 32
 * columns), and C[M x N] -= A[M x Kd] B[Kd x N] on the FP64 matrix cores.
 * No counterpart in the reference. */
int vjx_dense_lu(int32_t K, double* D_dev);
int vjx_dense_gemm_sub(int32_t M, int32_t N, int32_t Kd, const double* A_dev, int32_t lda,
                       const double* B_dev, int32_t ldb, double* C_dev, int32_t ldc);

/* Copy the device LU values (FP64 or FP32 widened) to the host. */
int vjx_lu_download(vjx_pattern* pat, double* lu_host);

/* ---------------------------------------------------------------------------
 * Circuit description and system build.
 * Replaces make_mna_build_system_fn / _build_system_csr_direct /
 * _build_residual (vajax/analysis/mna_builder.py:258-1078) and the vmapped
 * device eval contract (openvaf_jax/codegen/function_builder.py:967-996).
 * ------------------------------------------------------------------------- */

#define VJX_MODEL_RESISTOR 0
#define VJX_MODEL_CAPACITOR 1
#define VJX_MODEL_MOSFET 2
#define VJX_MODEL_DIODE 3 /* junction diode with $limit pnjlim: one limit state */
#define VJX_MODEL_EKV26 4 /* EKV 2.6 MOSFET with junctions (vajax/devices/models/ekv/ekv.va) */
#define VJX_MODEL_COUNT 5
#define VJX_MAX_MODELS 8

/* Independent-source waveforms (vajax/analysis/sources.py:25-294).  Each
 * source has VJX_SRC_NPAR parameters, stored column-major [VJX_SRC_NPAR][m]:
 *   column 0 the dc value, then by waveform
 *   pulse: 1 val0, 2 val1, 3 rise, 4 fall, 5 width, 6 period, 7 delay
 *   sine:  1 sinedc, 2 ampl, 3 freq, 4 phase
 *   exp:   1 val0, 2 val1, 3 delay, 4 td2, 5 tau1, 6 tau2
 *   am/fm: 1 sinedc, 2 ampl, 3 freq, 4 phase (deg), 5 delay, 6 modfreq,
 *          7 modphase (deg), 8 modindex
 *   pwl:   1 pwlperiod; its points are pwl_t/pwl_v[pwl_ptr[i] .. pwl_ptr[i+1]),
 *          times already multiplied by stretch and values by scale plus
 *          offset (sources.py:81-106). */
#define VJX_WAVE_DC 0
#define VJX_WAVE_PULSE 1
#define VJX_WAVE_SINE 2
#define VJX_WAVE_PWL 3
#define VJX_WAVE_EXP 4
#define VJX_WAVE_AM 5
#define VJX_WAVE_FM 6
#define VJX_WAVE_COUNT 7
#define VJX_SRC_NPAR 10

/* One model group.  All arrays are column-major [k][count] (structure of
 * arrays): column k of v_node1 is the first node of branch voltage k, etc. */
typedef struct {
  int32_t model_id;
  int32_t count;
  const int32_t* v_node1; /* [n_v][count] */
  const int32_t* v_node2; /* [n_v][count] */
  const int32_t* jac_dst; /* [n_jac][count] contribution slot, -1 dropped */
  const int32_t* res_dst; /* [n_res][count] contribution slot, -1 ground */
  const double* params;   /* [n_par][count] model parameters / init cache */
} vjx_model_desc;

typedef struct {
  int32_t n_total;    /* nodes incl. ground */
  int32_t n_vsources; /* branch unknowns */
  int32_t n_isources;
  int32_t nse;        /* CSR nnz */
  const int32_t* indptr;  /* [n_aug+1] */
  const int32_t* indices; /* [nse] */
  /* per-slot constant added after device contributions: diag (gmin+gshunt)
   * flag and B/B^T value, in the reference's stamping order */
  const int32_t* slot_diag; /* [nse] 1 for node-row diagonals */
  const double* slot_b;     /* [nse] +-1 for B/B^T slots else 0 */
  const int32_t* vs_node_p; /* [m] */
  const int32_t* vs_node_n; /* [m] */
  const int32_t* vs_wave;   /* [m] VJX_WAVE_* */
  const double* vs_params;  /* [VJX_SRC_NPAR][m] */
  const int32_t* is_wave;   /* [k] */
  const double* is_params;  /* [VJX_SRC_NPAR][k] */
  /* pwl points of the voltage / current sources ([m+1] / [k+1] pointers;
   * all three NULL when no source of the kind is a pwl) */
  const int32_t* vs_pwl_ptr;
  const double* vs_pwl_t;
  const double* vs_pwl_v;
  const int32_t* is_pwl_ptr;
  const double* is_pwl_t;
  const double* is_pwl_v;
  const int32_t* is_dst;    /* [k][2] residual contribution slots of +I at p, -I at n
                               (interleaved: is_dst[2i], is_dst[2i+1]; -1 = ground) */
  int32_t n_jac_contrib;
  int32_t n_res_contrib;
  const int32_t* jac_seg;   /* [nse+1] */
  const int32_t* res_seg;   /* [n_unknowns+1] */
  /* branch-current KCL terms per node row, in segment_sum order
   * (all +I_k at p, then all -I_k at n): [n_unknowns+1] ptr, entries = k or ~k */
  const int32_t* br_ptr;
  const int32_t* br_ent;
  const int32_t* conv_mask; /* [n_unknowns] 1 = checked in residual test (may be NULL) */
  int32_t n_models;
  vjx_model_desc models[VJX_MAX_MODELS];
  /* noise-only internal nodes (solver_factories.py:37-137, 1268-1276): their
   * residual rows; in the Newton solve their rows and their columns in node
   * rows become the identity, their right-hand side 0, and they are held at
   * 0 V after each update.  n_noi may be 0 (noi_rows NULL). */
  int32_t n_noi;
  const int32_t* noi_rows;
} vjx_circuit_desc;

int vjx_circuit_create(const vjx_circuit_desc* desc, vjx_circuit** out);
void vjx_circuit_destroy(vjx_circuit* c);
/* Sparse LU handle bound to the circuit's CSR pattern (owned by the circuit). */
vjx_pattern* vjx_circuit_pattern(vjx_circuit* c);

/* Limit states of the circuit's limited devices (one per $limit call), flat
 * over model groups in order, instances, states (mna_builder.py:387-401). */
int vjx_circuit_limit_states(const vjx_circuit* c);
/* Eval workgroup size of model group `group` (desc order): 256 unless the
 * transient autotuner (vjx_tran_config.autotune) found 128 or 64 lanes 10 %
 * faster for that group; negative on a bad argument.  No counterpart in the
 * reference (a launch-plan knob). */
int vjx_circuit_eval_block(const vjx_circuit* c, int32_t group);

/* build_system (mna_builder.py:528-596): evaluate all devices at X and
 * assemble csr_data, f (augmented), Q, I_vsource and max_res_contrib; with
 * limited devices also limit_state_out from limit_state_in (NULL: zeros) and
 * the lim_rhs corrections of f (mna_builder.py:1039, mna.py:673-719).
 * Device pointers; csr_data_dev may be NULL to skip the Jacobian,
 * limit_state_out_dev may be NULL when the circuit has no limit states.
 * gmin is both the node-row diagonal regularisation and the device simparam
 * $gmin: the reference's build closure fixes the former at the options' gmin
 * (mna_builder.py:384), so pass that value here. */
int vjx_build_system(vjx_circuit* c, const double* X_dev, const double* vsrc_dev,
                     const double* isrc_dev, const double* Q_prev_dev, const double* dQdt_prev_dev,
                     const double* Q_prev2_dev, double c0, double c1, double d1, double c2,
                     double gmin, double gshunt, double* csr_data_dev, double* f_dev, double* Q_dev,
                     double* I_vsource_dev, double* max_res_contrib_dev,
                     const double* limit_state_in_dev, double* limit_state_out_dev);

/* Fused eval + stamp of one model group of a circuit (the per-group device
 * loop of build_system, mna_builder.py:927-1010, through the precomputed
 * slot map): the devices of group model_index (position in the circuit's
 * model list; -1 = every group and the current sources, i.e. build_system
 * with no charge history) are evaluated at X and their Jacobian (resistive +
 * c0 x reactive), residual, charge and max-contribution entries stamped;
 * the other groups and the current sources contribute nothing, while the
 * structural stamps remain (gmin on node diagonals, the voltage sources'
 * B / B^T entries and branch rows).  limit_state_out receives the group's
 * entries (others untouched).  csr_data_dev may be NULL. */
int vjx_eval_stamp(vjx_circuit* c, int32_t model_index, const double* X_dev,
                   const double* vsrc_dev, const double* isrc_dev, double c0, double gmin,
                   double* csr_data_dev, double* f_dev, double* Q_dev,
                   double* max_res_contrib_dev, const double* limit_state_in_dev,
                   double* limit_state_out_dev);

/* Vectorised device eval for one model group (the openvaf_jax eval_fn
 * contract, openvaf_jax/codegen/function_builder.py:967-996, restricted to
 * these models): branch voltages in, residuals and Jacobians out, row-major
 * per instance: res [count][n_res], jac [count][n_jac].  params as in
 * vjx_model_desc.  For a limited model (the diode) limit_state_in
 * [count][n_lim] (NULL: zeros) gives each $limit's old value, and
 * lim_rhs_resist / lim_rhs_react [count][n_res] and limit_state_out
 * [count][n_lim] receive the limiting outputs (each may be NULL); the
 * limiting arguments are ignored for the other models. */
int vjx_model_eval(int32_t model_id, int32_t count, const double* vbranch_dev /*[n_v][count]*/,
                   const double* params_dev, const double* limit_state_in_dev,
                   double* res_resist_dev, double* res_react_dev, double* jac_resist_dev,
                   double* jac_react_dev, double* lim_rhs_resist_dev, double* lim_rhs_react_dev,
                   double* limit_state_out_dev);

/* The openvaf_jax eval_fn contract in full (openvaf_jax/codegen/
 * function_builder.py:967-996, vmapped with in_axes = (None, 0, None, 0,
 * None, 0) at vajax/analysis/openvaf_models.py:996-997): one model group's
 * inputs come split into arrays shared by every device and per-device rows,
 * as vajax/analysis/openvaf_models.py:668-1056 prepares them.  The layout
 * says where the eval finds what it reads: each branch voltage is a column of
 * device_params (the reference's voltage_positions) and each value of the
 * model's eval input block (the rows of vjx_model_desc.params: the init
 * cache) comes from shared_cache (src >= 0) or a device_cache column
 * (src = ~column).  simparams[mfactor_index] is $mfactor, which scales every
 * residual, Jacobian and lim_rhs output (-1: 1); the models here read no
 * other simparam and no model parameter outside their cache, so
 * shared_params may be NULL.  Host layout arrays; device data arrays. */
typedef struct {
  int32_t n_device_params;          /* columns of device_params per device */
  int32_t n_device_cache;           /* columns of device_cache per device */
  const int32_t* voltage_positions; /* [n_v] device_params column of each branch voltage */
  const int32_t* cache_src;         /* [n_par] */
  int32_t mfactor_index;
} vjx_eval_split;

/* Outputs row-major per device ([count][n_res], [count][n_jac],
 * [count][n_lim]); all nine are written: lim_rhs zero for models without
 * $limit, small-signal residuals zero (these models have no small-signal-only
 * contributions), limit_state_out empty for models without limit states. */
int vjx_model_eval_split(int32_t model_id, int32_t count, const vjx_eval_split* split,
                         const double* shared_params_dev, const double* device_params_dev,
                         const double* shared_cache_dev, const double* device_cache_dev,
                         const double* simparams_dev, const double* limit_state_in_dev,
                         double* res_resist_dev, double* res_react_dev, double* jac_resist_dev,
                         double* jac_react_dev, double* lim_rhs_resist_dev,
                         double* lim_rhs_react_dev, double* ss_resist_dev, double* ss_react_dev,
                         double* limit_state_out_dev);

/* The limiting functions, elementwise on device arrays (vajax/analysis/
 * limiting.py:22-107): kind 0 = pnjlim(vnew, vold, vt = p1, vcrit = p2),
 * kind 1 = fetlim(vnew, vold, vto = p1). */
int vjx_limit_eval(int32_t kind, int32_t n, const double* vnew_dev, const double* vold_dev,
                   double p1, double p2, double* out_dev);

/* Evaluate source waveforms at time t into vals_dev. */
int vjx_eval_sources(vjx_circuit* c, double t, double* vsrc_dev, double* isrc_dev);

/* ---------------------------------------------------------------------------
 * Transient analysis on the GPU.
 * Replaces FullMNAStrategy.run and its while-loop body
 * (vajax/analysis/transient/full_mna.py:780-1196, 1840-2194), the fixed-step
 * loop (:1198-1441) and nr_solve (vajax/analysis/solver_factories.py:140-727).
 * ------------------------------------------------------------------------- */

#define VJX_METHOD_BE 0
#define VJX_METHOD_TRAP 1
#define VJX_METHOD_GEAR2 2

typedef struct {
  double t_stop;
  double dt;          /* initial step as requested (tran_fs applied inside) */
  int32_t max_steps;  /* output capacity incl. t=0 */
  int32_t method;     /* VJX_METHOD_* */
  int32_t fixed_step; /* 1: fixed-step loop */
  int32_t uic;        /* 1: icmode uic, else DC operating point */
  int32_t precision;  /* VJX_PREC_* for the LU */
  int32_t refine_steps;
  int32_t tran_itl;   /* NR iteration cap */
  int32_t use_graph;  /* capture the Newton iteration as a hipGraph */
  /* options (options.py / adaptive.py defaults shown) */
  double reltol;      /* 1e-3 */
  double abstol;      /* 1e-12 */
  double vntol;       /* 1e-6 */
  double gmin;        /* 1e-12 netlist option: diagonal regularisation (and $gmin) */
  double nr_damping;  /* 1.0 */
  double lte_ratio;   /* 3.5 */
  double redo_factor; /* 2.5 */
  double min_dt;      /* 1e-15 */
  double max_dt;      /* inf */
  int32_t tran_minpts;/* 50 */
  int32_t warmup_steps; /* 2 */
  int32_t max_order;  /* 2 */
  double grow_factor; /* 2.0 */
  double gshunt_init, gshunt_target;
  int32_t gshunt_steps; /* 5 */
  double tran_fs;     /* 0.25 */
  double tran_ft;     /* 0.25 */
  int32_t max_consecutive_rejects; /* 5 */
  const double* v_init_host; /* [n_total] initial node voltages for the DC solve */
  int32_t autotune;          /* 1: after the symbolic analysis, vjx_lu_autotune */
  double repivot_growth;     /* re-analyse when max |L| exceeds this (0: 1e12) */
  /* output rows kept on the device: 0 keeps all max_steps rows resident;
   * k > 0 keeps k rows and copies them to the host result buffers each time
   * they fill (the checkpoint mode of full_mna.py:1443-1699) */
  int32_t output_chunk;
  /* replay (adaptive loop only; n_replay > 1): take exactly the steps to
   * replay_times_host[1 .. n_replay-1] (accepted times of another run, [0] =
   * 0), with the adaptive loop's coefficients, predictor and history but no
   * accept/reject decisions; rejected_steps then counts replayed steps whose
   * Newton solve did not converge.  A test harness for comparing waveforms
   * on a common time grid; no counterpart in the reference. */
  const double* replay_times_host;
  int32_t n_replay;
  /* measurement window: statistics of the window below are gathered from
   * the first step attempt at which the accepted-step count reaches
   * window_from_step (0: the whole loop) */
  int32_t window_from_step;
} vjx_tran_config;

typedef struct {
  /* outputs, host buffers sized by the caller: [max_steps], [max_steps][n_out], ... */
  double* times;
  double* V_out;     /* node voltages (n_total columns) */
  double* I_out;     /* vsource currents (max(m,1) columns) */
  /* statistics */
  int32_t n_steps;   /* valid rows incl. t=0 */
  int32_t rejected_steps;
  int64_t total_nr_iterations;
  int64_t total_evals;
  int32_t dc_converged;
  int32_t dc_iterations;
  double min_dt_used, max_dt_used;
  double wall_time;  /* seconds, time loop only (setup and DC excluded) */
  double dc_time;
  int32_t repivots;
  /* the measurement window (window_from_step): accepted steps, rejected
   * attempts, Newton iterations and seconds between stream synchronisations
   * at its start and at the end of the loop */
  int32_t window_steps;
  int32_t window_rejected;
  int64_t window_nr_iterations;
  double window_time;
} vjx_tran_result;

int vjx_transient_run(vjx_circuit* c, const vjx_tran_config* cfg, vjx_tran_result* res);

/* One Newton solve at fixed integration coefficients: the reference's
 * nr_solve(X_init, vsource_vals, isource_vals, Q_prev, integ_c0, ..., gmin,
 * gshunt, integ_c1, integ_d1, dQdt_prev, integ_c2, Q_prev2, limit_state_in,
 * res_tol_floor) -> (X, iterations, converged, max_f, Q, dQdt, I_vsource,
 * limit_state, max_res_contrib) (vajax/analysis/solver_factories.py:604-727),
 * iterating build -> residual / delta / limit tests -> NOI -> LU refactor ->
 * solve -> damped update on the device until converged or cfg->tran_itl
 * iterations, then one more build at the final X for Q, I_vsource and
 * max_res_contrib.  Of cfg only gmin (the options' gmin: diagonal
 * regularisation of node rows, fixed as in mna_builder.py:384), reltol,
 * abstol, vntol, nr_damping, tran_itl, precision, refine_steps, use_graph,
 * autotune and repivot_growth are read.  args->gmin is the device simparam
 * $gmin (mna_builder.py:486, 649-651) that homotopy gdev stepping raises;
 * only models that read $simparam("gmin") see it, and none of the models
 * built here do.  All vectors are device buffers; NULL
 * inputs mean zeros (Q_prev, dQdt_prev, Q_prev2, limit_state_in) or abstol
 * (res_tol_floor). */
typedef struct {
  const double* X_init;          /* [n_total + m] (ground first) */
  const double* vsource_vals;    /* [m] */
  const double* isource_vals;    /* [k] */
  const double* Q_prev;          /* [n_unk] */
  const double* dQdt_prev;       /* [n_unk] */
  const double* Q_prev2;         /* [n_unk] */
  const double* limit_state_in;  /* [vjx_circuit_limit_states(c)] */
  const double* res_tol_floor;   /* [n_unk] */
  double c0, c1, d1, c2;
  double gmin, gshunt;
} vjx_newton_args;

typedef struct {
  double* X;               /* [n_total + m] */
  double* Q;               /* [n_unk] */
  double* dQdt;            /* [n_unk]: c0 Q + c1 Q_prev + d1 dQdt_prev + c2 Q_prev2 */
  double* I_vsource;       /* [max(m, 1)] */
  double* limit_state;     /* [n_lim] (may be NULL) */
  double* max_res_contrib; /* [n_unk] */
  int32_t iterations;
  int32_t converged;
  double max_f;
  int32_t repivots;
} vjx_newton_result;

int vjx_newton_solve(vjx_circuit* c, const vjx_tran_config* cfg, const vjx_newton_args* args,
                     vjx_newton_result* res);

/* Phase timing of the Newton iterations of vjx_transient_run, from HIP events
 * recorded on the library's stream (only when use_graph = 0).  Times are
 * accumulated until the next vjx_profile_reset. */
typedef struct {
  double build_ms;  /* device eval + stamping + residual */
  double lu_ms;     /* numeric refactorisation */
  double solve_ms;  /* triangular solves (+ refinement) */
  int64_t n_iterations;
  /* per kernel class, from event pairs around each launch: 0 device eval,
   * 1 stamping reductions + residual, 2 LU gather, 3 LU band tasks (or LU
   * level launches), 4 LU Schur steps, 5 LU scaling, 6 solve band tasks (or
   * solve level launches), 7 solve hub rows, 8 dense trailing block of the
   * refactorisation */
  double kernel_ms[9];
  int64_t kernel_launches[9];
} vjx_phase_times;

int vjx_profile_enable(int32_t enable);
int vjx_profile_reset(void);
int vjx_profile_read(vjx_phase_times* out);

#ifdef __cplusplus
}
#endif

#endif /* VJX_H */
