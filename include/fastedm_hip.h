/*
 * fastedm_hip.h — C ABI of the MI355X prediction core of fastEDM (libfastedm_hip.so).
 *
 * Plain pointers and sizes only; nothing is thrown across this boundary.  Every entry point
 * returns 0 on success, a positive EDM retcode (1 = BREAK_HIT) when interrupted through keep_going,
 * or a negative value on a device or argument error (the per-target rc outputs are then 8000).
 *
 * Conventions shared with the reference (EDM-Developers/fastEDM-r):
 *   missing value   FEDM_MISSING_D = 1e100        (src/manifold.h:6)
 *   missing integer FEDM_MISSING_I = INT_MIN      (src/manifold.h:8)
 *   retcodes        0 SUCCESS, 1 BREAK_HIT, 400 INVALID_ALGORITHM, 503 INSUFFICIENT_UNIQUE,
 *                   8000 UNKNOWN_ERROR            (src/common.h:3-13)
 */
#ifndef FASTEDM_HIP_H
#define FASTEDM_HIP_H

#include <limits.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FEDM_MISSING_D 1.0e+100
#define FEDM_MISSING_I INT_MIN

enum
{
  FEDM_SUCCESS = 0,
  FEDM_BREAK_HIT = 1,
  FEDM_INVALID_ALGORITHM = 400,
  FEDM_INSUFFICIENT_UNIQUE = 503,
  FEDM_UNKNOWN_ERROR = 8000
};
enum
{
  FEDM_SIMPLEX = 0,
  FEDM_SMAP = 1
};
enum
{
  FEDM_DIST_MAE = 0,
  FEDM_DIST_EUCLIDEAN = 1,
  FEDM_DIST_WASSERSTEIN = 2
};
enum
{
  FEDM_PLAIN = 0, /* FP64 throughout: predictions to 1e-10 of the CPU path, neighbour indices bit-exact */
  FEDM_MIXED = 1, /* FP32/matrix-core distance screening with an exact FP64 re-rank (neighbours and
                     predictions identical to plain); the S-map Gram and solve are FP64 in both forms */
  FEDM_PLAIN_TASKWISE = 2 /* command level only: plain, one edm_predict per task (the drop-in seam) */
};

/* A manifold as the reference's Manifold exposes it (src/manifold.h:159-286). */
typedef struct
{
  const double* data;         /* [num_points x E_actual] row-major */
  const double* targets;      /* [num_points] */
  const double* point_times;  /* [num_points] */
  const double* target_times; /* [num_points] */
  const int* panel;           /* [num_points] or NULL */
  int num_points, E_x, E_dt, E_extras, E_lagged_extras, E_actual;
} edm_manifold_view;

/* The fields of the reference's Options (src/common.h:64-91) that the predictor reads. */
typedef struct
{
  int algorithm;       /* FEDM_SIMPLEX / FEDM_SMAP; anything else gives rc INVALID_ALGORITHM */
  int distance;        /* FEDM_DIST_* */
  int k;               /* neighbours, already set by the task configurator (src/edm.cpp:155-163); <0 = all */
  int num_thetas;
  const double* thetas; /* [num_thetas] */
  const int* metrics;   /* 0 = Diff, 1 = CheckSame, expanded as src/distances.cpp:29-67 does, or NULL:
                           L^p: one per coordinate [E_actual]; Wasserstein: one per variable
                           [1 + (E_dt > 0) + number of extras] */
  double missing_distance;
  int panel_mode;
  double idw;              /* panel mismatch penalty; <0 = skip other panels */
  int num_panels;          /* dimension of id_weights */
  const float* id_weights; /* [num_panels x num_panels] keyed (library panel, prediction panel); panel ids
                              must then lie in [0, num_panels). NULL = none */
  int has_missing, has_categorical, use_only_past, force_compute, save_smap_coeffs, low_memory;
  double aspect_ratio; /* Wasserstein time scaling */
  int precision;       /* FEDM_PLAIN / FEDM_MIXED */
  int nthreads;        /* unused by the GPU path */
} edm_options;

typedef struct edm_ctx edm_ctx;

/* Device context: one HIP stream and a grow-only pool of device buffers reused across tasks.
 * Replaces the reference's per-task af::setDevice + toGPU uploads (src/edm.cpp:320-332). */
int edm_ctx_create(edm_ctx** ctx, int device);
void edm_ctx_destroy(edm_ctx* ctx);

/*
 * edm_predict — every prediction of one task on the GPU.
 * Replaces af_make_prediction (src/edm.cpp:1058-1061), called from edm_task (src/edm.cpp:401-408) in
 * place of the per-target make_prediction loop (src/edm.cpp:367-428); semantics are those of the
 * CPU path make_prediction (src/edm.cpp:565-814).
 *   pred   [num_thetas x Mp->num_points] row-major (MISSING_D where no prediction)
 *   rc     [num_thetas x Mp->num_points]
 *   coeffs [Mp->num_points x (E_actual+1)] or NULL (S-map coefficients of the last theta)
 *   k_used [Mp->num_points] (MISSING_I where no neighbours were used)
 *   nn_idx [Mp->num_points x nn_cap] or NULL: the chosen library positions in the order the
 *          predictor used them, -1 padded (not part of the reference's seam; used for parity checks)
 *   keep_going: polled between stages, may be NULL.
 * Returns 0, -22 for invalid arguments, or -95 (outputs filled, rc 8000) for a width this build does
 * not take: E_actual above 4096, an S-map with E_actual above 127 (128 unknowns), Wasserstein with
 * E_x above 64.
 */
int edm_predict(edm_ctx* ctx, const edm_manifold_view* M, const edm_manifold_view* Mp, const edm_options* o,
                double* pred, int* rc, double* coeffs, int* k_used, int* nn_idx, int nn_cap, int (*keep_going)(void));

/*
 * edm_predict_noalloc — edm_predict for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 *   M, Mp: every array in the views (data, targets, times, panel) is a DEVICE pointer;
 *   pred [num_thetas x Np], rc [num_thetas x Np], k_used [Np], coeffs [Np x (E_actual+1)] or NULL: DEVICE
 *   pointers; o->thetas, o->metrics and o->id_weights stay host arrays.
 * Results equal edm_predict's.  The kernels' scratch lives in a static 32 MiB device array (one copy per
 * device, held until the process ends), so the call allocates nothing; a task whose scratch would not
 * fit (about Np x k x 24 bytes on the list routes, at least Nl x 8 bytes on the dense ones) returns -12
 * with a message on stderr.  The matrix-core screening is not used (FP64 routes only).  Calls that
 * overlap on one device race on the shared array: make them from one thread or one context at a time.
 * Synchronous on ctx's stream.  Returns 0, -22 (invalid arguments), -95 (a width this build does not
 * take; outputs untouched), -12, or a HIP error.
 */
int edm_predict_noalloc(edm_ctx* ctx, const edm_manifold_view* M, const edm_manifold_view* Mp, const edm_options* o,
                        double* pred, int* rc, double* coeffs, int* k_used);

/*
 * edm_predict_batch — many tasks that index one shared point table, in one persistent launch per
 * stage (CCM library-size x replicate sweeps, E/theta grids).  Task t uses library rows
 * lib_rows[lib_begin[t] .. lib_begin[t]+lib_count[t]) (ascending, rows of L) and prediction rows
 * pred_rows[pred_begin[t] .. +pred_count[t]) (rows of P); ranges may be shared between tasks.
 * The per-task results are those edm_predict gives for the manifolds these rows select
 * (the reference's per-task edm_task, src/edm.cpp:285-548, run serially at src/edm.cpp:249-251).
 *   k[t]: neighbours of task t (as edm_options.k); o->k is ignored.
 *   pred/rc: task t writes [num_thetas x pred_count[t]] at out_begin[t]; k_used at kbegin[t]=out_begin[t]/num_thetas.
 * The library table L carries the library-side targets/times; P the prediction-side points.
 */
int edm_predict_batch(edm_ctx* ctx, const edm_manifold_view* L, const edm_manifold_view* P, const edm_options* o,
                      int num_tasks, const int* lib_begin, const int* lib_count, const int* lib_rows,
                      const int* pred_begin, const int* pred_count, const int* pred_rows, const int* k,
                      const int64_t* out_begin, double* pred, int* rc, int* k_used, int (*keep_going)(void));

/* Launch and route choices that leave every result unchanged (the autotuner's search space; no
 * reference counterpart).  fedm_default_knobs fills the defaults (FEDM_RANK_LTHR_SCALE and
 * FEDM_SMAP_LDS_ROWS in the environment override the first and the last). */
typedef struct
{
  double rank_scale;        /* the rank-table route takes libraries of >= rank_scale sqrt(k npool) rows (1.5) */
  int direct_blocks_per_cu; /* persistent direct kernel: blocks per CU, 0 = as many as fit */
  int direct_alt_waves;     /* persistent direct kernel: 1 = the instance compiled for fewer waves per SIMD */
  int list_blocks;          /* FP64 list route: library slices chosen for about this many blocks (1024) */
  int mfma_blocks;          /* matrix-core list route: the same (2048) */
  int use_mfma;             /* mixed form may screen on the matrix cores (E_actual >= 8) (1) */
  int smap_lds;             /* S-map solve keeps U and V in LDS for every width (0) */
  int direct_prune;         /* persistent route: libraries of 64..512 rows (points of <= 8 coordinates) take the
                               spatially pruned direct kernel (1) */
  int concurrent_routes;    /* persistent route: the direct kernels run on a second stream beside the
                               rank-table kernels (0) */
  int smap_block;           /* S-map solve of up to 24 unknowns: one row per lane, every lane of a system
                               making the 2x2 steps (0); blocks of four wavefronts, each 2x2 step made once
                               per system (1); two rows per lane, each 2x2 step made once per system, as
                               many systems per wavefront as keep the register-bound occupancy (2) or as
                               fit (3); (2) for 9..24 unknowns and launches of at least 65 536 systems,
                               else (0) (4).  All give the same bits. (4) */
  int mfma_seed;            /* matrix-core route: each target's screening threshold starts unbounded (0), or
                               from an exact k-th distance over library rows near it in a k-d order of the
                               library, 4 096 rows for points of more than 16 coordinates, else 256 (1); or (2) a screen against the k-th over 256 such rows records
                               candidates, one exact FP64 re-rank per target follows, and a target with more
                               than 256 candidates takes a k-th over 1 024 rows and the streaming kernel;
                               (3) (2) for points of up to 16 coordinates and up to 32 neighbours, else (1).
                               All give the same results. (3) */
  int mfma_prune;           /* two-phase form: the screen of a block of targets skips the 64-row chunks of the
                               library's k-d order whose bounding box lies beyond every target's threshold
                               (1), or screens every chunk (0).  Same results. (1) */
  int mfma_split;           /* two-phase form: blocks that share out the library chunks of each block of 128
                               targets (0: about 16 384 blocks in all, at most 16).  Same results. (0) */
} fedm_knobs;
void fedm_default_knobs(fedm_knobs* k);

/* The same batch in three steps, so that the device-resident part can be timed or replayed:
 * edm_batch_create uploads the tables, row lists and work queue once; edm_batch_run makes every
 * prediction on the device (inputs resident in HBM, results left there); edm_batch_fetch copies
 * pred/rc/k_used back in the layout of edm_predict_batch.  Arguments as edm_predict_batch. */
typedef struct edm_batch edm_batch;
int edm_batch_create(edm_ctx* ctx, const edm_manifold_view* L, const edm_manifold_view* P, const edm_options* o,
                     int num_tasks, const int* lib_begin, const int* lib_count, const int* lib_rows,
                     const int* pred_begin, const int* pred_count, const int* pred_rows, const int* k,
                     const int64_t* out_begin, edm_batch** out);
/* edm_batch_create with explicit knobs (NULL = the defaults); edm_batch_knobs reports a batch's. */
int edm_batch_create_knobs(edm_ctx* ctx, const edm_manifold_view* L, const edm_manifold_view* P,
                           const edm_options* o, int num_tasks, const int* lib_begin, const int* lib_count,
                           const int* lib_rows, const int* pred_begin, const int* pred_count, const int* pred_rows,
                           const int* k, const int64_t* out_begin, const fedm_knobs* knobs, edm_batch** out);
void edm_batch_knobs(const edm_batch* b, fedm_knobs* out);
int edm_batch_run(edm_batch* b);
int edm_batch_fetch(edm_batch* b, double* pred, int* rc, int* k_used);
void edm_batch_destroy(edm_batch* b);
/* keep_going for the next edm_batch_run (may be NULL): polled before each task on the task-by-task
 * route and before each stage launch on the device route, as edm_task polls it between targets
 * (src/edm.cpp:377, 419).  A stop returns FEDM_BREAK_HIT; the outputs then hold missing predictions,
 * rc BREAK_HIT and missing k_used until the next complete run. */
void edm_batch_set_keep_going(edm_batch* b, int (*keep_going)(void));
/* 1 when the batch runs as one persistent device launch (simplex, Euclidean, no missing/panel
 * options, E_actual <= 32, 1 <= k <= 32), 0 when it runs task by task. */
int edm_batch_on_device(const edm_batch* b);
/* How the batch runs: 2 one persistent launch per stage; 1 the resident task route (every other
 * option set: each task's manifolds gathered and uploaded once at create, its kernels launched in
 * turn on each run, nothing copied from the host); 0 task by task through edm_predict, with a host
 * gather and upload per task on each run (only when the inputs would exceed 16 GiB or a task needs a
 * width this build does not take). */
int edm_batch_route(const edm_batch* b);
/* With on != 0, the second and later runs of the batch replay its launches as one captured hipGraph
 * (the first run executes them directly; routes 1 and 2 only).  Stage timings are then not
 * recorded (edm_ctx_timings gives the whole run).  Returns 1 when graphs are on. */
int edm_batch_set_graph(edm_batch* b, int on);
/* S-map coefficients of the last theta (routes 1 and 0 save them when the options ask for
 * them): [prediction rows x (E_actual + 1)] per task at out_begin[t] / num_thetas * (E_actual + 1),
 * missing elsewhere. */
int edm_batch_fetch_coeffs(edm_batch* b, double* coeffs);

/* Smallest library size that a batch sends to the rank-table route for k neighbours over a pool of
 * npool rows (scale from FEDM_RANK_LTHR_SCALE, default 1.5 * sqrt(k * npool)). */
int fedm_rank_threshold(int kmax, int npool);

/* Diagnostic: the lane layout of the FP32 matrix-core instruction the mixed route uses (a 16x16x4
 * product with A[i][0] = i+1, B[0][j] = 100 (j+1)); out256[lane*4 + r] is D[4 (lane/16) + r][lane%16]
 * on a conforming device.  Returns 0 when the probe ran. */
int edm_selftest_mfma(float* out256);

/* Diagnostic: out[i] = the device exp the simplex and S-map weights use, at x[i] (FDLIBM's method; the
 * CPU oracle evaluates the same expression).  Returns 0 when the probe ran. */
int edm_selftest_exp(const double* x, double* out, int n);

/* Last stage timings of the most recent call on ctx (milliseconds, HIP events on the ctx streams).
 * After edm_predict or a resident batch run: [0] distance/selection kernel, [1] merge/predict kernel,
 * [2] S-map kernel, [3] whole device section.  On the dense route (k = all, or k > 64) with one block
 * of targets: [0] the distance rows, [1] the neighbour policy and wide selection, [2] the S-map
 * systems and solves; with several blocks, [1] spans them all.
 * After a persistent batch run: [0] the whole run, [1] the rank-table build, [2] the rank scan, [3] the
 * direct kernels. */
void edm_ctx_timings(edm_ctx* ctx, double* ms4);
/* The same with the persistent run's direct stage split: [4] the direct kernel, [5] the spatially
 * pruned direct kernel, [6] the exact re-selection, [7] 1 when the direct kernels ran on the second
 * stream beside the rank-table kernels (their spans then overlap), else 0. */
void edm_ctx_timings_ex(edm_ctx* ctx, double* ms8);

/* Diagnostic: how many (task, target) pairs the last batched run on ctx re-selected with the
 * sqrt-ordered list because two different squared distances near the k-th place could share a
 * rounded square root (no reference counterpart; the results do not depend on it). */
int edm_ctx_exact_redos(const edm_ctx* ctx);

/* ---------------------------------------------------------------------------------------------
 * Command level: the reference's run_command (src/r.cpp:107-468), i.e. what the R function edm()
 * (R/edm.R:232-393) calls, with the predictions made on the GPU.
 */
typedef struct
{
  int n;
  const double* t;
  const double* x;
  const double* y;    /* xmap variable, or NULL for explore */
  const double* co_x; /* copredict variable or NULL */
  const int* panel;   /* or NULL */
  int num_extras;
  const double* extras; /* [num_extras x n] */
  int num_E;
  const int* E;
  int tau;
  int num_thetas;
  const double* thetas;
  int num_libraries;
  const int* libraries;
  int k, algorithm, num_reps, p, crossfold, full, shuffle, seed;
  int save_final_predictions, save_final_copredictions, save_manifolds, save_smap_coeffs;
  int dt, reldt;
  double dt_weight;
  int allow_missing;
  double missing_distance, panel_weight;
  const double* panel_weights; /* [P x P] row-major, or NULL */
  int panel_weights_dim;
  int num_threads, low_memory, predict_with_past;
  int distance;        /* the R front end always passes FEDM_DIST_EUCLIDEAN (src/r.cpp:156) */
  double aspect_ratio;
  int precision;       /* FEDM_PLAIN / FEDM_MIXED */
  int (*keep_going)(void); /* or NULL: polled between groups, tasks and stage launches, as launch_tasks and
                              edm_task poll theirs (src/edm.cpp:134, 377, 419); a stop makes the run
                              return FEDM_BREAK_HIT and the result hold rc BREAK_HIT and no statistics */
} fedm_command;

typedef struct fedm_result fedm_result;
typedef struct fedm_plan fedm_plan;

/* The command in three steps: fedm_plan_create normalises the inputs, draws every library/prediction
 * split, builds one point table per (E, copredict) and uploads it with the tasks' row lists as
 * device batches; fedm_plan_run makes every prediction on the device (repeatable: inputs stay
 * resident); fedm_plan_result copies the predictions back and computes rho/MAE per task as
 * edm_task does (src/edm.cpp:440-467). */
int fedm_plan_create(edm_ctx* ctx, const fedm_command* c, fedm_plan** out);
int fedm_plan_run(fedm_plan* p);
int fedm_plan_result(fedm_plan* p, fedm_result** out);
int64_t fedm_plan_num_predictions(const fedm_plan* p);
int fedm_plan_device_batches(const fedm_plan* p);   /* groups that run as one persistent launch */
int fedm_plan_resident_batches(const fedm_plan* p); /* groups on the resident task route */
/* The plan's tasks in the reference's task order (src/edm.cpp:128-213), for parity checks (no reference
 * counterpart).  fedm_plan_task: info[7] = {E_actual, library, k, library rows, prediction rows,
 * copredict, E}; lib_rows / pred_rows (or NULL) receive the table rows the task uses.
 * fedm_plan_fetch copies every task's outputs of the last run back in task order: pred and rc
 * [num_thetas x prediction rows] per task, k_used [prediction rows] per task, concatenated. */
int fedm_plan_num_tasks(const fedm_plan* p);
/* Graph replay (edm_batch_set_graph) for every group of the plan; returns the groups that have it. */
int fedm_plan_set_graph(fedm_plan* p, int on);
/* The autotuner (no reference counterpart).  For every group of the plan it searches the knobs of the
 * group's route one at a time (coordinate descent): the persistent route's rank-table scale and direct
 * kernel instance and occupancy; the resident route's slice targets, matrix-core screening and S-map
 * layout.  Each candidate batch is built, run `trials` times after a warm-up (median device time), and
 * kept only if its outputs equal the default's bit for bit.  The fastest replaces the group's batch;
 * the choice is written to the tuning cache (FEDM_TUNE_CACHE, else $HOME/.cache/fastedm_hip/tune.txt)
 * under the group's workload class (route-independent: algorithm, options, E_actual, largest k,
 * thetas, and the log2 of the task count, longest library and prediction rows), which later plans
 * read when they create their batches.  A text report of every measured alternative goes to `report`.
 * Returns the number of groups tuned, or a negative error. */
int fedm_plan_tune(fedm_plan* p, int trials, char* report, int report_len);
/* The knobs group `group` (in (E, copredict) order of first use) of the plan runs with; -22 if it has no batch. */
int fedm_plan_group_knobs(const fedm_plan* p, int group, fedm_knobs* out);
int fedm_plan_task(const fedm_plan* p, int task, int* info, int* lib_rows, int* pred_rows);
int fedm_plan_fetch(fedm_plan* p, double* pred, int* rc, int* k_used);
void fedm_plan_destroy(fedm_plan* p);

/* fedm_plan_create + fedm_plan_run + fedm_plan_result. */

int fedm_run_command(edm_ctx* ctx, const fedm_command* c, fedm_result** out);

/* ---------------------------------------------------------------------------------------------
 * A task group as the reference's saveInputs dump records it (src/r.cpp:275-309): the manifold
 * generator's state after its normalisation (src/manifold.cpp:364-402), the options that the dump
 * keeps (src/common.cpp:62-106) and the task-grid arguments of launch_tasks (src/edm.cpp:90-215).
 * The dump does not keep the splitter seed (replays pass it), allowMissing (taken as
 * missing_distance != 0: the R front end sets a positive default whenever it allows missing values,
 * and otherwise usable points hold no missing coordinate), lowMemory and predictWithPast (taken as
 * off) or panel weight matrices (only idw applies).
 */
typedef struct
{
  int n;
  const double* t;          /* _t */
  const double* x;          /* _x */
  const double* xmap;       /* _xmap (when _xmap_mode), or NULL */
  const double* co_x;       /* _co_x (copredict mode), or NULL */
  const int* obs;           /* _observation_number */
  const int* panel;         /* _panelIDs (when _panel_mode), or NULL */
  const double* extras;     /* _extras, [num_extras x n] */
  int num_extras, num_extras_lagged; /* the first num_extras_lagged extras enter a point with E lags each
                                        (src/manifold.cpp:225-233) */
  int dt, reldt;            /* _dt, _reldt */
  double dt_weight;         /* _dtWeight */
  int tau, p;               /* _tau, _p */
  int algorithm;            /* opts.algorithm: FEDM_SIMPLEX / FEDM_SMAP */
  int distance;             /* opts.distance: FEDM_DIST_* */
  const double* thetas;     /* opts.thetas */
  int num_thetas;
  double missing_distance;  /* opts.missingdistance */
  int panel_mode;           /* opts.panelMode */
  double idw;               /* opts.idw */
  const int* extra_metrics; /* opts.metrics, one per extra: 0 Diff, 1 CheckSame; or NULL */
  double aspect_ratio;      /* opts.aspectRatio */
  int save_smap_coeffs;     /* saveSMAPCoeffs */
  const int* E;             /* Es */
  int num_E;
  const int* libraries;
  int num_libraries;
  int k, num_reps, crossfold, explore, full, shuffle;
  int save_final_predictions, save_final_copredictions, copredict_mode;
  const int* usable;        /* [n] */
  int seed;
  int precision;            /* FEDM_PLAIN / FEDM_MIXED / FEDM_PLAIN_TASKWISE */
} fedm_task_group;

/* Runs one dumped task group as launch_tasks would (no reference counterpart: the reference only
 * writes the dumps); results as fedm_run_command's. */
int fedm_run_task_group(edm_ctx* ctx, const fedm_task_group* g, fedm_result** out);
int fedm_result_rc(const fedm_result* r);
int fedm_result_num_stats(const fedm_result* r, int copredict);
void fedm_result_stats(const fedm_result* r, int copredict, int* E, int* library, double* theta, double* rho,
                       double* mae);
void fedm_result_k(const fedm_result* r, int* kmin, int* kmax);
double fedm_result_scalar(const fedm_result* r, int which); /* 0 missingdistance, 1 dtWeight */
/* which: 0 predictions, 1 copredictions, 2 coeffs; R's to_R_matrix layout (column-major). */
int fedm_result_matrix_shape(const fedm_result* r, int which, int* rows, int* cols);
void fedm_result_matrix(const fedm_result* r, int which, double* out);
int fedm_result_num_manifolds(const fedm_result* r);
void fedm_result_manifold_shape(const fedm_result* r, int i, int prediction_set, int* rows, int* cols);
void fedm_result_manifold(const fedm_result* r, int i, int prediction_set, double* out);
void fedm_result_free(fedm_result* r);

#ifdef __cplusplus
}
#endif
#endif
