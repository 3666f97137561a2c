/*
 * CPU oracle for the fastEDM prediction core — TEST INFRASTRUCTURE ONLY.
 *
 * This library is a C++17 restatement of the reference CPU path (EDM-Developers/fastEDM-r,
 * src/manifold.*, src/distances.cpp, src/edm.cpp, src/stats.cpp, src/library_prediction_split.h,
 * src/vendor/mersennetwister.h, src/r.cpp).  It exists to check the HIP product path and to time
 * the CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and only as the checker, never as the thing measured or shipped.
 *
 * The struct layouts below are byte-identical to the ones the product declares in
 * include/fastedm_hip.h so that a test can hand the same inputs to both sides.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
  const double* data;         /* [num_points x E_actual] row-major */
  const double* targets;      /* [num_points] */
  const double* point_times;  /* [num_points] */
  const double* target_times; /* [num_points] */
  const int* panel;           /* [num_points] or NULL */
  int num_points, E_x, E_dt, E_extras, E_lagged_extras, E_actual;
} orc_manifold_view;

typedef struct
{
  int algorithm; /* 0 simplex, 1 smap, other = invalid */
  int distance;  /* 0 mean absolute error, 1 euclidean, 2 wasserstein */
  int k;         /* already adjusted by the task configurator; <0 = all */
  int num_thetas;
  const double* thetas;
  const int* metrics; /* [E_actual] 0 = Diff, 1 = CheckSame */
  double missing_distance;
  int panel_mode;
  double idw;
  int num_panels;          /* dimension of id_weights */
  const float* id_weights; /* [num_panels x num_panels] (library panel, prediction panel) or NULL */
  int has_missing, has_categorical, use_only_past, force_compute, save_smap_coeffs, low_memory;
  double aspect_ratio;
  int precision; /* ignored by the oracle */
  int nthreads;
} orc_options;

typedef struct
{
  int n;
  const double* t;
  const double* x;
  const double* y;    /* xmap variable or NULL (explore) */
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
  int distance;
  double aspect_ratio;
  int precision;
  int num_extras_lagged; /* the generator's numExtrasLagged: the first ones of the extras get E lags; the R
                            front end fixes it at 0 (src/r.cpp:246), task-group dumps may carry more */
} orc_command;

/* Task-level predictor: the edm_task fan-out over every target of Mp (reference src/edm.cpp:367-432).
 * pred/rc are [num_thetas x Mp.num_points]; coeffs [Mp.num_points x (E_actual+1)] or NULL;
 * k_used [Mp.num_points]; nn_idx [Mp.num_points x nn_cap] or NULL (selected neighbour library
 * positions in the order the predictor consumed them, -1 padded). Output buffers are pre-filled here
 * exactly as edm_task pre-fills them. */
int orc_predict(const orc_manifold_view* M, const orc_manifold_view* Mp, const orc_options* o, double* pred, int* rc,
                double* coeffs, int* k_used, int* nn_idx, int nn_cap);

/* Command-level entry: the reference run_command (src/r.cpp:107-468). Returns an opaque result. */
/* Potential neighbours of target Mp_i before the k selection (src/distances.cpp:151-268, 490-527):
 * library indices and distances, in library order; returns their count (buffers hold M->num_points). */
int orc_distances(const orc_manifold_view* M, const orc_manifold_view* Mp, const orc_options* o, int Mp_i, int* inds,
                  double* dists);
void* orc_run_command(const orc_command* c);
int orc_result_rc(void* r);
int orc_result_num_stats(void* r, int copredict);
void orc_result_stats(void* r, int copredict, int* E, int* library, double* theta, double* rho, double* mae);
void orc_result_k(void* r, int* kmin, int* kmax);
double orc_result_scalar(void* r, int which); /* 0 missingdistance, 1 dtWeight */
/* which: 0 predictions, 1 copredictions, 2 coeffs. Writes rows*cols doubles in R's to_R_matrix layout
 * (column-major, MISSING for NA). Returns 0 if absent. */
int orc_result_matrix_shape(void* r, int which, int* rows, int* cols);
void orc_result_matrix(void* r, int which, double* out);
int orc_result_num_manifolds(void* r);
void orc_result_manifold_shape(void* r, int i, int prediction_set, int* rows, int* cols);
void orc_result_manifold(void* r, int i, int prediction_set, double* out);
void orc_result_free(void* r);

/* Building blocks exposed for the tests. */
void orc_mt64_stream(uint64_t seed, int n, double* out); /* MtRng64::getReal2 stream */
/* Library/prediction masks the splitter draws for a task grid (reference configure_tasks +
 * LibraryPredictionSetSplitter); out is [num_tasks x n_usable_rows] of 0/1 library then prediction. */
int orc_split_masks(int n, const int* usable, int explore, int full, int shuffle, int crossfold, int seed, int num_reps,
                    int num_E, int num_libraries, const int* libraries, int* lib_out, int* pred_out, int max_tasks);
int orc_manifold(const orc_command* c, int E, const int* filter, int prediction_set, int copredict, double* out,
                 double* targets, int max_points);
int orc_usable(const orc_command* c, int maxE, int copredict, int* usable_out);
/* The normalised generator state (src/manifold.h:38-127): series after replace_nan, extras, observation
 * numbers, flags {dt, reldt, xmap mode, panel mode}, the dt weight and the usable rows for maxE; returns
 * the usable count.  xmap, co_x and extras may be NULL. */
int orc_generator_state(const orc_command* c, int maxE, double* t, double* x, double* xmap, double* co_x,
                        double* extras, int* obs, int* flags, double* dt_weight, int* usable_out);
/* Eigen-JacobiSVD-style solve of (X^T X) ics = X^T y for a dense n x n system: A row-major. */
void orc_jacobi_solve(int n, const double* A, const double* b, double* x, double* singular_values, int* rank);
double orc_correlation(int n, const double* a, const double* b);
double orc_mae(int n, const double* a, const double* b);
/* Exact optimal-transport cost with uniform marginals for a row-major li x lj cost matrix. */
double orc_emd(const double* C, int li, int lj);

#ifdef __cplusplus
}
#endif
