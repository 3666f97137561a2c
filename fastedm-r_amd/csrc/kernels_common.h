// Shared device-side definitions of the fastEDM HIP kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cstdint>

namespace fedm {

constexpr double MISSING_D = 1.0e+100;
constexpr double INF_D = __builtin_huge_val();
// A squared kth distance times this factor bounds every squared distance whose correctly rounded
// square root can still be below the kth distance (margin 2^-49 >> the 2^-52 rounding).
constexpr double SQRT_GUARD = 1.0 + 1.7763568394002505e-15;

constexpr int LDS_CHUNK_DOUBLES = 4096; // 32 KiB library chunk per workgroup
constexpr int LDS_CHUNK_MAXPTS = 1024;
constexpr int MAX_SLICES = 64;
constexpr int DT_TGT = 64; // targets per block of the dense distance kernel
// S-map solve of up to 24 unknowns (fedm_knobs.smap_block): per wavefront with one row slot per lane,
// block-wide with the 2x2 steps made once per system, per wavefront with two row slots per lane and
// the 2x2 steps made once per system (as many systems as leave the register-bound occupancy, or as fit),
// or the two-slot solve for 9..24 unknowns when a launch has at least SMAP_PACK_MIN_ITEMS systems and
// the one-slot solve otherwise (the two-slot kernel runs two wavefronts per SIMD, so it needs many
// groups to fill the device)
enum { SMAP_SOLVE_ROWS = 0, SMAP_SOLVE_BLOCK = 1, SMAP_SOLVE_PACK = 2, SMAP_SOLVE_PACK_MAX = 3, SMAP_SOLVE_AUTO = 4 };
constexpr long long SMAP_PACK_MIN_ITEMS = 65536;
constexpr int SMAP_BIG_MAX = 128;   // widest S-map system (unknowns = E_actual + 1) the solve takes
constexpr int SCR_CAP = 256;        // candidates recorded per target by the matrix-core screen (knn_screen)
constexpr int SMAP_BIG_BLOCKS = 256; // blocks of the wide S-map kernel (one system each at a time)

enum
{
  RC_SUCCESS = 0,
  RC_BREAK_HIT = 1,
  RC_INVALID_ALGORITHM = 400,
  RC_INSUFFICIENT_UNIQUE = 503,
  RC_UNKNOWN_ERROR = 8000
};

// A target left without a prediction (too few valid neighbours, k = 0): every theta's prediction
// missing, the code at theta 0 and unknown-error at the others, k_used missing — the values the
// reference's outputs are filled with before its per-target calls (src/edm.cpp:338-432), so that the
// kernels leave no output of the target to an earlier fill.
__device__ __forceinline__ void write_unpredicted(double* pred, int* rc, int* k_used, int64_t ob, int T, int np,
                                                  int t, int code)
{
  for (int th = 0; th < T; th++) {
    pred[ob + (int64_t)th * np + t] = 1.0e+100;
    rc[ob + (int64_t)th * np + t] = th == 0 ? code : (int)RC_UNKNOWN_ERROR;
  }
  k_used[ob / T + t] = (int)(-2147483647 - 1);
}
enum
{
  ALG_SIMPLEX = 0,
  ALG_SMAP = 1
};

// Distance option bits (src/distances.cpp:151-229 template switches and Options fields).
enum
{
  F_MISSING = 1,
  F_CATEGORICAL = 2,
  F_SKIP_OTHER = 4,
  F_SKIP_FUTURE = 8,
  F_MAE = 16,
  F_IDW = 32,
  F_IDWEIGHTS = 64
};

struct LpParams
{
  const double* M;     // library points [NL x Ea]
  const double* Mp;    // prediction points [NP x Ea]
  const double* Mtt;   // library target times [NL] (skip-future only)
  const double* Mppt;  // prediction point times [NP]
  const int* Mpanel;   // [NL] or null
  const int* Mppanel;  // [NP] or null
  const int* metrics;  // [Ea]
  const float* id_weights;
  int NL, NP, Ea, num_panels;
  double missing_distance, idw;
  int flags;
};

// Wasserstein distances (src/distances.cpp:296-527).
struct WassParams
{
  const double* M;        // [NL x Ea]
  const double* Mp;       // [NP x Ea]
  const int* Mpanel;      // [NL] or null
  const int* Mppanel;     // [NP] or null
  const int* metrics;     // one per variable: x, dt (when E_dt > 0), then the extras
  const float* id_weights;
  int NL, NP, Ea, E, E_dt, rows, n_unlagged, num_panels;
  double missing_distance, idw, aspect_ratio;
  int panel_term;         // 0 none, 1 idw penalty, 2 id_weights table
  int skip_other;
};

struct FinishParams
{
  int NP, NL, k, T, force_compute, algorithm;
  const double* thetas;  // [T]
  const double* targets; // library targets [NL]
  double* pred;          // [T x NP]
  int* rc;               // [T x NP]
  int* k_used;           // [NP]
  int* nn_k;             // [NP]: >0 list length, -1 dense (all valid), 0 none
  int* nn_cnt;           // [NP]: valid count on the dense route
  int* nn_i;             // [NP x nn_stride]
  double* nn_s;          // [NP x nn_stride]
  int nn_stride;
  int* nn_out;           // [NP x nn_cap] or null
  int nn_cap;
};

// Lexicographic (s, idx) order of the reference's k-NN (src/edm.cpp:663-668).
__device__ __forceinline__ bool key_less(double s1, int i1, double s2, int i2)
{
  return s1 < s2 || (s1 == s2 && i1 < i2);
}

struct SmapParams
{
  int NP, NL, Ea, T;
  const double* M;
  const double* Mp;
  const double* targets;
  const double* thetas;
  const int* nn_k;
  const int* nn_cnt;
  const int* nn_i;   // list of target t at row t - nn_t0
  const double* nn_s;
  int nn_stride;
  int nn_t0;
  double* pred;
  int* rc;
  double* coeffs; // [NP x (Ea+1)] or null
  double* big;    // smap_big: [blocks x 2 n^2] scratch for the transposed U and V of each block's system
  int big_blocks; // blocks that scratch holds
  const double* gram; // gram_dense: [items x (n^2 + n)] X^T X and X^T y of the dense block's all-neighbour
                      // items (nn_k = -1), or null; the solve kernels copy those instead of forming them
};

// Matrix-core k-NN (knn_mfma): FP32 screening operands, the FP64 rows of the exact re-rank and the
// screening bound's constants.
struct MfmaParams
{
  const float* M32;   // [NL x KP] centred FP32 library rows, zero padded
  const float* Mn;    // [NL] FP32 squared norms of those rows
  const float* P32;   // [NP x KP] centred FP32 targets
  const float* Pn;    // [NP]
  const double* M;    // FP64 library [NL x Ea]
  const double* Mp;   // FP64 targets [NP x Ea]
  const double* ynorm; // [NP] |centred target| (FP64)
  double R;            // largest |centred library row|
  double k1, k3;       // bound: coefficient of sqrt(D) per unit rr, constant term per unit rr^2
  double c2;           // bound: coefficient of D
  const double* seed;  // [NP] an exact upper bound on each target's k-th smallest nonzero squared distance
                       // (+inf: none), from knn_seed; null: the screening starts unbounded
  int NL, NP, Ea;
};

// The library and targets in the k-d order of the seeded matrix-core route, as the two-phase screen
// (knn_screen) walks them: library rows in 64-row chunks, targets in 128-target blocks, each with a box.
struct KdScreen
{
  const float* m32;  // [NL x KP] the centred FP32 library rows in k-d order
  const float* mn;   // [NL] their FP32 squared norms
  const int* pos;    // [NL] library position of each row of the order
  const float* cbox; // [ceil(NL / 64) x 2 Ea] per chunk of the order: lower then upper bounds of the FP64
                     // coordinates, rounded outward to FP32
  const float* tbox; // [ceil(NP / 128) x 2 Ea] the same per block of the target order
  const int* tord;   // [NP] targets in k-d window order
  int prune;         // 1: a chunk whose box lies beyond every threshold of a block is not screened
};

} // namespace fedm

#include "jacobi_wave.h"
