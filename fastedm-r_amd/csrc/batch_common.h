// Parameter blocks of the batched launches (direct kernel, rank-table build and scan), shared by the
// host dispatcher and the kernels.
#pragma once
#include "kernels_common.h"
#include "fdlibm_exp.h"

namespace fedm {

struct BatchParams
{
  const double* L;      // library-side point table [NLrows x Ea] (FP64)
  const double* Lg;     // the direct kernel's tasks' library rows, gathered per task [.. x Ea] at gbeg[task]
  const double* Tg;     // their targets, gathered the same way
  const int64_t* gbeg;  // [tasks] first gathered row of a direct-kernel task
  const double* Pg;     // prediction rows in prediction-list order [pred_total x Ea]
  const float* L32;     // FP32 copy of L (MIXED)
  const double* Ltgt;   // library-side targets [NLrows]
  const double* P;      // prediction-side point table [NProws x Ea]
  const int* lib_begin; // [tasks]
  const int* lib_count;
  const int* lib_rows;
  const int* pred_begin;
  const int* pred_count;
  const int* pred_rows;
  const int* k;          // [tasks]
  const int64_t* out_begin;
  const int2* items;     // (task, first target)
  int num_items;
  int* counter;          // work-queue head
  const double* thetas;
  int T, Ea, force_compute;
  double c1, c2, c3;     // mixed-precision error bound B(D) = c1 sqrt(D) + c2 D + c3
  double c4;             // bound on the rounding of the direct kernel's expanded FP64 screen
  double* pred;
  int* rc;
  int* k_used;
  int2* redo;            // (task, target) whose squared-distance ranking may differ from the sqrt ranking
  int* redo_count;       // entries appended (may exceed redo_cap: then every target is redone)
  int redo_cap;
  // spatially pruned direct kernel (batch_simplex_pruned)
  const int* tperm;      // spatial order of each prediction list: tperm[pred_begin + j] = list position
  const int2* pitems;    // (task, first target in spatial order), grouped by XCD queue
  int num_pitems;
  int* pcounter;         // work-queue heads, one per queue, PR_CSTRIDE ints apart
  const int* pq_begin;   // [queues + 1] item range of each queue
  int pq_n;              // queues (1 .. PR_QUEUES)
};

// The pruned kernel's items are dealt to one queue per XCD, a task's items all to one queue, so that a
// task's scattered output writes meet in one L2 (blocks are dispatched to the XCDs in turn, so block b
// serves queue b % PR_QUEUES); a block whose queue is empty takes items from the others.
constexpr int PR_QUEUES = 8;
constexpr int PR_CSTRIDE = 32; // ints between queue heads (128 bytes: separate cache lines)

// Libraries of PR_MIN..PR_BCH rows take the pruned direct kernel (when the knob allows it).
constexpr int PR_MIN = 64;
constexpr int SUB0_ROWS = 16; // rows of a library's first block in the direct kernels
constexpr int PR_BCH = 512;
constexpr int PR_EMAX = 8;  // widest point it takes

constexpr int batch_waves(int kcap, int emax) { return kcap <= 4 && emax <= 4 ? 6 : 1; }
constexpr int batch_waves_alt(int kcap, int emax) { return kcap <= 4 && emax <= 4 ? 5 : 2; }
template <int KCAP, int EMAX, int WPE>
__global__ void batch_simplex(BatchParams B);
template <int KCAP, int EMAX>
__global__ void batch_simplex_pruned(BatchParams B);

constexpr int RANK_STORE = 1024; // longest stored prefix
constexpr int RANK_ZCAP = 8;     // rows at distance 0 kept per target
constexpr int MS_WAVES = 1;      // wavefronts per block of rank_scan_multi (its LDS, 8 KiB a wavefront, sets occupancy)

struct RankParams
{
  const double* L;   // point table [rows x Ea] FP64
  const float* L32;  // FP32 copy
  const double* P;   // prediction-side table
  const int* pool;   // pool rows [npool], in a k-d order of their points
  const float* pool32; // FP32 coordinates of the pool rows, in pool order, zero padded to EMAX [npool x EMAX]
  const float* cbox;   // per 64-row chunk of the pool order: Ea lower then Ea upper bounds of the FP64
                       // coordinates, rounded outward (null: every chunk is screened)
  int npool;
  const int* pred;   // prediction rows [np]
  int np, Ea;
  double c1, c2, c3; // FP32 screening bound
  int* rank_row;     // entry j of target t at t * tstride + j * jstride (table rows)
  double* rank_s;    // same layout
  int* rank_len;     // [np]
  int* zrow;         // [RANK_ZCAP x np]
  int* zlen;         // [np]: count, or -1 when more than RANK_ZCAP rows are at distance 0
  int tstride, jstride;
};

struct ScanParams
{
  const double* L;       // point table FP64
  const double* Ltgt;    // library-side targets by row
  const double* P;
  const int* pred;       // prediction rows of the shared list
  int np, Ea, T, force_compute, mask_words;
  const unsigned* masks; // [tasks x mask_words] library membership by table row
  const int* lib_begin;
  const int* lib_count;
  const int* lib_rows;
  const int* k;
  const int64_t* out_begin;
  const int2* items;     // (task, first target)
  int num_items;
  int* counter;
  const int* rank_row;
  const double* rank_s;
  const int* rank_len;
  const int* zrow;
  const int* zlen;
  int tstride, jstride;  // rank table layout
  const double* thetas;
  double* pred_out;
  int* rc;
  int* k_used;
  // one pass per target for every rank task (rank_scan_multi)
  const unsigned* member;    // [table rows x words]: bit b of word w set when rank task 32 w + b has the row
  int words;
  int parts;                 // wavefronts per target, each with a contiguous range of task words
  const int* rank_task;      // [num_rank] task of each rank task
  int num_rank;
  double* stage;             // [T x np x num_rank] predictions, target-major (rank_scatter moves them)
  int2* redo;                // (task, target) pairs left to the exact selection
  int* redo_count;
  int redo_cap;
};

// The work lists of a batch (direct, rank-table and pruned items: (task, first target)), for the exact
// re-selection when more pairs were handed over than its list holds.
struct RedoLists
{
  const int2* items[3];
  int counts[3];
};

// The simplex weight of a neighbour at exactly the smallest distance is exp(-theta * 1) (src/edm.cpp:
// 707-735: its distance over the smallest is x / x = 1 exactly), the same for every prediction of a call:
// a block makes exp(-theta) of the call's first ETH_N thetas once, into LDS (before a barrier), with the
// exp of the weights.
constexpr int ETH_N = 8;
__device__ __forceinline__ void theta_exps(const double* thetas, int T, double* eth)
{
  if ((int)threadIdx.x < min(T, ETH_N))
    eth[threadIdx.x] = fdlibm_exp(-thetas[threadIdx.x]);
}

} // namespace fedm
