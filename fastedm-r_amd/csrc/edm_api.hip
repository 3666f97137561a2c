// C ABI of the fastEDM prediction core: device context, edm_predict and edm_predict_batch.
// The dispatcher picks the kernel route per task (register top-k for k <= 64, dense distances for
// k = all), stages inputs into a grow-only device pool and copies the outputs back on one stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <cstdio>
#include <cstdlib>
#include <array>
#include <map>
#include <mutex>
#include <type_traits>
#include <cstring>
#include <vector>

#include "../../include/fastedm_hip.h"
#include "batch_common.h"
#include "fdlibm_exp.h"

namespace fedm {

template <int KCAP, int EMAX, bool GENERIC>
__global__ void knn_slices_f64(LpParams P, int k, int slice_len, double* out_s, int* out_i, int* out_cnt);
template <int EMAX, bool GENERIC>
__global__ void dense_dist_f64(LpParams P, int t0, int nt, int slice_len, double* D, int* out_cnt);
template <bool GENERIC>
__global__ void dense_dist_rows(LpParams P, int t0, int nt, int slice_len, double* D, int* out_cnt);
template <int KCAP>
__global__ void finish_lists(FinishParams F, const double* ps, const int* pi, const int* pcnt, int S);
__global__ void finish_dense(FinishParams F, const double* D, int t0, int nt, const int* pcnt, int S);
template <int NPL, int W_WAVES>
__global__ void wasserstein_dense(WassParams P, int t0, int nt, int slice_len, double* D, int* out_cnt);
__global__ void select_wide(FinishParams F, const double* D, int t0, int nt, unsigned long long* gkey, int* gidx,
                            int gstride);
int smap_launch(int n, hipStream_t s, const SmapParams& S, const double* D, int t0, int nt, bool lds_rows,
                int solve_mode);
size_t gram_dense_bytes(int n, int nt, int T);
int gram_dense_launch(int n, hipStream_t s, const SmapParams& S, const double* D, int t0, int nt, double* gw);

template <int KCAP, int KP>
__global__ void knn_mfma(MfmaParams P, int k, int slice_len, double* out_s, int* out_i, int* out_cnt, const int* tl,
                         const int* tn);
template <int KP>
__global__ void knn_seed(MfmaParams P, const double* kdm, const int* wst, const int* tord, const int* sgrp, int ngrp,
                         int win, int wide, int k, double* seed);
template <int KP, int R>
__global__ void knn_seed_wide(MfmaParams P, const double* kdm, const int* wst, int win, const int* fl, const int* fc,
                              int k, double* seed);
template <int KP>
__global__ void knn_screen(MfmaParams P, KdScreen K, int split, int* cand, int* ccnt);
template <int KCAP, int KP>
__global__ void knn_rerank(MfmaParams P, int k, int slice_len, const int* cand, const int* ccnt, double* out_s,
                           int* out_i, int* out_cnt, int* fflag, int* flist, int* fcnt);
template <int KCAP>
__global__ void merge_fallback(int NP, int S, int SF, const int* flist, const int* fcnt, const double* fs, const int* fi,
                               const int* fc, double* out_s, int* out_i, int* out_cnt);
__global__ void mfma_layout_probe(float* out);


template <typename T>
__global__ void fill_kernel(T* p, size_t n, T v)
{
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

} // namespace fedm

using namespace fedm;

struct DevBuf
{
  void* p = nullptr;
  size_t bytes = 0;
};

struct edm_ctx
{
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr; // the persistent route's direct kernels, beside the rank-table kernels
  hipEvent_t fork = nullptr, join = nullptr;
  hipEvent_t ev[12] = {};
  std::vector<DevBuf> bufs;
  double timings[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
  int last_error = 0;
  int last_redo = 0;
  unsigned char* scratch = nullptr; // this device's copy of the no-allocation scratch array
  int scratch_err = 0;

  void* get(int slot, size_t bytes)
  {
    if ((int)bufs.size() <= slot)
      bufs.resize(slot + 1);
    DevBuf& b = bufs[slot];
    if (b.bytes < bytes) {
      if (b.p)
        hipFree(b.p);
      b.p = nullptr;
      b.bytes = 0;
      size_t want = std::max(bytes, (size_t)256);
      if (hipMalloc(&b.p, want) != hipSuccess) {
        b.p = nullptr;
        return nullptr;
      }
      b.bytes = want;
    }
    return b.p;
  }
};

namespace fedm {
// Device buffer slots of the context pool.
enum
{
  B_M,
  B_MP,
  B_MT,
  B_MTT,
  B_MPPT,
  B_MPAN,
  B_MPPAN,
  B_METRICS,
  B_IDW,
  B_THETAS,
  B_PS,
  B_PI,
  B_PCNT,
  B_NNI,
  B_NNS,
  B_NNK,
  B_NNCNT,
  B_PRED,
  B_RC,
  B_KUSED,
  B_COEFFS,
  B_NNOUT,
  B_D,
  B_WKEY,
  B_WIDX,
  B_MIX0,
  B_MIX1,
  B_MIX2,
  B_MIX3,
  B_MIX4,
  B_SMAPBIG,
  B_GRAM,
  B_KDP,
  B_WST,
  B_TORD,
  B_SEED,
  B_CAND,
  B_CCNT,
  B_FLIST,
  B_FCNT,
  B_FFLAG,
  B_FPS,
  B_FPI,
  B_FPC,
  B_KD32,
  B_KDN,
  B_KDPOS,
  B_CBOX,
  B_TBOX,
  B_SGRP,
  B_NUM
};
} // namespace fedm

#define HIP_TRY(x)                                                                                               \
  do {                                                                                                           \
    hipError_t e_ = (x);                                                                                         \
    if (e_ != hipSuccess) {                                                                                      \
      fprintf(stderr, "fastedm_hip: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), __FILE__, __LINE__);   \
      return -(int)e_ - 1000;                                                                                    \
    }                                                                                                            \
  } while (0)

template <typename T>
static T* up(edm_ctx* c, int slot, const T* host, size_t n, int* err)
{
  if (host == nullptr || n == 0)
    return nullptr;
  T* d = (T*)c->get(slot, n * sizeof(T));
  if (!d) {
    *err = -2;
    return nullptr;
  }
  if (hipMemcpyAsync(d, host, n * sizeof(T), hipMemcpyHostToDevice, c->stream) != hipSuccess)
    *err = -3;
  return d;
}

template <typename T>
static void dfill(edm_ctx* c, T* p, size_t n, T v)
{
  if (!p || n == 0)
    return;
  int blocks = (int)std::min<size_t>(4096, (n + 255) / 256);
  hipLaunchKernelGGL(fill_kernel<T>, dim3(blocks), dim3(256), 0, c->stream, p, n, v);
}

static int emax_for(int Ea)
{
  return Ea <= 8 ? 8 : Ea <= 16 ? 16 : Ea <= 32 ? 32 : Ea <= 64 ? 64 : -1;
}
static int kcap_for(int k)
{
  return k <= 4 ? 4 : k <= 8 ? 8 : k <= 16 ? 16 : k <= 32 ? 32 : k <= 64 ? 64 : -1;
}

template <int KCAP, int EMAX>
static void launch_knn_e(bool generic, dim3 g, hipStream_t s, const LpParams& P, int k, int sl, double* ps, int* pi,
                         int* pc)
{
  if (generic)
    hipLaunchKernelGGL((knn_slices_f64<KCAP, EMAX, true>), g, dim3(256), 0, s, P, k, sl, ps, pi, pc);
  else
    hipLaunchKernelGGL((knn_slices_f64<KCAP, EMAX, false>), g, dim3(256), 0, s, P, k, sl, ps, pi, pc);
}

template <int KCAP>
static void launch_knn_k(int emax, bool generic, dim3 g, hipStream_t s, const LpParams& P, int k, int sl, double* ps,
                         int* pi, int* pc)
{
  switch (emax) {
    case 8: launch_knn_e<KCAP, 8>(generic, g, s, P, k, sl, ps, pi, pc); break;
    case 16: launch_knn_e<KCAP, 16>(generic, g, s, P, k, sl, ps, pi, pc); break;
    case 32: launch_knn_e<KCAP, 32>(generic, g, s, P, k, sl, ps, pi, pc); break;
    default: launch_knn_e<KCAP, 64>(generic, g, s, P, k, sl, ps, pi, pc); break;
  }
}

static void launch_knn(int kcap, int emax, bool generic, dim3 g, hipStream_t s, const LpParams& P, int k, int sl,
                       double* ps, int* pi, int* pc)
{
  switch (kcap) {
    case 4: launch_knn_k<4>(emax, generic, g, s, P, k, sl, ps, pi, pc); break;
    case 8: launch_knn_k<8>(emax, generic, g, s, P, k, sl, ps, pi, pc); break;
    case 16: launch_knn_k<16>(emax, generic, g, s, P, k, sl, ps, pi, pc); break;
    case 32: launch_knn_k<32>(emax, generic, g, s, P, k, sl, ps, pi, pc); break;
    default: launch_knn_k<64>(emax, generic, g, s, P, k, sl, ps, pi, pc); break;
  }
}

template <int KCAP>
static void launch_mfma_k(int kp, dim3 g, hipStream_t s, const MfmaParams& P, int k, int sl, double* ps, int* pi,
                          int* pc, const int* tl, const int* tn)
{
  if (kp == 16)
    hipLaunchKernelGGL((knn_mfma<KCAP, 16>), g, dim3(256), 0, s, P, k, sl, ps, pi, pc, tl, tn);
  else if (kp == 32)
    hipLaunchKernelGGL((knn_mfma<KCAP, 32>), g, dim3(256), 0, s, P, k, sl, ps, pi, pc, tl, tn);
  else
    hipLaunchKernelGGL((knn_mfma<KCAP, 64>), g, dim3(256), 0, s, P, k, sl, ps, pi, pc, tl, tn);
}

static void launch_seed(int kp, hipStream_t s, const MfmaParams& P, const double* kdp, const int* wst,
                        const int* tord, const int* sgrp, int ngrp, int win, int wide, int k, double* seed)
{
  const dim3 g((ngrp + 3) / 4);
  if (kp == 16)
    hipLaunchKernelGGL((knn_seed<16>), g, dim3(256), 0, s, P, kdp, wst, tord, sgrp, ngrp, win, wide, k, seed);
  else if (kp == 32)
    hipLaunchKernelGGL((knn_seed<32>), g, dim3(256), 0, s, P, kdp, wst, tord, sgrp, ngrp, win, wide, k, seed);
  else
    hipLaunchKernelGGL((knn_seed<64>), g, dim3(256), 0, s, P, kdp, wst, tord, sgrp, ngrp, win, wide, k, seed);
}

// The wider seed of the targets the two-phase form hands over (list fl, *fc entries on the device): 1 024
// rows each, one wavefront per target for up to 16 384 targets.
static void launch_seed_wide(int kp, hipStream_t s, const MfmaParams& P, const double* kdp, const int* wst, int win,
                             const int* fl, const int* fc, int k, double* seed)
{
  const dim3 g(std::max(1, std::min(4096, (P.NP + 3) / 4)));
  if (kp == 16)
    hipLaunchKernelGGL((knn_seed_wide<16, 16>), g, dim3(256), 0, s, P, kdp, wst, win, fl, fc, k, seed);
  else if (kp == 32)
    hipLaunchKernelGGL((knn_seed_wide<32, 16>), g, dim3(256), 0, s, P, kdp, wst, win, fl, fc, k, seed);
  else
    hipLaunchKernelGGL((knn_seed_wide<64, 16>), g, dim3(256), 0, s, P, kdp, wst, win, fl, fc, k, seed);
}

// Phases one and two of the two-phase matrix-core route: the seeded screen of every slice, then the exact
// re-rank of each (slice, target) whose candidates fit; the others land in (flist, fcnt).
template <int KP>
static void launch_two_phase_p(int kcap, hipStream_t s, const MfmaParams& P, const KdScreen& K, int split, int k,
                               int S, int sl, int* cand, int* ccnt, double* ps, int* pi, int* pc, int* ff, int* fl,
                               int* fc)
{
  hipLaunchKernelGGL((knn_screen<KP>), dim3((P.NP + 127) / 128, split), dim3(256), 0, s, P, K, split, cand, ccnt);
  const dim3 g((P.NP + 3) / 4, S);
  switch (kcap) {
    case 4: hipLaunchKernelGGL((knn_rerank<4, KP>), g, dim3(256), 0, s, P, k, sl, cand, ccnt, ps, pi, pc, ff, fl, fc); break;
    case 8: hipLaunchKernelGGL((knn_rerank<8, KP>), g, dim3(256), 0, s, P, k, sl, cand, ccnt, ps, pi, pc, ff, fl, fc); break;
    case 16: hipLaunchKernelGGL((knn_rerank<16, KP>), g, dim3(256), 0, s, P, k, sl, cand, ccnt, ps, pi, pc, ff, fl, fc); break;
    case 32: hipLaunchKernelGGL((knn_rerank<32, KP>), g, dim3(256), 0, s, P, k, sl, cand, ccnt, ps, pi, pc, ff, fl, fc); break;
    default: hipLaunchKernelGGL((knn_rerank<64, KP>), g, dim3(256), 0, s, P, k, sl, cand, ccnt, ps, pi, pc, ff, fl, fc); break;
  }
}

static void launch_two_phase(int kcap, int kp, hipStream_t s, const MfmaParams& P, const KdScreen& K, int split,
                             int k, int S, int sl, int* cand, int* ccnt, double* ps, int* pi, int* pc, int* ff,
                             int* fl, int* fc)
{
  if (kp == 16)
    launch_two_phase_p<16>(kcap, s, P, K, split, k, S, sl, cand, ccnt, ps, pi, pc, ff, fl, fc);
  else if (kp == 32)
    launch_two_phase_p<32>(kcap, s, P, K, split, k, S, sl, cand, ccnt, ps, pi, pc, ff, fl, fc);
  else
    launch_two_phase_p<64>(kcap, s, P, K, split, k, S, sl, cand, ccnt, ps, pi, pc, ff, fl, fc);
}

static void launch_merge_fallback(int kcap, hipStream_t s, int NP, int S, int SF, const int* fl, const int* fc,
                                  const double* fs, const int* fi, const int* fcn, double* ps, int* pi, int* pc)
{
  const dim3 g((NP + 255) / 256);
  switch (kcap) {
    case 4: hipLaunchKernelGGL((merge_fallback<4>), g, dim3(256), 0, s, NP, S, SF, fl, fc, fs, fi, fcn, ps, pi, pc); break;
    case 8: hipLaunchKernelGGL((merge_fallback<8>), g, dim3(256), 0, s, NP, S, SF, fl, fc, fs, fi, fcn, ps, pi, pc); break;
    case 16: hipLaunchKernelGGL((merge_fallback<16>), g, dim3(256), 0, s, NP, S, SF, fl, fc, fs, fi, fcn, ps, pi, pc); break;
    case 32: hipLaunchKernelGGL((merge_fallback<32>), g, dim3(256), 0, s, NP, S, SF, fl, fc, fs, fi, fcn, ps, pi, pc); break;
    default: hipLaunchKernelGGL((merge_fallback<64>), g, dim3(256), 0, s, NP, S, SF, fl, fc, fs, fi, fcn, ps, pi, pc); break;
  }
}

// The streaming matrix-core k-NN over every target, or (tl, tn) over the targets listed per slice.
static void launch_mfma(int kcap, int kp, dim3 g, hipStream_t s, const MfmaParams& P, int k, int sl, double* ps,
                        int* pi, int* pc, const int* tl = nullptr, const int* tn = nullptr)
{
  switch (kcap) {
    case 4: launch_mfma_k<4>(kp, g, s, P, k, sl, ps, pi, pc, tl, tn); break;
    case 8: launch_mfma_k<8>(kp, g, s, P, k, sl, ps, pi, pc, tl, tn); break;
    case 16: launch_mfma_k<16>(kp, g, s, P, k, sl, ps, pi, pc, tl, tn); break;
    case 32: launch_mfma_k<32>(kp, g, s, P, k, sl, ps, pi, pc, tl, tn); break;
    default: launch_mfma_k<64>(kp, g, s, P, k, sl, ps, pi, pc, tl, tn); break;
  }
}

// Dense distance rows of targets t0 .. t0 + nt - 1 over S library slices of sl points.
static void launch_dense(int emax, bool generic, int S, hipStream_t s, const LpParams& P, int t0, int nt, int sl,
                         double* D, int* pc)
{
  const dim3 g = emax == 0 ? dim3((nt + 255) / 256, S) : dim3(S, (nt + DT_TGT - 1) / DT_TGT);
#define FEDM_DENSE(EM)                                                                                           \
  if (generic)                                                                                                   \
    hipLaunchKernelGGL((dense_dist_f64<EM, true>), g, dim3(256), 0, s, P, t0, nt, sl, D, pc);                    \
  else                                                                                                           \
    hipLaunchKernelGGL((dense_dist_f64<EM, false>), g, dim3(256), 0, s, P, t0, nt, sl, D, pc);
  switch (emax) {
    case 0:
      // points wider than the register rows
      if (generic)
        hipLaunchKernelGGL((dense_dist_rows<true>), g, dim3(256), 0, s, P, t0, nt, sl, D, pc);
      else
        hipLaunchKernelGGL((dense_dist_rows<false>), g, dim3(256), 0, s, P, t0, nt, sl, D, pc);
      break;
    case 8: FEDM_DENSE(8) break;
    case 16: FEDM_DENSE(16) break;
    case 32: FEDM_DENSE(32) break;
    default: FEDM_DENSE(64) break;
  }
#undef FEDM_DENSE
}

static void launch_finish_lists(int kcap, dim3 g, hipStream_t s, const FinishParams& F, const double* ps, const int* pi,
                                const int* pc, int S)
{
  switch (kcap) {
    case 4: hipLaunchKernelGGL(finish_lists<4>, g, dim3(256), 0, s, F, ps, pi, pc, S); break;
    case 8: hipLaunchKernelGGL(finish_lists<8>, g, dim3(256), 0, s, F, ps, pi, pc, S); break;
    case 16: hipLaunchKernelGGL(finish_lists<16>, g, dim3(256), 0, s, F, ps, pi, pc, S); break;
    case 32: hipLaunchKernelGGL(finish_lists<32>, g, dim3(256), 0, s, F, ps, pi, pc, S); break;
    default: hipLaunchKernelGGL(finish_lists<64>, g, dim3(256), 0, s, F, ps, pi, pc, S); break;
  }
}

extern "C" void fedm_default_knobs(fedm_knobs* k)
{
  k->rank_scale = 1.5;
  if (const char* e = getenv("FEDM_RANK_LTHR_SCALE"))
    k->rank_scale = atof(e);
  k->direct_blocks_per_cu = 0;
  k->direct_alt_waves = 0;
  k->list_blocks = 1024;
  k->mfma_blocks = 2048;
  k->use_mfma = 1;
  // FEDM_SMAP_LDS_ROWS=1 keeps U and V of the S-map solve in LDS for every width (the kernel used above
  // 24 unknowns); the default keeps rows in registers where they fit.  Both give the same bits.
  const char* e = getenv("FEDM_SMAP_LDS_ROWS");
  k->smap_lds = e && e[0] == '1';
  k->direct_prune = 1;
  // measured: the two routes each fill the device, so running them side by side gains nothing (5.66 ms
  // against 5.55 ms on the 40-library CCM sweep); the tuner can still choose it
  k->concurrent_routes = 0;
  // FEDM_SMAP_BLOCK=0..4 picks the S-map solve of up to 24 unknowns (same bits; SMAP_SOLVE_*).  Measured
  // on the panel workload (n = 17, 796 995 systems): one row per lane 63.8 ms, block-wide 115.8 ms (the
  // 2x2 chain of wavefront 0 sits on the critical path of every rotation of the block, behind three
  // block barriers), two rows per lane 56.0 ms; at n = 11 and 7 500 systems two rows per lane is slower
  // (0.55 against 0.45 ms: two wavefronts per SIMD and too few groups to fill the device)
  const char* b = getenv("FEDM_SMAP_BLOCK");
  k->smap_block = b && b[0] >= '0' && b[0] <= '4' ? b[0] - '0' : SMAP_SOLVE_AUTO;
  // FEDM_MFMA_SEED=0 starts the matrix-core screening unbounded, 1 seeds the streaming kernel, 2 runs the
  // two-phase form, 3 chooses between 1 and 2 (same results)
  const char* sd = getenv("FEDM_MFMA_SEED");
  k->mfma_seed = sd && sd[0] >= '0' && sd[0] <= '3' ? sd[0] - '0' : 3;
  // FEDM_MFMA_PRUNE=0 screens every chunk of the library in the two-phase form (same results)
  const char* pr = getenv("FEDM_MFMA_PRUNE");
  k->mfma_prune = !(pr && pr[0] == '0');
  const char* sp = getenv("FEDM_MFMA_SPLIT");
  k->mfma_split = sp ? atoi(sp) : 0;
}

static fedm_knobs default_knobs()
{
  fedm_knobs k;
  fedm_default_knobs(&k);
  return k;
}

// Scratch of edm_predict_noalloc: the kernels' slice and neighbour lists, counts and dense distance
// rows, and the small host-side option arrays, for callers that cannot allocate device memory.  Each
// device holds its own copy of the array for the life of the process; its address on a device comes
// from hipGetSymbolAddress, looked up once per device when a context is created there.
namespace fedm {
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)32 << 20;
}
static __device__ __align__(256) unsigned char scratch[fedm::NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static std::map<int, unsigned char*> g_scratch_addr; // device -> address of `scratch` there

// With `device` current: the address of its copy of the scratch array (looked up on first use).
static int scratch_address(int device, unsigned char** out)
{
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_addr.find(device);
  if (it == g_scratch_addr.end()) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
      return -(int)e - 1000;
    if (!p)
      return -2;
    it = g_scratch_addr.emplace(device, (unsigned char*)p).first;
  }
  *out = it->second;
  return 0;
}

extern "C" {

int edm_ctx_create(edm_ctx** out, int device)
{
  *out = nullptr;
  edm_ctx* c = new edm_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return -1;
  }
  if (hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess) {
    hipStreamDestroy(c->stream);
    delete c;
    return -1;
  }
  for (auto& e : c->ev)
    hipEventCreate(&e);
  hipEventCreateWithFlags(&c->fork, hipEventDisableTiming);
  hipEventCreateWithFlags(&c->join, hipEventDisableTiming);
  c->bufs.resize(B_NUM);
  c->scratch_err = scratch_address(device, &c->scratch);
  *out = c;
  return 0;
}

void edm_ctx_destroy(edm_ctx* c)
{
  if (!c)
    return;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  for (auto& b : c->bufs)
    if (b.p)
      hipFree(b.p);
  hipStreamSynchronize(c->stream2);
  for (auto& e : c->ev)
    hipEventDestroy(e);
  hipEventDestroy(c->fork);
  hipEventDestroy(c->join);
  hipStreamDestroy(c->stream2);
  hipStreamDestroy(c->stream);
  delete c;
}

void edm_ctx_timings(edm_ctx* c, double* ms4)
{
  for (int i = 0; i < 4; i++)
    ms4[i] = c->timings[i];
}

void edm_ctx_timings_ex(edm_ctx* c, double* ms8)
{
  for (int i = 0; i < 8; i++)
    ms8[i] = c->timings[i];
}

int edm_ctx_exact_redos(const edm_ctx* c)
{
  return c->last_redo;
}

static float ev_ms(hipEvent_t a, hipEvent_t b)
{
  float ms = 0.f;
  hipEventElapsedTime(&ms, a, b);
  return ms;
}

static void host_fill_outputs(const edm_manifold_view* Mp, const edm_options* o, double* pred, int* rc,
                              double* coeffs, int* k_used, int* nn_idx, int nn_cap, int Ea, int rcv)
{
  const size_t Np = Mp->num_points, T = o->num_thetas;
  for (size_t i = 0; i < T * Np; i++) {
    pred[i] = FEDM_MISSING_D;
    rc[i] = rcv;
  }
  if (coeffs)
    for (size_t i = 0; i < Np * (size_t)(Ea + 1); i++)
      coeffs[i] = FEDM_MISSING_D;
  for (size_t i = 0; i < Np; i++)
    k_used[i] = FEDM_MISSING_I;
  if (nn_idx)
    for (size_t i = 0; i < Np * (size_t)nn_cap; i++)
      nn_idx[i] = -1;
}

} // extern "C"

// ------------------------------------------------------------------------------------------------
// One task on the device in four steps, shared by edm_predict, the batches' task route and
// edm_predict_noalloc:
//   plan    route, grid shapes and buffer sizes, from the host views (and the FP32 operands of the
//           matrix-core route, which need the host rows);
//   bind    a device pointer per buffer, from the context pool, a batch's arena or the static scratch;
//   upload  the host inputs into their buffers;
//   launch  the output fills and the kernels, on one stream.

namespace fedm {

enum TaskRoute
{
  R_MFMA,
  R_LIST,
  R_DENSE
};

// inputs come from the host views, outputs receive the results, scratch lives for one launch
enum BufRole
{
  ROLE_IN,
  ROLE_SCRATCH,
  ROLE_OUT
};

static BufRole buf_role(int slot)
{
  switch (slot) {
    case B_PS:
    case B_PI:
    case B_PCNT:
    case B_NNI:
    case B_NNS:
    case B_NNK:
    case B_NNCNT:
    case B_D:
    case B_WKEY:
    case B_WIDX:
    case B_SMAPBIG:
    case B_GRAM:
    case B_SEED:
    case B_CAND:
    case B_CCNT:
    case B_FLIST:
    case B_FCNT:
    case B_FFLAG:
    case B_FPS:
    case B_FPI:
    case B_FPC:
      return ROLE_SCRATCH;
    case B_PRED:
    case B_RC:
    case B_KUSED:
    case B_COEFFS:
    case B_NNOUT:
      return ROLE_OUT;
    default:
      return ROLE_IN;
  }
}

struct TaskPlan
{
  int NL = 0, NP = 0, Ea = 0, T = 0, n = 0, k = 0, emax = 0, nmet = 0, nn_cap = 0;
  int route = R_LIST;
  bool wass = false, smap = false, generic = false, wide = false;
  int kcap = 0, kp = 0, S = 1, sl = 0, bx = 0, nt = 0, gstride = 0;
  int wE = 0, wE_dt = 0, wrows = 0, wunl = 0; // Wasserstein layout
  bool smap_lds = false;                      // S-map solve with U and V in LDS
  int smap_mode = SMAP_SOLVE_ROWS;            // S-map solve of n <= 24: SMAP_SOLVE_*
  size_t bytes[B_NUM] = {};
  void* ptr[B_NUM] = {};
  // host copies the upload needs beyond the views
  std::vector<int> metrics;
  std::vector<float> m32, mn, p32, pn;
  std::vector<double> ynorm;
  std::vector<int> kdp, wst, tord; // matrix-core seed: library rows in k-d order, window starts, target order
  std::vector<int> sgrp;           // seed groups: starts in the target order of runs of up to 4 sharing a window
  std::vector<double> kdm;         // the library rows in that order (FP64)
  std::vector<float> kd32, kdn;    // two-phase screen: the centred FP32 rows and their norms in that order
  std::vector<float> cbox, tbox;   // two-phase screen: boxes of the library chunks and the target blocks
  int seed_win = 0;                // rows per seed window (0: no seed)
  bool two_phase = false;          // seeded screen, exact re-rank, streaming kernel for what overflows
  int sf = 1;                      // two-phase: library sub-slices of the streaming kernel on the overflow
  int split = 1;                   // two-phase: blocks sharing each target block's chunks
  LpParams P;
  WassParams W;
  FinishParams F;
  SmapParams Sm;
  MfmaParams MP;
  KdScreen KS{};
};

constexpr int SEED_WIN = 256; // library rows per seed window of the matrix-core route (knn_seed: 4 per lane)
// library rows the seed scans around each window before the seeded streaming kernel for points of more
// than 16 coordinates, where a window's k-th distance lies far above the library's (Lorenz-96, 40
// coordinates: select 22.6 -> 11.7 ms); up to 16 coordinates the window's k-th is already near it
// (panel, 16 coordinates, forced to the streaming form: 94.6 ms from the window, 100.5 from 4 096 rows)
constexpr int SEED_WIDE = 4096;

// Seed windows of the matrix-core route.  A k-d order of the library rows: a node splits at the median of
// its coordinate of largest spread, with every row equal to the median on one side (so a coordinate that
// is constant over most rows, such as a dt column, separates only the rows that differ from it and does
// not scatter equal rows over both sides); leaves hold at most 32 rows.  Each target descends the tree to
// a leaf and takes the `win` rows of the order centred on it.  The windows only tighten the screening's
// starting threshold; any window gives the same results.
static void kd_seed_windows(const double* L, int NL, const double* P, int NP, int Ea, int win, std::vector<int>& kdp,
                            std::vector<int>& wst, std::vector<int>& tord, std::vector<int>& sgrp)
{
  constexpr int LEAF = 32;
  struct Node
  {
    int dim; // -1: leaf
    double val;
    int le;   // 1: rows <= val go left, 0: rows < val go left
    int a, b; // rows kdp[a .. b)
    int left, right;
  };
  kdp.resize(NL);
  std::iota(kdp.begin(), kdp.end(), 0);
  std::vector<Node> nodes;
  nodes.push_back({ -1, 0.0, 0, 0, NL, -1, -1 });
  std::vector<int> stack{ 0 };
  std::vector<double> lo(Ea), hi(Ea);
  std::vector<int> dims(Ea);
  while (!stack.empty()) {
    const int ni = stack.back();
    stack.pop_back();
    const int a = nodes[ni].a, b = nodes[ni].b;
    if (b - a <= LEAF)
      continue;
    std::fill(lo.begin(), lo.end(), INFINITY);
    std::fill(hi.begin(), hi.end(), -INFINITY);
    for (int i = a; i < b; i++) {
      const double* r = L + (size_t)kdp[i] * Ea;
      for (int d = 0; d < Ea; d++) {
        lo[d] = std::min(lo[d], r[d]);
        hi[d] = std::max(hi[d], r[d]);
      }
    }
    std::iota(dims.begin(), dims.end(), 0);
    std::stable_sort(dims.begin(), dims.end(), [&](int x, int y) { return hi[x] - lo[x] > hi[y] - lo[y]; });
    for (int d : dims) {
      if (!(hi[d] > lo[d]))
        break;
      auto key = [&](int r) { return L[(size_t)r * Ea + d]; };
      const int m = a + (b - a) / 2;
      std::nth_element(kdp.begin() + a, kdp.begin() + m, kdp.begin() + b, [&](int x, int y) { return key(x) < key(y); });
      const double v = key(kdp[m]);
      auto it = std::partition(kdp.begin() + a, kdp.begin() + b, [&](int r) { return key(r) < v; });
      int le = 0;
      if (it == kdp.begin() + a) {
        it = std::partition(kdp.begin() + a, kdp.begin() + b, [&](int r) { return key(r) <= v; });
        le = 1;
      }
      const int s = (int)(it - kdp.begin());
      if (s == a || s == b)
        continue;
      const int l = (int)nodes.size();
      nodes.push_back({ -1, 0.0, 0, a, s, -1, -1 });
      nodes.push_back({ -1, 0.0, 0, s, b, -1, -1 });
      nodes[ni].dim = d;
      nodes[ni].val = v;
      nodes[ni].le = le;
      nodes[ni].left = l;
      nodes[ni].right = l + 1;
      stack.push_back(l);
      stack.push_back(l + 1);
      break;
    }
  }
  const int w = std::min(win, NL);
  wst.resize(NP);
  for (int t = 0; t < NP; t++) {
    const double* y = P + (size_t)t * Ea;
    int ni = 0;
    while (nodes[ni].dim >= 0) {
      const Node& nd = nodes[ni];
      const double v = y[nd.dim];
      ni = (nd.le ? v <= nd.val : v < nd.val) ? nd.left : nd.right;
    }
    const int c = (nodes[ni].a + nodes[ni].b) / 2;
    wst[t] = std::max(0, std::min(NL - w, c - w / 2));
  }
  tord.resize(NP);
  std::iota(tord.begin(), tord.end(), 0);
  std::stable_sort(tord.begin(), tord.end(), [&](int x, int y) { return wst[x] < wst[y]; });
  // runs of up to 4 consecutive targets with one window (the seed kernel's groups), closed by NP
  sgrp.clear();
  for (int i = 0; i < NP; i++)
    if (i == 0 || wst[tord[i]] != wst[tord[sgrp.back()]] || i - sgrp.back() == 4)
      sgrp.push_back(i);
  sgrp.push_back(NP);
}

// The two-phase screen's operands in the k-d orders (kdp for the library, tord for the targets): the
// centred FP32 rows and norms of the library in its order, and the boxes of its 64-row chunks and of the
// 128-target blocks of the target order, from the FP64 coordinates rounded outward to FP32.
static void build_screen_order(TaskPlan& tp, const double* L, const double* P, bool prune)
{
  const int NL = tp.NL, NP = tp.NP, Ea = tp.Ea, kp = tp.kp;
  tp.kd32.resize((size_t)NL * kp);
  tp.kdn.resize(NL);
  for (int i = 0; i < NL; i++) {
    std::copy(tp.m32.begin() + (size_t)tp.kdp[i] * kp, tp.m32.begin() + (size_t)(tp.kdp[i] + 1) * kp,
              tp.kd32.begin() + (size_t)i * kp);
    tp.kdn[i] = tp.mn[tp.kdp[i]];
  }
  auto boxes = [Ea](const double* X, const int* order, int n, int per, std::vector<float>& out) {
    const int nb = (n + per - 1) / per;
    out.assign((size_t)nb * 2 * Ea, 0.f);
    for (int b = 0; b < nb; b++) {
      float* lo = out.data() + (size_t)b * 2 * Ea;
      float* hi = lo + Ea;
      for (int d = 0; d < Ea; d++) {
        double l = INFINITY, h = -INFINITY;
        for (int i = b * per; i < std::min(n, (b + 1) * per); i++) {
          const double v = X[(size_t)order[i] * Ea + d];
          l = std::min(l, v);
          h = std::max(h, v);
        }
        float fl = (float)l, fh = (float)h;
        if ((double)fl > l)
          fl = std::nextafter(fl, -INFINITY);
        if ((double)fh < h)
          fh = std::nextafter(fh, INFINITY);
        lo[d] = fl;
        hi[d] = fh;
      }
    }
  };
  boxes(L, tp.kdp.data(), NL, 64, tp.cbox);
  boxes(P, tp.tord.data(), NP, 128, tp.tbox);
  tp.KS.prune = prune ? 1 : 0;
}

// Chooses the route and sizes every buffer.  Returns 0, -22 (invalid arguments) or -95 (a width or
// option this build does not take; the caller fills the outputs).  `mfma_ok` allows the matrix-core
// route (it needs the host rows); `max_slices` caps the library slices of the list routes.
static int plan_task(const edm_manifold_view* M, const edm_manifold_view* Mp, const edm_options* o, bool want_coeffs,
                     int nn_cap, bool mfma_ok, int max_slices, TaskPlan& tp, size_t dense_budget = (size_t)1 << 30,
                     const fedm_knobs* kn = nullptr)
{
  const fedm_knobs knobs = kn ? *kn : default_knobs();
  mfma_ok = mfma_ok && knobs.use_mfma;
  tp.smap_lds = knobs.smap_lds != 0;
  tp.smap_mode = knobs.smap_block;
  const int NL = M->num_points, NP = Mp->num_points, Ea = M->E_actual, T = o->num_thetas;
  if (Ea != Mp->E_actual || T <= 0 || Ea <= 0 || NL < 0 || NP < 0)
    return -22;
  tp.NL = NL;
  tp.NP = NP;
  tp.Ea = Ea;
  tp.T = T;
  tp.n = Ea + 1;
  tp.k = o->k;
  tp.nn_cap = nn_cap;
  tp.emax = emax_for(Ea);
  const int k = o->k;
  tp.wass = o->distance == FEDM_DIST_WASSERSTEIN;
  tp.smap = o->algorithm == FEDM_SMAP;
  // Wasserstein point layout [x lags | dt lags | lagged extras | unlagged extras] (src/distances.cpp:318-326)
  const int wE = M->E_x;
  const int wrows = wE > 0 ? 1 + (M->E_dt > 0) + M->E_lagged_extras / wE : 0;
  const int wunl = M->E_extras - M->E_lagged_extras;
  bool wass_ok = wE >= 1 && wE <= 64 && (M->E_dt == 0 || M->E_dt == wE) && M->E_lagged_extras % wE == 0 &&
                 wunl >= 0 && wrows * wE + wunl == Ea && o->missing_distance >= 0.0;
  if (tp.wass && wass_ok && o->panel_mode && o->idw <= 0 && o->id_weights)
    for (int a = 0; a < o->num_panels * o->num_panels; a++)
      wass_ok = wass_ok && o->id_weights[a] >= 0.0f;
  // points wider than 64 coordinates take the dense route with the row-wise distance kernel (emax 0);
  // the S-map solve takes at most SMAP_BIG_MAX unknowns
  if (tp.emax < 0 && Ea <= LDS_CHUNK_DOUBLES)
    tp.emax = 0;
  if (tp.emax < 0 || (tp.wass && !wass_ok) || (tp.smap && tp.n > SMAP_BIG_MAX))
    return -95;
  tp.wE = wE;
  tp.wE_dt = M->E_dt;
  tp.wrows = wrows;
  tp.wunl = wunl;

  // metrics: one per coordinate for L^p, one per variable for Wasserstein (src/distances.cpp:29-67)
  tp.nmet = tp.wass ? wrows + wunl : Ea;
  tp.metrics.assign(tp.nmet, 0);
  if (o->metrics)
    std::copy(o->metrics, o->metrics + tp.nmet, tp.metrics.begin());

  LpParams& P = tp.P;
  std::memset(&P, 0, sizeof(P));
  P.NL = NL;
  P.NP = NP;
  P.Ea = Ea;
  P.num_panels = o->num_panels;
  P.missing_distance = o->missing_distance;
  P.idw = o->idw;
  const bool lazy = o->low_memory != 0;
  int f = 0;
  if (o->has_missing || lazy)
    f |= F_MISSING;
  if (o->has_categorical || lazy)
    f |= F_CATEGORICAL;
  if (o->panel_mode && o->idw < 0)
    f |= F_SKIP_OTHER;
  if (o->use_only_past && !lazy)
    f |= F_SKIP_FUTURE;
  if (o->distance == FEDM_DIST_MAE)
    f |= F_MAE;
  if (o->panel_mode && o->idw > 0)
    f |= F_IDW;
  else if (o->panel_mode && o->id_weights && o->num_panels > 0)
    f |= F_IDWEIGHTS;
  P.flags = f;
  tp.generic = f != 0;

  size_t* by = tp.bytes;
  by[B_M] = sizeof(double) * (size_t)NL * Ea;
  by[B_MP] = sizeof(double) * (size_t)NP * Ea;
  by[B_MT] = M->targets ? sizeof(double) * (size_t)NL : 0;
  by[B_MTT] = M->target_times ? sizeof(double) * (size_t)NL : 0;
  by[B_MPPT] = Mp->point_times ? sizeof(double) * (size_t)NP : 0;
  if (o->panel_mode) {
    by[B_MPAN] = M->panel ? sizeof(int) * (size_t)NL : 0;
    by[B_MPPAN] = Mp->panel ? sizeof(int) * (size_t)NP : 0;
  }
  by[B_METRICS] = sizeof(int) * (size_t)tp.nmet;
  by[B_IDW] = (o->id_weights && o->num_panels > 0) ? sizeof(float) * (size_t)o->num_panels * o->num_panels : 0;
  by[B_THETAS] = sizeof(double) * (size_t)T;
  by[B_PRED] = sizeof(double) * (size_t)T * NP;
  by[B_RC] = sizeof(int) * (size_t)T * NP;
  by[B_KUSED] = sizeof(int) * (size_t)NP;
  by[B_COEFFS] = want_coeffs ? sizeof(double) * (size_t)NP * tp.n : 0;
  by[B_NNOUT] = nn_cap > 0 ? sizeof(int) * (size_t)NP * nn_cap : 0;
  by[B_NNK] = sizeof(int) * (size_t)NP;
  by[B_NNCNT] = sizeof(int) * (size_t)NP;

  const bool list_route = !tp.wass && k > 0 && k <= 64 && tp.emax > 0;
  if (tp.smap && tp.n > 64)
    by[B_SMAPBIG] = sizeof(double) * 2 * (size_t)tp.n * tp.n * SMAP_BIG_BLOCKS;
  double amax = 0.0;
  const bool want_mfma = mfma_ok && list_route && o->precision == FEDM_MIXED && !tp.generic && Ea >= 8;
  if (want_mfma) {
    for (size_t i = 0; i < (size_t)NL * Ea; i++)
      amax = std::max(amax, std::fabs(M->data[i]));
    for (size_t i = 0; i < (size_t)NP * Ea; i++)
      amax = std::max(amax, std::fabs(Mp->data[i]));
  }
  if (want_mfma && amax < 1e30) {
    tp.route = R_MFMA;
    tp.kcap = kcap_for(k);
    tp.kp = Ea <= 16 ? 16 : Ea <= 32 ? 32 : 64;
    tp.bx = (NP + 63) / 64;
    int S = std::max(1, std::min(MAX_SLICES, (std::max(knobs.mfma_blocks, 1) + tp.bx - 1) / tp.bx));
    S = std::max(1, std::min(S, NL / 256));
    tp.S = std::max(1, std::min(S, max_slices));
    tp.sl = (NL + tp.S - 1) / tp.S;
    const int kp = tp.kp;
    // centred FP32 operands, their FP32 norms and the per-target screening bound
    std::vector<double> ctr(Ea, 0.0);
    for (int j = 0; j < NL; j++)
      for (int d = 0; d < Ea; d++)
        ctr[d] += M->data[(size_t)j * Ea + d];
    for (int d = 0; d < Ea; d++)
      ctr[d] /= std::max(NL, 1);
    tp.m32.assign((size_t)NL * kp, 0.f);
    tp.mn.assign(NL, 0.f);
    tp.p32.assign((size_t)NP * kp, 0.f);
    tp.pn.assign(NP, 0.f);
    tp.ynorm.assign(NP, 0.0);
    double R = 0.0;
    for (int j = 0; j < NL; j++) {
      double r2 = 0.0;
      float n2 = 0.f;
      for (int d = 0; d < Ea; d++) {
        const double v = M->data[(size_t)j * Ea + d] - ctr[d];
        const float fv = (float)v;
        tp.m32[(size_t)j * kp + d] = fv;
        n2 += fv * fv;
        r2 += v * v;
      }
      tp.mn[j] = n2;
      R = std::max(R, std::sqrt(r2));
    }
    for (int i = 0; i < NP; i++) {
      double r2 = 0.0;
      float n2 = 0.f;
      for (int d = 0; d < Ea; d++) {
        const double v = Mp->data[(size_t)i * Ea + d] - ctr[d];
        const float fv = (float)v;
        tp.p32[(size_t)i * kp + d] = fv;
        n2 += fv * fv;
        r2 += v * v;
      }
      tp.pn[i] = n2;
      tp.ynorm[i] = std::sqrt(r2);
    }
    const double u = 5.9604644775390625e-08;
    MfmaParams& MP = tp.MP;
    std::memset(&MP, 0, sizeof(MP));
    // the screening bound per target (kernel): 2 (k1 rr sqrt(D) + c2 D + k3 rr^2), rr >= |x| + |y|:
    //   k1 rr sqrt(D) + k1^2 rr^2 / 4: the FP32 rounding of the centred coordinates (|dx| <= u rr);
    //   c2 D: the rounding of the centring and the FP64 evaluation of the floor;
    //   k3 rr^2: the two FP32 norms (gamma_Ea (|x|^2 + |y|^2)) and the accumulation of the KP rounded
    //   products onto -||x||^2 / 2 (gamma_(KP+1) (||x||^2 / 2 + |x| |y|), doubled in the distance), together
    //   within gamma_(KP+1) (2 |x|^2 + |y|^2 + 2 |x| |y|) <= 2 gamma_(KP+1) rr^2 (Ea <= KP), with
    //   gamma_m = m u / (1 - m u) <= 1.02 m u, plus u^2 rr^2 from the inputs; B = 2 (...) covers it with
    //   k3 = (KP + 3) 1.02 u + u^2, as 2 (KP + 3) 1.02 >= 2.04 (KP + 1)
    MP.R = R;
    MP.k1 = 2.02 * u;
    MP.k3 = (kp + 3) * 1.02 * u + u * u;
    MP.c2 = 4.0 * u;
    MP.NL = NL;
    MP.NP = NP;
    MP.Ea = Ea;
    by[B_MIX0] = sizeof(float) * tp.m32.size();
    by[B_MIX1] = sizeof(float) * tp.mn.size();
    by[B_MIX2] = sizeof(float) * tp.p32.size();
    by[B_MIX3] = sizeof(float) * tp.pn.size();
    by[B_MIX4] = sizeof(double) * tp.ynorm.size();
    if (knobs.mfma_seed && NL > 0 && NP > 0) {
      kd_seed_windows(M->data, NL, Mp->data, NP, Ea, SEED_WIN, tp.kdp, tp.wst, tp.tord, tp.sgrp);
      tp.seed_win = std::min(NL, SEED_WIN);
      tp.kdm.resize((size_t)NL * Ea);
      for (int i = 0; i < NL; i++)
        std::copy(M->data + (size_t)tp.kdp[i] * Ea, M->data + (size_t)(tp.kdp[i] + 1) * Ea, tp.kdm.data() + (size_t)i * Ea);
      by[B_KDP] = sizeof(double) * tp.kdm.size();
      by[B_WST] = sizeof(int) * tp.wst.size();
      by[B_TORD] = sizeof(int) * tp.tord.size();
      by[B_SGRP] = sizeof(int) * tp.sgrp.size();
      by[B_SEED] = sizeof(double) * (size_t)NP;
      // auto (3): the two-phase form where the seeds are near the final k-th distance, that is for points of
      // up to 16 coordinates and up to 32 neighbours (panel workload, E_a = 16, k = 18: select 39.6 -> 35.5
      // ms); with 40 coordinates and k = 42 (Lorenz-96) nearly every target overflows the candidate space
      // and the seeded streaming kernel is 3x faster
      const bool two = knobs.mfma_seed == 2 || (knobs.mfma_seed == 3 && tp.kp == 16 && tp.kcap <= 32);
      if (two) {
        tp.two_phase = true;
        // one slice: the screen shares the library's chunks out over `split` blocks per target block instead
        tp.S = 1;
        tp.sl = NL;
        const int nb = (NP + 127) / 128;
        tp.split = knobs.mfma_split > 0 ? std::min(knobs.mfma_split, 64)
                                         : std::max(1, std::min(16, (16384 + nb - 1) / nb));
        build_screen_order(tp, M->data, Mp->data, knobs.mfma_prune != 0);
        by[B_KD32] = sizeof(float) * tp.kd32.size();
        by[B_KDN] = sizeof(float) * tp.kdn.size();
        by[B_KDPOS] = sizeof(int) * tp.kdp.size();
        by[B_CBOX] = sizeof(float) * tp.cbox.size();
        by[B_TBOX] = sizeof(float) * tp.tbox.size();
        by[B_CAND] = sizeof(int) * (size_t)NP * SCR_CAP;
        by[B_CCNT] = sizeof(int) * (size_t)NP;
        by[B_FLIST] = sizeof(int) * (size_t)NP;
        by[B_FCNT] = sizeof(int);
        by[B_FFLAG] = sizeof(int) * (size_t)NP;
        // the overflowing targets are few: split the library into up to 16 sub-slices for them, within a
        // 1 GiB budget for their lists
        const size_t per = (size_t)NP * tp.kcap * 12;
        tp.sf = (int)std::max<size_t>(1, std::min<size_t>({ 16, ((size_t)1 << 30) / std::max<size_t>(per, 1),
                                                            (size_t)std::max(1, NL / 256) }));
        by[B_FPS] = sizeof(double) * (size_t)tp.sf * NP * tp.kcap;
        by[B_FPI] = sizeof(int) * (size_t)tp.sf * NP * tp.kcap;
        by[B_FPC] = sizeof(int) * (size_t)tp.sf * NP;
      }
    }
    by[B_PS] = sizeof(double) * (size_t)tp.S * NP * tp.kcap;
    by[B_PI] = sizeof(int) * (size_t)tp.S * NP * tp.kcap;
    by[B_PCNT] = sizeof(int) * (size_t)tp.S * NP;
    by[B_NNI] = sizeof(int) * (size_t)NP * tp.kcap;
    by[B_NNS] = sizeof(double) * (size_t)NP * tp.kcap;
  } else if (list_route) {
    tp.route = R_LIST;
    tp.kcap = kcap_for(k);
    tp.bx = (NP + 255) / 256;
    int S = std::max(1, std::min(MAX_SLICES, (std::max(knobs.list_blocks, 1) + tp.bx - 1) / tp.bx));
    S = std::max(1, std::min(S, NL / 256));
    tp.S = std::max(1, std::min(S, max_slices));
    tp.sl = (NL + tp.S - 1) / tp.S;
    by[B_PS] = sizeof(double) * (size_t)tp.S * NP * tp.kcap;
    by[B_PI] = sizeof(int) * (size_t)tp.S * NP * tp.kcap;
    by[B_PCNT] = sizeof(int) * (size_t)tp.S * NP;
    by[B_NNI] = sizeof(int) * (size_t)NP * tp.kcap;
    by[B_NNS] = sizeof(double) * (size_t)NP * tp.kcap;
  } else {
    // dense distance rows: every valid neighbour (k < 0), or an exact selection wider than the
    // register lists (k > 64) made from the rows by select_wide
    tp.route = R_DENSE;
    tp.wide = (tp.wass || tp.emax == 0) ? k > 0 : k > 64;
    const int n2 =
      tp.wide ? (int)std::min<long long>(1ll << 30, 1ll << (64 - __builtin_clzll((unsigned long long)k - 1))) : 0;
    tp.gstride = tp.wide && n2 > 4096 ? n2 : 0;
    // all-neighbour S-map systems of up to 64 unknowns are formed ahead of the solve by gram_dense
    const bool gram = tp.smap && !tp.wide && tp.n <= 64;
    const size_t per_target = sizeof(double) * (size_t)std::max(NL, 1) +
                              (tp.wide ? (sizeof(int) + sizeof(double)) * (size_t)k : 0) +
                              (sizeof(unsigned long long) + sizeof(int)) * (size_t)tp.gstride +
                              (gram ? gram_dense_bytes(tp.n, 1, T) : 0);
    tp.nt = (int)std::max<size_t>(1, std::min<size_t>((size_t)std::max(NP, 1), dense_budget / per_target));
    if (!tp.wass && tp.emax > 0) {
      // dense_dist_f64: library slices of whole 256-point tiles, one block per slice and DT_TGT targets
      const int tiles = std::max(1, (NL + 255) / 256);
      tp.S = std::max(1, std::min(std::min(MAX_SLICES, max_slices), tiles));
      tp.sl = (tiles + tp.S - 1) / tp.S * 256;
      tp.S = std::max(1, (NL + tp.sl - 1) / tp.sl);
    } else {
      const int bxmax = tp.wass ? tp.nt : (tp.nt + 255) / 256;
      int S = std::max(1, std::min(MAX_SLICES, (1024 + bxmax - 1) / bxmax));
      S = std::max(1, std::min(S, NL / (tp.wass ? 64 : 256)));
      tp.S = std::max(1, std::min(S, max_slices));
      tp.sl = (NL + tp.S - 1) / tp.S;
    }
    by[B_D] = sizeof(double) * (size_t)tp.nt * std::max(NL, 1);
    by[B_PCNT] = sizeof(int) * (size_t)tp.S * NP;
    by[B_NNI] = tp.wide ? sizeof(int) * (size_t)tp.nt * k : 0;
    by[B_NNS] = tp.wide ? sizeof(double) * (size_t)tp.nt * k : 0;
    by[B_WKEY] = tp.gstride ? sizeof(unsigned long long) * (size_t)tp.nt * tp.gstride : 0;
    by[B_WIDX] = tp.gstride ? sizeof(int) * (size_t)tp.nt * tp.gstride : 0;
    by[B_GRAM] = gram ? gram_dense_bytes(tp.n, tp.nt, T) : 0;
  }
  return 0;
}

// Gives every buffer of the plan a device pointer: get(slot, role, bytes) returns one (nullptr on
// failure); then the kernels' parameter blocks are filled in.
template <class Get>
static int bind_task(TaskPlan& tp, const edm_options* o, Get&& get)
{
  for (int s = 0; s < B_NUM; s++) {
    tp.ptr[s] = nullptr;
    if (tp.bytes[s] == 0)
      continue;
    tp.ptr[s] = get(s, buf_role(s), tp.bytes[s]);
    if (!tp.ptr[s])
      return -2;
  }
  const int NL = tp.NL, NP = tp.NP, Ea = tp.Ea;
  LpParams& P = tp.P;
  P.M = (const double*)tp.ptr[B_M];
  P.Mp = (const double*)tp.ptr[B_MP];
  P.Mtt = (const double*)tp.ptr[B_MTT];
  P.Mppt = (const double*)tp.ptr[B_MPPT];
  P.Mpanel = (const int*)tp.ptr[B_MPAN];
  P.Mppanel = (const int*)tp.ptr[B_MPPAN];
  P.metrics = (const int*)tp.ptr[B_METRICS];
  P.id_weights = (const float*)tp.ptr[B_IDW];
  const double* dthetas = (const double*)tp.ptr[B_THETAS];
  const double* dtargets = (const double*)tp.ptr[B_MT];

  WassParams& W = tp.W;
  std::memset(&W, 0, sizeof(W));
  if (tp.wass) {
    W.M = P.M;
    W.Mp = P.Mp;
    W.Mpanel = P.Mpanel;
    W.Mppanel = P.Mppanel;
    W.metrics = P.metrics;
    W.id_weights = P.id_weights;
    W.NL = NL;
    W.NP = NP;
    W.Ea = Ea;
    W.E = tp.wE;
    W.E_dt = tp.wE_dt;
    W.rows = tp.wrows;
    W.n_unlagged = tp.wunl;
    W.num_panels = o->num_panels;
    W.missing_distance = o->missing_distance;
    W.idw = o->idw;
    W.aspect_ratio = o->aspect_ratio;
    W.panel_term = !o->panel_mode ? 0 : o->idw > 0 ? 1 : P.id_weights ? 2 : 0;
    W.skip_other = o->panel_mode && o->idw < 0;
  }

  FinishParams& F = tp.F;
  std::memset(&F, 0, sizeof(F));
  F.NP = NP;
  F.NL = NL;
  F.k = tp.k;
  F.T = tp.T;
  F.force_compute = o->force_compute;
  F.algorithm = o->algorithm;
  F.thetas = dthetas;
  F.targets = dtargets;
  F.pred = (double*)tp.ptr[B_PRED];
  F.rc = (int*)tp.ptr[B_RC];
  F.k_used = (int*)tp.ptr[B_KUSED];
  F.nn_k = (int*)tp.ptr[B_NNK];
  F.nn_cnt = (int*)tp.ptr[B_NNCNT];
  F.nn_out = (int*)tp.ptr[B_NNOUT];
  F.nn_cap = F.nn_out ? tp.nn_cap : 0;
  F.nn_i = (int*)tp.ptr[B_NNI];
  F.nn_s = (double*)tp.ptr[B_NNS];
  F.nn_stride = tp.route == R_DENSE ? (tp.wide ? tp.k : 0) : tp.kcap;

  SmapParams& Sm = tp.Sm;
  std::memset(&Sm, 0, sizeof(Sm));
  Sm.NP = NP;
  Sm.NL = NL;
  Sm.Ea = Ea;
  Sm.T = tp.T;
  Sm.M = P.M;
  Sm.Mp = P.Mp;
  Sm.targets = dtargets;
  Sm.thetas = dthetas;
  Sm.nn_k = F.nn_k;
  Sm.nn_cnt = F.nn_cnt;
  Sm.pred = F.pred;
  Sm.rc = F.rc;
  Sm.coeffs = o->save_smap_coeffs ? (double*)tp.ptr[B_COEFFS] : nullptr;
  Sm.nn_i = F.nn_i;
  Sm.nn_s = F.nn_s;
  Sm.nn_stride = F.nn_stride;
  Sm.big = (double*)tp.ptr[B_SMAPBIG];
  Sm.big_blocks = Sm.big ? SMAP_BIG_BLOCKS : 0;
  Sm.gram = nullptr;

  if (tp.route == R_MFMA) {
    MfmaParams& MP = tp.MP;
    MP.M32 = (const float*)tp.ptr[B_MIX0];
    MP.Mn = (const float*)tp.ptr[B_MIX1];
    MP.P32 = (const float*)tp.ptr[B_MIX2];
    MP.Pn = (const float*)tp.ptr[B_MIX3];
    MP.ynorm = (const double*)tp.ptr[B_MIX4];
    MP.M = P.M;
    MP.Mp = P.Mp;
    MP.seed = tp.seed_win > 0 ? (const double*)tp.ptr[B_SEED] : nullptr;
    KdScreen& KS = tp.KS;
    KS.m32 = (const float*)tp.ptr[B_KD32];
    KS.mn = (const float*)tp.ptr[B_KDN];
    KS.pos = (const int*)tp.ptr[B_KDPOS];
    KS.cbox = (const float*)tp.ptr[B_CBOX];
    KS.tbox = (const float*)tp.ptr[B_TBOX];
    KS.tord = (const int*)tp.ptr[B_TORD];
  }
  return 0;
}

// Copies the host inputs of the task into its bound buffers (slots of role ROLE_IN whose pointer is
// set; `skip_views` leaves out the manifold arrays, which a device-side caller passes itself).
static int upload_task(const TaskPlan& tp, const edm_manifold_view* M, const edm_manifold_view* Mp,
                       const edm_options* o, hipStream_t s, bool skip_views, const bool* skip = nullptr)
{
  auto cp = [&](int slot, const void* host) {
    if (!tp.ptr[slot] || !host || !tp.bytes[slot] || (skip && skip[slot]))
      return hipSuccess;
    return hipMemcpyAsync(tp.ptr[slot], host, tp.bytes[slot], hipMemcpyHostToDevice, s);
  };
  hipError_t e = hipSuccess;
  if (!skip_views) {
    const void* src[] = { M->data, Mp->data, M->targets, M->target_times, Mp->point_times, M->panel, Mp->panel };
    const int slots[] = { B_M, B_MP, B_MT, B_MTT, B_MPPT, B_MPAN, B_MPPAN };
    for (int i = 0; i < 7 && e == hipSuccess; i++)
      e = cp(slots[i], src[i]);
  }
  if (e == hipSuccess)
    e = cp(B_METRICS, tp.metrics.data());
  if (e == hipSuccess)
    e = cp(B_IDW, o->id_weights);
  if (e == hipSuccess)
    e = cp(B_THETAS, o->thetas);
  if (tp.route == R_MFMA) {
    if (e == hipSuccess)
      e = cp(B_MIX0, tp.m32.data());
    if (e == hipSuccess)
      e = cp(B_MIX1, tp.mn.data());
    if (e == hipSuccess)
      e = cp(B_MIX2, tp.p32.data());
    if (e == hipSuccess)
      e = cp(B_MIX3, tp.pn.data());
    if (e == hipSuccess)
      e = cp(B_MIX4, tp.ynorm.data());
    if (e == hipSuccess)
      e = cp(B_KDP, tp.kdm.data());
    if (e == hipSuccess)
      e = cp(B_WST, tp.wst.data());
    if (e == hipSuccess)
      e = cp(B_TORD, tp.tord.data());
    if (e == hipSuccess)
      e = cp(B_SGRP, tp.sgrp.data());
    if (e == hipSuccess)
      e = cp(B_KD32, tp.kd32.data());
    if (e == hipSuccess)
      e = cp(B_KDN, tp.kdn.data());
    if (e == hipSuccess)
      e = cp(B_KDPOS, tp.kdp.data());
    if (e == hipSuccess)
      e = cp(B_CBOX, tp.cbox.data());
    if (e == hipSuccess)
      e = cp(B_TBOX, tp.tbox.data());
  }
  return e == hipSuccess ? 0 : -3;
}

// Enqueues the task's fills and kernels on stream s.  ev (or nullptr) receives four stage events
// [start, after select, after merge, after S-map].  keep_going is polled between target blocks of the
// dense route.
static int launch_task(TaskPlan& tp, hipStream_t s, hipEvent_t* ev, int (*keep_going)(void))
{
  const int NP = tp.NP, T = tp.T;
  FinishParams& F = tp.F;
  SmapParams& Sm = tp.Sm;
  auto fill = [s](auto* p, size_t n, auto v) {
    using V = std::remove_pointer_t<decltype(p)>;
    if (!p || n == 0)
      return;
    int blocks = (int)std::min<size_t>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(fill_kernel<V>, dim3(blocks), dim3(256), 0, s, p, n, (V)v);
  };
  fill(F.pred, (size_t)T * NP, FEDM_MISSING_D);
  fill(F.rc, (size_t)T * NP, (int)FEDM_UNKNOWN_ERROR);
  fill(F.k_used, (size_t)NP, (int)FEDM_MISSING_I);
  fill((double*)tp.ptr[B_COEFFS], (size_t)NP * tp.n, FEDM_MISSING_D);
  fill(F.nn_out, (size_t)NP * tp.nn_cap, -1);
  fill(F.nn_k, (size_t)NP, 0);
  if (ev)
    HIP_TRY(hipEventRecord(ev[0], s));
  double* ps = (double*)tp.ptr[B_PS];
  int* pi = (int*)tp.ptr[B_PI];
  int* pc = (int*)tp.ptr[B_PCNT];
  if (tp.route != R_DENSE) {
    if (tp.route == R_MFMA) {
      if (tp.MP.seed)
        launch_seed(tp.kp, s, tp.MP, (const double*)tp.ptr[B_KDP], (const int*)tp.ptr[B_WST],
                    (const int*)tp.ptr[B_TORD], (const int*)tp.ptr[B_SGRP], (int)tp.sgrp.size() - 1, tp.seed_win,
                    tp.two_phase || tp.kp == 16 ? tp.seed_win : SEED_WIDE,
                    tp.k, (double*)tp.ptr[B_SEED]);
      if (tp.MP.seed && tp.two_phase) {
        int* fc = (int*)tp.ptr[B_FCNT];
        int* fl = (int*)tp.ptr[B_FLIST];
        int* ff = (int*)tp.ptr[B_FFLAG];
        fill(fc, (size_t)1, 0);
        fill(ff, (size_t)NP, 0);
        fill((int*)tp.ptr[B_CCNT], (size_t)NP, 0);
        launch_two_phase(tp.kcap, tp.kp, s, tp.MP, tp.KS, tp.split, tp.k, tp.S, tp.sl, (int*)tp.ptr[B_CAND],
                         (int*)tp.ptr[B_CCNT],
                         ps, pi, pc, ff, fl, fc);
        double* fs = (double*)tp.ptr[B_FPS];
        int* fi = (int*)tp.ptr[B_FPI];
        int* fcn = (int*)tp.ptr[B_FPC];
        // the targets that overflowed: a wider seed, the streaming kernel over sf sub-slices of the library,
        // then merged into the route's per-slice lists
        launch_seed_wide(tp.kp, s, tp.MP, (const double*)tp.ptr[B_KDP], (const int*)tp.ptr[B_WST], tp.seed_win, fl, fc,
                         tp.k, (double*)tp.ptr[B_SEED]);
        launch_mfma(tp.kcap, tp.kp, dim3(tp.bx, tp.sf), s, tp.MP, tp.k, (tp.NL + tp.sf - 1) / tp.sf, fs, fi, fcn, fl, fc);
        launch_merge_fallback(tp.kcap, s, NP, tp.S, tp.sf, fl, fc, fs, fi, fcn, ps, pi, pc);
      } else {
        launch_mfma(tp.kcap, tp.kp, dim3(tp.bx, tp.S), s, tp.MP, tp.k, tp.sl, ps, pi, pc);
      }
    }
    else
      launch_knn(tp.kcap, tp.emax, tp.generic, dim3(tp.bx, tp.S), s, tp.P, tp.k, tp.sl, ps, pi, pc);
    HIP_TRY(hipGetLastError());
    if (ev)
      HIP_TRY(hipEventRecord(ev[1], s));
    launch_finish_lists(tp.kcap, dim3((NP + 255) / 256), s, F, ps, pi, pc, tp.S);
    HIP_TRY(hipGetLastError());
    if (ev)
      HIP_TRY(hipEventRecord(ev[2], s));
    if (tp.smap) {
      const int r = smap_launch(tp.n, s, Sm, nullptr, 0, NP, tp.smap_lds, tp.smap_mode);
      if (r)
        return r;
      HIP_TRY(hipGetLastError());
    }
    if (ev)
      HIP_TRY(hipEventRecord(ev[3], s));
    return 0;
  }
  double* D = (double*)tp.ptr[B_D];
  unsigned long long* gkey = (unsigned long long*)tp.ptr[B_WKEY];
  int* gidx = (int*)tp.ptr[B_WIDX];
  // stage events: with one block of targets, [0, 1] the distances, [1, 2] the neighbour policy (and
  // the wide selection), [2, 3] the S-map systems and solves; with several, [1, 2] spans the loop
  const bool one = NP <= tp.nt;
  if (ev && !one)
    HIP_TRY(hipEventRecord(ev[1], s));
  for (int t0 = 0; t0 < NP; t0 += tp.nt) {
    const int cnt = std::min(tp.nt, NP - t0);
    if (tp.wass)
      if (tp.wE <= 32)
        hipLaunchKernelGGL((wasserstein_dense<1, 4>), dim3(cnt, tp.S), dim3(256), 0, s, tp.W, t0, cnt, tp.sl, D, pc);
      else
        hipLaunchKernelGGL((wasserstein_dense<2, 2>), dim3(cnt, tp.S), dim3(128), 0, s, tp.W, t0, cnt, tp.sl, D, pc);
    else
      launch_dense(tp.emax, tp.generic, tp.S, s, tp.P, t0, cnt, tp.sl, D, pc);
    HIP_TRY(hipGetLastError());
    if (ev && one)
      HIP_TRY(hipEventRecord(ev[1], s));
    hipLaunchKernelGGL(finish_dense, dim3((cnt + 255) / 256), dim3(256), 0, s, F, D, t0, cnt, pc, tp.S);
    HIP_TRY(hipGetLastError());
    if (tp.wide) {
      hipLaunchKernelGGL(select_wide, dim3(cnt), dim3(256), 0, s, F, D, t0, cnt, gkey, gidx, tp.gstride);
      HIP_TRY(hipGetLastError());
    }
    if (ev && one)
      HIP_TRY(hipEventRecord(ev[2], s));
    Sm.nn_t0 = t0;
    if (tp.smap && tp.ptr[B_GRAM]) {
      Sm.gram = (const double*)tp.ptr[B_GRAM];
      const int r = gram_dense_launch(tp.n, s, Sm, D, t0, cnt, (double*)tp.ptr[B_GRAM]);
      if (r)
        return r;
    }
    if (tp.smap) {
      const int r = smap_launch(tp.n, s, Sm, D, t0, cnt, tp.smap_lds, tp.smap_mode);
      if (r)
        return r;
      HIP_TRY(hipGetLastError());
    }
    if (keep_going && !keep_going())
      break;
  }
  if (ev) {
    if (!one)
      HIP_TRY(hipEventRecord(ev[2], s));
    HIP_TRY(hipEventRecord(ev[3], s));
  }
  return 0;
}

} // namespace fedm

extern "C" {

int edm_predict(edm_ctx* c, const edm_manifold_view* M, const edm_manifold_view* Mp, const edm_options* o,
                double* pred, int* rc, double* coeffs, int* k_used, int* nn_idx, int nn_cap, int (*keep_going)(void))
{
  const int NP = Mp->num_points, Ea = M->E_actual;
  if (!c || Ea != Mp->E_actual || o->num_thetas <= 0 || Ea <= 0)
    return -22;
  if (NP == 0)
    return 0;
  if (keep_going && !keep_going()) {
    host_fill_outputs(Mp, o, pred, rc, coeffs, k_used, nn_idx, nn_cap, Ea, FEDM_UNKNOWN_ERROR);
    for (int i = 0; i < NP; i++)
      rc[i] = FEDM_BREAK_HIT;
    return FEDM_BREAK_HIT;
  }
  TaskPlan tp;
  const int pr = plan_task(M, Mp, o, coeffs != nullptr, nn_idx ? nn_cap : 0, true, MAX_SLICES, tp);
  if (pr == -95)
    host_fill_outputs(Mp, o, pred, rc, coeffs, k_used, nn_idx, nn_cap, Ea, FEDM_UNKNOWN_ERROR);
  if (pr != 0)
    return pr;
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  int r = bind_task(tp, o, [c](int slot, BufRole, size_t bytes) { return c->get(slot, bytes); });
  if (r == 0)
    r = upload_task(tp, M, Mp, o, s, false);
  if (r == 0)
    r = launch_task(tp, s, c->ev, keep_going);
  if (r != 0)
    return r;
  const size_t T = o->num_thetas;
  HIP_TRY(hipMemcpyAsync(pred, tp.F.pred, sizeof(double) * T * NP, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(rc, tp.F.rc, sizeof(int) * T * NP, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(k_used, tp.F.k_used, sizeof(int) * (size_t)NP, hipMemcpyDeviceToHost, s));
  if (coeffs)
    HIP_TRY(hipMemcpyAsync(coeffs, tp.ptr[B_COEFFS], tp.bytes[B_COEFFS], hipMemcpyDeviceToHost, s));
  if (tp.F.nn_out)
    HIP_TRY(hipMemcpyAsync(nn_idx, tp.F.nn_out, tp.bytes[B_NNOUT], hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  c->timings[0] = ev_ms(c->ev[0], c->ev[1]);
  c->timings[1] = ev_ms(c->ev[1], c->ev[2]);
  c->timings[2] = ev_ms(c->ev[2], c->ev[3]);
  c->timings[3] = ev_ms(c->ev[0], c->ev[3]);
  if (keep_going && !keep_going()) {
    for (int i = 0; i < NP; i++)
      if (k_used[i] == FEDM_MISSING_I)
        rc[i] = FEDM_BREAK_HIT;
    return FEDM_BREAK_HIT;
  }
  return 0;
}
int edm_predict_noalloc(edm_ctx* c, const edm_manifold_view* M, const edm_manifold_view* Mp, const edm_options* o,
                        double* pred, int* rc, double* coeffs, int* k_used)
{
  if (!c || !M || !Mp || !o || !pred || !rc || !k_used || M->E_actual != Mp->E_actual || o->num_thetas <= 0 ||
      M->E_actual <= 0)
    return -22;
  if (Mp->num_points == 0)
    return 0;
  if (!c->scratch)
    return c->scratch_err ? c->scratch_err : -2;
  const size_t cap = NOALLOC_SCRATCH_BYTES;
  auto rounded = [](size_t b) { return (b + 255) / 256 * 256; };
  auto in_scratch = [](int slot) {
    return buf_role(slot) == ROLE_SCRATCH || slot == B_METRICS || slot == B_IDW || slot == B_THETAS;
  };
  // the most library slices (and, on the dense route, targets per block) whose buffers fit
  TaskPlan tp;
  size_t need = 0;
  for (int slices = MAX_SLICES;; slices /= 2) {
    tp = TaskPlan();
    const int r = plan_task(M, Mp, o, coeffs != nullptr, 0, false, slices, tp, cap / 2);
    if (r != 0)
      return r;
    need = 0;
    for (int q = 0; q < B_NUM; q++)
      if (in_scratch(q))
        need += rounded(tp.bytes[q]);
    if (need <= cap)
      break;
    if (slices == 1) {
      fprintf(stderr,
              "fastedm_hip: edm_predict_noalloc: this task needs %zu bytes of scratch, more than the %zu bytes "
              "the library holds for it (fewer targets or a smaller k per call)\n",
              need, cap);
      return -12;
    }
  }
  unsigned char* base = c->scratch;
  size_t off = 0;
  const void* view[B_NUM] = {};
  view[B_M] = M->data;
  view[B_MP] = Mp->data;
  view[B_MT] = M->targets;
  view[B_MTT] = M->target_times;
  view[B_MPPT] = Mp->point_times;
  view[B_MPAN] = M->panel;
  view[B_MPPAN] = Mp->panel;
  view[B_PRED] = pred;
  view[B_RC] = rc;
  view[B_KUSED] = k_used;
  view[B_COEFFS] = coeffs;
  hipSetDevice(c->device);
  int r = bind_task(tp, o, [&](int slot, BufRole, size_t bytes) -> void* {
    if (in_scratch(slot)) {
      void* p = base + off;
      off += rounded(bytes);
      return p;
    }
    return const_cast<void*>(view[slot]);
  });
  if (r == 0)
    r = upload_task(tp, M, Mp, o, c->stream, true);
  if (r == 0)
    r = launch_task(tp, c->stream, c->ev, nullptr);
  if (r != 0)
    return r;
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->timings[0] = ev_ms(c->ev[0], c->ev[1]);
  c->timings[1] = ev_ms(c->ev[1], c->ev[2]);
  c->timings[2] = ev_ms(c->ev[2], c->ev[3]);
  c->timings[3] = ev_ms(c->ev[0], c->ev[3]);
  return 0;
}

} // extern "C"

// ------------------------------------------------------------------------------------------------
// Batches of tasks over shared point tables.

namespace fedm {
template <int KCAP>
__global__ void batch_redo(BatchParams B, RedoLists L);

template <int EMAX>
__global__ void rank_build(RankParams R);
template <int KCAP, int WPL>
__global__ void rank_scan_multi(ScanParams S);
__global__ void rank_scatter(ScanParams S);
template <int KCAP, int EMAX>
__global__ void rank_scan(ScanParams S);
} // namespace fedm

constexpr int CTR_INTS = (3 + PR_QUEUES) * PR_CSTRIDE;

struct edm_batch
{
  edm_ctx* ctx = nullptr;
  bool device_route = false;
  int kcap = 0, emax = 0, mixed = 0, T = 0, Ea = 0, tasks = 0;
  int64_t total_out = 0, total_k = 0;
  std::vector<void*> owned;
  BatchParams B;
  int grid = 0;
  int pgrid = 0; // pruned direct kernel
  // rank-table route (one shared prediction list whose tasks draw libraries from one pool)
  bool rank_route = false;
  RankParams R;
  ScanParams S;
  int scan_grid = 0;
  size_t scan_lds = 0;
  int rank_tasks = 0;
  bool multi_scan = false; // one pass per target over every rank task (rank_scan_multi)
  size_t multi_lds = 0;
  float last_ms = 0.f;
  int (*keep_going)(void) = nullptr; // polled before each task or stage launch
  int redo_last = 0; // targets handed to the exact selection in the last run
  int* counters = nullptr;       // the per-run counters (CTR_INTS ints, see edm_batch_create)
  int* handed_host = nullptr;    // pinned copy of the hand-over count, written at the end of each run
  bool outputs_filled = false; // the outputs have held their fill values once (first run)
  bool split_last = false;     // the last run put the direct kernels on the second stream
  // host copies for the task-by-task route
  edm_manifold_view Lv, Pv;
  edm_options o;
  std::vector<double> Ldata, Ltgt, Ltt, Pdata, Ppt, thetas;
  std::vector<int> Lpan, Ppan, lb, lc, lr, pb, pc, pr, kk, metrics;
  std::vector<int64_t> ob;
  std::vector<double> hpred;
  std::vector<int> hrc, hku;
  std::vector<double> hcoeffs;
  // resident task route: every task planned, gathered and uploaded once at create, its kernels
  // launched in turn on each run (or replayed as one captured graph)
  bool resident = false;
  fedm_knobs kn;             // the launch and route choices this batch was built with
  std::vector<TaskPlan> plans;
  double* coeffs = nullptr;  // [total_k x (Ea+1)] when S-map coefficients are saved
  bool use_graph = false;    // capture the steady-state run into a hipGraph and replay it
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int runs = 0;              // completed runs (the graph captures the second one)

  template <typename T>
  T* dev(const T* host, size_t n)
  {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n * sizeof(T), 256)) != hipSuccess)
      return nullptr;
    owned.push_back(p);
    if (host && n && hipMemcpyAsync(p, host, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
      return nullptr;
    return (T*)p;
  }
  ~edm_batch()
  {
    if (ctx)
      hipStreamSynchronize(ctx->stream);
    if (graph_exec)
      hipGraphExecDestroy(graph_exec);
    if (graph)
      hipGraphDestroy(graph);
    for (void* p : owned)
      hipFree(p);
    if (handed_host)
      hipHostFree(handed_host);
  }
};

template <int KCAP, int EMAX>
static const void* batch_kernel_ke(bool alt)
{
  return alt ? (const void*)batch_simplex<KCAP, EMAX, batch_waves_alt(KCAP, EMAX)>
             : (const void*)batch_simplex<KCAP, EMAX, batch_waves(KCAP, EMAX)>;
}
template <int KCAP>
static const void* batch_kernel_k(int demax, bool alt)
{
  switch (demax) {
    case 2: return batch_kernel_ke<KCAP, 2>(alt);
    case 3: return batch_kernel_ke<KCAP, 3>(alt);
    case 4: return batch_kernel_ke<KCAP, 4>(alt);
    case 8: return batch_kernel_ke<KCAP, 8>(alt);
    case 16: return batch_kernel_ke<KCAP, 16>(alt);
    default: return batch_kernel_ke<KCAP, 32>(alt);
  }
}
template <int KCAP>
static const void* rank_scan_multi_k(int words)
{
  return words <= 64 ? (const void*)rank_scan_multi<KCAP, 1>
                     : words <= 128 ? (const void*)rank_scan_multi<KCAP, 2> : (const void*)rank_scan_multi<KCAP, 4>;
}
static const void* rank_scan_multi_kernel(int kcap, int words)
{
  switch (kcap) {
    case 4: return rank_scan_multi_k<4>(words);
    case 8: return rank_scan_multi_k<8>(words);
    case 16: return rank_scan_multi_k<16>(words);
    default: return rank_scan_multi_k<32>(words);
  }
}
template <int KCAP>
static const void* batch_pruned_k(int demax)
{
  switch (demax) {
    case 2: return (const void*)batch_simplex_pruned<KCAP, 2>;
    case 3: return (const void*)batch_simplex_pruned<KCAP, 3>;
    case 4: return (const void*)batch_simplex_pruned<KCAP, 4>;
    default: return (const void*)batch_simplex_pruned<KCAP, 8>;
  }
}
static const void* batch_pruned_kernel(int kcap, int demax)
{
  switch (kcap) {
    case 4: return batch_pruned_k<4>(demax);
    case 8: return batch_pruned_k<8>(demax);
    case 16: return batch_pruned_k<16>(demax);
    default: return batch_pruned_k<32>(demax);
  }
}
static const void* batch_redo_kernel(int kcap)
{
  switch (kcap) {
    case 4: return (const void*)batch_redo<4>;
    case 8: return (const void*)batch_redo<8>;
    case 16: return (const void*)batch_redo<16>;
    default: return (const void*)batch_redo<32>;
  }
}
// The direct kernel's padded dimension: exact for E_actual 2 and 3, else the next of 4, 8, 16, 32.
static int direct_emax(int Ea)
{
  return Ea <= 2 ? 2 : Ea == 3 ? 3 : Ea <= 4 ? 4 : Ea <= 8 ? 8 : Ea <= 16 ? 16 : 32;
}
static const void* batch_kernel(int kcap, int Ea, bool alt)
{
  switch (kcap) {
    case 4: return batch_kernel_k<4>(direct_emax(Ea), alt);
    case 8: return batch_kernel_k<8>(direct_emax(Ea), alt);
    case 16: return batch_kernel_k<16>(direct_emax(Ea), alt);
    default: return batch_kernel_k<32>(direct_emax(Ea), alt);
  }
}

// Smallest library that takes the rank route: scanning a ranked prefix costs about k * pool / L
// entries per target, the direct scan L pairs; the scale is a knob (default 1.5, FEDM_RANK_LTHR_SCALE).
static int rank_threshold(int kmax, int npool, double scale)
{
  return (int)std::max(scale * std::sqrt((double)std::max(kmax, 1) * npool), 2.0 * (kmax + RANK_ZCAP));
}

extern "C" int edm_selftest_mfma(float* out256)
{
  float* d = nullptr;
  if (hipMalloc(&d, 256 * sizeof(float)) != hipSuccess)
    return -2;
  hipLaunchKernelGGL(mfma_layout_probe, dim3(1), dim3(64), 0, 0, d);
  hipError_t e = hipMemcpy(out256, d, 256 * sizeof(float), hipMemcpyDeviceToHost);
  hipFree(d);
  return e == hipSuccess ? 0 : -3;
}

namespace fedm {
__global__ void exp_probe(const double* x, double* out, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    out[i] = fdlibm_exp(x[i]);
}
} // namespace fedm

extern "C" int edm_selftest_exp(const double* x, double* out, int n)
{
  if (n <= 0)
    return 0;
  double *dx = nullptr, *dy = nullptr;
  if (hipMalloc(&dx, sizeof(double) * n) != hipSuccess)
    return -2;
  if (hipMalloc(&dy, sizeof(double) * n) != hipSuccess) {
    hipFree(dx);
    return -2;
  }
  hipError_t e = hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(exp_probe, dim3((n + 255) / 256), dim3(256), 0, 0, dx, dy, n);
    e = hipMemcpy(out, dy, sizeof(double) * n, hipMemcpyDeviceToHost);
  }
  hipFree(dx);
  hipFree(dy);
  return e == hipSuccess ? 0 : -3;
}

extern "C" int fedm_rank_threshold(int kmax, int npool)
{
  return rank_threshold(kmax, npool, default_knobs().rank_scale);
}

static const void* rank_build_kernel(int emax)
{
  switch (emax) {
    case 4: return (const void*)rank_build<4>;
    case 8: return (const void*)rank_build<8>;
    case 16: return (const void*)rank_build<16>;
    default: return (const void*)rank_build<32>;
  }
}
template <int KCAP>
static const void* rank_scan_k(int emax)
{
  switch (emax) {
    case 4: return (const void*)rank_scan<KCAP, 4>;
    case 8: return (const void*)rank_scan<KCAP, 8>;
    case 16: return (const void*)rank_scan<KCAP, 16>;
    default: return (const void*)rank_scan<KCAP, 32>;
  }
}
static const void* rank_scan_kernel(int kcap, int emax)
{
  switch (kcap) {
    case 4: return rank_scan_k<4>(emax);
    case 8: return rank_scan_k<8>(emax);
    case 16: return rank_scan_k<16>(emax);
    default: return rank_scan_k<32>(emax);
  }
}

// A task's library and prediction manifolds, gathered from the batch's host tables by row lists.
struct GatheredTask
{
  std::vector<double> md, mt, mtt, mpd, mppt;
  std::vector<int> mpan, mppan;
  edm_manifold_view M, Mp;
};

static void gather_task(const edm_batch* b, int tk, GatheredTask& g)
{
  const int Ea = b->Ea, nl = b->lc[tk], np = b->pc[tk];
  g.md.resize((size_t)nl * Ea);
  g.mt.resize(nl);
  g.mtt.resize(nl);
  g.mpd.resize((size_t)np * Ea);
  g.mppt.resize(np);
  g.mpan.clear();
  g.mppan.clear();
  for (int r = 0; r < nl; r++) {
    const int row = b->lr[b->lb[tk] + r];
    std::copy(b->Ldata.begin() + (size_t)row * Ea, b->Ldata.begin() + (size_t)(row + 1) * Ea,
              g.md.begin() + (size_t)r * Ea);
    g.mt[r] = b->Ltgt[row];
    g.mtt[r] = b->Ltt.empty() ? 0.0 : b->Ltt[row];
    if (!b->Lpan.empty())
      g.mpan.push_back(b->Lpan[row]);
  }
  for (int r = 0; r < np; r++) {
    const int row = b->pr[b->pb[tk] + r];
    std::copy(b->Pdata.begin() + (size_t)row * Ea, b->Pdata.begin() + (size_t)(row + 1) * Ea,
              g.mpd.begin() + (size_t)r * Ea);
    g.mppt[r] = b->Ppt.empty() ? 0.0 : b->Ppt[row];
    if (!b->Ppan.empty())
      g.mppan.push_back(b->Ppan[row]);
  }
  g.M = b->Lv;
  g.Mp = b->Pv;
  g.M.data = g.md.data();
  g.M.targets = g.mt.data();
  g.M.target_times = g.mtt.data();
  g.M.point_times = nullptr;
  g.M.panel = g.mpan.empty() ? nullptr : g.mpan.data();
  g.M.num_points = nl;
  g.Mp.data = g.mpd.data();
  g.Mp.targets = nullptr;
  g.Mp.point_times = g.mppt.data();
  g.Mp.target_times = nullptr;
  g.Mp.panel = g.mppan.empty() ? nullptr : g.mppan.data();
  g.Mp.num_points = np;
}

static edm_options task_options(const edm_batch* b, int tk)
{
  edm_options to = b->o;
  to.k = b->kk[tk];
  to.thetas = b->thetas.data();
  to.metrics = b->metrics.empty() ? nullptr : b->metrics.data();
  return to;
}

// The resident task route for batches the persistent kernels do not take (S-map, missing values,
// panels, other distances, wide k): every task is planned, gathered and uploaded once, here; the
// prediction side of tasks that share a prediction list is uploaded once; the scratch of the
// kernels is one set of buffers sized for the largest task, since tasks run one after another on
// the stream.  Results go straight to the batch's output arrays.  Returns 0, 1 when the route does
// not apply (a task this build cannot run, or inputs beyond the budget: the batch then gathers and
// uploads task by task on each run, as before), or a negative error.
static int build_resident(edm_batch* b)
{
  edm_ctx* c = b->ctx;
  const int Ea = b->Ea, T = b->T, C = Ea + 1;
  const bool coeffs = b->o.save_smap_coeffs && b->o.algorithm == FEDM_SMAP;
  {
    size_t in_bytes = 0;
    std::map<std::pair<int, int>, int> lists;
    for (int tk = 0; tk < b->tasks; tk++) {
      in_bytes += (size_t)b->lc[tk] * (sizeof(double) * (Ea + 2) + sizeof(int));
      if (lists.emplace(std::make_pair(b->pb[tk], b->pc[tk]), tk).second)
        in_bytes += (size_t)b->pc[tk] * (sizeof(double) * (Ea + 1) + sizeof(int));
    }
    if (in_bytes > ((size_t)16 << 30))
      return 1;
  }
  double* pred = b->dev<double>(nullptr, (size_t)b->total_out);
  int* rc = b->dev<int>(nullptr, (size_t)b->total_out);
  int* ku = b->dev<int>(nullptr, (size_t)b->total_k);
  b->coeffs = coeffs ? b->dev<double>(nullptr, (size_t)b->total_k * C) : nullptr;
  if (!pred || !rc || !ku || (coeffs && !b->coeffs))
    return -2;
  // tasks without targets launch nothing: their outputs keep these values
  dfill(c, pred, (size_t)b->total_out, FEDM_MISSING_D);
  dfill(c, rc, (size_t)b->total_out, (int)FEDM_UNKNOWN_ERROR);
  dfill(c, ku, (size_t)b->total_k, (int)FEDM_MISSING_I);
  if (b->coeffs)
    dfill(c, b->coeffs, (size_t)b->total_k * C, FEDM_MISSING_D);
  b->plans.assign(b->tasks, TaskPlan());
  size_t smax[B_NUM] = {};
  std::map<std::pair<int, int>, std::array<void*, 3>> shared;
  const int pslots[3] = { B_MP, B_MPPT, B_MPPAN };
  GatheredTask g;
  for (int tk = 0; tk < b->tasks; tk++) {
    TaskPlan& tp = b->plans[tk];
    if (b->pc[tk] == 0)
      continue;
    gather_task(b, tk, g);
    const edm_options to = task_options(b, tk);
    int r = plan_task(&g.M, &g.Mp, &to, coeffs, 0, true, MAX_SLICES, tp, (size_t)1 << 30, &b->kn);
    if (r != 0)
      return r == -95 ? 1 : r;
    const int64_t ob = b->ob[tk];
    const auto key = std::make_pair(b->pb[tk], b->pc[tk]);
    const auto it = shared.find(key);
    bool skip[B_NUM] = {};
    r = bind_task(tp, &to, [&](int slot, BufRole role, size_t bytes) -> void* {
      if (role == ROLE_SCRATCH) {
        smax[slot] = std::max(smax[slot], bytes);
        return (void*)&smax[slot]; // a placeholder until the shared scratch exists (never launched)
      }
      if (role == ROLE_OUT) {
        switch (slot) {
          case B_PRED: return pred + ob;
          case B_RC: return rc + ob;
          case B_KUSED: return ku + ob / T;
          case B_COEFFS: return b->coeffs ? b->coeffs + (size_t)(ob / T) * C : nullptr;
          default: return nullptr;
        }
      }
      for (int q = 0; q < 3; q++)
        if (slot == pslots[q] && it != shared.end() && it->second[q]) {
          skip[slot] = true;
          return it->second[q];
        }
      return b->dev<char>(nullptr, bytes);
    });
    if (r != 0)
      return r;
    if (it == shared.end())
      shared[key] = { tp.ptr[B_MP], tp.ptr[B_MPPT], tp.ptr[B_MPPAN] };
    r = upload_task(tp, &g.M, &g.Mp, &to, c->stream, false, skip);
    if (r != 0)
      return r;
    HIP_TRY(hipStreamSynchronize(c->stream));
    // the host copies of the matrix-core operands are uploaded: drop them
    std::vector<float>().swap(tp.m32);
    std::vector<float>().swap(tp.mn);
    std::vector<float>().swap(tp.p32);
    std::vector<float>().swap(tp.pn);
    std::vector<double>().swap(tp.ynorm);
  }
  void* sp[B_NUM] = {};
  for (int q = 0; q < B_NUM; q++)
    if (smax[q]) {
      sp[q] = b->dev<char>(nullptr, smax[q]);
      if (!sp[q])
        return -2;
    }
  for (int tk = 0; tk < b->tasks; tk++) {
    TaskPlan& tp = b->plans[tk];
    if (tp.NP == 0)
      continue;
    void* old[B_NUM];
    std::copy(tp.ptr, tp.ptr + B_NUM, old);
    const edm_options to = task_options(b, tk);
    const int r = bind_task(tp, &to, [&](int slot, BufRole role, size_t) { return role == ROLE_SCRATCH ? sp[slot] : old[slot]; });
    if (r != 0)
      return r;
  }
  b->B.pred = pred;
  b->B.rc = rc;
  b->B.k_used = ku;
  b->resident = true;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

static int batch_break(edm_batch* b);

// One run of the resident route: each task's fills and kernels in task order, or, with graphs on
// and after one eager run, the captured sequence replayed as a single graph launch.
static int run_resident(edm_batch* b)
{
  edm_ctx* c = b->ctx;
  hipStream_t s = c->stream;
  if (b->keep_going && !b->keep_going())
    return batch_break(b);
  HIP_TRY(hipEventRecord(c->ev[4], s));
  if (b->use_graph && b->runs >= 1) {
    if (!b->graph_exec) {
      HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
      int r = 0;
      for (int tk = 0; tk < b->tasks && r == 0; tk++)
        if (b->plans[tk].NP > 0)
          r = launch_task(b->plans[tk], s, nullptr, nullptr);
      hipGraph_t gr = nullptr;
      const hipError_t e = hipStreamEndCapture(s, &gr);
      if (r != 0 || e != hipSuccess) {
        if (gr)
          hipGraphDestroy(gr);
        return r != 0 ? r : -(int)e - 1000;
      }
      b->graph = gr;
      HIP_TRY(hipGraphInstantiate(&b->graph_exec, gr, nullptr, nullptr, 0));
    }
    HIP_TRY(hipGraphLaunch(b->graph_exec, s));
  } else {
    int active = 0;
    for (int tk = 0; tk < b->tasks; tk++)
      active += b->plans[tk].NP > 0;
    // a single task records its stage events, as edm_predict does
    hipEvent_t* ev = active == 1 ? c->ev : nullptr;
    for (int tk = 0; tk < b->tasks; tk++) {
      if (b->plans[tk].NP == 0)
        continue;
      const int r = launch_task(b->plans[tk], s, ev, nullptr);
      if (r != 0)
        return r;
      if (b->keep_going) {
        // with a callback, the host waits for each task so that a stop takes effect between tasks
        HIP_TRY(hipStreamSynchronize(s));
        if (!b->keep_going())
          return batch_break(b);
      }
    }
  }
  HIP_TRY(hipEventRecord(c->ev[5], s));
  HIP_TRY(hipEventSynchronize(c->ev[5]));
  b->runs++;
  b->last_ms = ev_ms(c->ev[4], c->ev[5]);
  int active = 0;
  for (int tk = 0; tk < b->tasks; tk++)
    active += b->plans[tk].NP > 0;
  if (active == 1 && !(b->use_graph && b->graph_exec)) {
    // [0] distance/selection, [1] merge, [2] S-map, [3] whole run (edm_predict's layout)
    c->timings[0] = ev_ms(c->ev[0], c->ev[1]);
    c->timings[1] = ev_ms(c->ev[1], c->ev[2]);
    c->timings[2] = ev_ms(c->ev[2], c->ev[3]);
  } else {
    c->timings[0] = b->last_ms;
    c->timings[1] = c->timings[2] = 0.0;
  }
  c->timings[3] = b->last_ms;
  c->last_redo = 0;
  return 0;
}

extern "C" {

int edm_batch_create(edm_ctx* c, const edm_manifold_view* L, const edm_manifold_view* P, const edm_options* o,
                     int num_tasks, const int* lib_begin, const int* lib_count, const int* lib_rows,
                     const int* pred_begin, const int* pred_count, const int* pred_rows, const int* k,
                     const int64_t* out_begin, edm_batch** out)
{
  return edm_batch_create_knobs(c, L, P, o, num_tasks, lib_begin, lib_count, lib_rows, pred_begin, pred_count,
                                pred_rows, k, out_begin, nullptr, out);
}

void edm_batch_knobs(const edm_batch* b, fedm_knobs* out)
{
  *out = b->kn;
}

int edm_batch_create_knobs(edm_ctx* c, const edm_manifold_view* L, const edm_manifold_view* P, const edm_options* o,
                           int num_tasks, const int* lib_begin, const int* lib_count, const int* lib_rows,
                           const int* pred_begin, const int* pred_count, const int* pred_rows, const int* k,
                           const int64_t* out_begin, const fedm_knobs* knobs, edm_batch** out)
{
  *out = nullptr;
  if (!c || num_tasks < 0 || L->E_actual != P->E_actual || o->num_thetas <= 0)
    return -22;
  hipSetDevice(c->device);
  auto b = new edm_batch();
  b->ctx = c;
  b->kn = knobs ? *knobs : default_knobs();
  b->tasks = num_tasks;
  b->T = o->num_thetas;
  b->Ea = L->E_actual;
  const int Ea = b->Ea;
  int kmax = 0, kmin = INT_MAX;
  size_t lib_total = 0, pred_total = 0;
  for (int t = 0; t < num_tasks; t++) {
    kmax = std::max(kmax, k[t]);
    kmin = std::min(kmin, k[t]);
    lib_total = std::max(lib_total, (size_t)lib_begin[t] + lib_count[t]);
    pred_total = std::max(pred_total, (size_t)pred_begin[t] + pred_count[t]);
    b->total_out = std::max(b->total_out, out_begin[t] + (int64_t)b->T * pred_count[t]);
  }
  b->total_k = b->total_out / b->T;
  // host copies (every route keeps them: they make the task-by-task route and fetch possible)
  b->Lv = *L;
  b->Pv = *P;
  b->o = *o;
  b->Ldata.assign(L->data, L->data + (size_t)L->num_points * Ea);
  b->Ltgt.assign(L->targets, L->targets + L->num_points);
  if (L->target_times)
    b->Ltt.assign(L->target_times, L->target_times + L->num_points);
  if (L->panel)
    b->Lpan.assign(L->panel, L->panel + L->num_points);
  b->Pdata.assign(P->data, P->data + (size_t)P->num_points * Ea);
  if (P->point_times)
    b->Ppt.assign(P->point_times, P->point_times + P->num_points);
  if (P->panel)
    b->Ppan.assign(P->panel, P->panel + P->num_points);
  b->thetas.assign(o->thetas, o->thetas + o->num_thetas);
  if (o->metrics)
    b->metrics.assign(o->metrics, o->metrics + Ea);
  b->lb.assign(lib_begin, lib_begin + num_tasks);
  b->lc.assign(lib_count, lib_count + num_tasks);
  b->lr.assign(lib_rows, lib_rows + lib_total);
  b->pb.assign(pred_begin, pred_begin + num_tasks);
  b->pc.assign(pred_count, pred_count + num_tasks);
  b->pr.assign(pred_rows, pred_rows + pred_total);
  b->kk.assign(k, k + num_tasks);
  b->ob.assign(out_begin, out_begin + num_tasks);

  const bool generic = o->has_missing || o->has_categorical || o->panel_mode || o->use_only_past || o->low_memory ||
                       o->distance != FEDM_DIST_EUCLIDEAN;
  const int emax = Ea <= 4 ? 4 : Ea <= 8 ? 8 : Ea <= 16 ? 16 : Ea <= 32 ? 32 : -1;
  const int kcap = kmax <= 4 ? 4 : kmax <= 8 ? 8 : kmax <= 16 ? 16 : kmax <= 32 ? 32 : -1;
  // largest coordinate magnitude over the rows the tasks use (scales the FP32 screening bound)
  double A = 0;
  {
    std::vector<char> usedL(L->num_points, 0), usedP(P->num_points, 0);
    for (size_t i = 0; i < lib_total; i++)
      usedL[lib_rows[i]] = 1;
    for (size_t i = 0; i < pred_total; i++)
      usedP[pred_rows[i]] = 1;
    for (int r = 0; r < L->num_points; r++)
      if (usedL[r])
        for (int j = 0; j < Ea; j++)
          A = std::max(A, std::fabs(b->Ldata[(size_t)r * Ea + j]));
    for (int r = 0; r < P->num_points; r++)
      if (usedP[r])
        for (int j = 0; j < Ea; j++)
          A = std::max(A, std::fabs(b->Pdata[(size_t)r * Ea + j]));
  }
  b->device_route = o->algorithm == FEDM_SIMPLEX && !generic && emax > 0 && kcap > 0 && kmin >= 1 && A < 1e30;
  if (!b->device_route) {
    const char* env = getenv("FEDM_RESIDENT");
    const int r = (env && env[0] == '0') ? 1 : build_resident(b);
    if (r < 0) {
      delete b;
      return r;
    }
    *out = b;
    return 0;
  }
  b->kcap = kcap;
  b->emax = emax;
  b->mixed = o->precision == FEDM_MIXED;

  BatchParams& B = b->B;
  std::memset(&B, 0, sizeof(B));
  B.L = b->dev(b->Ldata.data(), b->Ldata.size());
  if (b->mixed) {
    // FP32 copy for screening; rows outside every task hold MISSING and saturate to infinity
    std::vector<float> l32(b->Ldata.begin(), b->Ldata.end());
    B.L32 = b->dev(l32.data(), l32.size());
    hipStreamSynchronize(c->stream);
  }
  B.Ltgt = b->dev(b->Ltgt.data(), b->Ltgt.size());
  B.P = b->dev(b->Pdata.data(), b->Pdata.size());
  B.lib_begin = b->dev(lib_begin, num_tasks);
  B.lib_count = b->dev(lib_count, num_tasks);
  B.lib_rows = b->dev(lib_rows, lib_total);
  B.pred_begin = b->dev(pred_begin, num_tasks);
  B.pred_count = b->dev(pred_count, num_tasks);
  B.pred_rows = b->dev(pred_rows, pred_total);
  B.k = b->dev(k, num_tasks);
  B.out_begin = b->dev(out_begin, num_tasks);
  // Rank-table route: the largest set of tasks sharing one prediction list, when their libraries are
  // large enough that scanning a ranked prefix beats scanning the library (k * pool / L < L).
  std::vector<char> is_rank(num_tasks, 0);
  int rank_pb = -1, rank_pc = 0;
  std::vector<int> pool;
  {
    std::map<std::pair<int, int>, std::vector<int>> by_list;
    for (int t = 0; t < num_tasks; t++)
      by_list[{ pred_begin[t], pred_count[t] }].push_back(t);
    size_t best = 0;
    std::pair<int, int> key{ -1, 0 };
    for (auto& kv : by_list)
      if (kv.second.size() > best) {
        best = kv.second.size();
        key = kv.first;
      }
    const auto& members = by_list[key];
    std::vector<char> inpool(L->num_points, 0);
    for (int t : members)
      for (int i = 0; i < lib_count[t]; i++)
        inpool[lib_rows[lib_begin[t] + i]] = 1;
    for (int r = 0; r < L->num_points; r++)
      if (inpool[r])
        pool.push_back(r);
    const double npool = (double)pool.size();
    const int lthr = rank_threshold(kmax, (int)pool.size(), b->kn.rank_scale);
    double work = 0;
    int nrank = 0;
    // library positions must follow table rows (ties between equal distances go by position)
    auto ascending = [&](int t) {
      for (int i = 1; i < lib_count[t]; i++)
        if (lib_rows[lib_begin[t] + i] <= lib_rows[lib_begin[t] + i - 1])
          return false;
      return true;
    };
    std::vector<int> eligible;
    for (int t : members)
      if (lib_count[t] >= lthr && ascending(t)) {
        work += lib_count[t];
        nrank++;
        eligible.push_back(t);
      }
    const char* env = getenv("FEDM_RANK_ROUTE");
    const bool allowed = !(env && env[0] == '0');
    if (allowed && nrank > 0 && work >= 4.0 * npool && npool >= 4096 && L->num_points <= (1 << 18) &&
        key.second > 0) {
      b->rank_route = true;
      rank_pb = key.first;
      rank_pc = key.second;
      for (int t : eligible)
        is_rank[t] = 1;
      b->rank_tasks = nrank;
    }
  }
  // work queues: (task, first target) items, heaviest library first
  std::vector<int> order(num_tasks);
  for (int t = 0; t < num_tasks; t++)
    order[t] = t;
  std::stable_sort(order.begin(), order.end(), [&](int a, int q) { return lib_count[a] > lib_count[q]; });
  std::vector<int2> items, ritems, pitems;
  // libraries of PR_MIN..PR_BCH rows with points of at most PR_EMAX coordinates: the pruned direct kernel
  const bool prune_ok = b->kn.direct_prune != 0 && direct_emax(Ea) <= PR_EMAX;
  const int pr_min = PR_MIN;
  std::vector<char> is_pruned(num_tasks, 0);
  for (int t : order) {
    is_pruned[t] = prune_ok && !is_rank[t] && lib_count[t] >= pr_min && lib_count[t] <= PR_BCH;
    for (int t0 = 0; t0 < pred_count[t]; t0 += 256)
      (is_rank[t] ? ritems : is_pruned[t] ? pitems : items).push_back(make_int2(t, t0));
  }
  // every per-run counter in one buffer, each on a cache line of its own, so that one fill zeroes them:
  // the direct kernel's queue head, the hand-over count, the rank scan's queue head, the pruned kernel's
  // PR_QUEUES queue heads
  b->counters = b->dev<int>(nullptr, (size_t)CTR_INTS);
  if (hipHostMalloc((void**)&b->handed_host, sizeof(int)) != hipSuccess)
    b->handed_host = nullptr;
  B.counter = b->counters;
  // targets whose squared-distance ranking may differ from the sqrt ranking (rare; see kernels_batch.hip)
  B.redo_cap = (int)std::min<int64_t>(b->total_k, 1 << 20);
  B.redo = b->dev<int2>(nullptr, (size_t)std::max(B.redo_cap, 1));
  B.redo_count = b->counters + PR_CSTRIDE;
  B.thetas = b->dev(o->thetas, (size_t)o->num_thetas);
  B.T = o->num_thetas;
  B.Ea = Ea;
  B.force_compute = o->force_compute;
  const double u = 5.9604644775390625e-08; // 2^-24
  const double Ap = A * (1 + u);
  B.c1 = (4 * u * Ap + 4 * u * u * Ap) * std::sqrt((double)Ea);
  B.c2 = 2.01 * u + 1.01 * Ea * u / (1 - Ea * u);
  B.c3 = 4.04 * Ea * u * u * Ap * Ap;
  {
    // |(|x|^2 - 2 x.y) + |y|^2 - sum (x_j - y_j)^2| <= (3 Ea + 2) 2^-53 (|x| + |y|)^2 for the FP64 forms
    // (|x|^2 and |y|^2 summed, the cross terms by fused multiply-adds, the direct sum in coordinate
    // order), and (|x| + |y|)^2 <= 4 Ea A^2; the threshold's own rounding adds at most a few 2^-53
    // (|x| + |y|)^2 more
    const double u64 = 1.1102230246251565e-16; // 2^-53
    const double A2 = A * A * (1 + 1e-12);
    B.c4 = 4.0 * (Ea + 4) * u64 * 4.0 * Ea * A2;
  }
  B.pred = b->dev<double>(nullptr, (size_t)b->total_out);
  B.rc = b->dev<int>(nullptr, (size_t)b->total_out);
  B.k_used = b->dev<int>(nullptr, (size_t)b->total_k);
  if (b->rank_route) {
    if (!b->mixed) {
      // the table build screens in FP32; the plain form keeps its FP64-only route
      b->rank_route = false;
      items.insert(items.end(), ritems.begin(), ritems.end());
      ritems.clear();
    }
  }
  B.items = b->dev(items.data(), items.size());
  B.num_items = (int)items.size();
  if (b->rank_route == false)
    for (int t = 0; t < num_tasks; t++)
      is_rank[t] = 0;
  {
    // deal the pruned tasks to PR_QUEUES queues (largest work first, to the lightest queue), keeping the
    // items of a task together and in the order above within each queue
    const int nq = PR_QUEUES;
    B.pq_n = nq;
    std::vector<double> load(nq, 0.0);
    std::vector<int> qof(num_tasks, 0);
    for (int t : order)
      if (is_pruned[t]) {
        const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        qof[t] = q;
        load[q] += (double)lib_count[t] * pred_count[t];
      }
    std::vector<int2> sorted;
    std::vector<int> qb(nq + 1, 0);
    for (int q = 0; q < nq; q++) {
      qb[q] = (int)sorted.size();
      for (const int2& it : pitems)
        if (qof[it.x] == q)
          sorted.push_back(it);
    }
    qb[nq] = (int)sorted.size();
    pitems.swap(sorted);
    B.pq_begin = b->dev(qb.data(), qb.size());
  }
  B.pitems = b->dev(pitems.data(), pitems.size());
  B.num_pitems = (int)pitems.size();
  B.pcounter = b->counters + 3 * PR_CSTRIDE;
  {
    // spatial order of every prediction list a pruned task uses: Morton order of the first (up to three)
    // coordinates, quantised to 21 bits over the list's range
    std::vector<int> tperm(std::max<size_t>(pred_total, 1), 0);
    std::map<std::pair<int, int>, int> done;
    for (int t = 0; t < num_tasks; t++) {
      if (!is_pruned[t] || !done.emplace(std::make_pair(pred_begin[t], pred_count[t]), t).second)
        continue;
      const int pb0 = pred_begin[t], n = pred_count[t], dims = std::min(Ea, 3);
      double lo[3] = { 0, 0, 0 }, hi[3] = { 0, 0, 0 };
      for (int d = 0; d < dims; d++) {
        lo[d] = hi[d] = b->Pdata[(size_t)pred_rows[pb0] * Ea + d];
        for (int i = 0; i < n; i++) {
          const double v = b->Pdata[(size_t)pred_rows[pb0 + i] * Ea + d];
          lo[d] = std::min(lo[d], v);
          hi[d] = std::max(hi[d], v);
        }
      }
      std::vector<std::pair<uint64_t, int>> key(n);
      for (int i = 0; i < n; i++) {
        uint64_t code = 0;
        unsigned qd[3] = { 0, 0, 0 };
        for (int d = 0; d < dims; d++) {
          const double v = b->Pdata[(size_t)pred_rows[pb0 + i] * Ea + d];
          const double f = hi[d] > lo[d] ? (v - lo[d]) / (hi[d] - lo[d]) : 0.0;
          qd[d] = (unsigned)std::min(2097151.0, std::max(0.0, f * 2097151.0));
        }
        for (int bit = 20; bit >= 0; bit--)
          for (int d = 0; d < dims; d++)
            code = (code << 1) | ((qd[d] >> bit) & 1u);
        key[i] = { code, i };
      }
      std::stable_sort(key.begin(), key.end());
      for (int i = 0; i < n; i++)
        tperm[pb0 + i] = key[i].second;
    }
    B.tperm = b->dev(tperm.data(), tperm.size());
  }
  {
    // the direct kernel's tasks: library rows and targets gathered contiguously per task, and the
    // prediction rows in list order, so that its loads need no row-list indirection
    std::vector<int64_t> gbeg(num_tasks, 0);
    int64_t gtot = 0;
    for (int t = 0; t < num_tasks; t++)
      if (!(b->rank_route && is_rank[t])) {
        gbeg[t] = gtot;
        gtot += lib_count[t];
      }
    std::vector<double> lg((size_t)gtot * Ea), tg((size_t)gtot);
    for (int t = 0; t < num_tasks; t++)
      if (!(b->rank_route && is_rank[t]))
        for (int i = 0; i < lib_count[t]; i++) {
          const int row = lib_rows[lib_begin[t] + i];
          std::copy(b->Ldata.begin() + (size_t)row * Ea, b->Ldata.begin() + (size_t)(row + 1) * Ea,
                    lg.begin() + (size_t)(gbeg[t] + i) * Ea);
          tg[(size_t)(gbeg[t] + i)] = b->Ltgt[row];
        }
    std::vector<double> pg((size_t)pred_total * Ea);
    for (size_t i = 0; i < pred_total; i++)
      std::copy(b->Pdata.begin() + (size_t)pred_rows[i] * Ea, b->Pdata.begin() + (size_t)(pred_rows[i] + 1) * Ea,
                pg.begin() + i * Ea);
    B.Lg = b->dev(lg.data(), lg.size());
    B.Tg = b->dev(tg.data(), tg.size());
    B.gbeg = b->dev(gbeg.data(), gbeg.size());
    B.Pg = b->dev(pg.data(), pg.size());
    hipStreamSynchronize(c->stream);
  }
  if (b->rank_route) {
    RankParams& R = b->R;
    std::memset(&R, 0, sizeof(R));
    R.L = B.L;
    R.L32 = B.L32;
    R.P = B.P;
    {
      // the pool in the k-d order of the seed windows (its 64-row chunks then cover compact regions), with
      // each chunk's box: rank_build screens only the chunks that may hold a row within its cut
      std::vector<double> pts(pool.size() * (size_t)Ea);
      for (size_t q = 0; q < pool.size(); q++)
        std::copy(b->Ldata.begin() + (size_t)pool[q] * Ea, b->Ldata.begin() + (size_t)(pool[q] + 1) * Ea,
                  pts.begin() + q * Ea);
      std::vector<int> kdp, wst, tord, sgrp;
      kd_seed_windows(pts.data(), (int)pool.size(), nullptr, 0, Ea, SEED_WIN, kdp, wst, tord, sgrp);
      std::vector<int> kpool(pool.size());
      for (size_t q = 0; q < pool.size(); q++)
        kpool[q] = pool[kdp[q]];
      pool.swap(kpool);
      const size_t nch = (pool.size() + 63) / 64;
      std::vector<float> cbox(nch * 2 * Ea);
      for (size_t c = 0; c < nch; c++)
        for (int j = 0; j < Ea; j++) {
          double l = INFINITY, h = -INFINITY;
          for (size_t q = 64 * c; q < std::min(pool.size(), 64 * c + 64); q++) {
            const double v = b->Ldata[(size_t)pool[q] * Ea + j];
            l = std::min(l, v);
            h = std::max(h, v);
          }
          float fl = (float)l, fh = (float)h;
          if ((double)fl > l)
            fl = std::nextafter(fl, -INFINITY);
          if ((double)fh < h)
            fh = std::nextafter(fh, INFINITY);
          cbox[c * 2 * Ea + j] = fl;
          cbox[c * 2 * Ea + Ea + j] = fh;
        }
      R.cbox = b->dev(cbox.data(), cbox.size());
    }
    R.pool = b->dev(pool.data(), pool.size());
    {
      // the pool rows' FP32 coordinates in pool order, padded to the kernel's width (16-byte rows)
      std::vector<float> p32(pool.size() * (size_t)emax, 0.0f);
      for (size_t q = 0; q < pool.size(); q++)
        for (int j = 0; j < Ea; j++)
          p32[q * emax + j] = (float)b->Ldata[(size_t)pool[q] * Ea + j];
      R.pool32 = b->dev(p32.data(), p32.size());
      hipStreamSynchronize(c->stream);
    }
    R.npool = (int)pool.size();
    R.pred = B.pred_rows + rank_pb;
    R.np = rank_pc;
    R.Ea = Ea;
    R.c1 = B.c1;
    R.c2 = B.c2;
    R.c3 = B.c3;
    // one pass per target over every rank task needs a hit list of KCAP prefix positions per
    // (target, task); beyond 4 GiB of it the per-(task, target) scan is used
    std::vector<int> rank_task;
    for (int t = 0; t < num_tasks; t++)
      if (is_rank[t])
        rank_task.push_back(t);
    // per wavefront: task words and KCAP hit positions per task (LDS); the task words of a target are
    // split over `parts` wavefronts so that one holds at most 12 KiB.  The predictions are staged
    // target-major ([T x targets x rank tasks] doubles, at most 4 GiB).
    const size_t words = (rank_task.size() + 31) / 32;
    const size_t per_word = sizeof(unsigned) * (1 + 16 * (size_t)kcap);
    const size_t stage_bytes = sizeof(double) * (size_t)B.T * rank_pc * rank_task.size();
    int parts = (int)std::max<size_t>(1, (words * per_word + 12 * 1024 - 1) / (12 * 1024));
    parts = (int)std::min<size_t>(parts, words);
    const size_t wp = (words + parts - 1) / parts;
    const size_t wave_lds = wp * per_word;
    b->multi_scan = rank_task.size() <= 8192 && wave_lds * 4 <= 128 * 1024 && stage_bytes <= ((size_t)4 << 30);
    R.tstride = b->multi_scan ? RANK_STORE : 1;
    R.jstride = b->multi_scan ? 1 : rank_pc;
    R.rank_row = b->dev<int>(nullptr, (size_t)RANK_STORE * rank_pc);
    R.rank_s = b->dev<double>(nullptr, (size_t)RANK_STORE * rank_pc);
    R.rank_len = b->dev<int>(nullptr, (size_t)rank_pc);
    R.zrow = b->dev<int>(nullptr, (size_t)RANK_ZCAP * rank_pc);
    R.zlen = b->dev<int>(nullptr, (size_t)rank_pc);
    ScanParams& S = b->S;
    std::memset(&S, 0, sizeof(S));
    S.L = B.L;
    S.Ltgt = B.Ltgt;
    S.P = B.P;
    S.pred = R.pred;
    S.np = rank_pc;
    S.Ea = Ea;
    S.T = B.T;
    S.force_compute = B.force_compute;
    S.mask_words = (L->num_points + 31) / 32;
    // library membership bitmasks of the rank-route tasks, indexed by task
    std::vector<unsigned> masks((size_t)num_tasks * S.mask_words, 0u);
    for (int t = 0; t < num_tasks; t++)
      if (is_rank[t])
        for (int i = 0; i < lib_count[t]; i++) {
          const int r = lib_rows[lib_begin[t] + i];
          masks[(size_t)t * S.mask_words + (r >> 5)] |= 1u << (r & 31);
        }
    S.masks = b->dev(masks.data(), masks.size());
    S.lib_begin = B.lib_begin;
    S.lib_count = B.lib_count;
    S.lib_rows = B.lib_rows;
    S.k = B.k;
    S.out_begin = B.out_begin;
    S.items = b->dev(ritems.data(), ritems.size());
    S.num_items = (int)ritems.size();
    S.counter = b->counters + 2 * PR_CSTRIDE;
    S.rank_row = R.rank_row;
    S.rank_s = R.rank_s;
    S.rank_len = R.rank_len;
    S.zrow = R.zrow;
    S.zlen = R.zlen;
    S.tstride = R.tstride;
    S.jstride = R.jstride;
    S.thetas = B.thetas;
    S.pred_out = B.pred;
    S.rc = B.rc;
    S.k_used = B.k_used;
    S.redo = B.redo;
    S.redo_count = B.redo_count;
    S.redo_cap = B.redo_cap;
    b->scan_lds = sizeof(unsigned) * (size_t)S.mask_words;
    if (b->multi_scan) {
      // membership of table rows in the rank tasks' libraries, one bit per rank task
      S.num_rank = (int)rank_task.size();
      S.words = (S.num_rank + 31) / 32;
      S.parts = parts;
      S.stage = b->dev<double>(nullptr, stage_bytes / sizeof(double));
      std::vector<unsigned> member((size_t)L->num_points * S.words, 0u);
      for (int r = 0; r < S.num_rank; r++) {
        const int t = rank_task[r];
        for (int i = 0; i < lib_count[t]; i++)
          member[(size_t)lib_rows[lib_begin[t] + i] * S.words + (r >> 5)] |= 1u << (r & 31);
      }
      S.member = b->dev(member.data(), member.size());
      S.rank_task = b->dev(rank_task.data(), rank_task.size());
      b->multi_lds = wave_lds * MS_WAVES;
      if (b->multi_lds > 48 * 1024)
        hipFuncSetAttribute(rank_scan_multi_kernel(kcap, S.words), hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)b->multi_lds);
      hipStreamSynchronize(c->stream);
    }
    hipStreamSynchronize(c->stream);
  }
  for (void* p : b->owned)
    if (!p) {
      delete b;
      return -2;
    }
  int dev = 0, ncu = 0, per = 0;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, batch_kernel(kcap, Ea, b->kn.direct_alt_waves != 0), 256, 0);
  if (b->kn.direct_blocks_per_cu > 0)
    per = std::min(std::max(per, 1), b->kn.direct_blocks_per_cu);
  b->grid = std::max(1, std::min(B.num_items, ncu * std::max(per, 1)));
  if (B.num_pitems > 0) {
    int pper = 0;
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&pper, batch_pruned_kernel(kcap, direct_emax(Ea)), 256, 0);
    b->pgrid = std::max(1, std::min(B.num_pitems, ncu * std::max(pper, 1)));
  }
  if (b->rank_route) {
    int sper = 0;
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&sper, rank_scan_kernel(kcap, emax), 256, b->scan_lds);
    b->scan_grid = std::max(1, std::min(b->S.num_items, ncu * std::max(sper, 1)));
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    delete b;
    return -3;
  }
  *out = b;
  return 0;
}

void edm_batch_set_keep_going(edm_batch* b, int (*keep_going)(void))
{
  b->keep_going = keep_going;
}

// A stop requested through keep_going: every output takes the interrupted values (prediction missing,
// BREAK_HIT, k_used missing) and the next run fills the outputs afresh before its stages.
static int batch_break(edm_batch* b)
{
  edm_ctx* c = b->ctx;
  if (!b->device_route && !b->resident) {
    std::fill(b->hpred.begin(), b->hpred.end(), FEDM_MISSING_D);
    std::fill(b->hrc.begin(), b->hrc.end(), (int)FEDM_BREAK_HIT);
    std::fill(b->hku.begin(), b->hku.end(), (int)FEDM_MISSING_I);
    return FEDM_BREAK_HIT;
  }
  dfill(c, b->B.pred, (size_t)b->total_out, FEDM_MISSING_D);
  dfill(c, b->B.rc, (size_t)b->total_out, (int)FEDM_BREAK_HIT);
  dfill(c, b->B.k_used, (size_t)b->total_k, (int)FEDM_MISSING_I);
  b->outputs_filled = false;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return FEDM_BREAK_HIT;
}

// The persistent route's stages on the context stream: work-queue counters, the outputs' fill values
// (first run only), the rank-table build and scan, the direct kernel and the exact re-selection.
// `timed` records the stage events and polls keep_going between stages (not under graph capture).
static int enqueue_device(edm_batch* b, bool timed)
{
  edm_ctx* c = b->ctx;
  BatchParams& B = b->B;
  hipStream_t s = c->stream;
  // the direct kernels (short libraries, VALU-bound) on a second stream beside the rank-table kernels
  // (long libraries, HBM-bound): the two task sets write disjoint outputs and meet at the re-selection
  const bool split = b->kn.concurrent_routes != 0 && b->rank_route && (B.num_items > 0 || B.num_pitems > 0);
  hipStream_t sd = split ? c->stream2 : s;
  b->split_last = split;
  auto stop = [b, timed] { return timed && b->keep_going && !b->keep_going(); };
  auto brk = [&] {
    if (split)
      hipStreamWaitEvent(s, c->join, 0);
    return batch_break(b);
  };
  dfill(c, b->counters, (size_t)CTR_INTS, 0);
  if (!b->outputs_filled) {
    // Every run writes every output of every target (a prediction, or the fill values themselves for a
    // target left unpredicted), so the outputs are filled on the first run only, as a safety net: a
    // refill would cost each run about 16 bytes of HBM writes per prediction.
    dfill(c, B.pred, (size_t)b->total_out, FEDM_MISSING_D);
    dfill(c, B.rc, (size_t)b->total_out, (int)FEDM_UNKNOWN_ERROR);
    dfill(c, B.k_used, (size_t)b->total_k, (int)FEDM_MISSING_I);
    b->outputs_filled = true;
  }
  if (timed)
    HIP_TRY(hipEventRecord(c->ev[4], s));
  const void* fn = batch_kernel(b->kcap, b->Ea, b->kn.direct_alt_waves != 0);
  void* args[] = { &B };
  auto direct = [&]() -> int {
    if (B.num_items > 0)
      HIP_TRY(hipLaunchKernel(fn, dim3(b->grid), dim3(256), args, 0, sd));
    if (timed)
      HIP_TRY(hipEventRecord(c->ev[8], sd));
    if (B.num_pitems > 0)
      HIP_TRY(hipLaunchKernel(batch_pruned_kernel(b->kcap, direct_emax(b->Ea)), dim3(b->pgrid), dim3(256), args, 0,
                              sd));
    if (timed)
      HIP_TRY(hipEventRecord(c->ev[9], sd));
    return 0;
  };
  if (split) {
    HIP_TRY(hipEventRecord(c->fork, s));
    HIP_TRY(hipStreamWaitEvent(sd, c->fork, 0));
    if (timed)
      HIP_TRY(hipEventRecord(c->ev[10], sd));
    if (const int r = direct())
      return r;
    HIP_TRY(hipEventRecord(c->join, sd));
  }
  if (b->rank_route) {
    void* rargs[] = { &b->R };
    HIP_TRY(hipLaunchKernel(rank_build_kernel(b->emax), dim3(b->R.np), dim3(256), rargs, 0, s));
    if (timed)
      HIP_TRY(hipEventRecord(c->ev[6], s));
    if (stop())
      return brk();
    void* sargs[] = { &b->S };
    if (b->multi_scan) {
      HIP_TRY(hipLaunchKernel(rank_scan_multi_kernel(b->kcap, b->S.words),
                              dim3(((size_t)b->S.np * b->S.parts + MS_WAVES - 1) / MS_WAVES), dim3(64 * MS_WAVES), sargs,
                              b->multi_lds, s));
      HIP_TRY(hipLaunchKernel((const void*)rank_scatter, dim3((b->S.num_rank + 63) / 64, (b->S.np + 63) / 64, b->S.T),
                              dim3(256), sargs, 0, s));
    } else if (b->S.num_items > 0)
      HIP_TRY(hipLaunchKernel(rank_scan_kernel(b->kcap, b->emax), dim3(b->scan_grid), dim3(256), sargs,
                              b->scan_lds, s));
  } else if (timed) {
    HIP_TRY(hipEventRecord(c->ev[6], s));
  }
  if (timed)
    HIP_TRY(hipEventRecord(c->ev[7], s));
  if (stop())
    return brk();
  if (split) {
    HIP_TRY(hipStreamWaitEvent(s, c->join, 0));
    if (timed)
      HIP_TRY(hipEventRecord(c->ev[11], s));
  } else if (const int r = direct()) {
    return r;
  }
  {
    // targets handed over by the direct kernel or the rank scan: exact selection; with more hand-overs
    // than the list holds, every target again (one launch either way, so the host does not wait for the
    // count mid-step)
    RedoLists L{ { B.items, b->rank_route ? b->S.items : nullptr, B.pitems },
                 { B.num_items, b->rank_route ? b->S.num_items : 0, B.num_pitems } };
    void* rargs[] = { &B, &L };
    HIP_TRY(hipLaunchKernel(batch_redo_kernel(b->kcap), dim3(std::max(b->grid, 256)), dim3(256), rargs, 0, s));
  }
  return 0;
}

int edm_batch_run(edm_batch* b)
{
  edm_ctx* c = b->ctx;
  hipSetDevice(c->device);
  auto stop = [b] { return b->keep_going && !b->keep_going(); };
  if (b->resident)
    return run_resident(b);
  if (!b->device_route) {
    // task-by-task: gather each task's manifolds and predict them with edm_predict (uploads per task)
    b->hpred.assign((size_t)b->total_out, FEDM_MISSING_D);
    b->hrc.assign((size_t)b->total_out, FEDM_UNKNOWN_ERROR);
    b->hku.assign((size_t)b->total_k, FEDM_MISSING_I);
    const bool want_coeffs = b->o.save_smap_coeffs && b->o.algorithm == FEDM_SMAP;
    const size_t C = (size_t)b->Ea + 1;
    b->hcoeffs.assign(want_coeffs ? (size_t)b->total_k * C : 0, FEDM_MISSING_D);
    GatheredTask g;
    for (int tk = 0; tk < b->tasks; tk++) {
      if (stop())
        return batch_break(b);
      gather_task(b, tk, g);
      const edm_options to = task_options(b, tk);
      const int64_t ob = b->ob[tk];
      const int r = edm_predict(c, &g.M, &g.Mp, &to, b->hpred.data() + ob, b->hrc.data() + ob,
                                want_coeffs ? b->hcoeffs.data() + (size_t)(ob / b->T) * C : nullptr,
                                b->hku.data() + ob / b->T, nullptr, 0, b->keep_going);
      if (r == FEDM_BREAK_HIT)
        return batch_break(b);
      if (r != 0)
        return r;
    }
    return 0;
  }
  if (stop())
    return batch_break(b);
  hipStream_t s = c->stream;
  const bool graphed = b->use_graph && b->outputs_filled;
  if (graphed) {
    // the steady-state run (outputs filled once already) captured once and replayed
    if (!b->graph_exec) {
      HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
      const int r = enqueue_device(b, false);
      hipGraph_t gr = nullptr;
      const hipError_t e = hipStreamEndCapture(s, &gr);
      if (r != 0 || e != hipSuccess) {
        if (gr)
          hipGraphDestroy(gr);
        return r != 0 ? r : -(int)e - 1000;
      }
      b->graph = gr;
      HIP_TRY(hipGraphInstantiate(&b->graph_exec, gr, nullptr, nullptr, 0));
    }
    HIP_TRY(hipEventRecord(c->ev[4], s));
    HIP_TRY(hipGraphLaunch(b->graph_exec, s));
  } else {
    const int r = enqueue_device(b, true);
    if (r != 0)
      return r;
  }
  // the hand-over count comes back with the run (pinned memory, so the one wait below covers it)
  if (b->handed_host)
    HIP_TRY(hipMemcpyAsync(b->handed_host, b->B.redo_count, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipEventRecord(c->ev[5], s));
  HIP_TRY(hipEventSynchronize(c->ev[5]));
  {
    int handed = 0;
    if (b->handed_host)
      handed = *b->handed_host;
    else
      HIP_TRY(hipMemcpy(&handed, b->B.redo_count, sizeof(int), hipMemcpyDeviceToHost));
    b->redo_last = handed;
    c->last_redo = handed;
  }
  b->runs++;
  b->last_ms = ev_ms(c->ev[4], c->ev[5]);
  // [0] whole batch, [1] rank-table build, [2] rank scan, [3] direct kernels, [4] the unpruned direct
  // kernel, [5] the pruned one, [6] the exact re-selection (stages: eager runs only); with the routes
  // concurrent, [1]-[2] and [3]-[5] overlap in time and each kernel shares the device with the others
  const bool sp = b->split_last;
  c->timings[0] = b->last_ms;
  c->timings[1] = graphed ? 0.0 : ev_ms(c->ev[4], c->ev[6]);
  c->timings[2] = graphed ? 0.0 : ev_ms(c->ev[6], c->ev[7]);
  c->timings[3] = graphed ? 0.0 : ev_ms(c->ev[sp ? 10 : 7], c->ev[9]);
  c->timings[4] = graphed ? 0.0 : ev_ms(c->ev[sp ? 10 : 7], c->ev[8]);
  c->timings[5] = graphed ? 0.0 : ev_ms(c->ev[8], c->ev[9]);
  c->timings[6] = graphed ? 0.0 : ev_ms(c->ev[sp ? 11 : 9], c->ev[5]);
  c->timings[7] = graphed ? 0.0 : (sp ? 1.0 : 0.0);
  return 0;
}

int edm_batch_fetch(edm_batch* b, double* pred, int* rc, int* k_used)
{
  edm_ctx* c = b->ctx;
  hipSetDevice(c->device);
  if (!b->device_route && !b->resident) {
    std::copy(b->hpred.begin(), b->hpred.end(), pred);
    std::copy(b->hrc.begin(), b->hrc.end(), rc);
    std::copy(b->hku.begin(), b->hku.end(), k_used);
    return 0;
  }
  HIP_TRY(hipMemcpyAsync(pred, b->B.pred, sizeof(double) * b->total_out, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(rc, b->B.rc, sizeof(int) * b->total_out, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(k_used, b->B.k_used, sizeof(int) * b->total_k, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int edm_batch_on_device(const edm_batch* b)
{
  return b->device_route ? 1 : 0;
}

int edm_batch_route(const edm_batch* b)
{
  return b->device_route ? 2 : b->resident ? 1 : 0;
}

int edm_batch_set_graph(edm_batch* b, int on)
{
  b->use_graph = on != 0 && (b->device_route || b->resident);
  if (!b->use_graph && b->graph_exec) {
    hipGraphExecDestroy(b->graph_exec);
    hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
  }
  return b->use_graph ? 1 : 0;
}

int edm_batch_fetch_coeffs(edm_batch* b, double* coeffs)
{
  edm_ctx* c = b->ctx;
  const size_t n = (size_t)b->total_k * (b->Ea + 1);
  if (!b->coeffs) {
    if (b->hcoeffs.size() == n)
      std::copy(b->hcoeffs.begin(), b->hcoeffs.end(), coeffs);
    else
      std::fill(coeffs, coeffs + n, FEDM_MISSING_D);
    return 0;
  }
  hipSetDevice(c->device);
  HIP_TRY(hipMemcpyAsync(coeffs, b->coeffs, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

void edm_batch_destroy(edm_batch* b)
{
  delete b;
}

int edm_predict_batch(edm_ctx* c, const edm_manifold_view* L, const edm_manifold_view* P, const edm_options* o,
                      int num_tasks, const int* lib_begin, const int* lib_count, const int* lib_rows,
                      const int* pred_begin, const int* pred_count, const int* pred_rows, const int* k,
                      const int64_t* out_begin, double* pred, int* rc, int* k_used, int (*keep_going)(void))
{
  if (keep_going && !keep_going())
    return FEDM_BREAK_HIT;
  edm_batch* b = nullptr;
  int r = edm_batch_create(c, L, P, o, num_tasks, lib_begin, lib_count, lib_rows, pred_begin, pred_count, pred_rows,
                           k, out_begin, &b);
  if (r != 0)
    return r;
  edm_batch_set_keep_going(b, keep_going);
  r = edm_batch_run(b);
  if (r == 0 || r == FEDM_BREAK_HIT) {
    const int q = edm_batch_fetch(b, pred, rc, k_used);
    r = q != 0 ? q : r;
  }
  edm_batch_destroy(b);
  return r;
}

} // extern "C"
