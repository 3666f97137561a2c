// C ABI of the fastEDM prediction core: device context, edm_predict and edm_predict_batch.
// The dispatcher picks the kernel route per task (register top-k for k <= 64, dense distances for
// k = all), stages inputs into a grow-only device pool and copies the outputs back on one stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <cstring>
#include <vector>

#include "../../include/fastedm_hip.h"
#include "batch_common.h"

namespace fedm {

template <int KCAP, int EMAX, bool GENERIC>
__global__ void knn_slices_f64(LpParams P, int k, int slice_len, double* out_s, int* out_i, int* out_cnt);
template <int EMAX, bool GENERIC>
__global__ void dense_dist_f64(LpParams P, int t0, int nt, int slice_len, double* D, int* out_cnt);
template <int KCAP>
__global__ void finish_lists(FinishParams F, const double* ps, const int* pi, const int* pcnt, int S);
__global__ void finish_dense(FinishParams F, const double* D, int t0, int nt, const int* pcnt, int S);
__global__ void wasserstein_dense(WassParams P, int t0, int nt, int slice_len, double* D, int* out_cnt);
__global__ void select_wide(FinishParams F, const double* D, int t0, int nt, unsigned long long* gkey, int* gidx,
                            int gstride);
int smap_launch(int n, hipStream_t s, const SmapParams& S, const double* D, int t0, int nt, bool lds_rows);

template <int KCAP, int KP>
__global__ void knn_mfma(MfmaParams P, int k, int slice_len, double* out_s, int* out_i, int* out_cnt);
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
  hipEvent_t ev[8] = {};
  std::vector<DevBuf> bufs;
  double timings[4] = { 0, 0, 0, 0 };
  int last_error = 0;
  int last_redo = 0;

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
                          int* pc)
{
  if (kp == 16)
    hipLaunchKernelGGL((knn_mfma<KCAP, 16>), g, dim3(256), 0, s, P, k, sl, ps, pi, pc);
  else if (kp == 32)
    hipLaunchKernelGGL((knn_mfma<KCAP, 32>), g, dim3(256), 0, s, P, k, sl, ps, pi, pc);
  else
    hipLaunchKernelGGL((knn_mfma<KCAP, 64>), g, dim3(256), 0, s, P, k, sl, ps, pi, pc);
}

static void launch_mfma(int kcap, int kp, dim3 g, hipStream_t s, const MfmaParams& P, int k, int sl, double* ps,
                        int* pi, int* pc)
{
  switch (kcap) {
    case 4: launch_mfma_k<4>(kp, g, s, P, k, sl, ps, pi, pc); break;
    case 8: launch_mfma_k<8>(kp, g, s, P, k, sl, ps, pi, pc); break;
    case 16: launch_mfma_k<16>(kp, g, s, P, k, sl, ps, pi, pc); break;
    case 32: launch_mfma_k<32>(kp, g, s, P, k, sl, ps, pi, pc); break;
    default: launch_mfma_k<64>(kp, g, s, P, k, sl, ps, pi, pc); break;
  }
}

static void launch_dense(int emax, bool generic, dim3 g, hipStream_t s, const LpParams& P, int t0, int nt, int sl,
                         double* D, int* pc)
{
#define FEDM_DENSE(EM)                                                                                           \
  if (generic)                                                                                                   \
    hipLaunchKernelGGL((dense_dist_f64<EM, true>), g, dim3(256), 0, s, P, t0, nt, sl, D, pc);                    \
  else                                                                                                           \
    hipLaunchKernelGGL((dense_dist_f64<EM, false>), g, dim3(256), 0, s, P, t0, nt, sl, D, pc);
  switch (emax) {
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

// S-map systems of targets [t0, t0 + nt) at every theta: a persistent grid of one-wavefront blocks.
// FEDM_SMAP_LDS_ROWS=1 keeps U and V in LDS for every n (the kernel used above 24 unknowns).
static int launch_smap(edm_ctx* c, int n, hipStream_t s, const SmapParams& S, const double* D, int t0, int nt)
{
  (void)c;
  const char* env = getenv("FEDM_SMAP_LDS_ROWS");
  return smap_launch(n, s, S, D, t0, nt, env && env[0] == '1');
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
  for (auto& e : c->ev)
    hipEventCreate(&e);
  c->bufs.resize(B_NUM);
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
  for (auto& e : c->ev)
    hipEventDestroy(e);
  hipStreamDestroy(c->stream);
  delete c;
}

void edm_ctx_timings(edm_ctx* c, double* ms4)
{
  for (int i = 0; i < 4; i++)
    ms4[i] = c->timings[i];
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

int edm_predict(edm_ctx* c, const edm_manifold_view* M, const edm_manifold_view* Mp, const edm_options* o,
                double* pred, int* rc, double* coeffs, int* k_used, int* nn_idx, int nn_cap, int (*keep_going)(void))
{
  const int NL = M->num_points, NP = Mp->num_points, Ea = M->E_actual, T = o->num_thetas, n = Ea + 1;
  if (!c || Ea != Mp->E_actual || T <= 0 || Ea <= 0)
    return -22;
  if (NP == 0)
    return 0;
  if (keep_going && !keep_going()) {
    host_fill_outputs(Mp, o, pred, rc, coeffs, k_used, nn_idx, nn_cap, Ea, FEDM_UNKNOWN_ERROR);
    for (int i = 0; i < NP; i++)
      rc[i] = FEDM_BREAK_HIT;
    return FEDM_BREAK_HIT;
  }
  const int emax = emax_for(Ea);
  const int k = o->k;
  const bool wass = o->distance == FEDM_DIST_WASSERSTEIN;
  const bool list_route = !wass && k > 0 && k <= 64;
  // Wasserstein point layout [x lags | dt lags | lagged extras | unlagged extras] (src/distances.cpp:318-326)
  const int wE = M->E_x;
  const int wrows = wE > 0 ? 1 + (M->E_dt > 0) + M->E_lagged_extras / wE : 0;
  const int wunl = M->E_extras - M->E_lagged_extras;
  bool wass_ok = wE >= 1 && wE <= 32 && (M->E_dt == 0 || M->E_dt == wE) && M->E_lagged_extras % wE == 0 &&
                 wunl >= 0 && wrows * wE + wunl == Ea && o->missing_distance >= 0.0;
  if (wass && wass_ok && o->panel_mode && o->idw <= 0 && o->id_weights)
    for (int a = 0; a < o->num_panels * o->num_panels; a++)
      wass_ok = wass_ok && o->id_weights[a] >= 0.0f;
  if (emax < 0 || (wass && !wass_ok)) {
    host_fill_outputs(Mp, o, pred, rc, coeffs, k_used, nn_idx, nn_cap, Ea, FEDM_UNKNOWN_ERROR);
    return -95; // not supported on this route
  }
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  int err = 0;

  // metrics: one per coordinate for L^p, one per variable for Wasserstein (src/distances.cpp:29-67)
  const int nmet = wass ? wrows + wunl : Ea;
  std::vector<int> metrics(nmet, 0);
  if (o->metrics)
    std::copy(o->metrics, o->metrics + nmet, metrics.begin());

  LpParams P;
  std::memset(&P, 0, sizeof(P));
  P.M = up(c, B_M, M->data, (size_t)NL * Ea, &err);
  P.Mp = up(c, B_MP, Mp->data, (size_t)NP * Ea, &err);
  const double* dtargets = up(c, B_MT, M->targets, (size_t)NL, &err);
  P.Mtt = up(c, B_MTT, M->target_times, (size_t)NL, &err);
  P.Mppt = up(c, B_MPPT, Mp->point_times, (size_t)NP, &err);
  if (o->panel_mode) {
    P.Mpanel = up(c, B_MPAN, M->panel, (size_t)NL, &err);
    P.Mppanel = up(c, B_MPPAN, Mp->panel, (size_t)NP, &err);
  }
  P.metrics = up(c, B_METRICS, metrics.data(), (size_t)nmet, &err);
  if (o->id_weights && o->num_panels > 0)
    P.id_weights = up(c, B_IDW, o->id_weights, (size_t)o->num_panels * o->num_panels, &err);
  const double* dthetas = up(c, B_THETAS, o->thetas, (size_t)T, &err);
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
  else if (o->panel_mode && P.id_weights)
    f |= F_IDWEIGHTS;
  P.flags = f;
  const bool generic = f != 0;

  WassParams W;
  std::memset(&W, 0, sizeof(W));
  if (wass) {
    W.M = P.M;
    W.Mp = P.Mp;
    W.Mpanel = P.Mpanel;
    W.Mppanel = P.Mppanel;
    W.metrics = P.metrics;
    W.id_weights = P.id_weights;
    W.NL = NL;
    W.NP = NP;
    W.Ea = Ea;
    W.E = wE;
    W.E_dt = M->E_dt;
    W.rows = wrows;
    W.n_unlagged = wunl;
    W.num_panels = o->num_panels;
    W.missing_distance = o->missing_distance;
    W.idw = o->idw;
    W.aspect_ratio = o->aspect_ratio;
    W.panel_term = !o->panel_mode ? 0 : o->idw > 0 ? 1 : P.id_weights ? 2 : 0;
    W.skip_other = o->panel_mode && o->idw < 0;
  }

  double* dpred = (double*)c->get(B_PRED, sizeof(double) * (size_t)T * NP);
  int* drc = (int*)c->get(B_RC, sizeof(int) * (size_t)T * NP);
  int* dku = (int*)c->get(B_KUSED, sizeof(int) * (size_t)NP);
  double* dco = coeffs ? (double*)c->get(B_COEFFS, sizeof(double) * (size_t)NP * n) : nullptr;
  int* dnno = (nn_idx && nn_cap > 0) ? (int*)c->get(B_NNOUT, sizeof(int) * (size_t)NP * nn_cap) : nullptr;
  int* nnk = (int*)c->get(B_NNK, sizeof(int) * (size_t)NP);
  int* nncnt = (int*)c->get(B_NNCNT, sizeof(int) * (size_t)NP);
  if (err || !dpred || !drc || !dku || !nnk || !nncnt || (coeffs && !dco) || (nn_idx && nn_cap > 0 && !dnno))
    return err ? err : -2;
  dfill(c, dpred, (size_t)T * NP, FEDM_MISSING_D);
  dfill(c, drc, (size_t)T * NP, (int)FEDM_UNKNOWN_ERROR);
  dfill(c, dku, (size_t)NP, (int)FEDM_MISSING_I);
  dfill(c, dco, (size_t)NP * n, FEDM_MISSING_D);
  dfill(c, dnno, (size_t)NP * nn_cap, -1);
  dfill(c, nnk, (size_t)NP, 0);

  FinishParams F;
  std::memset(&F, 0, sizeof(F));
  F.NP = NP;
  F.NL = NL;
  F.k = k;
  F.T = T;
  F.force_compute = o->force_compute;
  F.algorithm = o->algorithm;
  F.thetas = dthetas;
  F.targets = dtargets;
  F.pred = dpred;
  F.rc = drc;
  F.k_used = dku;
  F.nn_k = nnk;
  F.nn_cnt = nncnt;
  F.nn_out = dnno;
  F.nn_cap = dnno ? nn_cap : 0;

  SmapParams Sm;
  std::memset(&Sm, 0, sizeof(Sm));
  Sm.NP = NP;
  Sm.NL = NL;
  Sm.Ea = Ea;
  Sm.T = T;
  Sm.M = P.M;
  Sm.Mp = P.Mp;
  Sm.targets = dtargets;
  Sm.thetas = dthetas;
  Sm.nn_k = nnk;
  Sm.nn_cnt = nncnt;
  Sm.pred = dpred;
  Sm.rc = drc;
  Sm.coeffs = (o->save_smap_coeffs && dco) ? dco : nullptr;

  const bool smap = o->algorithm == FEDM_SMAP;
  HIP_TRY(hipEventRecord(c->ev[0], s));

  // matrix-core screening pays once a point has a few dozen coordinates
  double amax = 0.0;
  const bool want_mfma = list_route && o->precision == FEDM_MIXED && !generic && Ea >= 8;
  if (want_mfma) {
    for (size_t i = 0; i < (size_t)NL * Ea; i++)
      amax = std::max(amax, std::fabs(M->data[i]));
    for (size_t i = 0; i < (size_t)NP * Ea; i++)
      amax = std::max(amax, std::fabs(Mp->data[i]));
  }
  const bool mfma = want_mfma && amax < 1e30;
  if (list_route && mfma) {
    const int kcap = kcap_for(k);
    const int kp = Ea <= 16 ? 16 : Ea <= 32 ? 32 : 64;
    const int bx = (NP + 63) / 64;
    int S = std::max(1, std::min(MAX_SLICES, (2048 + bx - 1) / bx));
    S = std::max(1, std::min(S, NL / 256));
    const int sl = (NL + S - 1) / S;
    // centred FP32 operands, their FP32 norms and the per-target screening bound
    std::vector<double> ctr(Ea, 0.0);
    for (int j = 0; j < NL; j++)
      for (int d = 0; d < Ea; d++)
        ctr[d] += M->data[(size_t)j * Ea + d];
    for (int d = 0; d < Ea; d++)
      ctr[d] /= std::max(NL, 1);
    std::vector<float> m32((size_t)NL * kp, 0.f), mn(NL), p32((size_t)NP * kp, 0.f), pn(NP);
    std::vector<double> ynorm(NP);
    double R = 0.0;
    for (int j = 0; j < NL; j++) {
      double r2 = 0.0;
      float n2 = 0.f;
      for (int d = 0; d < Ea; d++) {
        const double v = M->data[(size_t)j * Ea + d] - ctr[d];
        const float f = (float)v;
        m32[(size_t)j * kp + d] = f;
        n2 += f * f;
        r2 += v * v;
      }
      mn[j] = n2;
      R = std::max(R, std::sqrt(r2));
    }
    const double u = 5.9604644775390625e-08;
    for (int i = 0; i < NP; i++) {
      double r2 = 0.0;
      float n2 = 0.f;
      for (int d = 0; d < Ea; d++) {
        const double v = Mp->data[(size_t)i * Ea + d] - ctr[d];
        const float f = (float)v;
        p32[(size_t)i * kp + d] = f;
        n2 += f * f;
        r2 += v * v;
      }
      pn[i] = n2;
      ynorm[i] = std::sqrt(r2);
    }
    MfmaParams MP;
    MP.M32 = up(c, B_MIX0, m32.data(), m32.size(), &err);
    MP.Mn = up(c, B_MIX1, mn.data(), mn.size(), &err);
    MP.P32 = up(c, B_MIX2, p32.data(), p32.size(), &err);
    MP.Pn = up(c, B_MIX3, pn.data(), pn.size(), &err);
    // the screening bound per target (kernel): 2 (k1 rr sqrt(D) + c2 D + k3 rr^2), rr >= |x| + |y|
    MP.ynorm = up(c, B_MIX4, ynorm.data(), ynorm.size(), &err);
    MP.R = R;
    MP.k1 = 2.02 * u;
    MP.k3 = (kp + 3) * 1.02 * u + u * u;
    MP.c2 = 4.0 * u;
    MP.M = P.M;
    MP.Mp = P.Mp;
    MP.NL = NL;
    MP.NP = NP;
    MP.Ea = Ea;
    double* ps = (double*)c->get(B_PS, sizeof(double) * (size_t)S * NP * kcap);
    int* pi = (int*)c->get(B_PI, sizeof(int) * (size_t)S * NP * kcap);
    int* pc = (int*)c->get(B_PCNT, sizeof(int) * (size_t)S * NP);
    F.nn_i = (int*)c->get(B_NNI, sizeof(int) * (size_t)NP * kcap);
    F.nn_s = (double*)c->get(B_NNS, sizeof(double) * (size_t)NP * kcap);
    F.nn_stride = kcap;
    if (err || !ps || !pi || !pc || !F.nn_i || !F.nn_s)
      return err ? err : -2;
    HIP_TRY(hipEventRecord(c->ev[0], s));
    launch_mfma(kcap, kp, dim3(bx, S), s, MP, k, sl, ps, pi, pc);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(c->ev[1], s));
    launch_finish_lists(kcap, dim3((NP + 255) / 256), s, F, ps, pi, pc, S);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(c->ev[2], s));
    if (smap) {
      Sm.nn_i = F.nn_i;
      Sm.nn_s = F.nn_s;
      Sm.nn_stride = kcap;
      int r = launch_smap(c, n, s, Sm, nullptr, 0, NP);
      if (r)
        return r;
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(c->ev[3], s));
  } else if (list_route) {
    const int kcap = kcap_for(k);
    const int bx = (NP + 255) / 256;
    int S = std::max(1, std::min(MAX_SLICES, (1024 + bx - 1) / bx));
    S = std::max(1, std::min(S, NL / 256));
    const int sl = (NL + S - 1) / S;
    double* ps = (double*)c->get(B_PS, sizeof(double) * (size_t)S * NP * kcap);
    int* pi = (int*)c->get(B_PI, sizeof(int) * (size_t)S * NP * kcap);
    int* pc = (int*)c->get(B_PCNT, sizeof(int) * (size_t)S * NP);
    F.nn_i = (int*)c->get(B_NNI, sizeof(int) * (size_t)NP * kcap);
    F.nn_s = (double*)c->get(B_NNS, sizeof(double) * (size_t)NP * kcap);
    F.nn_stride = kcap;
    if (!ps || !pi || !pc || !F.nn_i || !F.nn_s)
      return -2;
    launch_knn(kcap, emax, generic, dim3(bx, S), s, P, k, sl, ps, pi, pc);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(c->ev[1], s));
    launch_finish_lists(kcap, dim3(bx), s, F, ps, pi, pc, S);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(c->ev[2], s));
    if (smap) {
      Sm.nn_i = F.nn_i;
      Sm.nn_s = F.nn_s;
      Sm.nn_stride = kcap;
      int r = launch_smap(c, n, s, Sm, nullptr, 0, NP);
      if (r)
        return r;
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(c->ev[3], s));
  } else {
    // dense distance rows: every valid neighbour (k < 0), or an exact selection wider than the
    // register lists (k > 64) made from the rows by select_wide
    const bool wide = wass ? k > 0 : k > 64;
    const int n2 = wide ? (int)std::min<long long>(1ll << 30, 1ll << (64 - __builtin_clzll((unsigned long long)k - 1))) : 0;
    const int gstride = wide && n2 > 4096 ? n2 : 0;
    const size_t per_target = sizeof(double) * (size_t)std::max(NL, 1) +
                              (wide ? (sizeof(int) + sizeof(double)) * (size_t)k : 0) +
                              (sizeof(unsigned long long) + sizeof(int)) * (size_t)gstride;
    const size_t budget = (size_t)1 << 30;
    int nt = (int)std::max<size_t>(1, std::min<size_t>((size_t)NP, budget / per_target));
    const int bxmax = wass ? nt : (nt + 255) / 256;
    int S = std::max(1, std::min(MAX_SLICES, (1024 + bxmax - 1) / bxmax));
    S = std::max(1, std::min(S, NL / (wass ? 64 : 256)));
    const int sl = (NL + S - 1) / S;
    double* D = (double*)c->get(B_D, sizeof(double) * (size_t)nt * std::max(NL, 1));
    int* pc = (int*)c->get(B_PCNT, sizeof(int) * (size_t)S * NP);
    F.nn_i = wide ? (int*)c->get(B_NNI, sizeof(int) * (size_t)nt * k) : nullptr;
    F.nn_s = wide ? (double*)c->get(B_NNS, sizeof(double) * (size_t)nt * k) : nullptr;
    F.nn_stride = wide ? k : 0;
    unsigned long long* gkey =
      gstride ? (unsigned long long*)c->get(B_WKEY, sizeof(unsigned long long) * (size_t)nt * gstride) : nullptr;
    int* gidx = gstride ? (int*)c->get(B_WIDX, sizeof(int) * (size_t)nt * gstride) : nullptr;
    if (!D || !pc || (wide && (!F.nn_i || !F.nn_s)) || (gstride && (!gkey || !gidx)))
      return -2;
    Sm.nn_i = F.nn_i;
    Sm.nn_s = F.nn_s;
    Sm.nn_stride = F.nn_stride;
    HIP_TRY(hipEventRecord(c->ev[1], s));
    for (int t0 = 0; t0 < NP; t0 += nt) {
      const int cnt = std::min(nt, NP - t0);
      if (wass)
        hipLaunchKernelGGL(wasserstein_dense, dim3(cnt, S), dim3(256), 0, s, W, t0, cnt, sl, D, pc);
      else
        launch_dense(emax, generic, dim3((cnt + 255) / 256, S), s, P, t0, cnt, sl, D, pc);
      HIP_TRY(hipGetLastError());
      hipLaunchKernelGGL(finish_dense, dim3((cnt + 255) / 256), dim3(256), 0, s, F, D, t0, cnt, pc, S);
      HIP_TRY(hipGetLastError());
      if (wide) {
        hipLaunchKernelGGL(select_wide, dim3(cnt), dim3(256), 0, s, F, D, t0, cnt, gkey, gidx, gstride);
        HIP_TRY(hipGetLastError());
      }
      Sm.nn_t0 = t0;
      if (smap) {
        int r = launch_smap(c, n, s, Sm, D, t0, cnt);
        if (r)
          return r;
        HIP_TRY(hipGetLastError());
      }
      if (keep_going && !keep_going())
        break;
    }
    HIP_TRY(hipEventRecord(c->ev[2], s));
    HIP_TRY(hipEventRecord(c->ev[3], s));
  }
  HIP_TRY(hipMemcpyAsync(pred, dpred, sizeof(double) * (size_t)T * NP, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(rc, drc, sizeof(int) * (size_t)T * NP, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(k_used, dku, sizeof(int) * (size_t)NP, hipMemcpyDeviceToHost, s));
  if (coeffs)
    HIP_TRY(hipMemcpyAsync(coeffs, dco, sizeof(double) * (size_t)NP * n, hipMemcpyDeviceToHost, s));
  if (dnno)
    HIP_TRY(hipMemcpyAsync(nn_idx, dnno, sizeof(int) * (size_t)NP * nn_cap, hipMemcpyDeviceToHost, s));
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

} // extern "C"

// ------------------------------------------------------------------------------------------------
// Batches of tasks over shared point tables.

namespace fedm {
template <int KCAP, int EMAX>
__global__ void batch_simplex(BatchParams B);
template <int KCAP>
__global__ void batch_redo(BatchParams B, const int2* items, int num_items);

template <int EMAX>
__global__ void rank_build(RankParams R);
template <int KCAP, int WPL>
__global__ void rank_scan_multi(ScanParams S);
__global__ void rank_scatter(ScanParams S);
template <int KCAP, int EMAX>
__global__ void rank_scan(ScanParams S);
} // namespace fedm

struct edm_batch
{
  edm_ctx* ctx = nullptr;
  bool device_route = false;
  int kcap = 0, emax = 0, mixed = 0, T = 0, Ea = 0, tasks = 0;
  int64_t total_out = 0, total_k = 0;
  std::vector<void*> owned;
  BatchParams B;
  int grid = 0;
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
  bool outputs_filled = false; // the outputs have held their fill values once (first run)
  // host copies for the task-by-task route
  edm_manifold_view Lv, Pv;
  edm_options o;
  std::vector<double> Ldata, Ltgt, Ltt, Pdata, Ppt, thetas;
  std::vector<int> Lpan, Ppan, lb, lc, lr, pb, pc, pr, kk, metrics;
  std::vector<int64_t> ob;
  std::vector<double> hpred;
  std::vector<int> hrc, hku;

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
    for (void* p : owned)
      hipFree(p);
  }
};

template <int KCAP>
static const void* batch_kernel_k(int demax)
{
  switch (demax) {
    case 2: return (const void*)batch_simplex<KCAP, 2>;
    case 3: return (const void*)batch_simplex<KCAP, 3>;
    case 4: return (const void*)batch_simplex<KCAP, 4>;
    case 8: return (const void*)batch_simplex<KCAP, 8>;
    case 16: return (const void*)batch_simplex<KCAP, 16>;
    default: return (const void*)batch_simplex<KCAP, 32>;
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
static const void* batch_kernel(int kcap, int Ea)
{
  switch (kcap) {
    case 4: return batch_kernel_k<4>(direct_emax(Ea));
    case 8: return batch_kernel_k<8>(direct_emax(Ea));
    case 16: return batch_kernel_k<16>(direct_emax(Ea));
    default: return batch_kernel_k<32>(direct_emax(Ea));
  }
}

// Smallest library that takes the rank route: scanning a ranked prefix costs about k * pool / L
// entries per target, the direct scan L pairs; FEDM_RANK_LTHR_SCALE overrides the scale (tuning).
static int rank_threshold(int kmax, int npool)
{
  double scale = 1.5;
  if (const char* e = getenv("FEDM_RANK_LTHR_SCALE"))
    scale = atof(e);
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

extern "C" int fedm_rank_threshold(int kmax, int npool)
{
  return rank_threshold(kmax, npool);
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

extern "C" {

int edm_batch_create(edm_ctx* c, const edm_manifold_view* L, const edm_manifold_view* P, const edm_options* o,
                     int num_tasks, const int* lib_begin, const int* lib_count, const int* lib_rows,
                     const int* pred_begin, const int* pred_count, const int* pred_rows, const int* k,
                     const int64_t* out_begin, edm_batch** out)
{
  *out = nullptr;
  if (!c || num_tasks < 0 || L->E_actual != P->E_actual || o->num_thetas <= 0)
    return -22;
  hipSetDevice(c->device);
  auto b = new edm_batch();
  b->ctx = c;
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
    const int lthr = rank_threshold(kmax, (int)pool.size());
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
  std::vector<int2> items, ritems;
  for (int t : order)
    for (int t0 = 0; t0 < pred_count[t]; t0 += 256)
      (is_rank[t] ? ritems : items).push_back(make_int2(t, t0));
  B.counter = b->dev<int>(nullptr, 1);
  // targets whose squared-distance ranking may differ from the sqrt ranking (rare; see kernels_batch.hip)
  B.redo_cap = (int)std::min<int64_t>(b->total_k, 1 << 16);
  B.redo = b->dev<int2>(nullptr, (size_t)std::max(B.redo_cap, 1));
  B.redo_count = b->dev<int>(nullptr, 1);
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
    S.counter = b->dev<int>(nullptr, 1);
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
  hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, batch_kernel(kcap, Ea), 256, 0);
  b->grid = std::max(1, std::min(B.num_items, ncu * std::max(per, 1)));
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
  if (!b->device_route) {
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

int edm_batch_run(edm_batch* b)
{
  edm_ctx* c = b->ctx;
  hipSetDevice(c->device);
  auto stop = [b] { return b->keep_going && !b->keep_going(); };
  if (!b->device_route) {
    // task-by-task: gather each task's manifolds and predict them with edm_predict
    const int Ea = b->Ea;
    b->hpred.assign((size_t)b->total_out, FEDM_MISSING_D);
    b->hrc.assign((size_t)b->total_out, FEDM_UNKNOWN_ERROR);
    b->hku.assign((size_t)b->total_k, FEDM_MISSING_I);
    for (int tk = 0; tk < b->tasks; tk++) {
      if (stop())
        return batch_break(b);
      const int nl = b->lc[tk], np = b->pc[tk];
      std::vector<double> md((size_t)nl * Ea), mt(nl), mtt(nl), mpd((size_t)np * Ea), mppt(np);
      std::vector<int> mpan, mppan;
      for (int r = 0; r < nl; r++) {
        const int row = b->lr[b->lb[tk] + r];
        std::copy(b->Ldata.begin() + (size_t)row * Ea, b->Ldata.begin() + (size_t)(row + 1) * Ea,
                  md.begin() + (size_t)r * Ea);
        mt[r] = b->Ltgt[row];
        mtt[r] = b->Ltt.empty() ? 0.0 : b->Ltt[row];
        if (!b->Lpan.empty())
          mpan.push_back(b->Lpan[row]);
      }
      for (int r = 0; r < np; r++) {
        const int row = b->pr[b->pb[tk] + r];
        std::copy(b->Pdata.begin() + (size_t)row * Ea, b->Pdata.begin() + (size_t)(row + 1) * Ea,
                  mpd.begin() + (size_t)r * Ea);
        mppt[r] = b->Ppt.empty() ? 0.0 : b->Ppt[row];
        if (!b->Ppan.empty())
          mppan.push_back(b->Ppan[row]);
      }
      edm_manifold_view M = b->Lv, Mp = b->Pv;
      M.data = md.data();
      M.targets = mt.data();
      M.target_times = mtt.data();
      M.point_times = nullptr;
      M.panel = mpan.empty() ? nullptr : mpan.data();
      M.num_points = nl;
      Mp.data = mpd.data();
      Mp.targets = nullptr;
      Mp.point_times = mppt.data();
      Mp.target_times = nullptr;
      Mp.panel = mppan.empty() ? nullptr : mppan.data();
      Mp.num_points = np;
      edm_options to = b->o;
      to.k = b->kk[tk];
      to.thetas = b->thetas.data();
      to.metrics = b->metrics.empty() ? nullptr : b->metrics.data();
      const int64_t ob = b->ob[tk];
      const int r = edm_predict(c, &M, &Mp, &to, b->hpred.data() + ob, b->hrc.data() + ob, nullptr,
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
  BatchParams& B = b->B;
  hipStream_t s = c->stream;
  dfill(c, B.counter, 1, 0);
  if (!b->outputs_filled) {
    // Every run writes every output of every target (a prediction, or the fill values themselves for a
    // target left unpredicted), so the outputs are filled on the first run only, as a safety net: a
    // refill would cost each run about 16 bytes of HBM writes per prediction.
    dfill(c, B.pred, (size_t)b->total_out, FEDM_MISSING_D);
    dfill(c, B.rc, (size_t)b->total_out, (int)FEDM_UNKNOWN_ERROR);
    dfill(c, B.k_used, (size_t)b->total_k, (int)FEDM_MISSING_I);
    b->outputs_filled = true;
  }
  dfill(c, B.redo_count, 1, 0);
  HIP_TRY(hipEventRecord(c->ev[4], s));
  if (b->rank_route) {
    dfill(c, b->S.counter, 1, 0);
    void* rargs[] = { &b->R };
    HIP_TRY(hipLaunchKernel(rank_build_kernel(b->emax), dim3(b->R.np), dim3(256), rargs, 0, s));
    HIP_TRY(hipEventRecord(c->ev[6], s));
    if (stop())
      return batch_break(b);
    void* sargs[] = { &b->S };
    if (b->multi_scan) {
      HIP_TRY(hipLaunchKernel(rank_scan_multi_kernel(b->kcap, b->S.words),
                              dim3(((size_t)b->S.np * b->S.parts + MS_WAVES - 1) / MS_WAVES), dim3(64 * MS_WAVES), sargs,
                              b->multi_lds, s));
      HIP_TRY(hipLaunchKernel((const void*)rank_scatter, dim3((b->S.num_rank + 63) / 64, (b->S.np + 63) / 64, b->S.T),
                              dim3(256), sargs, 0, s));
    }
    else if (b->S.num_items > 0)
      HIP_TRY(hipLaunchKernel(rank_scan_kernel(b->kcap, b->emax), dim3(b->scan_grid), dim3(256), sargs,
                              b->scan_lds, s));
  } else {
    HIP_TRY(hipEventRecord(c->ev[6], s));
  }
  HIP_TRY(hipEventRecord(c->ev[7], s));
  if (stop())
    return batch_break(b);
  const void* fn = batch_kernel(b->kcap, b->Ea);
  void* args[] = { &B };
  if (B.num_items > 0)
    HIP_TRY(hipLaunchKernel(fn, dim3(b->grid), dim3(256), args, 0, s));
  {
    // targets handed over by the direct kernel or the rank scan: exact selection
    const int2* none = nullptr;
    int zero = 0;
    void* rargs[] = { &B, &none, &zero };
    HIP_TRY(hipLaunchKernel(batch_redo_kernel(b->kcap), dim3(256), dim3(256), rargs, 0, s));
    // more hand-overs than the list holds: every target again with the exact selection (these
    // launches return at once otherwise, so the host does not wait for the count mid-step)
    const int2* items[2] = { B.items, b->rank_route ? b->S.items : nullptr };
    int counts[2] = { B.num_items, b->rank_route ? b->S.num_items : 0 };
    for (int q = 0; q < 2; q++)
      if (counts[q] > 0) {
        void* aargs[] = { &B, &items[q], &counts[q] };
        HIP_TRY(hipLaunchKernel(batch_redo_kernel(b->kcap), dim3(b->grid), dim3(256), aargs, 0, s));
      }
  }
  HIP_TRY(hipEventRecord(c->ev[5], s));
  HIP_TRY(hipEventSynchronize(c->ev[5]));
  {
    int handed = 0;
    HIP_TRY(hipMemcpy(&handed, B.redo_count, sizeof(int), hipMemcpyDeviceToHost));
    b->redo_last = handed;
    c->last_redo = handed;
  }
  b->last_ms = ev_ms(c->ev[4], c->ev[5]);
  // [0] whole batch, [1] rank-table build, [2] rank scan, [3] direct kernel
  c->timings[0] = b->last_ms;
  c->timings[1] = ev_ms(c->ev[4], c->ev[6]);
  c->timings[2] = ev_ms(c->ev[6], c->ev[7]);
  c->timings[3] = ev_ms(c->ev[7], c->ev[5]);
  return 0;
}

int edm_batch_fetch(edm_batch* b, double* pred, int* rc, int* k_used)
{
  edm_ctx* c = b->ctx;
  hipSetDevice(c->device);
  if (!b->device_route) {
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
