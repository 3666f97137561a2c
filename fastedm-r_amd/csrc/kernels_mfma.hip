// Mixed-precision k-NN for one task: FP32 matrix-core distance tiles with an exact FP64 re-rank.
//
// Distances are screened with the expansion ||x||^2 + ||y||^2 - 2 x.y on centred FP32 coordinates;
// the dot products of a 16-target x 16-library tile come from v_mfma_f32_16x16x4f32 (KP/4
// instructions per tile).  A pair is a candidate when its FP32 distance is within a rigorous bound of
// the target's current k-th distance (bound: FP32 rounding of the centred inputs, of the norms and dot
// products, and of the final combination).  Candidates are queued per target in LDS and recomputed in
// FP64 in the reference's operation order (src/distances.cpp:186-216) by the target's owner lane,
// which keeps the exact (sqrt distance, position) list (src/edm.cpp:652-681).  The per-slice lists have
// the layout of knn_slices_f64, so the merge, simplex and S-map stages are shared with the plain route.
#include "kernels_common.h"

namespace fedm {

typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float up_f32(double x)
{
  float f = (float)x;
  if ((double)f < x)
    f = __int_as_float(__float_as_int(f) + 1);
  return f;
}

// FP64 squared distance in the reference's order (src/distances.cpp:186-216); the loads of 8 coordinates
// of both rows are issued together (16 VGPR pairs in flight, which leaves room for four wavefronts per
// SIMD).
template <int KP>
__device__ __forceinline__ double dist_f64(const double* x, const double* y, int Ea)
{
  constexpr int LB = 8;
  double acc = 0.0;
#pragma unroll 1
  for (int d0 = 0; d0 < KP; d0 += LB) {
    double xv[LB], yv[LB];
#pragma unroll
    for (int u = 0; u < LB; u++) {
      xv[u] = d0 + u < Ea ? x[d0 + u] : 0.0;
      yv[u] = d0 + u < Ea ? y[d0 + u] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < LB; u++)
      if (d0 + u < Ea) {
        const double dd = xv[u] - yv[u];
        acc += dd * dd;
      }
  }
  return acc;
}

// Makes one lane's LDS writes visible to the other lanes of its wavefront.
__device__ __forceinline__ void lds_wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

constexpr int PAR = 4; // queued candidates per target recomputed by the whole wavefront at once

// Wavefronts per SIMD the register budget aims at: the blocks of 4 wavefronts the static LDS of
// knn_mfma<KCAP, KP> (about 256 KP + 768 KCAP + 11 008 bytes) leaves room for on a CU, at most 4 (3 for
// KP = 64, whose operands need more registers).
constexpr int mfma_waves(int kcap, int kp)
{
  return (163840 / (256 * kp + 768 * kcap + 11008)) >= 4 ? (kp == 64 ? 3 : 4)
                                                         : 163840 / (256 * kp + 768 * kcap + 11008);
}

template <int KCAP, int KP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(mfma_waves(KCAP, KP)))) void knn_mfma(MfmaParams P, int k, int slice_len, double* __restrict__ out_s,
                                               int* __restrict__ out_i, int* __restrict__ out_cnt)
{
  constexpr int MCH = 64; // library points per LDS chunk (16-point tiles); LDS for four blocks per CU
  constexpr int F4 = MCH * KP / 4 / 256;    // float4 of the next chunk each thread prefetches
  __shared__ float Bs[MCH * KP];
  __shared__ float Bn[MCH];
  __shared__ float thr_s[64];
  __shared__ int qcnt[64];
  constexpr int QCAP = 2 * MCH;             // queue slots per target: the candidates of two chunks
  __shared__ unsigned char queue[64 * QCAP];
  // Per-target arrays are entry-major (entry e of target t at e * 64 + t): the 16 lanes that work on one
  // wavefront's targets touch consecutive words, without bank conflicts.
  __shared__ double Ls[64 * KCAP]; // exact lists of the block's targets, by (sqrt distance, position)
  __shared__ int Li[64 * KCAP];
  __shared__ double accq[64 * PAR]; // FP64 distances of each target's first PAR queued candidates

  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int g = lane >> 4;  // lane group: supplies K index g of each 4-wide MFMA step
  const int l16 = lane & 15;
  const int tbase = blockIdx.x * 64 + w * 16;
  const int lo = blockIdx.y * slice_len;
  const int hi = min(P.NL, lo + slice_len);

  // A operand: target row (tbase + l16), coordinates 16q + 4g .. +3
  v4f a[KP / 16];
  {
    const int t = tbase + l16;
#pragma unroll
    for (int q = 0; q < KP / 16; q++) {
      if (t < P.NP)
        a[q] = *(const v4f*)(P.P32 + (size_t)t * KP + 16 * q + 4 * g);
      else
        a[q] = v4f{ 0.f, 0.f, 0.f, 0.f };
    }
  }
  // output rows of this lane: targets tbase + 4g + r
  float ny[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int t = tbase + 4 * g + r;
    ny[r] = t < P.NP ? P.Pn[t] : 0.f;
  }
  // owner lanes 0..15 keep the exact lists of targets tbase + lane (LDS rows)
  const bool owner = lane < 16 && (tbase + lane) < P.NP;
  const int ot = tbase + lane;
  double* ls = Ls + (w * 16 + (lane & 15));
  int* li = Li + (w * 16 + (lane & 15));
  if (lane < 16)
    for (int j = 0; j < KCAP; j++) {
      ls[j * 64] = INF_D;
      li[j * 64] = INT_MAX;
    }
  double s_kth = INF_D, thr = INF_D;
  int zeros = 0;
  const double yn = owner ? P.ynorm[ot] : 0.0;
  if (lane < 16) {
    // the screening threshold starts from the seed's bound when there is one: every pair that can
    // enter the list has sqrt(D) <= sqrt(seed), so D <= seed' = sqrt(seed)^2 (1 + 2^-49), and the
    // FP32 value of such a pair is within B(seed') of D (the bound of the threshold updates below)
    float ts = __builtin_huge_valf();
    if (owner && P.seed) {
      const double a = P.seed[ot];
      if (a < INF_D) {
        const double s0 = sqrt(a);
        const double T = s0 * s0 * SQRT_GUARD;
        const double sq = sqrt(T);
        const double rr = fmin(P.R + yn, 2.0 * yn + sq) * (1.0 + 1e-6);
        ts = up_f32(T + 2.0 * (P.k1 * rr * sq + P.c2 * T + P.k3 * rr * rr));
      }
    }
    thr_s[w * 16 + lane] = ts;
    qcnt[w * 16 + lane] = 0;
  }

  // the next chunk's FP32 rows and norms, loaded while the current chunk is processed
  v4f nb[F4];
  float nn = 0.f;
  auto prefetch = [&](int c0) {
    const int cn = min(MCH, hi - c0);
#pragma unroll
    for (int v = 0; v < F4; v++) {
      const int E = threadIdx.x + 256 * v;
      const int j = E / (KP / 4), sub = E - j * (KP / 4);
      nb[v] = j < cn ? *(const v4f*)(P.M32 + (size_t)(c0 + j) * KP + 4 * sub) : v4f{ 0.f, 0.f, 0.f, 0.f };
    }
    nn = (threadIdx.x < MCH && (int)threadIdx.x < cn) ? P.Mn[c0 + threadIdx.x] : 0.f;
  };
  if (lo < hi)
    prefetch(lo);

  // The queues are drained every second chunk (and after the last): a queued entry is the position
  // in the pair of chunks that starts at cbase.
  int half = 0, cbase = lo;
  for (int c0 = lo; c0 < hi; c0 += MCH) {
    const int cn = min(MCH, hi - c0);
    __syncthreads();
#pragma unroll
    for (int v = 0; v < F4; v++)
      *(v4f*)(Bs + 4 * (threadIdx.x + 256 * v)) = nb[v];
    if (threadIdx.x < MCH)
      Bn[threadIdx.x] = nn;
    if (c0 + MCH < hi)
      prefetch(c0 + MCH);
    __syncthreads();
    float th[4];
#pragma unroll
    for (int r = 0; r < 4; r++)
      th[r] = thr_s[w * 16 + 4 * g + r];
    bool tok[4]; // this lane's output rows are real targets
#pragma unroll
    for (int r = 0; r < 4; r++)
      tok[r] = tbase + 4 * g + r < P.NP;
    // TB tiles at a time: their operand loads and their independent MFMA chains are issued together,
    // then the 4 TB results of each lane are screened without branches; only a wavefront in which
    // some pair passes takes the (divergent) queue appends
    constexpr int TB = KP == 64 ? 2 : 4;
    for (int jt0 = 0; jt0 < MCH / 16; jt0 += TB) {
      v4f b[TB][KP / 16];
#pragma unroll
      for (int u = 0; u < TB; u++)
#pragma unroll
        for (int q = 0; q < KP / 16; q++)
          b[u][q] = *(const v4f*)(Bs + ((jt0 + u) * 16 + l16) * KP + 16 * q + 4 * g);
      v4f acc[TB];
#pragma unroll
      for (int u = 0; u < TB; u++)
        acc[u] = v4f{ 0.f, 0.f, 0.f, 0.f };
#pragma unroll
      for (int q = 0; q < KP / 16; q++)
#pragma unroll
        for (int kk = 0; kk < 4; kk++)
#pragma unroll
          for (int u = 0; u < TB; u++)
            acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[q][kk], b[u][q][kk], acc[u], 0, 0, 0);
      float d2[TB][4];
      bool any = false; // some value passes the threshold (validity checked on the slow path)
#pragma unroll
      for (int u = 0; u < TB; u++) {
        const float nx = Bn[(jt0 + u) * 16 + l16];
#pragma unroll
        for (int r = 0; r < 4; r++) {
          d2[u][r] = __builtin_fmaf(-2.0f, acc[u][r], ny[r] + nx);
          any = any | (d2[u][r] <= th[r]);
        }
      }
      if (__ballot(any)) {
#pragma unroll
        for (int u = 0; u < TB; u++) {
          const int j = (jt0 + u) * 16 + l16;
#pragma unroll
          for (int r = 0; r < 4; r++)
            if ((d2[u][r] <= th[r]) & tok[r] & (j < cn)) {
              const int ti = 4 * g + r;
              const int slot = atomicAdd(&qcnt[w * 16 + ti], 1);
              queue[slot * 64 + w * 16 + ti] = (unsigned char)(half * MCH + j);
            }
        }
      }
    }
    const bool drain = half == 1 || c0 + MCH >= hi;
    half ^= 1;
    if (!drain)
      continue;
    half = 0;
    const int cb = cbase;
    cbase = c0 + MCH;
    // the queues and lists belong to this wavefront: a wave-level sync suffices
    lds_wave_sync();
    // exact FP64 recomputation: lane (target lane % 16, slot lane / 16) takes queued candidate `slot`
    // of its target, so the first PAR candidates of all 16 targets cost one round of loads
    {
      const int ti = lane & 15, sl = lane >> 4;
      const int qi = w * 16 + ti;
      if (tbase + ti < P.NP && sl < qcnt[qi]) {
        const int j = queue[sl * 64 + qi];
        accq[sl * 64 + qi] = dist_f64<KP>(P.M + (size_t)(cb + j) * P.Ea, P.Mp + (size_t)(tbase + ti) * P.Ea, P.Ea);
      }
    }
    lds_wave_sync();
    // owner lanes: the candidates in queue order (later ones recomputed here), into the exact list
    if (owner) {
      const int nq = qcnt[w * 16 + lane];
      const double* y = P.Mp + (size_t)ot * P.Ea;
      bool moved_kth = false;
      for (int e = 0; e < nq; e++) {
        const int j = queue[e * 64 + w * 16 + lane];
        const double acc =
          e < PAR ? accq[e * 64 + w * 16 + lane] : dist_f64<KP>(P.M + (size_t)(cb + j) * P.Ea, y, P.Ea);
        if (acc == 0.0 || acc == MISSING_D) {
          zeros++;
          continue;
        }
        if (acc < thr) {
          const double s = sqrt(acc);
          const int ci = cb + j;
          // a tie with the k-th entry enters when its position is smaller, so the list does not depend
          // on the order in which candidates arrive
          if (s < s_kth || (s == s_kth && ci < li[(k - 1) * 64])) {
            // sorted insertion among the first k entries (entry k - 1 drops out), walking down four
            // entries per round of LDS reads; the new k-th entry is known from the first round
            int q = k - 1;
            double kth_next = s;
            bool first = true;
            while (q > 0) {
              double v[4];
              int iv[4];
#pragma unroll
              for (int u = 0; u < 4; u++) {
                const int r = max(q - 1 - u, 0);
                v[u] = ls[r * 64];
                iv[u] = li[r * 64];
              }
              if (first && (s < v[0] || (s == v[0] && ci < iv[0])))
                kth_next = v[0];
              first = false;
              int moved = 0;
              bool more = true;
#pragma unroll
              for (int u = 0; u < 4; u++) {
                more = more && q - 1 - u >= 0 && (s < v[u] || (s == v[u] && ci < iv[u]));
                if (more) {
                  ls[(q - u) * 64] = v[u];
                  li[(q - u) * 64] = iv[u];
                  moved++;
                }
              }
              q -= moved;
              if (moved < 4)
                break;
            }
            ls[q * 64] = s;
            li[q * 64] = ci;
            s_kth = kth_next;
            moved_kth = true;
            if (s_kth != INF_D)
              thr = s_kth * s_kth * SQRT_GUARD;
          }
        }
      }
      qcnt[w * 16 + lane] = 0;
      if (moved_kth && thr != INF_D) {
        // FP32 screening threshold: thr plus the bound B(thr) = 2 (c1 sqrt(thr) + c2 thr + c3) on
        // |FP32 d2 - D| for pairs with D <= thr, where c1 = k1 rr and c3 = k3 rr^2 for any rr >=
        // |x| + |y| (centred rows).  Such a pair has |x| <= |y| + sqrt(D), so rr = 2 |y| + sqrt(thr) bounds
        // it as well as R + |y| does, and more tightly once thr is small.
        const double sq = sqrt(thr);
        const double rr = fmin(P.R + yn, 2.0 * yn + sq) * (1.0 + 1e-6);
        thr_s[w * 16 + lane] = up_f32(thr + 2.0 * (P.k1 * rr * sq + P.c2 * thr + P.k3 * rr * rr));
      }
    }
  }
  if (owner) {
    const size_t base = ((size_t)blockIdx.y * P.NP + ot) * KCAP;
#pragma unroll
    for (int j = 0; j < KCAP; j++) {
      out_s[base + j] = ls[j * 64];
      out_i[base + j] = li[j * 64];
    }
    out_cnt[(size_t)blockIdx.y * P.NP + ot] = (hi > lo ? hi - lo : 0) - zeros;
  }
}

// Seed of knn_mfma's screening threshold: for each target, the k-th smallest nonzero exact squared distance
// over a window of `win` (<= 256) library rows that lie near the target in a k-d order of the library
// (rows kdp[wst[t]] .. kdp[wst[t] + win - 1]), or +inf when the window has fewer than k.  The k-th over
// a subset is never below the k-th over the whole library, so the seed bounds the true k-th distance
// from above.  Distances are summed in the reference's order (src/distances.cpp:186-216), as in the
// re-rank.  One wavefront per target, four window rows per lane; targets are taken in window order
// (tord), so consecutive wavefronts read overlapping rows.
template <int KP>
__global__ __launch_bounds__(256) void knn_seed(MfmaParams P, const int* __restrict__ kdp, const int* __restrict__ wst,
                                               const int* __restrict__ tord, int win, int k,
                                               double* __restrict__ seed)
{
  const int lane = threadIdx.x & 63;
  const int wi = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (wi >= P.NP)
    return;
  const int t = tord[wi];
  const int a = wst[t];
  const double* y = P.Mp + (size_t)t * P.Ea;
  double v[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int j = lane + 64 * r;
    double acc = INF_D;
    if (j < win) {
      const double d = dist_f64<KP>(P.M + (size_t)kdp[a + j] * P.Ea, y, P.Ea);
      acc = (d != 0.0 && d != MISSING_D) ? d : INF_D;
    }
    v[r] = acc;
  }
  // sort the lane's four values, then take k wave-wide minima of the lanes' heads
  auto cs = [](double& x, double& z) {
    const double lo = fmin(x, z), hi = fmax(x, z);
    x = lo;
    z = hi;
  };
  cs(v[0], v[1]);
  cs(v[2], v[3]);
  cs(v[0], v[2]);
  cs(v[1], v[3]);
  cs(v[1], v[2]);
  double kth = INF_D;
  int h = 0;
  for (int i = 0; i < k; i++) {
    const double head = h == 0 ? v[0] : h == 1 ? v[1] : h == 2 ? v[2] : h == 3 ? v[3] : INF_D;
    double m = head;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
      m = fmin(m, __shfl_xor(m, off));
    kth = m;
    if (m == INF_D)
      break;
    const unsigned long long b = __ballot(head == m);
    if (lane == (int)__ffsll((long long)b) - 1)
      h++;
  }
  if (lane == 0)
    seed[t] = kth;
}

#define FEDM_SEED(KPv)                                                                                           \
  template __global__ void knn_seed<KPv>(MfmaParams, const int*, const int*, const int*, int, int, double*);
FEDM_SEED(16)
FEDM_SEED(32)
FEDM_SEED(64)

// Lane-layout self-test of v_mfma_f32_16x16x4f32 as knn_mfma uses it: A[i][k] = i + 1 for k = 0,
// B[k][j] = 100 (j + 1) for k = 0, zero elsewhere; out[lane * 4 + r] must be D[4 g + r][lane % 16].
__global__ void mfma_layout_probe(float* out)
{
  const int lane = threadIdx.x;
  const int g = lane >> 4, l16 = lane & 15;
  const float av = g == 0 ? (float)(l16 + 1) : 0.f;
  const float bv = g == 0 ? 100.f * (float)(l16 + 1) : 0.f;
  v4f acc = { 0.f, 0.f, 0.f, 0.f };
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
#pragma unroll
  for (int r = 0; r < 4; r++)
    out[lane * 4 + r] = acc[r];
}

#define FEDM_MFMA(KC, KPv) template __global__ void knn_mfma<KC, KPv>(MfmaParams, int, int, double*, int*, int*);
#define FEDM_MFMA_K(KC) FEDM_MFMA(KC, 16) FEDM_MFMA(KC, 32) FEDM_MFMA(KC, 64)
FEDM_MFMA_K(4)
FEDM_MFMA_K(8)
FEDM_MFMA_K(16)
FEDM_MFMA_K(32)
FEDM_MFMA_K(64)

} // namespace fedm
