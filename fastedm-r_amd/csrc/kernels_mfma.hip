// Mixed-precision k-NN for one task: FP32 matrix-core distance tiles with an exact FP64 re-rank.
//
// Distances are screened with the expansion ||x||^2 + ||y||^2 - 2 x.y on centred FP32 coordinates:
// a 16-target x 16-library tile of x.y - ||x||^2 / 2 comes from v_mfma_f32_16x16x4f32 (KP/4 instructions
// per tile, the accumulator started at -||x||^2 / 2 of the tile's library point), and a pair is a
// candidate when that value is at least the target's screening floor (||y||^2 - th) / 2, rounded down,
// where th is the target's current k-th squared distance plus a rigorous bound (FP32 rounding of the
// centred inputs, of the norms, and of the accumulation of KP products and the starting value); one
// compare per pair.  Candidates are queued per target in LDS and recomputed in
// FP64 in the reference's operation order (src/distances.cpp:186-216) by the target's owner lane,
// which keeps the exact (sqrt distance, position) list (src/edm.cpp:652-681).  The per-slice lists have
// the layout of knn_slices_f64, so the merge, simplex and S-map stages are shared with the plain route.
#include "kernels_common.h"

namespace fedm {

typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float down_f32(double x)
{
  float f = (float)x;
  if ((double)f > x)
    f = nextafterf(f, -__builtin_huge_valf());
  return f;
}

// Screening floor of a target with FP32 squared norm ny, for squared distances up to T: a pair whose
// exact squared distance is at most T has x.y - ||x||^2 / 2 (as the matrix cores accumulate it) at least
// (ny - T - B(T)) / 2, B(T) = 2 (k1 rr sqrt(T) + c2 T + k3 rr^2) for any rr >= |x| + |y| of the centred rows
// (rr is taken at least 1e-12, so that the flush of subnormal FP32 products stays inside the bound).
__device__ __forceinline__ float screen_floor(const MfmaParams& P, double T, double yn, float ny)
{
  if (!(T < INF_D))
    return -__builtin_huge_valf();
  const double sq = sqrt(T);
  const double rr = fmax(fmin(P.R + yn, 2.0 * yn + sq) * (1.0 + 1e-6), 1e-12);
  const double th = T + 2.0 * (P.k1 * rr * sq + P.c2 * T + P.k3 * rr * rr);
  // (ny - th) / 2, lowered past the rounding of its FP64 evaluation, then rounded down to FP32
  const double v = 0.5 * ((double)ny - th);
  return down_f32(v - fabs(v) * 0x1p-50 - 0x1p-1000);
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
                                               int* __restrict__ out_i, int* __restrict__ out_cnt,
                                               const int* __restrict__ tl, const int* __restrict__ tn)
{
  constexpr int MCH = 64; // library points per LDS chunk (16-point tiles); LDS for four blocks per CU
  constexpr int F4 = MCH * KP / 4 / 256;    // float4 of the next chunk each thread prefetches
  __shared__ float Bs[MCH * KP];
  __shared__ float Bc[MCH];  // -||x||^2 / 2 of the chunk's library points (the accumulators' start)
  __shared__ float flo_s[64]; // screening floor of each target of the block
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
  const int lo = blockIdx.y * slice_len;
  const int hi = min(P.NL, lo + slice_len);
  // Targets: block row x of the wavefront is target blockIdx.x * 64 + 16 w + x, or, with a target list
  // (tl, *tn entries, the same for every slice), that entry of the list.
  const int nlist = tl ? *tn : P.NP;
  if (tl && blockIdx.x * 64 >= nlist)
    return;
  auto target_of = [&](int x) {
    const int e = blockIdx.x * 64 + w * 16 + x;
    return e < nlist ? (tl ? tl[e] : e) : -1;
  };

  // A operand: target row target_of(l16), coordinates 16q + 4g .. +3
  v4f a[KP / 16];
  {
    const int t = target_of(l16);
#pragma unroll
    for (int q = 0; q < KP / 16; q++) {
      if (t >= 0)
        a[q] = *(const v4f*)(P.P32 + (size_t)t * KP + 16 * q + 4 * g);
      else
        a[q] = v4f{ 0.f, 0.f, 0.f, 0.f };
    }
  }
  // output rows of this lane: block rows 4g + r
  // owner lanes 0..15 keep the exact lists of the targets of block rows lane (LDS rows)
  const int ot = lane < 16 ? target_of(lane) : -1;
  const bool owner = ot >= 0;
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
  const float nyo = owner ? P.Pn[ot] : 0.f;
  if (lane < 16) {
    // the screening starts from the seed's bound when there is one: every pair that can enter the list
    // has sqrt(D) <= sqrt(seed), so D <= sqrt(seed)^2 (1 + 2^-49)
    float fl = -__builtin_huge_valf();
    if (owner && P.seed) {
      const double a = P.seed[ot];
      if (a < INF_D) {
        const double s0 = sqrt(a);
        fl = screen_floor(P, s0 * s0 * SQRT_GUARD, yn, nyo);
      }
    }
    flo_s[w * 16 + lane] = fl;
    qcnt[w * 16 + lane] = 0;
  }

  // Library chunks are staged through registers two chunks ahead (the rows of chunk c + 2 are loaded
  // while chunk c is screened), in two register sets that alternate with the chunks.
  v4f nbA[F4], nbB[F4];
  float nnA = 0.f, nnB = 0.f;
  auto prefetch = [&](v4f(&nb)[F4], float& nn, int c0) {
    const int cn = min(MCH, hi - c0);
#pragma unroll
    for (int v = 0; v < F4; v++) {
      const int E = threadIdx.x + 256 * v;
      const int j = E / (KP / 4), sub = E - j * (KP / 4);
      nb[v] = j < cn ? *(const v4f*)(P.M32 + (size_t)(c0 + j) * KP + 4 * sub) : v4f{ 0.f, 0.f, 0.f, 0.f };
    }
    nn = (threadIdx.x < MCH && (int)threadIdx.x < cn) ? -0.5f * P.Mn[c0 + threadIdx.x] : 0.f;
  };
  if (lo < hi)
    prefetch(nbA, nnA, lo);
  if (lo + MCH < hi)
    prefetch(nbB, nnB, lo + MCH);

  // The queues are drained every second chunk (and after the last): a queued entry is the position
  // in the pair of chunks that starts at cbase.
  int half = 0, cbase = lo;
  auto chunk = [&](int c0, v4f(&nb)[F4], float& nn) {
    const int cn = min(MCH, hi - c0);
    __syncthreads();
#pragma unroll
    for (int v = 0; v < F4; v++)
      *(v4f*)(Bs + 4 * (threadIdx.x + 256 * v)) = nb[v];
    if (threadIdx.x < MCH)
      Bc[threadIdx.x] = nn;
    if (c0 + 2 * MCH < hi)
      prefetch(nb, nn, c0 + 2 * MCH);
    __syncthreads();
    float fl[4];
#pragma unroll
    for (int r = 0; r < 4; r++)
      fl[r] = flo_s[w * 16 + 4 * g + r];
    bool tok[4]; // this lane's output rows are real targets
#pragma unroll
    for (int r = 0; r < 4; r++)
      tok[r] = target_of(4 * g + r) >= 0;
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
      for (int u = 0; u < TB; u++) {
        const float c = Bc[(jt0 + u) * 16 + l16];
        acc[u] = v4f{ c, c, c, c };
      }
#pragma unroll
      for (int q = 0; q < KP / 16; q++)
#pragma unroll
        for (int kk = 0; kk < 4; kk++)
#pragma unroll
          for (int u = 0; u < TB; u++)
            acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[q][kk], b[u][q][kk], acc[u], 0, 0, 0);
      // some value reaches its floor (validity is checked on the slow path)
      bool any = false;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        float m = acc[0][r];
#pragma unroll
        for (int u = 1; u < TB; u++)
          m = fmaxf(m, acc[u][r]);
        any = any | (m >= fl[r]);
      }
      if (__ballot(any)) {
#pragma unroll
        for (int u = 0; u < TB; u++) {
          const int j = (jt0 + u) * 16 + l16;
#pragma unroll
          for (int r = 0; r < 4; r++)
            if ((acc[u][r] >= fl[r]) & tok[r] & (j < cn)) {
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
      return;
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
      const int tt = target_of(ti);
      if (tt >= 0 && sl < qcnt[qi]) {
        const int j = queue[sl * 64 + qi];
        accq[sl * 64 + qi] = dist_f64<KP>(P.M + (size_t)(cb + j) * P.Ea, P.Mp + (size_t)tt * P.Ea, P.Ea);
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
        // screening floor for pairs with D <= thr: a pair that can still enter has |x| <= |y| + sqrt(D),
        // so rr = 2 |y| + sqrt(thr) bounds |x| + |y| as well as R + |y| does, and more tightly once thr
        // is small (screen_floor)
        const float f = screen_floor(P, thr, yn, nyo);
        flo_s[w * 16 + lane] = fmaxf(flo_s[w * 16 + lane], f);
      }
    }
    };
  for (int c0 = lo; c0 < hi; c0 += 2 * MCH) {
    chunk(c0, nbA, nnA);
    if (c0 + MCH < hi)
      chunk(c0 + MCH, nbB, nnB);
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
// over rows of a k-d order of the library near the target, or +inf when fewer than k are nonzero.  The
// target's window is rows wst[t] .. wst[t] + win - 1 of that order (win <= 256); with wide > win the rows
// scanned are `wide` rows centred on that window.  Lane l takes rows l, l + 64, ... and keeps, per target,
// its four smallest values; the seed is the k-th smallest of what the lanes keep.  That is the k-th
// smallest over a subset of the library, so it bounds the true k-th distance from above (at wide = win
// every lane holds at most four rows, so it is the window's k-th).  Distances are summed in the
// reference's order (src/distances.cpp:186-216), operation for operation as dist_f64 sums them in the
// re-rank.  kdm holds the library rows in the k-d order, so a window is contiguous.  One wavefront per
// group of up to SEED_G consecutive targets of the window order that share a window (groups start at
// sgrp[g]; sgrp[ngrp] = NP): each lane loads its rows once and sums their distances to every target of the
// group (the targets' coordinates in LDS), then each target takes its k-th by k rounds of a wavefront
// minimum over the lanes' kept values.
constexpr int SEED_G = 4;
template <int KP>
__global__ __launch_bounds__(256) void knn_seed(MfmaParams P, const double* __restrict__ kdm,
                                               const int* __restrict__ wst, const int* __restrict__ tord,
                                               const int* __restrict__ sgrp, int ngrp, int win, int wide, int k,
                                               double* __restrict__ seed)
{
  __shared__ double ys[4][SEED_G][KP];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int gi = blockIdx.x * 4 + wv;
  if (gi >= ngrp)
    return;
  const int s0 = sgrp[gi];
  const int cnt = min(SEED_G, sgrp[gi + 1] - s0);
  const int Ea = P.Ea;
  const int nr = max(win, min(wide, P.NL));
  const int a = nr == win ? wst[tord[s0]] : max(0, min(P.NL - nr, wst[tord[s0]] + win / 2 - nr / 2));
  for (int e = lane; e < SEED_G * KP; e += 64) {
    const int j = e / KP, d = e - j * KP;
    ys[wv][j][d] = (j < cnt && d < Ea) ? P.Mp[(size_t)tord[s0 + j] * Ea + d] : 0.0;
  }
  lds_wave_sync();
  auto cs = [](double& x, double& z) {
    const double lo = fmin(x, z), hi = fmax(x, z);
    x = lo;
    z = hi;
  };
  // per target, the lane's four smallest values in increasing order
  double v[SEED_G][4];
#pragma unroll
  for (int j = 0; j < SEED_G; j++)
#pragma unroll
    for (int r = 0; r < 4; r++)
      v[j][r] = INF_D;
  for (int row = lane; row < nr; row += 64) {
    double acc[SEED_G];
#pragma unroll
    for (int j = 0; j < SEED_G; j++)
      acc[j] = 0.0;
    const double* x = kdm + (size_t)(a + row) * Ea;
    constexpr int LB = 8;
#pragma unroll 1
    for (int d0 = 0; d0 < KP; d0 += LB) {
      double xv[LB];
#pragma unroll
      for (int u = 0; u < LB; u++)
        xv[u] = d0 + u < Ea ? x[d0 + u] : 0.0;
#pragma unroll
      for (int u = 0; u < LB; u++)
        if (d0 + u < Ea)
#pragma unroll
          for (int j = 0; j < SEED_G; j++) {
            const double dd = xv[u] - ys[wv][j][d0 + u];
            acc[j] += dd * dd;
          }
    }
#pragma unroll
    for (int j = 0; j < SEED_G; j++) {
      v[j][3] = fmin(v[j][3], (acc[j] != 0.0 && acc[j] != MISSING_D) ? acc[j] : INF_D);
      cs(v[j][2], v[j][3]);
      cs(v[j][1], v[j][2]);
      cs(v[j][0], v[j][1]);
    }
  }
  // per target: k wave-wide minima of the lanes' heads
#pragma unroll
  for (int j = 0; j < SEED_G; j++) {
    if (j >= cnt)
      break;
    double kth = INF_D;
    int h = 0;
    for (int i = 0; i < k; i++) {
      const double head = h == 0 ? v[j][0] : h == 1 ? v[j][1] : h == 2 ? v[j][2] : h == 3 ? v[j][3] : INF_D;
      double m = head;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1)
        m = fmin(m, __shfl_xor(m, off));
      kth = m;
      if (m == INF_D)
        break;
      const unsigned long long bl = __ballot(head == m);
      if (lane == (int)__ffsll((long long)bl) - 1)
        h++;
    }
    if (lane == 0)
      seed[tord[s0 + j]] = kth;
  }
}

#define FEDM_SEED(KPv)                                                                                           \
  template __global__ void knn_seed<KPv>(MfmaParams, const double*, const int*, const int*, const int*, int, int, int, \
                                         int, double*);
FEDM_SEED(16)
FEDM_SEED(32)
FEDM_SEED(64)

// A wider seed for the targets the two-phase form hands to the streaming kernel (list fl, *fc entries): the
// k-th smallest nonzero exact squared distance over 64 R rows of the k-d order centred on the target's
// window (rows wst[t] .. wst[t] + win - 1), kept when below the target's seed.  Both are k-th smallest
// values over subsets of the library, so the smaller still bounds the true k-th distance from above, and
// the streaming kernel's screen starts from a floor nearer to it.  Distances are summed as in knn_seed.  One wavefront per target (grid-stride);
// lane l holds rows l, l + 64, ... of the window, and the k-th smallest is taken by k rounds of a
// wavefront minimum over the lanes' smallest values not yet taken.
template <int KP, int R>
__global__ __launch_bounds__(256) void knn_seed_wide(MfmaParams P, const double* __restrict__ kdm,
                                                    const int* __restrict__ wst, int win, const int* __restrict__ fl,
                                                    const int* __restrict__ fc, int k, double* __restrict__ seed)
{
  __shared__ double ys[4][KP];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nlist = *fc;
  const int Ea = P.Ea;
  static_assert(R <= 64, "one bit per row of a lane");
  constexpr int WW = 64 * R;
  const int ww = min(WW, P.NL);
  for (int e = blockIdx.x * 4 + wv; e < nlist; e += gridDim.x * 4) {
    const int t = fl[e];
    const int a = max(0, min(P.NL - ww, wst[t] + win / 2 - ww / 2));
    for (int d = lane; d < KP; d += 64)
      ys[wv][d] = d < Ea ? P.Mp[(size_t)t * Ea + d] : 0.0;
    lds_wave_sync();
    double v[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int row = lane + 64 * r;
      double acc = 0.0;
      if (row < ww) {
        const double* x = kdm + (size_t)(a + row) * Ea;
        constexpr int LB = 8;
#pragma unroll 1
        for (int d0 = 0; d0 < KP; d0 += LB) {
          double xv[LB];
#pragma unroll
          for (int u = 0; u < LB; u++)
            xv[u] = d0 + u < Ea ? x[d0 + u] : 0.0;
#pragma unroll
          for (int u = 0; u < LB; u++)
            if (d0 + u < Ea) {
              const double dd = xv[u] - ys[wv][d0 + u];
              acc += dd * dd;
            }
        }
      }
      v[r] = (row < ww && acc != 0.0 && acc != MISSING_D) ? acc : INF_D;
    }
    // k rounds: the lanes' smallest untaken values, their wavefront minimum, taken by the first lane holding it
    unsigned long long taken = 0ull;
    double kth = INF_D;
    for (int i = 0; i < k; i++) {
      double h = INF_D;
      int hr = 0;
#pragma unroll
      for (int r = 0; r < R; r++)
        if (!((taken >> r) & 1ull) && v[r] < h) {
          h = v[r];
          hr = r;
        }
      double m = h;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1)
        m = fmin(m, __shfl_xor(m, off));
      kth = m;
      if (m == INF_D)
        break;
      const unsigned long long bl = __ballot(h == m);
      if (lane == (int)__ffsll((long long)bl) - 1)
        taken |= 1ull << hr;
    }
    if (lane == 0 && kth < seed[t])
      seed[t] = kth;
    lds_wave_sync();
  }
}

#define FEDM_SEEDW(KPv, Rv)                                                                                      \
  template __global__ void knn_seed_wide<KPv, Rv>(MfmaParams, const double*, const int*, int, const int*, const int*,  \
                                                  int, double*);
FEDM_SEEDW(16, 16)
FEDM_SEEDW(32, 16)
FEDM_SEEDW(64, 16)

// Two-phase form of the seeded matrix-core route.  Phase one (knn_screen) screens the (target, library
// point) pairs against the target's fixed seeded floor and records the positions that pass, up to SCR_CAP
// per target; phase two (knn_rerank) recomputes the recorded pairs in FP64 and keeps the exact (sqrt
// distance, position) list.  The seed bounds the target's k-th distance over the whole library, so every
// pair of the final list, and every zero-distance pair, passes the screen, and the count of valid pairs is
// the library size less the zeros among the candidates.  A target with more than SCR_CAP candidates, or
// without a seed, is handed to the streaming kernel (knn_mfma on a target list).

// Phase one.  A block of four wavefronts takes 128 consecutive targets of the k-d target order and one of
// `split` interleaved parts of the library's 64-row chunks in its k-d order (chunk c belongs to part c mod
// split, so a block whose targets need most of the library shares it out over `split` blocks).  It first
// lists the chunks of its part (1 024 at a time) whose box may hold a pair within the largest threshold of
// its targets: the squared gap between the chunk's box and the block's box, evaluated in FP64 on the
// outward-rounded FP32 bounds and lowered by 2^-40 of itself (more than the rounding of that sum and of the
// re-rank's FP64 distance), bounds every pair of the chunk from below, so a chunk left out holds no pair
// that can pass the exact test of any target of the block (no list pair, no zero distance).  Each wavefront
// then screens two 16-target tiles against the listed chunks, double-buffered in LDS (one barrier per
// chunk), five wavefronts per SIMD (four at 64 coordinates, where the LDS of five blocks would not fit).  A
// pair passes when x.y - ||x||^2 / 2 (the matrix cores' accumulation, started at -||x||^2 / 2) reaches the
// target's floor.  The passing pairs of one target row and one 64-row step take their slots with one
// atomic add on the target's count ccnt (zeroed before the launch), which ends as the number of pairs that
// passed.
// Largest of four FP32 values with v_max3_f32 and v_max_f32 (the values are finite or infinite, never NaN,
// so no canonicalisation is needed).
__device__ __forceinline__ float max4(float a, float b, float c, float d)
{
  float m;
  asm("v_max3_f32 %0, %1, %2, %3\n\tv_max_f32 %0, %0, %4" : "=&v"(m) : "v"(a), "v"(b), "v"(c), "v"(d));
  return m;
}

template <int KP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KP == 64 ? 4 : 5))) void knn_screen(MfmaParams P,
                                                                                                KdScreen K,
                                                                                                int split,
                                                                                                int* __restrict__ cand,
                                                                                                int* __restrict__ ccnt)
{
  constexpr int MCH = 64, TW = 2, CLW = 1024;
  constexpr int F4 = MCH * KP / 4 / 256;
  constexpr int TB = KP == 64 ? 2 : 4; // library tiles per MFMA batch
  __shared__ float Bs[2][MCH * KP];
  __shared__ float Bc[2][MCH];
  __shared__ int tid[64 * TW];
  __shared__ int clist[CLW];
  __shared__ float tb[2 * KP];
  __shared__ int ncl;
  __shared__ unsigned long long tmax;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, l16 = lane & 15;
  const int tblk = blockIdx.x * 64 * TW; // first slot of the block in the target order
  const int tw0 = tblk + w * 16 * TW;    // first slot of the wavefront
  const int Ea = P.Ea;
  auto target = [&](int x) { return x < P.NP ? K.tord[x] : -1; };

  v4f a[TW][KP / 16];
#pragma unroll
  for (int i = 0; i < TW; i++) {
    const int t = target(tw0 + 16 * i + l16);
#pragma unroll
    for (int q = 0; q < KP / 16; q++)
      a[i][q] = t >= 0 ? *(const v4f*)(P.P32 + (size_t)t * KP + 16 * q + 4 * g) : v4f{ 0.f, 0.f, 0.f, 0.f };
  }
  // floors of this lane's output rows: tile i, row 4g + r (-inf and no record for a target without a seed)
  float fl[TW][4];
  bool tok[TW][4];
#pragma unroll
  for (int i = 0; i < TW; i++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int t = target(tw0 + 16 * i + 4 * g + r);
      float f = -__builtin_huge_valf();
      if (t >= 0) {
        const double sd = P.seed[t];
        if (sd < INF_D) {
          const double s0 = sqrt(sd);
          f = screen_floor(P, s0 * s0 * SQRT_GUARD, P.ynorm[t], P.Pn[t]);
        }
      }
      fl[i][r] = f;
      tok[i][r] = f > -__builtin_huge_valf();
    }
  if (threadIdx.x == 0)
    tmax = 0ull;
  if (threadIdx.x < 2 * Ea)
    tb[threadIdx.x] = K.tbox[(size_t)blockIdx.x * 2 * Ea + threadIdx.x];
  __syncthreads();
  if (threadIdx.x < 64 * TW) {
    const int t = target(tblk + threadIdx.x);
    tid[threadIdx.x] = t;
    if (t >= 0 && P.seed[t] < INF_D) {
      const double s0 = sqrt(P.seed[t]);
      // the largest threshold of the block (non-negative doubles order as their bit patterns)
      atomicMax(&tmax, (unsigned long long)__double_as_longlong(s0 * s0 * SQRT_GUARD));
    }
  }

  v4f nb[F4];
  float nn = 0.f;
  auto load = [&](int c0) {
    const int cn = min(MCH, P.NL - c0);
#pragma unroll
    for (int v = 0; v < F4; v++) {
      const int E = threadIdx.x + 256 * v;
      const int j = E / (KP / 4), sub = E - j * (KP / 4);
      nb[v] = j < cn ? *(const v4f*)(K.m32 + (size_t)(c0 + j) * KP + 4 * sub) : v4f{ 0.f, 0.f, 0.f, 0.f };
    }
    nn = (threadIdx.x < MCH && (int)threadIdx.x < cn) ? -0.5f * K.mn[c0 + threadIdx.x] : 0.f;
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int v = 0; v < F4; v++)
      *(v4f*)(&Bs[buf][4 * (threadIdx.x + 256 * v)]) = nb[v];
    if (threadIdx.x < MCH)
      Bc[buf][threadIdx.x] = nn;
  };
  const int nch = (P.NL + MCH - 1) / MCH;
  for (int cw = blockIdx.y; cw < nch; cw += split * CLW) {
    // the chunks cw, cw + split, ... (up to CLW of them) that the block screens
    __syncthreads();
    if (threadIdx.x == 0)
      ncl = 0;
    __syncthreads();
    const double tm = __longlong_as_double((long long)tmax);
    for (int m = threadIdx.x; m < CLW; m += 256) {
      const int c = cw + m * split;
      if (c >= nch)
        break;
      bool keep = true;
      if (K.prune) {
        const float* bx = K.cbox + (size_t)c * 2 * Ea;
        double lb = 0.0;
        for (int d = 0; d < Ea; d++) {
          const double gp = fmax(0.0, fmax((double)bx[d] - (double)tb[Ea + d], (double)tb[d] - (double)bx[Ea + d]));
          lb += gp * gp;
        }
        keep = lb - lb * 0x1p-40 <= tm;
      }
      if (keep)
        clist[atomicAdd(&ncl, 1)] = c;
    }
    __syncthreads();
    const int n = ncl;
    if (n == 0)
      continue;
    load(clist[0] * MCH);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int li = 0; li < n; li++) {
      const int c0 = clist[li] * MCH;
      const int cn = min(MCH, P.NL - c0);
      const bool more = li + 1 < n;
      if (more)
        load(clist[li + 1] * MCH);
      for (int jt0 = 0; jt0 < MCH / 16; jt0 += TB) {
        v4f b[TB][KP / 16];
#pragma unroll
        for (int u = 0; u < TB; u++)
#pragma unroll
          for (int q = 0; q < KP / 16; q++)
            b[u][q] = *(const v4f*)(&Bs[buf][((jt0 + u) * 16 + l16) * KP + 16 * q + 4 * g]);
        v4f acc[TW][TB];
#pragma unroll
        for (int u = 0; u < TB; u++) {
          const float c = Bc[buf][(jt0 + u) * 16 + l16];
#pragma unroll
          for (int i = 0; i < TW; i++)
            acc[i][u] = v4f{ c, c, c, c };
        }
#pragma unroll
        for (int q = 0; q < KP / 16; q++)
#pragma unroll
          for (int kk = 0; kk < 4; kk++)
#pragma unroll
            for (int i = 0; i < TW; i++)
#pragma unroll
              for (int u = 0; u < TB; u++)
                acc[i][u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][q][kk], b[u][q][kk], acc[i][u], 0, 0, 0);
        bool any = false;
#pragma unroll
        for (int i = 0; i < TW; i++)
#pragma unroll
          for (int r = 0; r < 4; r++) {
            float m;
            if constexpr (TB == 4)
              m = max4(acc[i][0][r], acc[i][1][r], acc[i][2][r], acc[i][3][r]);
            else
              m = acc[i][0][r] > acc[i][1][r] ? acc[i][0][r] : acc[i][1][r];
            any = any | ((m >= fl[i][r]) & tok[i][r]);
          }
        if (__ballot(any)) {
          // lanes 16g .. 16g + 15 hold target row 4g + r of tile i against the 16 rows of each library tile:
          // one atomic add per (tile, row, group) reserves the slots of that target's passing pairs, then
          // each passing lane writes at the reserved base plus the passing pairs before it
          auto pass = [&](int i, int u, int r) {
            return (acc[i][u][r] >= fl[i][r]) & tok[i][r] & ((jt0 + u) * 16 + l16 < cn);
          };
          int base[TW][4];
#pragma unroll
          for (int i = 0; i < TW; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
              int cntg = 0;
#pragma unroll
              for (int u = 0; u < TB; u++)
                cntg += __builtin_popcount((unsigned)(__ballot(pass(i, u, r)) >> (16 * g)) & 0xffffu);
              base[i][r] = 0;
              if (l16 == 0 && cntg > 0)
                base[i][r] = atomicAdd(&ccnt[tid[w * 16 * TW + 16 * i + 4 * g + r]], cntg);
            }
#pragma unroll
          for (int i = 0; i < TW; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
              int off = __shfl(base[i][r], lane & ~15);
#pragma unroll
              for (int u = 0; u < TB; u++) {
                const unsigned gm = (unsigned)(__ballot(pass(i, u, r)) >> (16 * g)) & 0xffffu;
                const int slot = off + __builtin_popcount(gm & ((1u << l16) - 1u));
                if (((gm >> l16) & 1u) && slot < SCR_CAP)
                  cand[(size_t)tid[w * 16 * TW + 16 * i + 4 * g + r] * SCR_CAP + slot] =
                    K.pos[c0 + (jt0 + u) * 16 + l16];
                off += __builtin_popcount(gm);
              }
            }
        }
      }
      if (more)
        store(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
  }
}

// Phase two: one wavefront per (slice, target) whose candidates all fit.  The candidates are recomputed in
// FP64 in the reference's order (src/distances.cpp:186-216) and ordered by (sqrt distance, position), the
// exact list (src/edm.cpp:652-681): up to 64 candidates one per lane and a bitonic sort of the wavefront;
// more, up to four per lane, each lane's sorted and k rounds of a wavefront minimum.  The list is written
// in the layout of knn_mfma's per-slice lists (KCAP entries, +inf / INT_MAX after the k-th).  A target with a
// (slice, target) that does not fit is appended once (fflag) to the list for the streaming kernel (flist,
// fcnt).
template <int KCAP, int KP>
__global__ __launch_bounds__(256) void knn_rerank(MfmaParams P, int k, int slice_len, const int* __restrict__ cand,
                                                 const int* __restrict__ ccnt, double* __restrict__ out_s,
                                                 int* __restrict__ out_i, int* __restrict__ out_cnt,
                                                 int* __restrict__ fflag, int* __restrict__ flist,
                                                 int* __restrict__ fcnt)
{
  static_assert(SCR_CAP == 256, "four candidates per lane");
  const int lane = threadIdx.x & 63;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= P.NP)
    return;
  const size_t st = (size_t)blockIdx.y * P.NP + t;
  const int n = ccnt[st];
  if (n > SCR_CAP || !(P.seed[t] < INF_D)) {
    if (lane == 0 && atomicExch(&fflag[t], 1) == 0)
      flist[atomicAdd(fcnt, 1)] = t;
    return;
  }
  const int lo = blockIdx.y * slice_len;
  const int hi = min(P.NL, lo + slice_len);
  const double* y = P.Mp + (size_t)t * P.Ea;
  if (n <= 64) {
    // one candidate per lane: a bitonic sort of the wavefront by (sqrt distance, position), the positions
    // distinct (the empty entries, (inf, INT_MAX), alike), then lane o < k holds entry o of the list
    double ks = INF_D;
    int ki = INT_MAX, z = 0;
    if (lane < n) {
      const int pos = cand[st * SCR_CAP + lane];
      const double acc = dist_f64<KP>(P.M + (size_t)pos * P.Ea, y, P.Ea);
      if (acc == 0.0 || acc == MISSING_D) {
        z = 1;
      } else {
        ks = sqrt(acc);
        ki = pos;
      }
    }
#pragma unroll
    for (int size = 2; size <= 64; size <<= 1)
#pragma unroll
      for (int stride = size / 2; stride > 0; stride >>= 1) {
        const double os = __shfl_xor(ks, stride);
        const int oi = __shfl_xor(ki, stride);
        const bool keep_min = ((lane & stride) == 0) == ((lane & size) == 0 || size == 64);
        const bool take = keep_min ? key_less(os, oi, ks, ki) : key_less(ks, ki, os, oi);
        ks = take ? os : ks;
        ki = take ? oi : ki;
      }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
      z += __shfl_xor(z, off);
    const size_t base = st * KCAP;
    if (lane < KCAP) {
      out_s[base + lane] = lane < k ? ks : INF_D;
      out_i[base + lane] = lane < k ? ki : INT_MAX;
    }
    if (lane == 0)
      out_cnt[st] = (hi > lo ? hi - lo : 0) - z;
    return;
  }
  double s[4];
  int ix[4];
  int zeros = 0;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int e = lane + 64 * v;
    s[v] = INF_D;
    ix[v] = INT_MAX;
    if (e < n) {
      const int pos = cand[st * SCR_CAP + e];
      const double acc = dist_f64<KP>(P.M + (size_t)pos * P.Ea, y, P.Ea);
      if (acc == 0.0 || acc == MISSING_D) {
        zeros++;
      } else {
        s[v] = sqrt(acc);
        ix[v] = pos;
      }
    }
  }
  auto cs = [&](int x, int z) {
    const bool sw = key_less(s[z], ix[z], s[x], ix[x]);
    const double ts = s[x];
    const int ti = ix[x];
    s[x] = sw ? s[z] : s[x];
    ix[x] = sw ? ix[z] : ix[x];
    s[z] = sw ? ts : s[z];
    ix[z] = sw ? ti : ix[z];
  };
  cs(0, 1);
  cs(2, 3);
  cs(0, 2);
  cs(1, 3);
  cs(1, 2);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    zeros += __shfl_xor(zeros, off);
  const size_t base = st * KCAP;
  int h = 0;
  for (int o = 0; o < KCAP; o++) {
    double bs = INF_D;
    int bi = INT_MAX;
    if (o < k) {
      const double hs = h == 0 ? s[0] : h == 1 ? s[1] : h == 2 ? s[2] : h == 3 ? s[3] : INF_D;
      const int hi_ = h == 0 ? ix[0] : h == 1 ? ix[1] : h == 2 ? ix[2] : h == 3 ? ix[3] : INT_MAX;
      bs = hs;
      bi = hi_;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double os = __shfl_xor(bs, off);
        const int oi = __shfl_xor(bi, off);
        if (key_less(os, oi, bs, bi)) {
          bs = os;
          bi = oi;
        }
      }
      if (bi != INT_MAX && hi_ == bi)
        h++; // positions are distinct: only the lane that holds the minimum advances
    }
    if (lane == 0) {
      out_s[base + o] = bs;
      out_i[base + o] = bi;
    }
  }
  if (lane == 0)
    out_cnt[st] = (hi > lo ? hi - lo : 0) - zeros;
}

#define FEDM_SCREEN(KPv) template __global__ void knn_screen<KPv>(MfmaParams, KdScreen, int, int*, int*);
FEDM_SCREEN(16)
FEDM_SCREEN(32)
FEDM_SCREEN(64)
// The streaming kernel's lists of the listed targets over SF sub-slices of the library (fs, fi, fc: SF x NP
// lists and counts) merged into the per-slice lists of the main route: the whole list in slice 0, empty
// lists in the others (the merge of the slices then sees the same pairs).  One thread per listed target.
template <int KCAP>
__global__ __launch_bounds__(256) void merge_fallback(int NP, int S, int SF, const int* __restrict__ flist,
                                                     const int* __restrict__ fcnt, const double* __restrict__ fs,
                                                     const int* __restrict__ fi, const int* __restrict__ fc,
                                                     double* __restrict__ out_s, int* __restrict__ out_i,
                                                     int* __restrict__ out_cnt)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= *fcnt)
    return;
  const int t = flist[e];
  int head[MAX_SLICES];
  int cnt = 0;
  for (int f = 0; f < SF; f++) {
    head[f] = 0;
    cnt += fc[(size_t)f * NP + t];
  }
  for (int o = 0; o < KCAP; o++) {
    int best = -1;
    double bs = INF_D;
    int bi = INT_MAX;
    for (int f = 0; f < SF; f++) {
      if (head[f] >= KCAP)
        continue;
      const size_t b = ((size_t)f * NP + t) * KCAP + head[f];
      if (fi[b] != INT_MAX && key_less(fs[b], fi[b], bs, bi)) {
        bs = fs[b];
        bi = fi[b];
        best = f;
      }
    }
    if (best >= 0)
      head[best]++;
    out_s[(size_t)t * KCAP + o] = bs;
    out_i[(size_t)t * KCAP + o] = bi;
  }
  out_cnt[t] = cnt;
  for (int s = 1; s < S; s++) {
    for (int o = 0; o < KCAP; o++) {
      out_s[((size_t)s * NP + t) * KCAP + o] = INF_D;
      out_i[((size_t)s * NP + t) * KCAP + o] = INT_MAX;
    }
    out_cnt[(size_t)s * NP + t] = 0;
  }
}

#define FEDM_RERANK(KC, KPv)                                                                                     \
  template __global__ void knn_rerank<KC, KPv>(MfmaParams, int, int, const int*, const int*, double*, int*, int*, \
                                               int*, int*, int*);
#define FEDM_RERANK_K(KC) FEDM_RERANK(KC, 16) FEDM_RERANK(KC, 32) FEDM_RERANK(KC, 64)
FEDM_RERANK_K(4)
FEDM_RERANK_K(8)
FEDM_RERANK_K(16)
FEDM_RERANK_K(32)
FEDM_RERANK_K(64)
template __global__ void merge_fallback<4>(int, int, int, const int*, const int*, const double*, const int*, const int*,
                                           double*, int*, int*);
template __global__ void merge_fallback<8>(int, int, int, const int*, const int*, const double*, const int*, const int*,
                                           double*, int*, int*);
template __global__ void merge_fallback<16>(int, int, int, const int*, const int*, const double*, const int*,
                                            const int*, double*, int*, int*);
template __global__ void merge_fallback<32>(int, int, int, const int*, const int*, const double*, const int*,
                                            const int*, double*, int*, int*);
template __global__ void merge_fallback<64>(int, int, int, const int*, const int*, const double*, const int*,
                                            const int*, double*, int*, int*);

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

#define FEDM_MFMA(KC, KPv)                                                                                       \
  template __global__ void knn_mfma<KC, KPv>(MfmaParams, int, int, double*, int*, int*, const int*, const int*);
#define FEDM_MFMA_K(KC) FEDM_MFMA(KC, 16) FEDM_MFMA(KC, 32) FEDM_MFMA(KC, 64)
FEDM_MFMA_K(4)
FEDM_MFMA_K(8)
FEDM_MFMA_K(16)
FEDM_MFMA_K(32)
FEDM_MFMA_K(64)

} // namespace fedm
