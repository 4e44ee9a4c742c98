// Leaf QR body shared by the leaf kernel and the fused leaf + tree kernel.
#pragma once

#include "cqr_device.h"
#include "cqr_internal.h"

namespace cqr {

__device__ __forceinline__ int64_t leaf_off(int64_t l, int64_t m, int64_t L) { return (l * m) / L; }

// ---------------------------------------------------------------------------
// K1: leaf QR.  One workgroup per leaf; the leaf lives in registers: wave w
// owns columns w, w+WAVES, ... (CPW of them), lane l owns rows l, l+64, ...
// (RPL of them).  Per column step the owner wave forms the reflector
// (FP64 DPP wave reduction for the norm, dlarfg semantics), publishes v
// through LDS, and one barrier later every wave applies it to its columns.
// The dot products against the already-finished columns are the Gram
// entries V(:,j)ᵀv_i that give the compact-WY T at the end.
// ---------------------------------------------------------------------------
//
// With `lcnt` set, the leaf also publishes its R while it runs: column i of
// R (rows ≤ i) is final once its reflector exists, so the owner wave stores
// it at once with device-coherent stores; after every eighth column all
// waves wait for their stores and lcnt[leaf] = epoch·256 + (columns done)
// tells a consumer in the same launch that those columns may be read.
template <int WAVES, int CPW, int RPL>
__device__ __forceinline__ void leaf_qr_body(double* __restrict__ A, int64_t lda, int64_t m, int L, int b,
                                    double* __restrict__ Rout, double* __restrict__ Tout,
                                    double* __restrict__ tauout, const BatchProb* __restrict__ bp, int nprob,
                                    int gleaf, double* smem, uint64_t* __restrict__ lcnt, uint64_t epoch) {
  constexpr int NT = 64 * WAVES;
  int leaf = gleaf;
  if (bp) {                      // batched: A is the group's base, the problem's block at aoff
    const BatchProb P = bp[bp_find(bp, nprob, gleaf, BP_LEAF)];
    A += P.aoff;
    lda = P.m;
    m = P.m;
    L = P.L;
    leaf = gleaf - P.leaf0;
  }
  const int64_t r0 = leaf_off(leaf, m, L);
  const int r = (int)(leaf_off(leaf + 1, m, L) - r0);
  double* Al = A + r0;
  double* Rl = Rout + (int64_t)gleaf * b * b;
  double* Tl = Tout + (int64_t)gleaf * b * b;
  double* taul = tauout + (int64_t)gleaf * b;

  const int tid = threadIdx.x, lane = tid & 63, w = wave_id();
  double* G = smem;                       // b×b Gram, then T
  double* vbuf = G + b * b;               // 2 × 64·RPL
  double* stau = vbuf + 2 * 64 * RPL;     // b

  double a[CPW][RPL];
#pragma unroll
  for (int cl = 0; cl < CPW; ++cl) {
    const int col = w + WAVES * cl;
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      const int row = lane + 64 * s;
      a[cl][s] = (col < b && row < r) ? Al[row + col * lda] : 0.0;
    }
  }
  for (int cc = w; cc < b; cc += WAVES)
    for (int rr = lane; rr < b; rr += 64) G[rr + cc * b] = 0.0;
  __syncthreads();

  for (int i = 0; i < b; ++i) {
    const int ow = i % WAVES, ocl = i / WAVES;
    const int plane = i & 63, pslot = i >> 6;
    double* vb = vbuf + (i & 1) * (64 * RPL);
    if (w == ow) {
#pragma unroll
      for (int cl = 0; cl < CPW; ++cl) {
        if (cl != ocl) continue;
        double ss = 0.0, piv = 0.0;
#pragma unroll
        for (int s = 0; s < RPL; ++s) {
          const int row = lane + 64 * s;
          const double x = a[cl][s];
          if (row > i) ss += x * x;
          if (s == pslot) piv = x;
        }
        ss = wave_sum(ss);
        const double alpha = readlane_d(piv, plane);
        Reflector h;
        if (!dlarfg_fast(alpha, ss, h)) {
        double xnorm;
        if (ss_needs_rescale(ss)) {
          double mx = 0.0;
#pragma unroll
          for (int s = 0; s < RPL; ++s)
            if (lane + 64 * s > i) mx = fmax(mx, fabs(a[cl][s]));
          mx = wave_max(mx);
          if (mx == 0.0) {
            xnorm = 0.0;
          } else {
            double s2 = 0.0;
#pragma unroll
            for (int s = 0; s < RPL; ++s)
              if (lane + 64 * s > i) { const double t = a[cl][s] / mx; s2 += t * t; }
            xnorm = mx * sqrt(wave_sum(s2));
          }
        } else {
          xnorm = sqrt(ss);
        }
        h = dlarfg_scalar(alpha, xnorm);
        }
        if (h.tau != 0.0) {
#pragma unroll
          for (int s = 0; s < RPL; ++s)
            if (lane + 64 * s > i) {
              double x = a[cl][s];
              for (int k = 0; k < h.knt; ++k) x *= 0x1p969;
              a[cl][s] = x * h.scal;
            }
        }
#pragma unroll
        for (int s = 0; s < RPL; ++s) {
          const int row = lane + 64 * s;
          if (row == i) a[cl][s] = h.beta;
          vb[row] = row > i ? a[cl][s] : (row == i ? 1.0 : 0.0);
        }
        if (lane == 0) stau[i] = h.tau;
        if (lcnt) {
#pragma unroll
          for (int s = 0; s < RPL; ++s) {
            const int row = lane + 64 * s;
            if (row <= i) store_coherent(Rl + row + (int64_t)i * b, a[cl][s]);
          }
        }
      }
    }
    // before a block of eight columns is published, every owner's R stores
    // have completed
    if (lcnt && ((i & 7) == 7 || i + 1 == b)) wait_memory();
    __syncthreads();
    if (lcnt && tid == 0 && ((i & 7) == 7 || i + 1 == b)) flag_store(lcnt + gleaf, epoch * 256 + (uint64_t)(i + 1));
    const double tau = stau[i];
    double v[RPL];
#pragma unroll
    for (int s = 0; s < RPL; ++s) v[s] = vb[lane + 64 * s];
    double part[CPW];
#pragma unroll
    for (int cl = 0; cl < CPW; ++cl) {
      double acc = 0.0;
#pragma unroll
      for (int s = 0; s < RPL; ++s) acc += v[s] * a[cl][s];
      part[cl] = acc;
    }
    const double mine = xreduce<CPW>(part, lane);
    {
      // Gram entries of finished columns: one lane per reduced value
      const int col = w + WAVES * (lane & (CPW - 1));
      if (lane < CPW && col < i) G[col + i * b] = mine;
    }
#pragma unroll
    for (int cl = 0; cl < CPW; ++cl) {
      const int col = w + WAVES * cl;
      if (col > i && col < b) {
        const double f = tau * readlane_d(mine, cl);
#pragma unroll
        for (int s = 0; s < RPL; ++s) a[cl][s] -= f * v[s];
      }
    }
  }
  __syncthreads();
  tri_inv_upper_lds<NT>(G, b, stau, b);

#pragma unroll
  for (int cl = 0; cl < CPW; ++cl) {
    const int col = w + WAVES * cl;
    if (col >= b) continue;
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      const int row = lane + 64 * s;
      if (row < r) Al[row + col * lda] = a[cl][s];
      if (row < b) Rl[row + col * b] = row <= col ? a[cl][s] : 0.0;
    }
  }
  for (int cc = w; cc < b; cc += WAVES)
    for (int rr = lane; rr < b; rr += 64) Tl[rr + cc * b] = rr <= cc ? G[rr + cc * b] : 0.0;
  for (int p = tid; p < b; p += NT) taul[p] = stau[p];
}

}  // namespace cqr
