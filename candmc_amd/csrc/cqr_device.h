// Device helpers shared by the CA-QR kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

namespace cqr {

// ---------------------------------------------------------------------------
// Wave-level primitives
// ---------------------------------------------------------------------------
// Index of the calling wave within its workgroup, as a wave-uniform value
// (lets the compiler keep per-wave conditions in scalar branches).
__device__ inline int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// Value of x held by lane `l` (l wave-uniform), through scalar registers.
__device__ inline double readlane_d(double x, int l) {
  int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
  int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
  return __hiloint2double(hi, lo);
}

// One DPP move of a double (lanes whose source is outside the pattern or
// whose row is masked off receive 0).
template <int CTRL, int ROWMASK>
__device__ inline double dpp_d(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, ROWMASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, ROWMASK, 0xf, false);
  return __hiloint2double(hi, lo);
}

// Wave-wide sum with DPP (quad_perm, row mirrors, row broadcasts): no LDS
// traffic; the total lands in lane 63 and is returned to every lane.
__device__ inline double wave_sum(double x) {
  x += dpp_d<0xB1, 0xf>(x);   // quad_perm [1,0,3,2]
  x += dpp_d<0x4E, 0xf>(x);   // quad_perm [2,3,0,1]
  x += dpp_d<0x141, 0xf>(x);  // row_half_mirror
  x += dpp_d<0x140, 0xf>(x);  // row_mirror
  x += dpp_d<0x142, 0xa>(x);  // row_bcast:15 into rows 1 and 3
  x += dpp_d<0x143, 0xc>(x);  // row_bcast:31 into rows 2 and 3
  return readlane_d(x, 63);
}

__device__ inline double wave_max(double x) {
  x = fmax(x, dpp_d<0xB1, 0xf>(x));
  x = fmax(x, dpp_d<0x4E, 0xf>(x));
  x = fmax(x, dpp_d<0x141, 0xf>(x));
  x = fmax(x, dpp_d<0x140, 0xf>(x));
  x = fmax(x, dpp_d<0x142, 0xa>(x));
  x = fmax(x, dpp_d<0x143, 0xc>(x));
  return readlane_d(x, 63);
}

// Reduce N per-lane partial values (N a power of two ≤ 64) across the 64
// lanes of a wave with log2(64) exchange steps in total: the first log2(N)
// steps halve the number of values each lane carries (each lane sends the
// half it does not keep), the rest are plain butterflies.  On return every
// lane holds the full sum of value index ((lane >> log2(64/N)) & (N-1)).
template <int C, int O, int N>
__device__ inline void xreduce_stage(double (&v)[N], int lane) {
  if constexpr (C > 1) {
    const bool hi = (lane & O) != 0;
#pragma unroll
    for (int t = 0; t < C / 2; ++t) {
      const double send = hi ? v[t] : v[t + C / 2];
      const double keep = hi ? v[t + C / 2] : v[t];
      v[t] = keep + __shfl_xor(send, O, 64);
    }
    xreduce_stage<C / 2, O / 2, N>(v, lane);
  }
}
template <int O>
__device__ inline double xreduce_tail(double s) {
  if constexpr (O > 0) {
    s += __shfl_xor(s, O, 64);
    return xreduce_tail<O / 2>(s);
  } else {
    return s;
  }
}
template <int N>
__device__ inline double xreduce(double (&v)[N], int lane) {
  static_assert(N >= 1 && N <= 64 && (N & (N - 1)) == 0, "N power of two");
  xreduce_stage<N, 32, N>(v, lane);
  return xreduce_tail<32 / N>(v[0]);
}
template <int N>
__host__ __device__ constexpr int xreduce_shift() {
  return N >= 64 ? 0 : N >= 32 ? 1 : N >= 16 ? 2 : N >= 8 ? 3 : N >= 4 ? 4 : N >= 2 ? 5 : 6;
}

// ---------------------------------------------------------------------------
// LAPACK dlarfg semantics on lane-uniform scalars.
//   Given alpha and xnorm = ‖x‖₂ (x the part below the pivot), returns beta,
//   tau and the multiplier applied to x, following LAPACK's dlarfg:
//   tau = 0 when xnorm == 0; beta = −sign(dlapy2(alpha, xnorm), alpha);
//   tau = (beta − alpha)/beta; x *= 1/(alpha − beta); the safe-minimum
//   rescaling loop is reproduced with exact power-of-two scalings (knt > 0
//   means x must first be multiplied by rsafmn knt times).
// ---------------------------------------------------------------------------
struct Reflector {
  double beta, tau, scal;
  int knt;
};

__device__ inline double dlapy2(double x, double y) {
  const double xa = fabs(x), ya = fabs(y);
  const double w = fmax(xa, ya), z = fmin(xa, ya);
  if (z == 0.0 || w > 1.7976931348623157e308) return w;
  const double q = z / w;
  return w * sqrt(1.0 + q * q);
}

__device__ inline Reflector dlarfg_scalar(double alpha, double xnorm) {
  Reflector r;
  r.knt = 0;
  if (xnorm == 0.0) {
    r.beta = alpha;
    r.tau = 0.0;
    r.scal = 1.0;
    return r;
  }
  // dlamch('S')/dlamch('E') = 2^-1022 / 2^-53 = 2^-969
  const double safmin = 0x1p-969, rsafmn = 0x1p969;
  double beta = -copysign(dlapy2(alpha, xnorm), alpha);
  if (fabs(beta) < safmin) {
    do {
      r.knt++;
      xnorm *= rsafmn;
      beta *= rsafmn;
      alpha *= rsafmn;
    } while (fabs(beta) < safmin && r.knt < 20);
    beta = -copysign(dlapy2(alpha, xnorm), alpha);
  }
  r.tau = (beta - alpha) / beta;
  r.scal = 1.0 / (alpha - beta);
  for (int j = 0; j < r.knt; ++j) beta *= safmin;
  r.beta = beta;
  return r;
}

// Robust norm finish: `ss` is the plain sum of squares; when it may have
// under/overflowed the caller recomputes with scaling (rare, uniform path).
__device__ inline bool ss_needs_rescale(double ss) {
  return !(ss > 0x1p-960 && ss < 0x1p960);
}

// ---------------------------------------------------------------------------
// In-LDS inverse of an upper triangular b×b matrix (column-major, ld),
// whole workgroup.  On entry the strictly upper part of M holds N and
// d[i] the reciprocal diagonal; the matrix inverted is diag(1/d) + N, and
// the recurrence used (X[p][i] = −d_i·Σ_{p≤q<i} X[p][q]·N[q][i]) also
// yields LAPACK's compact-WY T when d = tau and N = striu(VᵀV) — including
// tau = 0.  Diagonal 16×16 blocks are done by one wave each (rows in lanes,
// wave-synchronous, so in-place is safe); blocks are then merged pairwise
// with X12 = −X11·N12·X22.  The strictly lower part is used as scratch.
// ---------------------------------------------------------------------------
template <int NT>
__device__ void tri_inv_upper_lds(double* M, int ld, const double* d, int b) {
  const int tid = threadIdx.x, lane = tid & 63, wave = wave_id();
  constexpr int NW = NT / 64;
  const int nblk = (b + 15) / 16;
  for (int blk = wave; blk < nblk; blk += NW) {
    const int p = blk * 16 + lane;
    const int end = min(b, blk * 16 + 16);
    if (lane < 16 && p < b) M[p + p * ld] = d[p];
    // The column loop is uniform across the wave: in iteration i every lane
    // reads column i (rows p..i-1, still N) before any lane writes it.
    for (int i = blk * 16 + 1; i < end; ++i) {
      const bool act = lane < 16 && p < i;
      double acc = 0.0;
      if (act)
        for (int q = p; q < i; ++q) acc += M[p + q * ld] * M[q + i * ld];
      if (act) M[p + i * ld] = -d[i] * acc;
    }
  }
  __syncthreads();
  for (int h = 16, lh = 4; h < b; h *= 2, ++lh) {
    // pairs of blocks [a, a+h) and [a+h, a+2h)
    // step 1: Xs = N12·X22 stored in the strictly lower mirror block
    for (int e = tid; e < (b / (2 * h) + 1) * h * h; e += NT) {
      const int pair = e >> (2 * lh), r = e & (h * h - 1);
      const int p = r & (h - 1), j = r >> lh;
      const int a = pair * 2 * h;
      const int h2 = min(h, b - a - h);
      if (a + h >= b || j >= h2) continue;
      double acc = 0.0;
      for (int q = 0; q <= j; ++q) acc += M[(a + p) + (a + h + q) * ld] * M[(a + h + q) + (a + h + j) * ld];
      M[(a + h + j) + (a + p) * ld] = acc;  // lower mirror, (row a+h+j, col a+p)
    }
    __syncthreads();
    // step 2: X12 = −X11·Xs
    for (int e = tid; e < (b / (2 * h) + 1) * h * h; e += NT) {
      const int pair = e >> (2 * lh), r = e & (h * h - 1);
      const int p = r & (h - 1), j = r >> lh;
      const int a = pair * 2 * h;
      const int h2 = min(h, b - a - h);
      if (a + h >= b || j >= h2) continue;
      double acc = 0.0;
      for (int q = p; q < h; ++q) acc += M[(a + p) + (a + q) * ld] * M[(a + h + j) + (a + q) * ld];
      M[(a + p) + (a + h + j) * ld] = -acc;
    }
    __syncthreads();
  }
}

}  // namespace cqr
