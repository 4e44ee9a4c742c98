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

// Reduce N per-lane partial values (N = 1, 2, 4, 8 or 16) across the 64
// lanes of a wave.  The first log2(N) steps halve the number of values each
// lane carries: lanes pair up through a DPP pattern that flips one lane bit
// (row mirror: bit 3, half-row mirror: bit 2, quad permutes: bits 1 and 0)
// and each lane sends the half its partner keeps.  The remaining lane bits
// are summed with DPP row rotations within a row of 16 and two cross-row
// exchanges.  On return lane l holds the full sum of value index l & (N−1);
// lane k (k < N) is the designated holder of index k.
template <int CTRL, int HB, int C, int N>
__device__ inline void xr_halve(double (&v)[N], int lane) {
  const bool hi = ((lane >> HB) & 1) != 0;
#pragma unroll
  for (int t = 0; t < C / 2; ++t) {
    const double send = hi ? v[t] : v[t + C / 2];
    const double keep = hi ? v[t + C / 2] : v[t];
    v[t] = keep + dpp_d<CTRL, 0xf>(send);
  }
}
// Sum over the four rows of 16 lanes, lane by lane, with the gfx950 lane
// swaps (VALU only, no LDS round trip): swapping odd rows of one copy with
// even rows of another and adding the two results gives every lane the sum
// of its row pair; the 32-lane swap then adds the two halves.
__device__ inline double xr_swap_sum16(double s) {
  const unsigned lo = (unsigned)__double2loint(s), hi = (unsigned)__double2hiint(s);
  const auto l = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
  const auto h = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
  return __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
}
__device__ inline double xr_swap_sum32(double s) {
  const unsigned lo = (unsigned)__double2loint(s), hi = (unsigned)__double2hiint(s);
  const auto l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
  const auto h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  return __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
}
__device__ inline double xr_cross_rows(double s) { return xr_swap_sum32(xr_swap_sum16(s)); }
template <int N>
__device__ inline double xreduce(double (&v)[N], int lane) {
  static_assert(N == 1 || N == 2 || N == 4 || N == 8 || N == 16, "N in {1,2,4,8,16}");
  if constexpr (N == 16) xr_halve<0x140, 3, 16, N>(v, lane);  // row_mirror
  if constexpr (N >= 8) xr_halve<0x141, 2, (N >= 16 ? 8 : N), N>(v, lane);  // row_half_mirror
  if constexpr (N >= 4) xr_halve<0x4E, 1, (N >= 8 ? 4 : N), N>(v, lane);    // quad_perm [2,3,0,1]
  if constexpr (N >= 2) xr_halve<0xB1, 0, 2, N>(v, lane);                    // quad_perm [1,0,3,2]
  double s = v[0];
  if constexpr (N == 1) s += dpp_d<0xB1, 0xf>(s);
  if constexpr (N <= 2) s += dpp_d<0x4E, 0xf>(s);
  if constexpr (N <= 4) s += dpp_d<0x124, 0xf>(s);  // row_ror:4
  if constexpr (N <= 8) s += dpp_d<0x128, 0xf>(s);  // row_ror:8
  return xr_cross_rows(s);
}

// ---------------------------------------------------------------------------
// Hand-off of data between workgroups of one launch without cache-wide
// fences.  Every value that crosses workgroups is written and read with
// device-scope (agent) relaxed atomics, which the hardware keeps coherent
// across the XCDs' L2 caches (write-through on store, no stale hit on
// load).  A producer stores its data, waits for all of its memory
// operations to complete (s_waitcnt 0 in every wave, then a barrier) and
// only then stores the flag; a consumer polls the flag, meets its
// workgroup at a barrier and then loads the data.  This replaces the
// release/acquire fences, which write back and invalidate whole L2 caches.
// ---------------------------------------------------------------------------
__device__ inline double load_coherent(const double* p) {
  const uint64_t u = __hip_atomic_load(reinterpret_cast<const uint64_t*>(p), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double((long long)u);
}
__device__ inline void store_coherent(double* p, double v) {
  __hip_atomic_store(reinterpret_cast<uint64_t*>(p), (uint64_t)__double_as_longlong(v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint64_t flag_load(const uint64_t* f) {
  return __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void flag_store(uint64_t* f, uint64_t v) {
  __hip_atomic_store(f, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// All of this wave's outstanding memory operations have completed.
__device__ inline void wait_memory() { __builtin_amdgcn_s_waitcnt(0); }
// Bounded wait (by one thread) for *f ≥ want, at most `limit` polls; false
// on timeout.
__device__ inline bool flag_wait(const uint64_t* f, uint64_t want, long limit) {
  long spins = 0;
  while (flag_load(f) < want) {
    __builtin_amdgcn_s_sleep(2);
    if (++spins > limit) return false;
  }
  return true;
}
// The poll bound of a launch: err[1] when the host set one (a diagnostic
// override that lets a test drive the timeout path), else 2²² polls.
__device__ inline long wait_limit(const int* err) {
  const int v = err ? err[1] : 0;
  return v > 0 ? (long)v : (1L << 22);
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

// 1/x from the hardware reciprocal refined by two Newton steps (within an
// ulp or two of the correctly rounded quotient; x normal and finite).
__device__ inline double recip_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// dlarfg when no rescaling can be needed (ss = ‖x‖² in (2^-960, 2^960),
// |alpha| < 2^480): |beta| = sqrt(alpha² + ss) ≥ 2^-480, so knt = 0; the
// two quotients use recip_nr.  Returns false (h untouched) otherwise.
__device__ inline bool dlarfg_fast(double alpha, double ss, Reflector& h) {
  if (!(ss > 0x1p-960 && ss < 0x1p960) || !(fabs(alpha) < 0x1p480)) return false;
  const double beta = -copysign(sqrt(fma(alpha, alpha, ss)), alpha);
  h.beta = beta;
  h.tau = (beta - alpha) * recip_nr(beta);
  h.scal = recip_nr(alpha - beta);
  h.knt = 0;
  return true;
}

// Robust norm finish: `ss` is the plain sum of squares; when it may have
// under/overflowed the caller recomputes with scaling (rare, uniform path).
__device__ inline bool ss_needs_rescale(double ss) {
  return !(ss > 0x1p-960 && ss < 0x1p960);
}

// ---------------------------------------------------------------------------
// One 16×16 tile of P = X·Y on FP64 MFMA (v_mfma_f64_16x16x4f64), whole
// wave.  x(m, k) and y(k, n) return the operands (0 outside the product);
// k runs over [k0, k1) in steps of 4.  Lane l, register v of the result
// holds P[l & 15][(l >> 4) + 4v].
// ---------------------------------------------------------------------------
typedef double dbl4 __attribute__((ext_vector_type(4)));

template <class FX, class FY>
__device__ inline dbl4 mfma_tile16(const FX& x, const FY& y, int k0, int k1, int lane) {
  dbl4 acc = {0.0, 0.0, 0.0, 0.0};
  const int r = lane & 15, q = lane >> 4;
  for (int k = k0; k < k1; k += 4) {
    const double xa = x(r, k + q);  // X[m = r][k]
    const double yb = y(k + q, r);  // Y[k][n = r]
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(yb, xa, acc, 0, 0, 0);
  }
  return acc;
}

// Same tile with the operands of 16 k steps fetched before their MFMAs, so
// that global-memory latency is paid once per 64 values of k, not per step.
template <class FX, class FY>
__device__ inline dbl4 mfma_tile16_k128(const FX& x, const FY& y, int k0, int k1, int lane) {
  const int r = lane & 15, q = lane >> 4;
  dbl4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int kc = k0; kc < k1; kc += 64) {   // chunks of 16 k steps, loaded up front
    double xa[16], yb[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int k = kc + 4 * s;
      const bool in = k < k1;
      xa[s] = in ? x(r, k + q) : 0.0;
      yb[s] = in ? y(k + q, r) : 0.0;
    }
#pragma unroll
    for (int s = 0; s < 16; ++s)
      if (kc + 4 * s < k1) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(yb[s], xa[s], acc, 0, 0, 0);
  }
  return acc;
}

// Same tile with the k steps dealt round-robin to four accumulators, so
// consecutive MFMAs are independent and issue back to back instead of each
// waiting for the previous result (the chains are summed at the end).  The
// accessors are called for every k of a 64-wide chunk, also k ≥ k1: they
// must stay in bounds there and return 0 (branch-free accessors let the
// 32 loads of a chunk issue together).
template <class FX, class FY>
__device__ inline dbl4 mfma_tile16_acc4(const FX& x, const FY& y, int k0, int k1, int lane) {
  const int r = lane & 15, q = lane >> 4;
  dbl4 acc[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) acc[a] = dbl4{0.0, 0.0, 0.0, 0.0};
  for (int kc = k0; kc < k1; kc += 64) {
    double xa[16], yb[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int k = kc + 4 * s;
      xa[s] = x(r, k + q);
      yb[s] = y(k + q, r);
    }
#pragma unroll
    for (int s = 0; s < 16; ++s)
      if (kc + 4 * s < k1) acc[s & 3] = __builtin_amdgcn_mfma_f64_16x16x4f64(yb[s], xa[s], acc[s & 3], 0, 0, 0);
  }
  return (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

// ---------------------------------------------------------------------------
// In-LDS inverse of an upper triangular b×b matrix (column-major, ld),
// whole workgroup.  On entry the strictly upper part of M holds N and
// d[i] the reciprocal diagonal; the matrix inverted is diag(1/d) + N, and
// the recurrence used (X[p][i] = −d_i·Σ_{p≤q<i} X[p][q]·N[q][i]) also
// yields LAPACK's compact-WY T when d = tau and N = striu(VᵀV) — including
// tau = 0.  Diagonal 16×16 blocks are done by one wave each (rows in lanes,
// wave-synchronous, so in-place is safe); blocks are then merged pairwise,
// X12 = −X11·(N12·X22), both products as 16×16 MFMA tiles spread over the
// waves.  The intermediate N12·X22 is kept in the strictly lower mirror
// block (rows of the second block, columns of the first), which no earlier
// merge touched; the strictly lower part may hold such scratch, so every
// read of a triangle masks it.
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
  for (int h = 16; h < b; h *= 2) {
    const int npairs = (b - h + 2 * h - 1) / (2 * h);   // pairs with a second block
    const int tm = h / 16, tn = (h + 15) / 16;
    const int ntiles = npairs * tm * tn;
    // step 1: P = N12·X22 → lower mirror M[a+h+j][a+p]
    for (int t = wave; t < ntiles; t += NW) {
      const int pr = t / (tm * tn), ti = (t / tn) % tm, tj = t % tn;
      const int a = pr * 2 * h, h2 = min(h, b - a - h);
      if (16 * tj >= h2) continue;
      const double* N12 = M + a + (a + h) * ld;        // (p, q) p < h, q < h2
      const double* X22 = M + (a + h) + (a + h) * ld;  // (q, j) q ≤ j < h2
      const int i0 = 16 * ti, j0 = 16 * tj;
      auto x = [&](int r, int k) { return k < h2 ? N12[(i0 + r) + k * ld] : 0.0; };
      auto y = [&](int k, int c) {
        const int j = j0 + c;
        return (k <= j && j < h2) ? X22[k + j * ld] : 0.0;
      };
      const dbl4 acc = mfma_tile16(x, y, 0, min(h2, j0 + 16), lane);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int pp = i0 + (lane & 15), j = j0 + (lane >> 4) + 4 * v;
        if (j < h2) M[(a + h + j) + (a + pp) * ld] = acc[v];
      }
    }
    __syncthreads();
    // step 2: X12 = −X11·P → M[a+p][a+h+j]
    for (int t = wave; t < ntiles; t += NW) {
      const int pr = t / (tm * tn), ti = (t / tn) % tm, tj = t % tn;
      const int a = pr * 2 * h, h2 = min(h, b - a - h);
      if (16 * tj >= h2) continue;
      const double* X11 = M + a + a * ld;              // (p, q) p ≤ q < h
      const int i0 = 16 * ti, j0 = 16 * tj;
      auto x = [&](int r, int k) {
        const int pp = i0 + r;
        return (pp <= k && k < h) ? X11[pp + k * ld] : 0.0;
      };
      auto y = [&](int k, int c) {
        const int j = j0 + c;
        return (k < h && j < h2) ? M[(a + h + j) + (a + k) * ld] : 0.0;
      };
      const dbl4 acc = mfma_tile16(x, y, i0, h, lane);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int pp = i0 + (lane & 15), j = j0 + (lane >> 4) + 4 * v;
        if (j < h2) M[(a + pp) + (a + h + j) * ld] = -acc[v];
      }
    }
    __syncthreads();
  }
}

}  // namespace cqr
