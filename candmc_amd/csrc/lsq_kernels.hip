// Kernels of the least-squares path: thin products with a few right-hand
// sides (A·X and Aᵀ·X with X of at most 16 columns, FP64, bandwidth-bound
// on A), element-wise updates and column norms.
#include "cqr_device.h"
#include "cqr_internal.h"

namespace cqr {

namespace {
constexpr int TN_ROWS = 64;   // k-slab of X staged in LDS (A·X)
constexpr int TT_ROWS = 32;   // row slab of A and X staged in LDS (Aᵀ·X)
}  // namespace

// C (m×nr) = alpha·A·X + beta·C, A m×k, X k×nr (nr ≤ NR); one row per thread.
// With P, workgroup row y covers columns [y·kchunk, (y+1)·kchunk) and stores
// its raw sums at P + (y·nr + j)·m (thin_n_reduce_kernel finishes C).
template <int NR>
__global__ void __launch_bounds__(256)
thin_gemm_n_kernel(const double* __restrict__ A, int64_t lda, int64_t m, int64_t k,
                   const double* __restrict__ X, int64_t ldx, int nr, double alpha, double beta,
                   double* __restrict__ C, int64_t ldc, int unit_top, int64_t kchunk, double* __restrict__ P) {
  __shared__ double Xs[TN_ROWS][NR];
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t kb = (int64_t)blockIdx.y * kchunk, ke = min<int64_t>(k, kb + kchunk);
  double acc[NR];
#pragma unroll
  for (int j = 0; j < NR; ++j) acc[j] = 0.0;
  for (int64_t k0 = kb; k0 < ke; k0 += TN_ROWS) {
    __syncthreads();
    for (int e = threadIdx.x; e < TN_ROWS * NR; e += 256) {
      const int kk = e % TN_ROWS, j = e / TN_ROWS;
      Xs[kk][j] = (k0 + kk < ke && j < nr) ? X[(k0 + kk) + (int64_t)j * ldx] : 0.0;
    }
    __syncthreads();
    if (r < unit_top) {        // unit-lower top rows: A(r, c) = 0 for c > r, 1 at c = r
      const int kend = (int)min<int64_t>(TN_ROWS, ke - k0);
      for (int kk = 0; kk < kend; ++kk) {
        const int64_t cidx = k0 + kk;
        const double a = cidx < r ? A[r + cidx * lda] : (cidx == r ? 1.0 : 0.0);
#pragma unroll
        for (int j = 0; j < NR; ++j) acc[j] += a * Xs[kk][j];
      }
    } else if (r < m) {
      const int kend = (int)min<int64_t>(TN_ROWS, ke - k0);
      const double* ap = A + r + k0 * lda;
      int kk = 0;
      for (; kk + 4 <= kend; kk += 4) {
        const double a0 = ap[(int64_t)kk * lda], a1 = ap[(int64_t)(kk + 1) * lda];
        const double a2 = ap[(int64_t)(kk + 2) * lda], a3 = ap[(int64_t)(kk + 3) * lda];
#pragma unroll
        for (int j = 0; j < NR; ++j)
          acc[j] += a0 * Xs[kk][j] + a1 * Xs[kk + 1][j] + a2 * Xs[kk + 2][j] + a3 * Xs[kk + 3][j];
      }
      for (; kk < kend; ++kk) {
        const double a = ap[(int64_t)kk * lda];
#pragma unroll
        for (int j = 0; j < NR; ++j) acc[j] += a * Xs[kk][j];
      }
    }
  }
  if (r < m) {
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      if (j >= nr) continue;
      if (P) {
        P[((int64_t)blockIdx.y * nr + j) * m + r] = acc[j];
      } else {
        double* cp = C + r + (int64_t)j * ldc;
        *cp = alpha * acc[j] + (beta != 0.0 ? beta * *cp : 0.0);
      }
    }
  }
}

// C = alpha·Σ_y P_y + beta·C for the column-split A·X (sum in y order).
__global__ void __launch_bounds__(256) thin_n_reduce_kernel(const double* __restrict__ P, int ny, int64_t m, int nr,
                                                            double alpha, double beta, double* __restrict__ C,
                                                            int64_t ldc) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= m * nr) return;
  const int64_t r = e % m, j = e / m;
  double sum = 0.0;
  for (int y = 0; y < ny; ++y) sum += P[((int64_t)y * nr + j) * m + r];
  double* cp = C + r + j * ldc;
  *cp = alpha * sum + (beta != 0.0 ? beta * *cp : 0.0);
}

// Partial P_z (k×nr, ld k) = A(rows of slab z, :)ᵀ·X(rows of slab z, :) for
// a block of 128 columns of A; thread (column c = tid & 127, half h) owns
// NR/2 right-hand sides.  Slabs are reduced afterwards in a fixed order.
template <int NR>
__global__ void __launch_bounds__(256)
thin_gemm_t_kernel(const double* __restrict__ A, int64_t lda, int64_t m, int64_t k,
                   const double* __restrict__ X, int64_t ldx, int nr, int64_t slab,
                   double* __restrict__ P, int unit_top) {
  constexpr int NH = NR > 1 ? NR / 2 : 1;
  __shared__ double As[TT_ROWS][128 + 1];
  __shared__ double Xs[TT_ROWS][NR];
  const int tid = threadIdx.x, c = tid & 127, h = tid >> 7;
  const int64_t c0 = (int64_t)blockIdx.x * 128, r0 = (int64_t)blockIdx.y * slab;
  const int64_t r1 = min<int64_t>(m, r0 + slab);
  double acc[NH];
#pragma unroll
  for (int j = 0; j < NH; ++j) acc[j] = 0.0;
  for (int64_t rb = r0; rb < r1; rb += TT_ROWS) {
    __syncthreads();
    {  // the slab's 16 loads per thread are issued before any is stored
      constexpr int PER = TT_ROWS * 128 / 256;
      double v[PER];
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int e = tid + 256 * u, rr = e % TT_ROWS, cc = e / TT_ROWS;
        const int64_t row = rb + rr, col = c0 + cc;
        const bool ok = row < r1 && col < k;
        v[u] = A[ok ? row + col * lda : 0];
      }
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int e = tid + 256 * u, rr = e % TT_ROWS, cc = e / TT_ROWS;
        const int64_t row = rb + rr, col = c0 + cc;
        double x = (row < r1 && col < k) ? v[u] : 0.0;
        if (row < unit_top) x = row < col ? 0.0 : (row == col ? 1.0 : x);
        As[rr][cc] = x;
      }
    }
    for (int e = tid; e < TT_ROWS * NR; e += 256) {
      const int rr = e % TT_ROWS, j = e / TT_ROWS;
      const int64_t row = rb + rr;
      Xs[rr][j] = (row < r1 && j < nr) ? X[row + (int64_t)j * ldx] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int rr = 0; rr < TT_ROWS; ++rr) {
      const double a = As[rr][c];
#pragma unroll
      for (int j = 0; j < NH; ++j) acc[j] += a * Xs[rr][(NR > 1 ? h * NH : 0) + j];
    }
  }
  const int64_t col = c0 + c;
  if (col < k && (NR > 1 || h == 0)) {
    double* pz = P + (int64_t)blockIdx.y * k * nr;
#pragma unroll
    for (int j = 0; j < NH; ++j) {
      const int jj = (NR > 1 ? h * NH : 0) + j;
      if (jj < nr) pz[col + (int64_t)jj * k] = acc[j];
    }
  }
}

// C = alpha·Σ_z P_z + beta·C (fixed order): 64 outputs per workgroup,
// sixteen groups of slabs per output, then the sixteen group sums in order.
__global__ void __launch_bounds__(1024) slab_reduce_kernel(const double* __restrict__ P, int nslab, int64_t k, int nr,
                                                           double alpha, double beta, double* __restrict__ C,
                                                           int64_t ldc) {
  __shared__ double part[16][64];
  const int t = threadIdx.x, el = t & 63, g = t >> 6;
  const int64_t kn = k * nr, e = (int64_t)blockIdx.x * 64 + el;
  double s = 0.0;
  if (e < kn)
    for (int z = g; z < nslab; z += 16) s += P[(int64_t)z * kn + e];
  part[g][el] = s;
  __syncthreads();
  if (g == 0 && e < kn) {
    double tot = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) tot += part[q][el];
    double* cp = C + (e % k) + (e / k) * ldc;
    *cp = alpha * tot + (beta != 0.0 ? beta * *cp : 0.0);
  }
}

// Z' = op(T)·Σ_z P_z in one workgroup, for the few-column apply of Q or Qᵀ
// (k ≤ 128 rows, nr ≤ 4 columns): the slab partials of Z = Yᵀ·B are summed
// in a fixed order (G groups of slabs, each in slab order, then the group
// sums in order), then Z'(i, j) = Σ_l op(Tt)(i, l)·Z(l, j) with Tt's lower
// triangle (op the transpose when ta) staged in LDS, summed in l order over
// two halves.
__global__ void __launch_bounds__(1024) slab_reduce_t_kernel(const double* __restrict__ P, int nslab, int k, int nr,
                                                             const double* __restrict__ Tt, int64_t ldt, int ta,
                                                             double* __restrict__ Zp, int64_t ldz) {
  extern __shared__ double sm[];
  double* ts = sm;             // op(Tt), k×k, ld k
  double* gs = ts + k * k;     // group sums, G·kn ≤ 1024
  double* zs = gs + 1024;      // Z, kn ≤ 512
  double* part = zs + 512;     // two halves of the l range, 2 × 512
  const int t = threadIdx.x, kn = k * nr, G = 1024 / kn;
  for (int e = t; e < k * k; e += 1024) {   // Tt in storage order (coalesced)
    const int row = e % k, col = e / k;
    const double v = row >= col ? Tt[row + (int64_t)col * ldt] : 0.0;
    if (ta) ts[col + row * k] = v;
    else ts[row + col * k] = v;
  }
  if (t < G * kn) {
    const int g = t / kn, e = t - g * kn;
    double sum = 0.0;
    for (int z0 = g; z0 < nslab; z0 += 16 * G) {
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int z = z0 + u * G;
        v[u] = P[(int64_t)(z < nslab ? z : 0) * kn + e];
      }
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (z0 + u * G < nslab) sum += v[u];
    }
    gs[g * kn + e] = sum;
  }
  __syncthreads();
  if (t < kn) {
    double tot = 0.0;
    for (int g = 0; g < G; ++g) tot += gs[g * kn + t];
    zs[t] = tot;
  }
  __syncthreads();
  const int h = t >> 9, o = t & 511;
  double acc = 0.0;
  if (o < kn) {
    const int i = o % k, j = o / k, l1 = min(k, h * 64 + 64);
    for (int l = h * 64; l < l1; ++l)
      if (ta ? l >= i : l <= i) acc += ts[i + l * k] * zs[l + j * k];
  }
  part[h * 512 + o] = acc;
  __syncthreads();
  if (t < kn) Zp[(t % k) + (int64_t)(t / k) * ldz] = part[t] + part[512 + t];
}

// One step of the blocked solve with R (upper triangle of a [Y\R] block, ld
// ldr) on nr ≤ NR columns of B, in 128-row blocks.  Workgroup q of [q0,
// q0 + grid) updates its block with block p of the solution,
//   trans = 0 (R⁻¹, blocks backward): y_q −= R(q, p)·x_p
//   trans = 1 (R⁻ᵀ, blocks forward):  y_q −= R(p, q)ᵀ·x_p,
// and the workgroup of block `solve` then finishes that block:
// x = op(Ri)·y with Ri its diagonal block's inverse (upper, ld its width).
template <int NR>
__global__ void __launch_bounds__(256)
trsm_step_kernel(const double* __restrict__ R, int64_t ldr, int64_t n, const double* __restrict__ Rinv,
                 double* __restrict__ B, int64_t ldb, int nr, int trans, int p, int q0, int solve) {
  constexpr int KC = 32;
  __shared__ double xs[128][NR + 1];
  __shared__ double ys[128][NR + 1];
  __shared__ double part[2][128][NR + 1];
  const int q = q0 + blockIdx.x, t = threadIdx.x;
  const int r = t & 127, h = t >> 7;   // row of the block, half of the k range
  const int64_t jp = (int64_t)p * 128, jq = (int64_t)q * 128;
  const int bp = (int)min<int64_t>(128, n - jp), bq = (int)min<int64_t>(128, n - jq);
  for (int e = t; e < 128 * NR; e += 256) {
    const int rr = e & 127, j = e >> 7;
    xs[rr][j] = (rr < bp && j < nr) ? B[jp + rr + (int64_t)j * ldb] : 0.0;
  }
  __syncthreads();
  double acc[NR];
#pragma unroll
  for (int j = 0; j < NR; ++j) acc[j] = 0.0;
  if (r < bq) {
    const int kb = h * 64, ke = min(bp, kb + 64);
    for (int k0 = kb; k0 < ke; k0 += KC) {
      double a[KC];
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = k0 + u;
        const int64_t idx = trans ? (jp + k) + (jq + r) * ldr : (jq + r) + (jp + k) * ldr;
        a[u] = R[k < ke ? idx : 0];
      }
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = k0 + u;
        if (k < ke) {
#pragma unroll
          for (int j = 0; j < NR; ++j) acc[j] += a[u] * xs[k][j];
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < NR; ++j) part[h][r][j] = acc[j];
  __syncthreads();
  if (h == 0 && r < bq) {
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      if (j >= nr) continue;
      double* yp = B + jq + r + (int64_t)j * ldb;
      const double y = *yp - (part[0][r][j] + part[1][r][j]);
      if (q == solve) ys[r][j] = y;
      else *yp = y;
    }
  }
  if (q != solve) return;                  // uniform across the workgroup
  __syncthreads();
  const double* Ri = Rinv + (int64_t)q * 128 * 128;
#pragma unroll
  for (int j = 0; j < NR; ++j) acc[j] = 0.0;
  if (r < bq) {
    const int kb = h * 64, ke = min(bq, kb + 64);
    for (int k0 = kb; k0 < ke; k0 += KC) {
      // op(Ri)(r, k): Ri upper (r ≤ k), or Riᵀ (k ≤ r); a chunk's loads first
      double a[KC];
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = k0 + u;
        const bool ok = k < ke && (trans ? k <= r : k >= r);
        a[u] = Ri[ok ? (trans ? k + (int64_t)r * bq : r + (int64_t)k * bq) : 0];
      }
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int k = k0 + u;
        const bool ok = k < ke && (trans ? k <= r : k >= r);
        if (ok) {
#pragma unroll
          for (int j = 0; j < NR; ++j) acc[j] += a[u] * ys[k][j];
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < NR; ++j) part[h][r][j] = acc[j];
  __syncthreads();
  if (h == 0 && r < bq) {
#pragma unroll
    for (int j = 0; j < NR; ++j)
      if (j < nr) B[jq + r + (int64_t)j * ldb] = part[0][r][j] + part[1][r][j];
  }
}

int launch_trsm_step(cqr_ctx* c, const double* R, int64_t ldr, int64_t n, const double* Rinv, double* B, int64_t ldb,
                     int nr, int trans, int p, int q0, int nq, int solve) {
  if (nq <= 0) return CQR_OK;
  if (nr < 1 || nr > 16) return set_err(c, CQR_ERR_UNSUPPORTED, "trsm step with %d columns", nr);
  if (nr == 1)
    hipLaunchKernelGGL(trsm_step_kernel<1>, dim3(nq), dim3(256), 0, c->stream, R, ldr, n, Rinv, B, ldb, nr, trans, p,
                       q0, solve);
  else if (nr <= 4)
    hipLaunchKernelGGL(trsm_step_kernel<4>, dim3(nq), dim3(256), 0, c->stream, R, ldr, n, Rinv, B, ldb, nr, trans, p,
                       q0, solve);
  else
    hipLaunchKernelGGL(trsm_step_kernel<16>, dim3(nq), dim3(256), 0, c->stream, R, ldr, n, Rinv, B, ldb, nr, trans,
                       p, q0, solve);
  return check_launch(c, "trsm_step");
}

// The whole blocked triangular solve B(0:n, :) := R⁻¹·B (trans = 0) or
// R⁻ᵀ·B (trans = 1) in one launch, one workgroup per 128-row block q.  A
// block waits for the blocks it depends on (q + 1.. for R⁻¹, ..q − 1 for
// R⁻ᵀ), one at a time in the order they finish, subtracts R(q, p)·x_p (or
// R(p, q)ᵀ·x_p) as each x_p is published, then forms x_q = op(R_qq⁻¹)·y_q
// and publishes it (device-coherent stores, then flags[q] = 1).  Workgroup
// ids follow the finishing order, so every block waits only on smaller ids
// and dispatch order alone guarantees progress.  R's block for the next
// dependency and R_qq⁻¹ are loaded into registers before each wait.  The
// flags must be zero at launch; a wait that times out sets err[0] and the
// block finishes (and publishes) with what it has.
template <int NR>
__global__ void __launch_bounds__(256)
trsm_chain_kernel(const double* __restrict__ R, int64_t ldr, int64_t n, const double* __restrict__ Rinv,
                  double* __restrict__ B, int64_t ldb, int nr, int trans, uint64_t* __restrict__ flags,
                  int* __restrict__ err) {
  __shared__ double xs[128][NR + 1];
  __shared__ double ys[128][NR + 1];
  __shared__ double part[2][128][NR + 1];
  __shared__ int s_err;
  const int nb = (int)((n + 127) / 128);
  const int q = trans ? (int)blockIdx.x : nb - 1 - (int)blockIdx.x;
  const int t = threadIdx.x, r = t & 127, h = t >> 7, kb = h * 64;   // row of the block, half of the k range
  const int64_t jq = (int64_t)q * 128;
  const int bq = (int)min<int64_t>(128, n - jq);
  for (int e = t; e < 128 * NR; e += 256) {
    const int rr = e & 127, j = e >> 7;
    ys[rr][j] = (rr < bq && j < nr) ? B[jq + rr + (int64_t)j * ldb] : 0.0;
  }
  if (t == 0) s_err = 0;
  // op(R_qq⁻¹)(r, k) for this thread's half of k: upper (k ≥ r), or the
  // transpose (k ≤ r)
  const double* Ri = Rinv + (int64_t)q * 128 * 128;
  double ri[64];
#pragma unroll
  for (int u = 0; u < 64; ++u) {
    const int k = kb + u;
    const bool ok = r < bq && k < bq && (trans ? k <= r : k >= r);
    ri[u] = Ri[ok ? (trans ? k + (int64_t)r * bq : r + (int64_t)k * bq) : 0];
  }
  const int ndep = trans ? q : nb - 1 - q;
  double a[64];
  auto load_block = [&](int i) {   // R's block against the i-th dependency
    const int p = trans ? i : nb - 1 - i;
    const int64_t jp = (int64_t)p * 128;
    const int bp = (int)min<int64_t>(128, n - jp);
#pragma unroll
    for (int u = 0; u < 64; ++u) {
      const int k = kb + u;
      const bool ok = r < bq && k < bp;
      const int64_t idx = trans ? (jp + k) + (jq + r) * ldr : (jq + r) + (jp + k) * ldr;
      a[u] = R[ok ? idx : 0];
    }
  };
  if (ndep > 0) load_block(0);
  for (int i = 0; i < ndep; ++i) {
    const int p = trans ? i : nb - 1 - i;
    const int64_t jp = (int64_t)p * 128;
    const int bp = (int)min<int64_t>(128, n - jp);
    if (t == 0 && !s_err && !flag_wait(flags + p, 1, wait_limit(err))) s_err = 1;
    __syncthreads();
    for (int e = t; e < 128 * NR; e += 256) {
      const int rr = e & 127, j = e >> 7;
      xs[rr][j] = (rr < bp && j < nr) ? load_coherent(B + jp + rr + (int64_t)j * ldb) : 0.0;
    }
    __syncthreads();
    double acc[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) acc[j] = 0.0;
#pragma unroll
    for (int u = 0; u < 64; ++u) {
      if (kb + u < bp) {
#pragma unroll
        for (int j = 0; j < NR; ++j) acc[j] += a[u] * xs[kb + u][j];
      }
    }
#pragma unroll
    for (int j = 0; j < NR; ++j) part[h][r][j] = acc[j];
    __syncthreads();
    if (h == 0 && r < bq) {
#pragma unroll
      for (int j = 0; j < NR; ++j) ys[r][j] -= part[0][r][j] + part[1][r][j];
    }
    if (i + 1 < ndep) load_block(i + 1);   // in flight during the next wait
  }
  __syncthreads();
  double acc[NR];
#pragma unroll
  for (int j = 0; j < NR; ++j) acc[j] = 0.0;
#pragma unroll
  for (int u = 0; u < 64; ++u) {
    const int k = kb + u;
    if (k < bq && r < bq && (trans ? k <= r : k >= r)) {
#pragma unroll
      for (int j = 0; j < NR; ++j) acc[j] += ri[u] * ys[k][j];
    }
  }
#pragma unroll
  for (int j = 0; j < NR; ++j) part[h][r][j] = acc[j];
  __syncthreads();
  if (h == 0 && r < bq) {
#pragma unroll
    for (int j = 0; j < NR; ++j)
      if (j < nr) store_coherent(B + jq + r + (int64_t)j * ldb, part[0][r][j] + part[1][r][j]);
  }
  wait_memory();   // this wave's x_q stores have reached memory
  __syncthreads();
  if (t == 0) {
    if (s_err) err[0] = 1;
    flag_store(flags + q, 1);
  }
}

int launch_trsm_chain(cqr_ctx* c, const double* R, int64_t ldr, int64_t n, const double* Rinv, double* B,
                      int64_t ldb, int nr, int trans) {
  const int64_t nb = (n + 127) / 128;
  if (nr < 1 || nr > 16 || nb < 1 || nb > CQR_TRSM_MAX_BLOCKS)
    return set_err(c, CQR_ERR_UNSUPPORTED, "trsm chain: %d columns, %lld blocks", nr, (long long)nb);
  CQR_TRY(launch_fill2d(c, reinterpret_cast<double*>(c->d_sflags), nb, nb, 1, 0.0));   // all-zero bits
  auto run = [&](auto k) -> int {
    SpinScope gate(c, (const void*)k, nb, 256, 0);
    CQR_TRY(gate.status());
    hipLaunchKernelGGL(k, dim3((unsigned)nb), dim3(256), 0, c->stream, R, ldr, n, Rinv, B, ldb, nr, trans,
                       c->d_sflags, c->d_terr);
    return CQR_OK;
  };
  if (nr == 1) CQR_TRY(run(trsm_chain_kernel<1>));
  else if (nr <= 4) CQR_TRY(run(trsm_chain_kernel<4>));
  else CQR_TRY(run(trsm_chain_kernel<16>));
  return check_launch(c, "trsm_chain");
}

// Y = alpha·X + beta·Y on an m×n block.
__global__ void axpby_kernel(int64_t m, int64_t n, double alpha, const double* __restrict__ X, int64_t ldx,
                             double beta, double* __restrict__ Y, int64_t ldy) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m * n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e % m, j = e / m;
    double* yp = Y + r + j * ldy;
    *yp = alpha * X[r + j * ldx] + (beta != 0.0 ? beta * *yp : 0.0);
  }
}

// out[j] = ‖X(:, j)‖₂ (scaled sum of squares in FP64), one workgroup per column.
__global__ void __launch_bounds__(256)
colnorm_kernel(const double* __restrict__ X, int64_t ldx, int64_t m, double* __restrict__ out) {
  __shared__ double part[4];
  __shared__ int pnan[4];
  const int j = blockIdx.x;
  const double* x = X + (int64_t)j * ldx;
  double mx = 0.0;
  int nan = 0;   // fmax drops NaN, so a NaN entry is tracked on its own
  for (int64_t r = threadIdx.x; r < m; r += 256) {
    const double a = fabs(x[r]);
    nan |= a != a;
    mx = fmax(mx, a);
  }
  mx = wave_max(mx);
  for (int o = 32; o >= 1; o >>= 1) nan |= __shfl_xor(nan, o);
  if ((threadIdx.x & 63) == 0) {
    part[threadIdx.x >> 6] = mx;
    pnan[threadIdx.x >> 6] = nan;
  }
  __syncthreads();
  mx = fmax(fmax(part[0], part[1]), fmax(part[2], part[3]));
  if (pnan[0] | pnan[1] | pnan[2] | pnan[3]) {
    if (threadIdx.x == 0) out[j] = NAN;
    return;
  }
  __syncthreads();
  double ss = 0.0;
  if (mx > 0.0)
    for (int64_t r = threadIdx.x; r < m; r += 256) { const double t = x[r] / mx; ss += t * t; }
  ss = wave_sum(ss);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = ss;
  __syncthreads();
  if (threadIdx.x == 0) out[j] = mx * sqrt(part[0] + part[1] + part[2] + part[3]);
}

// ---------------------------------------------------------------------------
int thin_gemm(cqr_ctx* c, int ta, int64_t m, int64_t k, double alpha, const double* A, int64_t lda,
              const double* X, int64_t ldx, int nr, double beta, double* C, int64_t ldc, double* scratch,
              int unit_top) {
  if (nr < 1 || nr > 16) return set_err(c, CQR_ERR_UNSUPPORTED, "thin product with %d columns", nr);
  if (m <= 0 || k <= 0) return CQR_OK;
  if (!ta) {
    // one row a thread; a tall and wide product with scratch also splits its
    // columns over thin_n_splits() workgroup rows, for more waves in flight
    const int ny = (scratch && unit_top == 0) ? thin_n_splits(m, k) : 1;
    const int64_t kchunk = ny > 1 ? ((k + ny - 1) / ny + TN_ROWS - 1) / TN_ROWS * TN_ROWS : k;
    double* P = ny > 1 ? scratch : nullptr;
    const dim3 grid((unsigned)((m + 255) / 256), (unsigned)ny);
    if (nr == 1) hipLaunchKernelGGL(thin_gemm_n_kernel<1>, grid, dim3(256), 0, c->stream, A, lda, m, k, X, ldx, nr, alpha, beta, C, ldc, unit_top, kchunk, P);
    else if (nr <= 4) hipLaunchKernelGGL(thin_gemm_n_kernel<4>, grid, dim3(256), 0, c->stream, A, lda, m, k, X, ldx, nr, alpha, beta, C, ldc, unit_top, kchunk, P);
    else hipLaunchKernelGGL(thin_gemm_n_kernel<16>, grid, dim3(256), 0, c->stream, A, lda, m, k, X, ldx, nr, alpha, beta, C, ldc, unit_top, kchunk, P);
    CQR_TRY(check_launch(c, "thin_gemm_n"));
    if (ny == 1) return CQR_OK;
    hipLaunchKernelGGL(thin_n_reduce_kernel, dim3((unsigned)((m * nr + 255) / 256)), dim3(256), 0, c->stream, P, ny, m, nr,
                       alpha, beta, C, ldc);
    return check_launch(c, "thin_n_reduce");
  }
  // C (k×nr) = alpha·Aᵀ·X + beta·C, A m×k; slabs of rows, then a fixed-order sum
  const int64_t slab = thin_t_slab_rows(m, k);
  const int nslab = (int)((m + slab - 1) / slab);
  const dim3 grid((unsigned)((k + 127) / 128), (unsigned)nslab);
  if (nr == 1) hipLaunchKernelGGL(thin_gemm_t_kernel<1>, grid, dim3(256), 0, c->stream, A, lda, m, k, X, ldx, nr, slab, scratch, unit_top);
  else if (nr <= 4) hipLaunchKernelGGL(thin_gemm_t_kernel<4>, grid, dim3(256), 0, c->stream, A, lda, m, k, X, ldx, nr, slab, scratch, unit_top);
  else hipLaunchKernelGGL(thin_gemm_t_kernel<16>, grid, dim3(256), 0, c->stream, A, lda, m, k, X, ldx, nr, slab, scratch, unit_top);
  CQR_TRY(check_launch(c, "thin_gemm_t"));
  if (!C) return CQR_OK;   // the caller sums the slab partials itself
  hipLaunchKernelGGL(slab_reduce_kernel, dim3((unsigned)((k * nr + 63) / 64)), dim3(1024), 0, c->stream, scratch, nslab, k,
                     nr, alpha, beta, C, ldc);
  return check_launch(c, "slab_reduce");
}

int launch_slab_reduce_t(cqr_ctx* c, const double* P, int nslab, int k, int nr, const double* Tt, int64_t ldt,
                         int ta, double* Zp, int64_t ldz) {
  if (k < 1 || k > 128 || nr < 1 || nr > 4 || nslab < 1)
    return set_err(c, CQR_ERR_UNSUPPORTED, "slab_reduce_t: k %d, %d columns", k, nr);
  const size_t lds = sizeof(double) * ((size_t)k * k + 1024 + 512 + 1024);
  CQR_TRY(set_max_lds(c, (const void*)slab_reduce_t_kernel, (int)(sizeof(double) * (128 * 128 + 2560))));
  hipLaunchKernelGGL(slab_reduce_t_kernel, dim3(1), dim3(1024), lds, c->stream, P, nslab, k, nr, Tt, ldt, ta, Zp, ldz);
  return check_launch(c, "slab_reduce_t");
}

// Rows per slab of the Aᵀ·X product: about 512 workgroups in total.
int64_t thin_t_slab_rows(int64_t m, int64_t k) {
  const int64_t cblocks = (k + 127) / 128;
  int64_t nslab = std::max<int64_t>(1, 512 / cblocks);
  int64_t slab = (m + nslab - 1) / nslab;
  slab = std::max<int64_t>(TT_ROWS, (slab + TT_ROWS - 1) / TT_ROWS * TT_ROWS);
  return slab;
}

// Column chunks of a plain A·X (at least 256 columns each): 8 from
// 8192×1024 on, else 1.
int thin_n_splits(int64_t m, int64_t k) { return (m >= 8192 && k >= 1024) ? 8 : 1; }

size_t thin_n_scratch_doubles(int64_t m, int64_t k, int nr) {
  const int ny = thin_n_splits(m, k);
  return ny > 1 ? (size_t)ny * m * nr : 0;
}

size_t thin_gemm_scratch_doubles(int64_t m, int64_t k, int nr) {
  const int64_t slab = thin_t_slab_rows(m, k);
  return (size_t)((m + slab - 1) / slab) * k * nr;
}

int launch_axpby(cqr_ctx* c, int64_t m, int64_t n, double alpha, const double* X, int64_t ldx, double beta,
                 double* Y, int64_t ldy) {
  if (m <= 0 || n <= 0) return CQR_OK;
  const int gx = (int)std::min<int64_t>(2048, (m * n + 255) / 256);
  hipLaunchKernelGGL(axpby_kernel, dim3(gx), dim3(256), 0, c->stream, m, n, alpha, X, ldx, beta, Y, ldy);
  return check_launch(c, "axpby");
}

// out[0] = max_i |R(i,i)|, out[1] = min_i |R(i,i)| over the n diagonal
// entries (a NaN in any propagates to both).  One workgroup.
__global__ void __launch_bounds__(256)
diag_range_kernel(const double* __restrict__ R, int64_t ldr, int64_t n, double* __restrict__ out) {
  __shared__ double pmax[4], pmin[4];
  __shared__ int pnan[4];
  double mx = 0.0, mn = INFINITY;
  int nan = 0;
  for (int64_t i = threadIdx.x; i < n; i += 256) {
    const double v = fabs(R[i + i * ldr]);
    nan |= v != v;
    mx = fmax(mx, v);
    mn = fmin(mn, v);
  }
  mx = wave_max(mx);
  mn = -wave_max(-mn);
  for (int o = 32; o >= 1; o >>= 1) nan |= __shfl_xor(nan, o);
  if ((threadIdx.x & 63) == 0) {
    pmax[threadIdx.x >> 6] = mx;
    pmin[threadIdx.x >> 6] = mn;
    pnan[threadIdx.x >> 6] = nan;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const bool any_nan = pnan[0] | pnan[1] | pnan[2] | pnan[3];
    out[0] = any_nan ? NAN : fmax(fmax(pmax[0], pmax[1]), fmax(pmax[2], pmax[3]));
    out[1] = any_nan ? NAN : fmin(fmin(pmin[0], pmin[1]), fmin(pmin[2], pmin[3]));
  }
}

int launch_diag_range(cqr_ctx* c, const double* R, int64_t ldr, int64_t n, double* out) {
  hipLaunchKernelGGL(diag_range_kernel, dim3(1), dim3(256), 0, c->stream, R, ldr, n, out);
  return check_launch(c, "diag_range");
}

int launch_colnorm(cqr_ctx* c, const double* X, int64_t ldx, int64_t m, int n, double* out) {
  if (n <= 0) return CQR_OK;
  hipLaunchKernelGGL(colnorm_kernel, dim3(n), dim3(256), 0, c->stream, X, ldx, m, out);
  return check_launch(c, "colnorm");
}

// Element-wise helpers of the batched path (kernels, so that they can live
// inside a captured graph).
__global__ void fill_u64_kernel(uint64_t* __restrict__ p, int64_t n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    p[e] = 0;
}

__global__ void convert_kernel(const void* __restrict__ src, void* __restrict__ dst, int64_t n, int to_f64) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    if (to_f64) static_cast<double*>(dst)[e] = (double)static_cast<const float*>(src)[e];
    else static_cast<float*>(dst)[e] = (float)static_cast<const double*>(src)[e];
  }
}

int fill_u64(cqr_ctx* c, uint64_t* p, int64_t n) {
  if (n <= 0) return CQR_OK;
  const int gx = (int)std::min<int64_t>(1024, (n + 255) / 256);
  hipLaunchKernelGGL(fill_u64_kernel, dim3(gx), dim3(256), 0, c->stream, p, n);
  return check_launch(c, "fill_u64");
}

int convert_f32_f64(cqr_ctx* c, const void* src, void* dst, int64_t n, int to_f64) {
  if (n <= 0) return CQR_OK;
  const int gx = (int)std::min<int64_t>(4096, (n + 255) / 256);
  hipLaunchKernelGGL(convert_kernel, dim3(gx), dim3(256), 0, c->stream, src, dst, n, to_f64);
  return check_launch(c, "convert");
}

}  // namespace cqr
