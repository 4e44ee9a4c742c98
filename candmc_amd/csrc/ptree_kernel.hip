// Column-pipelined TSQR reduction tree (all levels in one launch).
//
// The tree is the reference's binary schedule (alg/QR/tsqr/bitree_tsqr.cxx:
// 172-232): on a level with n live ranks, receiver z stacks its R on the R
// of sender z + ceil(n/2) and refactors the 2b×b triangle pair (dtpqrt with
// l = b, :20-34).  Column i of a node's output R depends only on columns
// 0..i of its two inputs, so every node runs left-looking and publishes each
// finished column through a monotone counter; a parent starts column i as
// soon as both children have published it.  All levels then advance
// together and the tree costs about (b + depth) column steps instead of
// b × depth.
//
// One workgroup (512 threads) per node.  Node ids follow forward level
// order, so every producer has a smaller block id than its consumer and is
// dispatched first; waits are bounded (a timeout raises a device flag and
// the node finishes with whatever it has).
//
// Per column i of a node with reflectors [e_p; v2_p], p < i, and compact-WY
// factor T (Q_node = I − V·T·Vᵀ, V = [I; V2]):
//   s = c_top(0:i) + V2(:,0:i)ᵀ·c_bot,  u = Tᵀ·s,
//   c_top(0:i) −= u,  c_bot −= V2(:,0:i)·u          (apply H_{i-1}···H_0)
//   dlarfg on (c_top(i), c_bot(0:i)) → β, τ_i, v2_i; R_out(:,i) = [c_top(0:i); β]
//   T(0:i, i) = −τ_i·T(0:i,0:i)·(V2(:,0:i)ᵀ·v2_i),  T(i,i) = τ_i.
#include "cqr_device.h"
#include "cqr_internal.h"
#include "leaf_body.h"

namespace cqr {

// Producer of rank r's current R when entering level `lv` (level sizes
// n_0 = L, n_{k+1} = ceil(n_k / 2)): the node of the last earlier level at
// which r was a receiver, or leaf r.  Returns a node id, or −1 for a leaf.
__device__ inline int producer_of(int r, int lv, int L) {
  int n = L, node0 = 0, best = -1;
  for (int k = 0; k < lv; ++k) {
    const int pairs = n / 2;
    if (r < pairs) best = node0 + r;
    node0 += pairs;
    n = n / 2 + (n & 1);
  }
  return best;
}

template <int CPW, int RPL>
__global__ void __launch_bounds__(512)
ptree_kernel(const double* __restrict__ leafR, double* __restrict__ nodeR, double* __restrict__ V2out,
             double* __restrict__ Tout, double* __restrict__ tauout, uint64_t* __restrict__ cnt,
             uint64_t epoch, int L, int b, int* __restrict__ err, const BatchProb* __restrict__ bp, const BatchSeg* __restrict__ seg, int nseg) {
  constexpr int WAVES = 8;
  int nid = blockIdx.x;
  const int lane = threadIdx.x & 63, w = wave_id();
  const int64_t bb = (int64_t)b * b;
  if (bp) {  // batched: this problem's leaves and nodes are contiguous slices
    const BatchSeg S = seg[seg_find(seg, nseg, nid)];
    const BatchProb P = bp[S.prob];
    nid = S.local0 + (nid - S.g0);
    L = P.L;
    leafR += (int64_t)P.leaf0 * bb;
    nodeR += (int64_t)P.node0 * bb;
    V2out += (int64_t)P.node0 * bb;
    Tout += (int64_t)P.node0 * bb;
    tauout += (int64_t)P.node0 * b;
    cnt += P.node0;
  }

  // locate this node: level lv, pair z, live ranks n
  int lv = 0, n = L, node0 = 0;
  while (nid >= node0 + n / 2) {
    node0 += n / 2;
    n = n / 2 + (n & 1);
    ++lv;
  }
  const int z = nid - node0, snd = z + (n + 1) / 2;
  const int ptop = producer_of(z, lv, L), pbot = producer_of(snd, lv, L);
  const double* Rt = ptop >= 0 ? nodeR + ptop * bb : leafR + z * bb;
  const double* Rb = pbot >= 0 ? nodeR + pbot * bb : leafR + snd * bb;
  double* Ro = nodeR + nid * bb;

  __shared__ double sbuf[128], ubuf[128], zbuf[128], stau[128];
  __shared__ double red[WAVES][128];
  __shared__ int s_err;

  double v2[CPW][RPL], t[CPW][RPL];
#pragma unroll
  for (int cl = 0; cl < CPW; ++cl)
#pragma unroll
    for (int s = 0; s < RPL; ++s) v2[cl][s] = t[cl][s] = 0.0;
  // Thread 0 waits until both inputs have published column i (bounded).
  auto wait_inputs = [&](int i) {
    const uint64_t want = epoch * 256 + (uint64_t)(i + 1);
    if (ptop >= 0 && !flag_wait(cnt + ptop, want, wait_limit(err))) s_err = 1;
    if (pbot >= 0 && !flag_wait(cnt + pbot, want, wait_limit(err))) s_err = 1;
  };
  // Column i of both inputs (rows ≤ i), device-coherent loads.
  double ctop[RPL], cbot[RPL], ntop[RPL], nbot[RPL];
  auto load_inputs = [&](int i, double (&top)[RPL], double (&bot)[RPL]) {
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      const int row = lane + 64 * s;
      const bool in = row <= i;
      top[s] = in ? load_coherent(Rt + row + (int64_t)i * b) : 0.0;
      bot[s] = in ? load_coherent(Rb + row + (int64_t)i * b) : 0.0;
    }
  };
  // z = V2(:, 0:c)ᵀ·x over this wave's columns q < c, reduced into buf[q]
  auto v2t_times = [&](const double (&x)[RPL], int c, double* buf) {
    double part[CPW];
#pragma unroll
    for (int cl = 0; cl < CPW; ++cl) {
      double a = 0.0;
#pragma unroll
      for (int s = 0; s < RPL; ++s) a += v2[cl][s] * x[s];
      part[cl] = a;
    }
    const double r = xreduce<CPW>(part, lane);
    const int q = w + WAVES * (lane & (CPW - 1));
    if (lane < CPW && q < c) buf[q] = r;
  };
  // T(0:c, c) = −τ·T(0:c,0:c)·z: this wave's partial sums over its columns
  auto t_partials = [&](int c) {
    double py[RPL];
#pragma unroll
    for (int s = 0; s < RPL; ++s) py[s] = 0.0;
    // T's columns q ≥ c are still zero and zbuf holds finite values, so the
    // sum needs no mask (and no branch per column)
#pragma unroll
    for (int cl = 0; cl < CPW; ++cl) {
      const double zq = zbuf[w + WAVES * cl];
#pragma unroll
      for (int s = 0; s < RPL; ++s) py[s] += t[cl][s] * zq;
    }
#pragma unroll
    for (int s = 0; s < RPL; ++s) red[w][lane + 64 * s] = py[s];
  };
  // the owner of column c sums the partials into its T column
  auto t_finish = [&](int c, double tau_c) {
    if (w == (c % WAVES)) {
#pragma unroll
      for (int s = 0; s < RPL; ++s) {
        const int row = lane + 64 * s;
        double acc = 0.0;
#pragma unroll
        for (int ww = 0; ww < WAVES; ++ww) acc += red[ww][row];
        const double tv = row < c ? -tau_c * acc : (row == c ? tau_c : 0.0);
        const int cc = c / WAVES;
#pragma unroll
        for (int cl = 0; cl < CPW; ++cl) t[cl][s] = cl == cc ? tv : t[cl][s];
      }
    }
  };

  for (int p = threadIdx.x; p < 128; p += 512) sbuf[p] = ubuf[p] = zbuf[p] = 0.0;
  if (threadIdx.x == 0) {
    s_err = 0;
    wait_inputs(0);
  }
  __syncthreads();
  if (!s_err) load_inputs(0, ctop, cbot);

  // Software pipeline: iteration i finishes the T column of i−1 alongside
  // the start of column i, so each column costs four barriers.
  double vprev[RPL];
#pragma unroll
  for (int s = 0; s < RPL; ++s) vprev[s] = 0.0;
  double tau_prev = 0.0;
  int done = 0;  // columns whose reflector exists
  for (int i = 0; i < b && !s_err; ++i) {
    // ---- A: z_{i−1} = V2ᵀ·v_{i−1} and s_i = V2ᵀ·c_bot (columns q < i−1, q < i)
    if (i > 0) v2t_times(vprev, i - 1, zbuf);
    v2t_times(cbot, i, sbuf);
    if (w == 0) wait_memory();  // column i−1 of R has reached memory
    __syncthreads();                                                       // B1
    if (threadIdx.x == 0) {
      if (i > 0) flag_store(cnt + nid, epoch * 256 + (uint64_t)i);
      if (i + 1 < b) wait_inputs(i + 1);
    }
    // ---- B: partial sums of T's column i−1
    if (i > 0) t_partials(i - 1);
    __syncthreads();                                                       // B2
    if (i + 1 < b && !s_err) load_inputs(i + 1, ntop, nbot);  // in flight from here on
    if (i > 0) t_finish(i - 1, tau_prev);
    // ---- C: u = Tᵀ·(c_top + V2ᵀ·c_bot)
    {
      double sv[RPL];
#pragma unroll
      for (int s = 0; s < RPL; ++s) {
        const int row = lane + 64 * s;
        sv[s] = row < i ? sbuf[row] + ctop[s] : 0.0;
      }
      double part[CPW];
#pragma unroll
      for (int cl = 0; cl < CPW; ++cl) {
        double a = 0.0;
#pragma unroll
        for (int s = 0; s < RPL; ++s) a += t[cl][s] * sv[s];
        part[cl] = a;
      }
      const double r = xreduce<CPW>(part, lane);
      const int q = w + WAVES * (lane & (CPW - 1));
      if (lane < CPW && q < i) ubuf[q] = r;
    }
    __syncthreads();                                                       // B3
    // ---- D: c_bot −= V2·u (partials per wave, summed through LDS), c_top −= u
    {
      double pb[RPL];
#pragma unroll
      for (int s = 0; s < RPL; ++s) pb[s] = 0.0;
      // V2's columns q ≥ i are still zero: no mask needed (see t_partials)
#pragma unroll
      for (int cl = 0; cl < CPW; ++cl) {
        const double uq = ubuf[w + WAVES * cl];
#pragma unroll
        for (int s = 0; s < RPL; ++s) pb[s] += v2[cl][s] * uq;
      }
#pragma unroll
      for (int s = 0; s < RPL; ++s) red[w][lane + 64 * s] = pb[s];
    }
    __syncthreads();                                                       // B4
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      const int row = lane + 64 * s;
      double acc = 0.0;
#pragma unroll
      for (int ww = 0; ww < WAVES; ++ww) acc += red[ww][row];
      cbot[s] -= acc;
      if (row < i) ctop[s] -= ubuf[row];
    }
    // ---- E: reflector for column i (every wave computes the same scalars)
    double ss = 0.0, piv = 0.0;
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      const int row = lane + 64 * s;
      if (row <= i) ss += cbot[s] * cbot[s];
      if (s == (i >> 6)) piv = ctop[s];
    }
    ss = wave_sum(ss);
    const double alpha = readlane_d(piv, i & 63);
    double xnorm;
    if (ss_needs_rescale(ss)) {
      double mx = 0.0;
#pragma unroll
      for (int s = 0; s < RPL; ++s)
        if (lane + 64 * s <= i) mx = fmax(mx, fabs(cbot[s]));
      mx = wave_max(mx);
      if (mx == 0.0) {
        xnorm = 0.0;
      } else {
        double s2 = 0.0;
#pragma unroll
        for (int s = 0; s < RPL; ++s)
          if (lane + 64 * s <= i) { const double tt = cbot[s] / mx; s2 += tt * tt; }
        xnorm = mx * sqrt(wave_sum(s2));
      }
    } else {
      xnorm = sqrt(ss);
    }
    const Reflector h = dlarfg_scalar(alpha, xnorm);
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      const int row = lane + 64 * s;
      double x = cbot[s];
      if (h.tau != 0.0 && row <= i) {
        for (int k = 0; k < h.knt; ++k) x *= 0x1p969;
        x *= h.scal;
      }
      vprev[s] = row <= i ? x : 0.0;
    }
    if (w == (i % WAVES)) {
      const int ci = i / WAVES;
#pragma unroll
      for (int cl = 0; cl < CPW; ++cl)
#pragma unroll
        for (int s = 0; s < RPL; ++s) v2[cl][s] = cl == ci ? vprev[s] : v2[cl][s];
    }
    // column i of R: [c_top(0:i); β], published after the next barrier
    if (w == 0) {
      for (int s = 0; s < RPL; ++s) {
        const int row = lane + 64 * s;
        if (row < b) store_coherent(Ro + row + (int64_t)i * b, row < i ? ctop[s] : (row == i ? h.beta : 0.0));
      }
    }
    if (threadIdx.x == 0) stau[i] = h.tau;
    tau_prev = h.tau;
    done = i + 1;
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      ctop[s] = ntop[s];
      cbot[s] = nbot[s];
    }
  }
  // ---- finish the last column: publish it and complete its T column
  if (done > 0) {
    const int c = done - 1;
    v2t_times(vprev, c, zbuf);
    if (w == 0) wait_memory();
    __syncthreads();
    if (threadIdx.x == 0) flag_store(cnt + nid, epoch * 256 + (uint64_t)done);
    t_partials(c);
    __syncthreads();
    t_finish(c, tau_prev);
  }
  __syncthreads();
  if (s_err && threadIdx.x == 0) *err = 1;
  // ---- V2, T, tau of this node
  double* V2 = V2out + nid * bb;
  double* Tn = Tout + nid * bb;
#pragma unroll
  for (int cl = 0; cl < CPW; ++cl) {
    const int col = w + WAVES * cl;
    if (col >= b) continue;
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      const int row = lane + 64 * s;
      if (row < b) {
        V2[row + (int64_t)col * b] = row <= col ? v2[cl][s] : 0.0;
        Tn[row + (int64_t)col * b] = row <= col ? t[cl][s] : 0.0;
      }
    }
  }
  for (int p = threadIdx.x; p < b; p += 512) tauout[(int64_t)nid * b + p] = stau[p];
}

// ---------------------------------------------------------------------------
// Blocked form of the same tree node (b a multiple of 8): the node advances
// eight columns at a time, left-looking.  For block [c0, c0+8):
//   S = C_top(0:c0) + V2(:,0:c0)ᵀ·C_bot,  U = T(0:c0,0:c0)ᵀ·S,
//   R(0:c0, block) = C_top(0:c0) − U,  C_bot −= V2(:,0:c0)·U
// (Qᵀ of the earlier columns applied to the block at once), then one wave
// factors the 8 columns (dlarfg per column, reflectors applied within the
// block), builds their 8×8 T and publishes the block of R.  T's columns of
// block k−1 above its diagonal block, T(0:c, blk) = −T(0:c,0:c)·(V2(:,0:c)ᵀ·
// V2_blk)·T_blk, are completed at the start of block k: V2ᵀ·V2_blk shares
// the MFMA tiles of V2ᵀ·C_bot (columns 8..15 of the 16-wide tile).
// V2 and T live in LDS as packed upper triangles; each product is one
// 16-row FP64 MFMA tile per wave, so a block costs six barriers where the
// column form spends four per column.
// ---------------------------------------------------------------------------
__device__ inline int pk_off(int q) { return q * (q + 1) / 2; }
// MFMA operand accessors load from a clamped (in-bounds) address
// unconditionally and mask the value, so no load sits behind a branch
__device__ inline double sel(bool ok, double v) { return ok ? v : 0.0; }

// LDS of one node (doubles): packed V2 and T, three 8-wide row blocks, the
// 8×8 diagonal input block and tau.
template <int BMAX>
constexpr int btree_smem_doubles() { return BMAX * (BMAX + 1) + 3 * BMAX * 8 + 64 + BMAX; }

// One tree node per workgroup (4·BMAX threads).  A producer that is a node
// is waited for through cnt; a leaf producer is complete at launch unless
// lcnt is set (fused launch), in which case lcnt[leaf] is waited for too.
template <int BMAX>
__device__ __forceinline__ void btree_body(const double* __restrict__ leafR, double* __restrict__ nodeR,
                                  double* __restrict__ V2out, double* __restrict__ Tout,
                                  double* __restrict__ tauout, uint64_t* __restrict__ cnt, uint64_t epoch, int L,
                                  int b, int* __restrict__ err, const BatchProb* __restrict__ bp, const BatchSeg* __restrict__ seg, int nseg, int nid,
                                  double* smem, uint64_t* __restrict__ lcnt) {
  constexpr int NT = 4 * BMAX, NB = 8, RPL = (BMAX + 63) / 64;
  constexpr int PACK = BMAX * (BMAX + 1) / 2;
  double* V2p = smem;                      // packed upper triangles of the node's V2 and T
  double* Tp = V2p + PACK;
  double* Cb = Tp + PACK;                  // bottom rows of the block, row-major [r][8]
  double* Xs = Cb + BMAX * NB;             // S
  double* Xg = Xs + BMAX * NB;             // V2ᵀ·V2_blk, then U
  double* Ctd = Xg + BMAX * NB;            // top 8×8 diagonal block of the input, row-major
  double* stau = Ctd + NB * NB;
  __shared__ int s_err;

  const int tid = threadIdx.x, lane = tid & 63, w = wave_id();
  const int64_t bb = (int64_t)b * b;
  if (bp) {
    const BatchSeg S = seg[seg_find(seg, nseg, nid)];
    const BatchProb P = bp[S.prob];
    nid = S.local0 + (nid - S.g0);
    L = P.L;
    leafR += (int64_t)P.leaf0 * bb;
    nodeR += (int64_t)P.node0 * bb;
    V2out += (int64_t)P.node0 * bb;
    Tout += (int64_t)P.node0 * bb;
    tauout += (int64_t)P.node0 * b;
    cnt += P.node0;
  }
  int lv = 0, n = L, node0 = 0;
  while (nid >= node0 + n / 2) {
    node0 += n / 2;
    n = n / 2 + (n & 1);
    ++lv;
  }
  const int z = nid - node0, snd = z + (n + 1) / 2;
  const int ptop = producer_of(z, lv, L), pbot = producer_of(snd, lv, L);
  const double* Rt = ptop >= 0 ? nodeR + ptop * bb : leafR + z * bb;
  const double* Rb = pbot >= 0 ? nodeR + pbot * bb : leafR + snd * bb;
  double* Ro = nodeR + nid * bb;

  // MFMA tile of this wave: rows t0..t0+15; in a result, lane l holds row
  // tq = t0 + (l & 15) and columns q4 + 4v (v = 0..3)
  const int r16 = lane & 15, q4 = lane >> 4;
  const int t0 = 16 * w, tq = t0 + r16;

  // T(0:cp, cp:cp+8) = −T(0:cp,0:cp)·G·T(cp:cp+8, cp:cp+8), G = V2(:,0:cp)ᵀ·V2(:,cp:cp+8) in Xg
  auto t12_block = [&](int cp) {
    if (t0 >= cp) return;
    auto x = [&](int m, int p) {
      const int q = t0 + m;
      const bool ok = q <= p && p < cp;
      return sel(ok, Tp[ok ? pk_off(p) + q : 0]);
    };
    auto y = [&](int p, int c) {
      const bool ok = c < 8 && p < cp;
      return sel(ok, Xg[ok ? p * 8 + c : 0]);
    };
    const dbl4 g2 = mfma_tile16_acc4(x, y, t0, cp, lane);
    // G2·T_blk with G2 straight from the accumulator (its layout is the A operand's)
    auto tb = [&](int kk, int c) {
      const bool ok = c < 8 && kk <= c;
      return sel(ok, Tp[ok ? pk_off(cp + c) + cp + kk : 0]);
    };
    dbl4 t12 = {0.0, 0.0, 0.0, 0.0};
    t12 = __builtin_amdgcn_mfma_f64_16x16x4f64(tb(q4, r16), g2[0], t12, 0, 0, 0);
    t12 = __builtin_amdgcn_mfma_f64_16x16x4f64(tb(4 + q4, r16), g2[1], t12, 0, 0, 0);
    if (tq < cp) {
      Tp[pk_off(cp + q4) + tq] = -t12[0];
      Tp[pk_off(cp + q4 + 4) + tq] = -t12[1];
    }
  };

  if (tid == 0) s_err = 0;
  const int nblk = b / NB;
  for (int k = 0; k < nblk; ++k) {
    const int c0 = k * NB, cp = c0 - NB;
    if (tid == 0) {
      const uint64_t want = epoch * 256 + (uint64_t)(c0 + NB);
      if (ptop >= 0) {
        if (!flag_wait(cnt + ptop, want, wait_limit(err))) s_err = 1;
      } else if (lcnt && !flag_wait(lcnt + z, want, wait_limit(err))) {
        s_err = 1;
      }
      if (pbot >= 0) {
        if (!flag_wait(cnt + pbot, want, wait_limit(err))) s_err = 1;
      } else if (lcnt && !flag_wait(lcnt + snd, want, wait_limit(err))) {
        s_err = 1;
      }
    }
    __syncthreads();                                                       // B0
    if (s_err) break;
    // ---- inputs of the block (rows ≤ column), zeros below the block's diagonal
    for (int e = tid; e < BMAX * NB; e += NT) {
      const int r = e >> 3, jj = e & 7;
      Cb[e] = (r < b && r <= c0 + jj) ? load_coherent(Rb + r + (int64_t)(c0 + jj) * b) : 0.0;
    }
    if (tid < NB * NB) {
      const int rr = tid >> 3, jj = tid & 7;
      Ctd[tid] = rr <= jj ? load_coherent(Rt + (c0 + rr) + (int64_t)(c0 + jj) * b) : 0.0;
    }
    double ct0 = 0.0, ct1 = 0.0;
    if (tq < c0) {
      ct0 = load_coherent(Rt + tq + (int64_t)(c0 + q4) * b);
      ct1 = load_coherent(Rt + tq + (int64_t)(c0 + q4 + 4) * b);
    }
    {
      const int rows = b - c0 - NB;
      for (int e = tid; e < rows * NB; e += NT) {
        const int jj = e / rows, r = c0 + NB + (e - jj * rows);
        Ro[r + (int64_t)(c0 + jj) * b] = 0.0;
      }
    }
    __syncthreads();                                                       // B1
    if (k > 0) {
      // S = C_top + V2ᵀ·C_bot (tile columns 0..7) and V2ᵀ·V2_{k−1} (columns 8..15)
      if (t0 < c0) {
        auto x = [&](int m, int r) {
          const int q = t0 + m;
          const bool ok = r <= q && q < c0;
          return sel(ok, V2p[ok ? pk_off(q) + r : 0]);
        };
        auto y = [&](int r, int c) {
          const int col = cp + c - 8;
          const bool ok = r < c0 && (c < 8 || r <= col);
          const double* a = c < 8 ? Cb + r * 8 + c : V2p + pk_off(col) + r;
          return sel(ok, *(ok ? a : Cb));
        };
        const dbl4 acc = mfma_tile16_acc4(x, y, 0, min(t0 + 16, c0), lane);
        if (tq < c0) {
          Xs[tq * 8 + q4] = ct0 + acc[0];
          Xs[tq * 8 + q4 + 4] = ct1 + acc[1];
          Xg[tq * 8 + q4] = acc[2];
          Xg[tq * 8 + q4 + 4] = acc[3];
        }
      }
      __syncthreads();                                                     // B2
      // T columns of block k−1 above its diagonal block
      if (cp > 0) t12_block(cp);
      __syncthreads();                                                     // B3
      // U = Tᵀ·S; R(0:c0, block) = C_top − U
      if (t0 < c0) {
        auto x = [&](int m, int p) {
          const int q = t0 + m;
          const bool ok = p <= q && q < c0;
          return sel(ok, Tp[ok ? pk_off(q) + p : 0]);
        };
        auto y = [&](int p, int c) {
          const bool ok = c < 8 && p < c0;
          return sel(ok, Xs[ok ? p * 8 + c : 0]);
        };
        const dbl4 u = mfma_tile16_acc4(x, y, 0, min(t0 + 16, c0), lane);
        if (tq < c0) {
          Xg[tq * 8 + q4] = u[0];
          Xg[tq * 8 + q4 + 4] = u[1];
          store_coherent(Ro + tq + (int64_t)(c0 + q4) * b, ct0 - u[0]);
          store_coherent(Ro + tq + (int64_t)(c0 + q4 + 4) * b, ct1 - u[1]);
        }
      }
      __syncthreads();                                                     // B4
      // C_bot −= V2·U
      if (t0 < c0) {
        auto x = [&](int m, int q) {
          const int r = t0 + m;
          const bool ok = r <= q && q < c0;
          return sel(ok, V2p[ok ? pk_off(q) + r : 0]);
        };
        auto y = [&](int q, int c) {
          const bool ok = c < 8 && q < c0;
          return sel(ok, Xg[ok ? q * 8 + c : 0]);
        };
        const dbl4 d = mfma_tile16_acc4(x, y, t0, c0, lane);
        if (tq < c0) {
          Cb[tq * 8 + q4] -= d[0];
          Cb[tq * 8 + q4 + 4] -= d[1];
        }
      }
      __syncthreads();                                                     // B5
    }
    // ---- one wave factors the 8 columns
    if (w == 0) {
      double cb[RPL][NB];
#pragma unroll
      for (int s = 0; s < RPL; ++s) {
        const int r = lane + 64 * s;
#pragma unroll
        for (int jj = 0; jj < NB; ++jj) cb[s][jj] = r < BMAX ? Cb[r * 8 + jj] : 0.0;
      }
      double trow[NB];   // row `lane` of the block's T (lanes < 8)
      double crow[NB];   // row jj of the input's top diagonal block
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        trow[i] = 0.0;
        crow[i] = Ctd[i];
      }
#pragma unroll
      for (int jj = 0; jj < NB; ++jj) {
        double crn[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i) crn[i] = jj + 1 < NB ? Ctd[(jj + 1) * 8 + i] : 0.0;
        // one reduction of x·c_i over the 8 columns of the block, x = column
        // jj below the pivot: i ≠ jj feeds the reflector's application and
        // T, i = jj is ‖x‖²; none of it depends on the reflector itself
        double dv[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
          double a = 0.0;
#pragma unroll
          for (int s = 0; s < RPL; ++s) a += cb[s][i] * cb[s][jj];
          dv[i] = a;
        }
        const double red = xreduce<NB>(dv, lane);
        const double ss = readlane_d(red, jj);
        const double alpha = crow[jj];
        Reflector h;
        if (!dlarfg_fast(alpha, ss, h)) {
          double xnorm;
          if (ss_needs_rescale(ss)) {
            double mx = 0.0;
#pragma unroll
            for (int s = 0; s < RPL; ++s) mx = fmax(mx, fabs(cb[s][jj]));
            mx = wave_max(mx);
            if (mx == 0.0) {
              xnorm = 0.0;
            } else {
              double s2 = 0.0;
#pragma unroll
              for (int s = 0; s < RPL; ++s) { const double tt = cb[s][jj] / mx; s2 += tt * tt; }
              xnorm = mx * sqrt(wave_sum(s2));
            }
          } else {
            xnorm = sqrt(ss);
          }
          h = dlarfg_scalar(alpha, xnorm);
        }
        // v = x scaled as dlarfg scales it; the reduced dot products of x take the same factors
        double d[NB];   // i > jj: v·c_i; i < jj: v_i·v
#pragma unroll
        for (int i = 0; i < NB; ++i) d[i] = i == jj ? 0.0 : readlane_d(red, i);
        if (h.knt == 0) {
          const double f = h.tau != 0.0 ? h.scal : 1.0;
#pragma unroll
          for (int s = 0; s < RPL; ++s) cb[s][jj] *= f;
#pragma unroll
          for (int i = 0; i < NB; ++i) d[i] *= f;
        } else {
          auto scale = [&](double x) {
            if (h.tau != 0.0) {
              for (int kk = 0; kk < h.knt; ++kk) x *= 0x1p969;
              x *= h.scal;
            }
            return x;
          };
#pragma unroll
          for (int s = 0; s < RPL; ++s) cb[s][jj] = scale(cb[s][jj]);
#pragma unroll
          for (int i = 0; i < NB; ++i) d[i] = scale(d[i]);
        }
        double rval = lane == jj ? h.beta : 0.0;
#pragma unroll
        for (int i = jj + 1; i < NB; ++i) {
          const double wv = crow[i] + d[i];
#pragma unroll
          for (int s = 0; s < RPL; ++s) cb[s][i] -= h.tau * wv * cb[s][jj];
          rval = lane == i ? crow[i] - h.tau * wv : rval;
        }
        if (lane < NB) store_coherent(Ro + (c0 + jj) + (int64_t)(c0 + lane) * b, rval);
        double tsum = 0.0;
#pragma unroll
        for (int i = 0; i < jj; ++i) tsum += trow[i] * d[i];
        trow[jj] = lane < jj ? -h.tau * tsum : (lane == jj ? h.tau : 0.0);
        if (lane == 0) stau[c0 + jj] = h.tau;
#pragma unroll
        for (int i = 0; i < NB; ++i) crow[i] = crn[i];
      }
#pragma unroll
      for (int s = 0; s < RPL; ++s) {
        const int r = lane + 64 * s;
#pragma unroll
        for (int jj = 0; jj < NB; ++jj)
          if (r <= c0 + jj) V2p[pk_off(c0 + jj) + r] = cb[s][jj];
      }
      if (lane < NB) {
#pragma unroll
        for (int jj = 0; jj < NB; ++jj)
          if (lane <= jj) Tp[pk_off(c0 + jj) + c0 + lane] = trow[jj];
      }
    }
    wait_memory();   // this wave's R stores have reached memory
    __syncthreads();                                                       // B6
    if (tid == 0) flag_store(cnt + nid, epoch * 256 + (uint64_t)(c0 + NB));
  }
  // ---- T columns of the last block above its diagonal block
  const int cl = (nblk - 1) * NB;
  if (!s_err && cl > 0) {
    if (t0 < cl) {
      auto x = [&](int m, int r) {
        const int q = t0 + m;
        const bool ok = r <= q && q < cl;
        return sel(ok, V2p[ok ? pk_off(q) + r : 0]);
      };
      auto y = [&](int r, int c) {
        const int col = cl + c;
        const bool ok = c < 8 && r <= col;
        return sel(ok, V2p[ok ? pk_off(col) + r : 0]);
      };
      const dbl4 g = mfma_tile16_acc4(x, y, 0, min(t0 + 16, cl), lane);
      if (tq < cl) {
        Xg[tq * 8 + q4] = g[0];
        Xg[tq * 8 + q4 + 4] = g[1];
      }
    }
    __syncthreads();
    t12_block(cl);
  }
  __syncthreads();
  if (s_err) {
    if (tid == 0) *err = 1;
    return;
  }
  // ---- V2, T (full b×b, zeros below the diagonal) and tau of this node
  double* V2 = V2out + nid * bb;
  double* Tn = Tout + nid * bb;
  for (int e = tid; e < b * b; e += NT) {
    const int col = e / b, row = e - col * b;
    const bool up = row <= col;
    V2[e] = up ? V2p[pk_off(col) + row] : 0.0;
    Tn[e] = up ? Tp[pk_off(col) + row] : 0.0;
  }
  for (int p = tid; p < b; p += NT) tauout[(int64_t)nid * b + p] = stau[p];
}

template <int BMAX>
__global__ void __launch_bounds__(4 * BMAX)
btree_kernel(const double* __restrict__ leafR, double* __restrict__ nodeR, double* __restrict__ V2out,
             double* __restrict__ Tout, double* __restrict__ tauout, uint64_t* __restrict__ cnt,
             uint64_t epoch, int L, int b, int* __restrict__ err, const BatchProb* __restrict__ bp, const BatchSeg* __restrict__ seg, int nseg) {
  extern __shared__ double smem[];
  btree_body<BMAX>(leafR, nodeR, V2out, Tout, tauout, cnt, epoch, L, b, err, bp, seg, nseg, blockIdx.x, smem, nullptr);
}

// Leaves and tree in one launch (b in (64, 128], leaves of at most 256
// rows): workgroups [0, L) factor the leaves and publish their R eight
// columns at a time, workgroups [L, L + nnodes) are the tree nodes, which
// start on a leaf's first columns while the leaf is still running.  Block
// ids order every producer before its consumers, so dispatch order alone
// guarantees progress.
template <int RPL>
__global__ void __launch_bounds__(512)
leaf_tree_kernel(double* __restrict__ A, int64_t lda, int64_t m, int L, int b, double* __restrict__ leafR,
                 double* __restrict__ leafT, double* __restrict__ leafTau, double* __restrict__ nodeR,
                 double* __restrict__ V2out, double* __restrict__ Tout, double* __restrict__ tauout,
                 uint64_t* __restrict__ cnt, uint64_t* __restrict__ lcnt, uint64_t epoch, int* __restrict__ err) {
  extern __shared__ double smem[];
  if ((int)blockIdx.x < L)
    leaf_qr_body<8, 16, RPL>(A, lda, m, L, b, leafR, leafT, leafTau, nullptr, 0, blockIdx.x, smem, lcnt, epoch);
  else
    btree_body<128>(leafR, nodeR, V2out, Tout, tauout, cnt, epoch, L, b, err, nullptr, nullptr, 0, blockIdx.x - L, smem,
                    lcnt);
}

template <int BMAX>
static int btree_launch(cqr_ctx* c, const double* leafR, double* nodeR, double* V2, double* T, double* tau,
                        uint64_t* cnt, int L, int nnodes, int b, uint64_t epoch, const BatchProb* bp, const BatchSeg* seg, int nseg) {
  const size_t lds = sizeof(double) * btree_smem_doubles<BMAX>();
  auto k = btree_kernel<BMAX>;
  CQR_TRY(set_max_lds(c, (const void*)k, (int)lds));
  hipEvent_t ev = prof_begin(c);
  {
    SpinScope gate(c, (const void*)k, nnodes, 4 * BMAX, lds);
    CQR_TRY(gate.status());
    hipLaunchKernelGGL(k, dim3(nnodes), dim3(4 * BMAX), lds, c->stream, leafR, nodeR, V2, T, tau, cnt, epoch, L, b,
                       c->d_terr, bp, seg, nseg);
  }
  prof_end(c, ev, PK_TREE, nnodes * (2.0 / 3.0) * b * (double)b * b, nnodes * 40.0 * b * b);
  return check_launch(c, "btree");
}

template <int CPW, int RPL>
static int ptree_launch(cqr_ctx* c, const double* leafR, double* nodeR, double* V2, double* T, double* tau,
                        uint64_t* cnt, int L, int nnodes, int b, uint64_t epoch, const BatchProb* bp, const BatchSeg* seg, int nseg) {
  hipEvent_t ev = prof_begin(c);
  {
    SpinScope gate(c, (const void*)ptree_kernel<CPW, RPL>, nnodes, 512, 0);
    CQR_TRY(gate.status());
    hipLaunchKernelGGL((ptree_kernel<CPW, RPL>), dim3(nnodes), dim3(512), 0, c->stream, leafR, nodeR, V2, T, tau,
                       cnt, epoch, L, b, c->d_terr, bp, seg, nseg);
  }
  prof_end(c, ev, PK_TREE, nnodes * (2.0 / 3.0) * b * (double)b * b, nnodes * 40.0 * b * b);
  return check_launch(c, "ptree");
}

static int ptree_dispatch(cqr_ctx* c, const double* leafR, double* nodeR, double* V2, double* T, double* tau,
                          uint64_t* cnt, int L, int nnodes, int b, uint64_t epoch, const BatchProb* bp, const BatchSeg* seg, int nseg) {
  if (b % 8 == 0) {
    if (b <= 32) return btree_launch<32>(c, leafR, nodeR, V2, T, tau, cnt, L, nnodes, b, epoch, bp, seg, nseg);
    if (b <= 64) return btree_launch<64>(c, leafR, nodeR, V2, T, tau, cnt, L, nnodes, b, epoch, bp, seg, nseg);
    if (b <= 128) return btree_launch<128>(c, leafR, nodeR, V2, T, tau, cnt, L, nnodes, b, epoch, bp, seg, nseg);
  }
  if (b <= 16) return ptree_launch<2, 1>(c, leafR, nodeR, V2, T, tau, cnt, L, nnodes, b, epoch, bp, seg, nseg);
  if (b <= 32) return ptree_launch<4, 1>(c, leafR, nodeR, V2, T, tau, cnt, L, nnodes, b, epoch, bp, seg, nseg);
  if (b <= 64) return ptree_launch<8, 1>(c, leafR, nodeR, V2, T, tau, cnt, L, nnodes, b, epoch, bp, seg, nseg);
  return ptree_launch<16, 2>(c, leafR, nodeR, V2, T, tau, cnt, L, nnodes, b, epoch, bp, seg, nseg);
}

template <int RPL>
static int leaf_tree_launch(cqr_ctx* c, double* A, int64_t lda, int64_t m, int L, int b, double* leafR,
                            double* leafT, double* leafTau, double* nodeR, double* V2, double* T, double* tau,
                            int nnodes) {
  const size_t lds = sizeof(double) * std::max<size_t>(btree_smem_doubles<128>(),
                                                      (size_t)b * b + 2 * 64 * RPL + 128);
  auto k = leaf_tree_kernel<RPL>;
  CQR_TRY(set_max_lds(c, (const void*)k, (int)lds));
  const uint64_t epoch = ++c->tree_epoch;
  hipEvent_t ev = prof_begin(c);
  {
    SpinScope gate(c, (const void*)k, L + nnodes, 512, lds);
    CQR_TRY(gate.status());
    hipLaunchKernelGGL(k, dim3(L + nnodes), dim3(512), lds, c->stream, A, lda, m, L, b, leafR, leafT, leafTau, nodeR,
                       V2, T, tau, c->d_cnt, c->d_cnt + CQR_MAX_TREE_NODES, epoch, c->d_terr);
  }
  prof_end(c, ev, PK_TREE,
           2.0 * m * b * b - (2.0 / 3.0) * L * (double)b * b * b + nnodes * (2.0 / 3.0) * b * (double)b * b,
           16.0 * m * b + nnodes * 40.0 * b * b);
  return check_launch(c, "leaf_tree");
}

// Leaf QR of every row block, then the reduction tree; one launch when the
// shapes allow the fused kernel, else the two kernels in stream order.
int leaf_tree(cqr_ctx* c, double* A, int64_t lda, int64_t m, int L, int b, double* leafR, double* leafT,
              double* leafTau, double* nodeR, double* V2, double* T, double* tau, int nnodes) {
  const int64_t rows = (m + L - 1) / L;
  // fused only while every leaf and node workgroup is resident at once (one
  // per CU, for the LDS): beyond that the nodes start only as leaves retire
  // and the two launches in order are faster (65536×2048: 32.6 → 30.4 ms)
  const bool fuse = nnodes > 0 && b > 64 && b <= 128 && b % 8 == 0 && rows <= 256 && L <= CQR_MAX_TREE_NODES &&
                    nnodes <= CQR_MAX_TREE_NODES && L + nnodes <= c->num_cu;
  if (!fuse) {
    CQR_TRY(leaf_qr(c, A, lda, m, L, b, leafR, leafT, leafTau));
    return launch_ptree(c, leafR, nodeR, V2, T, tau, L, nnodes, b);
  }
  if (rows <= 128) return leaf_tree_launch<2>(c, A, lda, m, L, b, leafR, leafT, leafTau, nodeR, V2, T, tau, nnodes);
  return leaf_tree_launch<4>(c, A, lda, m, L, b, leafR, leafT, leafTau, nodeR, V2, T, tau, nnodes);
}

int launch_ptree(cqr_ctx* c, const double* leafR, double* nodeR, double* V2, double* T, double* tau,
                 int L, int nnodes, int b) {
  if (nnodes <= 0) return CQR_OK;
  if (nnodes > CQR_MAX_TREE_NODES) return set_err(c, CQR_ERR_UNSUPPORTED, "tree of %d nodes", nnodes);
  return ptree_dispatch(c, leafR, nodeR, V2, T, tau, c->d_cnt, L, nnodes, b, ++c->tree_epoch, nullptr, nullptr, 0);
}

// Batched: all trees of a group in one launch; cnt (one word per node) must
// hold values below 256 (the batch driver zeroes it with a kernel before
// every run, so the launch is graph-replayable with epoch 1).
int launch_ptree_batched(cqr_ctx* c, const double* leafR, double* nodeR, double* V2, double* T, double* tau,
                         uint64_t* cnt, int total_nodes, int b, const BatchProb* bp, const BatchSeg* seg, int nseg) {
  if (total_nodes <= 0) return CQR_OK;
  return ptree_dispatch(c, leafR, nodeR, V2, T, tau, cnt, 0, total_nodes, b, 1, bp, seg, nseg);
}

}  // namespace cqr
