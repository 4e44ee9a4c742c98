// construct_Q1 in one launch: the explicit first b columns of the TSQR Q
// (alg/QR/tsqr/construct_tsqr_Q.cxx:34-242).
//
// The reference walks the tree from the root down: the root starts from
// C = [I; 0]; every node (receiver z, sender s) turns the b×b block of its
// receiver into [C_z; C_s] = Q_node·[C_z; 0] with Q_node = I − [I;V2]·T·[I;V2]ᵀ
// (the dtprfb of :195-208):
//     X = T·C_z,   C_s = −V2·X,   C_z = C_z − X;
// and every leaf l finally expands its block with its own reflectors (the
// dormqr of :229-232):
//     Q1_l = [C_l; 0] − V_l·(T_l·(V_topᵀ·C_l)).
// Columns never mix, so the work splits into (node, 16-column block) and
// (leaf, 16-column block) items.  Items are numbered root first, then down
// the tree, then the leaves: an item only waits for the item of its parent
// node and column block, which has a smaller block id and is therefore
// dispatched earlier.  Blocks C pass between items through device-coherent
// atomics and a flag per (node, column block) (see cqr_device.h).  Waits
// are bounded; a timeout raises a flag that the host reads at the next
// synchronisation.
//
// All blocks C are upper triangular (the root's identity is, and T, V2 keep
// it so), so column block [j0, j0+16) only has rows below R = j0 + 16.
#include "cqr_device.h"
#include "cqr_internal.h"

namespace cqr {

namespace {

constexpr int CW = 16;     // columns per work item
constexpr int LDS_LD = CW + 1;

// The node consuming rank r's block above level `lv` (lv = −1 for a leaf):
// the first level k > lv at which r is a receiver (node z = r) or a sender
// (node z = r − ⌈n_k/2⌉); −1 when none exists (the root, or a one-leaf tree).
__device__ inline int parent_node(int r, int lv, int L) {
  int n = L, node0 = 0;
  for (int k = 0; n > 1; ++k) {
    const int pairs = n / 2, half = (n + 1) / 2;
    if (k > lv) {
      if (r < pairs) return node0 + r;
      if (r >= half && r < n) return node0 + (r - half);
    }
    node0 += pairs;
    n = half;
  }
  return -1;
}

}  // namespace

template <int NWV>
__global__ void __launch_bounds__(64 * NWV)
q1_kernel(const double* __restrict__ nodeV2, const double* __restrict__ nodeT, const double* __restrict__ leafT,
          const double* __restrict__ V, int64_t ldv, int64_t m, double* __restrict__ C,
          double* __restrict__ Q, int64_t ldq, uint64_t* __restrict__ flags, uint64_t epoch, int L,
          int nnodes, int b, int* __restrict__ err, const BatchProb* __restrict__ bp,
          const BatchSeg* __restrict__ seg, int nseg) {
  __shared__ double S0[128][LDS_LD];
  __shared__ double S1[128][LDS_LD];
  __shared__ int s_ok;

  const int ncb = (b + CW - 1) / CW;
  int item = blockIdx.x;
  if (bp) {  // batched: V is the group's base; the problem's slices of every buffer
    const int64_t bb0 = (int64_t)b * b;
    const BatchSeg S = seg[seg_find(seg, nseg, item)];
    const BatchProb P = bp[S.prob];
    item = S.local0 + (item - S.g0);
    L = P.L;
    nnodes = P.L - 1;
    nodeV2 += (int64_t)P.node0 * bb0;
    nodeT += (int64_t)P.node0 * bb0;
    leafT += (int64_t)P.leaf0 * bb0;
    V += P.aoff;
    ldv = P.m;
    m = P.m;
    C += (int64_t)P.leaf0 * bb0;
    Q += P.yoff;
    ldq = P.m;
    flags += (int64_t)P.node0 * ncb;
  }
  if (!bp) {
    // The leaf items of eight consecutive leaves are dealt so that leaf z's
    // items run on XCD z mod 8 (blocks go to XCD blockIdx mod 8), the XCD
    // whose L2 holds its V and T from the leaf factorization (leaf z is
    // block z there).  A partial last group keeps the linear order.
    const int s0 = nnodes * ncb, rel = item - s0;
    if (rel >= 0) {
      const int grp = rel / (8 * ncb);
      if (8 * (grp + 1) <= L) item = s0 + (8 * grp + (item & 7)) * ncb + (rel - grp * 8 * ncb) / 8;
    }
  }
  const int cb = item % ncb;
  const int j0 = cb * CW, nj = min(CW, b - j0);
  const int R = min(b, j0 + CW);
  constexpr int NT = 64 * NWV, CPT = 16 * 128 / NT;   // threads; C columns loaded per thread
  const int tid = threadIdx.x, i = tid & 127, jq = tid / 128;
  const int64_t bb = (int64_t)b * b;

  const bool is_node = item < nnodes * ncb;
  int z, parent, nid = -1, snd = -1;
  if (is_node) {
    nid = nnodes - 1 - item / ncb;
    int lv = 0, n = L, node0 = 0;  // level, live ranks and first node of nid's level
    while (nid >= node0 + n / 2) {
      node0 += n / 2;
      n = n / 2 + (n & 1);
      ++lv;
    }
    z = nid - node0;
    snd = z + (n + 1) / 2;
    parent = parent_node(z, lv, L);
  } else {
    z = item / ncb - nnodes;
    parent = parent_node(z, -1, L);
  }

  // ---- input block: C_z[:, j0:j0+16] (identity at the root)
  if (tid == 0) s_ok = 1;
  if (parent >= 0 && tid == 0) {
    if (!flag_wait(flags + (int64_t)parent * ncb + cb, epoch, wait_limit(err))) s_ok = 0;
  }
  __syncthreads();
  if (!s_ok) {
    if (tid == 0) *err = 1;
    return;
  }
  double* Cz = C + (int64_t)z * bb;
  {
    // thread (k = tid & 127, group jq) loads CPT columns of row k, all in flight
    double cv[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      const int jj = jq * CPT + j;
      double v = 0.0;
      if (i < R && jj < nj)
        v = parent >= 0 ? load_coherent(Cz + i + (int64_t)(j0 + jj) * b) : (i == j0 + jj ? 1.0 : 0.0);
      cv[j] = v;
    }
#pragma unroll
    for (int j = 0; j < CPT; ++j) S0[i][jq * CPT + j] = cv[j];
  }
  __syncthreads();

  // Products on FP64 MFMA (16×16 tiles, a wave per tile row): the 16-column
  // block is one tile wide; wave w takes tile rows w, w+NWV, ...  Fragment
  // layout as in mfma_tile16: lane l returns row (l & 15), columns
  // (l >> 4) + 4v.
  const int lane = tid & 63, w = tid >> 6;
  const int tr = lane & 15, tc = lane >> 4;
  auto fromS = [&](const double (*S)[LDS_LD], int K) {
    return [=](int k, int c) { return k < K ? S[k][c] : 0.0; };
  };
  const int nrt = (R + 15) / 16;  // tile rows of an R-row result
  if (is_node) {
    const double* Tn = nodeT + (int64_t)nid * bb;
    const double* V2 = nodeV2 + (int64_t)nid * bb;
    // X = T·C_z (T upper: row tile t needs k ≥ 16t)
    for (int t = w; t < nrt; t += NWV) {
      const int m0 = 16 * t;
      auto xT = [&](int r, int k) { return (m0 + r < b && k < R) ? Tn[(m0 + r) + (int64_t)k * b] : 0.0; };
      const dbl4 acc = mfma_tile16_k128(xT, fromS(S0, R), m0, R, lane);
#pragma unroll
      for (int v = 0; v < 4; ++v) S1[m0 + tr][tc + 4 * v] = m0 + tr < R ? acc[v] : 0.0;
    }
    for (int t = nrt + w; t < 8; t += NWV)   // rows ≥ R of X are zero
#pragma unroll
      for (int v = 0; v < 4; ++v) S1[16 * t + tr][tc + 4 * v] = 0.0;
    __syncthreads();
    // C_s = −V2·X, C_z −= X
    double* Cs = C + (int64_t)snd * bb;
    for (int t = w; t * 16 < b; t += NWV) {
      const int m0 = 16 * t;
      dbl4 acc = {0.0, 0.0, 0.0, 0.0};
      if (t < nrt) {
        auto xV = [&](int r, int k) { return (m0 + r < b && k < R) ? V2[(m0 + r) + (int64_t)k * b] : 0.0; };
        acc = mfma_tile16_k128(xV, fromS(S1, R), m0, R, lane);
      }
      const int row = m0 + tr;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int jj = tc + 4 * v;
        if (row < b && jj < nj) {
          store_coherent(Cs + row + (int64_t)(j0 + jj) * b, row < R ? -acc[v] : 0.0);
          store_coherent(Cz + row + (int64_t)(j0 + jj) * b, row < R ? S0[row][jj] - S1[row][jj] : 0.0);
        }
      }
    }
    wait_memory();
    __syncthreads();
    if (tid == 0) flag_store(flags + (int64_t)nid * ncb + cb, epoch);
    return;
  }

  // ---- leaf z: rows [r0, r1) of V = [V\R] (unit lower in its top b×b)
  const int64_t r0 = ((int64_t)z * m) / L, r1 = ((int64_t)(z + 1) * m) / L;
  const double* Vl = V + r0;
  const int64_t ml = r1 - r0;
  // Z = V_topᵀ·C_l: Z[i] = Σ_{i≤k<R} V_top[k][i]·C[k], R rows
  for (int t = w; t < nrt; t += NWV) {
    const int m0 = 16 * t;
    auto xVt = [&](int r, int k) {
      const int col = m0 + r;
      return (k < R && col < b) ? (k > col ? Vl[k + (int64_t)col * ldv] : (k == col ? 1.0 : 0.0)) : 0.0;
    };
    const dbl4 acc = mfma_tile16_k128(xVt, fromS(S0, R), m0, R, lane);
#pragma unroll
    for (int v = 0; v < 4; ++v) S1[m0 + tr][tc + 4 * v] = m0 + tr < R ? acc[v] : 0.0;
  }
  __syncthreads();
  // M = T_l·Z, overwriting Z in S1 once every tile is computed (wave w
  // holds tile rows w, w+NWV, ... of the eight)
  const double* Tl = leafT + (int64_t)z * bb;
  constexpr int TPW = 8 / NWV;
  dbl4 macc[TPW];
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    const int t = w + NWV * q, m0 = 16 * t;
    macc[q] = dbl4{0.0, 0.0, 0.0, 0.0};
    if (t < nrt) {
      auto xT = [&](int r, int k) { return (m0 + r < b && k < R) ? Tl[(m0 + r) + (int64_t)k * b] : 0.0; };
      macc[q] = mfma_tile16_k128(xT, fromS(S1, R), m0, R, lane);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    const int t = w + NWV * q, m0 = 16 * t;
    if (t < nrt)
#pragma unroll
      for (int v = 0; v < 4; ++v) S1[m0 + tr][tc + 4 * v] = m0 + tr < R ? macc[q][v] : 0.0;
  }
  __syncthreads();
  // Q1_l = [C_l; 0] − V_l·M over the leaf rows (V_l unit lower on top)
  for (int64_t t = w; t * 16 < ml; t += NWV) {
    const int64_t m0 = 16 * t;
    auto xV = [&](int r, int k) {
      const int64_t row = m0 + r;
      if (row >= ml || k >= R || k > row) return 0.0;
      return k == row ? 1.0 : Vl[row + (int64_t)k * ldv];
    };
    const int kmax = (int)min<int64_t>(R, m0 + 16);
    const dbl4 acc = mfma_tile16_k128(xV, fromS(S1, R), 0, kmax, lane);
    const int64_t row = m0 + tr;
    if (row < ml) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int jj = tc + 4 * v;
        if (jj < nj) Q[r0 + row + (int64_t)(j0 + jj) * ldq] = (row < R ? S0[row][jj] : 0.0) - acc[v];
      }
    }
  }
}

int launch_q1(cqr_ctx* c, const double* nodeV2, const double* nodeT, const double* leafT, const double* V,
              int64_t ldv, int64_t m, double* C, double* Q, int64_t ldq, int L, int nnodes, int b) {
  if (b > 128 || b < 1) return set_err(c, CQR_ERR_UNSUPPORTED, "construct_Q1: b = %d", b);
  const int ncb = (b + CW - 1) / CW;
  if ((int64_t)nnodes * ncb > CQR_MAX_TREE_NODES * 8)
    return set_err(c, CQR_ERR_UNSUPPORTED, "construct_Q1: tree of %d nodes", nnodes);
  const uint64_t epoch = ++c->q1_epoch;
  hipEvent_t ev = prof_begin(c);
  // Deep trees take eight waves an item (each wave one tile row of a
  // product: the chain of tree levels is latency-bound); shallow ones four,
  // two items a CU, so that the launch needs half the device's room beside
  // other such launches (a served solve's).  The batched launch takes four.
  const int64_t items = (int64_t)(nnodes + L) * ncb;
  const bool deep = L >= 64;
  {
    SpinScope gate(c, deep ? (const void*)q1_kernel<8> : (const void*)q1_kernel<4>, items, deep ? 512 : 256, 0);
    CQR_TRY(gate.status());
    if (deep)
      hipLaunchKernelGGL(q1_kernel<8>, dim3((unsigned)items), dim3(512), 0, c->stream, nodeV2, nodeT, leafT, V, ldv, m,
                         C, Q, ldq, c->d_qflags, epoch, L, nnodes, b, c->d_terr, nullptr, nullptr, 0);
    else
      hipLaunchKernelGGL(q1_kernel<4>, dim3((unsigned)items), dim3(256), 0, c->stream, nodeV2, nodeT, leafT, V, ldv, m,
                         C, Q, ldq, c->d_qflags, epoch, L, nnodes, b, c->d_terr, nullptr, nullptr, 0);
  }
  prof_end(c, ev, PK_BGEMM, 4.0 * nnodes * (double)b * b * b + 4.0 * (double)m * b * b,
           8.0 * (3.0 * nnodes * b * b + 2.0 * m * b + 2.0 * L * b * b));
  return check_launch(c, "q1_kernel");
}

// Batched: every problem of a group, items in the order of the segment
// table (node levels from the top across all problems, then the leaves);
// flags must be zero before the launch (epoch 1, graph-replayable).
int launch_q1_batched(cqr_ctx* c, const double* nodeV2, const double* nodeT, const double* leafT, const double* Vbase,
                      double* C, double* Y, uint64_t* flags, int total_items, int b, const BatchProb* bp,
                      const BatchSeg* seg, int nseg) {
  if (total_items <= 0) return CQR_OK;
  hipEvent_t ev = prof_begin(c);
  {
    SpinScope gate(c, (const void*)q1_kernel<4>, total_items, 256, 0);
    CQR_TRY(gate.status());
    hipLaunchKernelGGL(q1_kernel<4>, dim3((unsigned)total_items), dim3(256), 0, c->stream, nodeV2, nodeT, leafT, Vbase,
                       (int64_t)0, (int64_t)0, C, Y, (int64_t)0, flags, (uint64_t)1, 0, 0, b, c->d_terr, bp, seg, nseg);
  }
  prof_end(c, ev, PK_BGEMM, 0.0, 0.0);
  return check_launch(c, "q1_kernel_batched");
}

int q1_items_per_problem(int L, int b) { return L * ((b + CW - 1) / CW) + (L - 1) * ((b + CW - 1) / CW); }

}  // namespace cqr
