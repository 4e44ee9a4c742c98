// Host-only check of the batched dispatch order (candmc_amd/csrc/batch.cpp,
// build_order): the tree launch and the construct_Q1 launch map their
// blocks through segment tables, and a block may only wait on blocks with
// smaller ids (the hardware dispatches blocks in id order, so a wait on a
// later block could never end).  For groups of problems with assorted leaf
// counts this checks that every (problem, node) and (problem, item) gets
// exactly one block, that the segments tile the block range, and that every
// dependency comes earlier: a tree node after the nodes producing its two
// inputs, a construct_Q1 item after its parent node's item of the same
// column block; and that the Y product's row tiles cover every problem's
// rows below its top b once.  The dependencies are derived here from the binary tree's
// pairing (receiver z, sender z + ceil(n/2) on a level of n live ranks,
// alg/QR/tsqr/bitree_tsqr.cxx:172-208).  Built by tests/test_batch_order.py.
#include "../../candmc_amd/csrc/batch.cpp"

#include <cstdio>
#include <cstdlib>
#include <map>
#include <utility>

namespace {

int fails = 0;
#define CHECK(cond, ...)                 \
  do {                                   \
    if (!(cond)) {                       \
      std::printf("FAIL: " __VA_ARGS__); \
      std::printf("\n");                 \
      ++fails;                           \
    }                                    \
  } while (0)

int at(std::map<std::pair<int, int>, int>& m, int p, int i) {
  auto f = m.find(std::make_pair(p, i));
  return f == m.end() ? 1 << 30 : f->second;
}

struct Tree {
  // level k: pairs[k] nodes from node0[k]; live ranks n[k]
  std::vector<int> node0, pairs, n;
  explicit Tree(int L) {
    for (int nn = L, first = 0; nn > 1; nn = nn / 2 + nn % 2) {
      node0.push_back(first);
      pairs.push_back(nn / 2);
      n.push_back(nn);
      first += nn / 2;
    }
  }
  int levels() const { return (int)node0.size(); }
  // node holding rank r's R when entering level k (-1: still the leaf)
  int producer(int r, int k) const {
    int best = -1;
    for (int j = 0; j < k; ++j)
      if (r < pairs[j]) best = node0[j] + r;
    return best;
  }
  // first node after level `lv` (lv = -1: the leaf) that consumes rank r
  int consumer(int r, int lv) const {
    for (int k = lv + 1; k < levels(); ++k) {
      const int half = (n[k] + 1) / 2;
      if (r < pairs[k]) return node0[k] + r;
      if (r >= half && r < n[k]) return node0[k] + (r - half);
    }
    return -1;
  }
  int level_of(int nid) const {
    int k = 0;
    while (nid >= node0[k] + pairs[k]) ++k;
    return k;
  }
};

void check_group(const std::vector<int>& Ls, int b) {
  Group g;
  g.b = b;
  int leaf0 = 0, node0 = 0, item0 = 0;
  for (int L : Ls) {
    BatchProb P{};
    P.L = L;
    P.m = 100 * L + b - 50;   // some problems of one panel (m ≤ b), most taller
    P.leaf0 = leaf0;
    P.node0 = node0;
    P.item0 = item0;
    g.hp.push_back(P);
    leaf0 += L;
    node0 += L - 1;
    item0 += q1_items_per_problem(L, b);
  }
  build_order(g);
  const int ncb = (b + 15) / 16, np = (int)Ls.size();
  // tree launch
  std::map<std::pair<int, int>, int> tblock;   // (problem, node) -> block id
  int next = 0;
  for (int s = 0; s < g.nseg_tree; ++s) {
    const BatchSeg& S = g.hseg[s];
    CHECK(S.g0 == next, "tree segment %d starts at %d, expected %d", s, S.g0, next);
    for (int j = 0; j < S.n; ++j) {
      const bool fresh = tblock.emplace(std::make_pair(S.prob, S.local0 + j), S.g0 + j).second;
      CHECK(fresh, "tree node (%d, %d) dispatched twice", S.prob, S.local0 + j);
    }
    next = S.g0 + S.n;
  }
  CHECK(next == node0, "tree blocks %d, nodes %d", next, node0);
  for (int p = 0; p < np; ++p) {
    const Tree t(Ls[p]);
    for (int nid = 0; nid < Ls[p] - 1; ++nid) {
      auto it = tblock.find(std::make_pair(p, nid));
      CHECK(it != tblock.end(), "tree node (%d, %d) never dispatched", p, nid);
      if (it == tblock.end()) continue;
      const int k = t.level_of(nid), z = nid - t.node0[k], snd = z + (t.n[k] + 1) / 2;
      for (int r : {z, snd}) {
        const int pr = t.producer(r, k);
        if (pr >= 0) CHECK(at(tblock, p, pr) < it->second, "tree node (%d, %d) waits on a later block", p, nid);
      }
    }
  }
  // construct_Q1 launch
  std::map<std::pair<int, int>, int> qblock;   // (problem, local item) -> block id
  next = 0;
  for (int s = g.nseg_tree; s < g.nseg_tree + g.nseg_q1; ++s) {
    const BatchSeg& S = g.hseg[s];
    CHECK(S.g0 == next, "q1 segment %d starts at %d, expected %d", s, S.g0, next);
    for (int j = 0; j < S.n; ++j) {
      const bool fresh = qblock.emplace(std::make_pair(S.prob, S.local0 + j), S.g0 + j).second;
      CHECK(fresh, "q1 item (%d, %d) dispatched twice", S.prob, S.local0 + j);
    }
    next = S.g0 + S.n;
  }
  CHECK(next == item0, "q1 blocks %d, items %d", next, item0);
  // Y product: every row tile below the top b rows of every problem, once
  std::map<std::pair<int, int>, int> yblock;
  next = 0;
  for (int s = g.nseg_tree + g.nseg_q1; s < (int)g.hseg.size(); ++s) {
    const BatchSeg& S = g.hseg[s];
    CHECK(S.g0 == next, "Y segment %d starts at %d, expected %d", s, S.g0, next);
    for (int j = 0; j < S.n; ++j)
      CHECK(yblock.emplace(std::make_pair(S.prob, S.local0 + j), S.g0 + j).second, "Y tile dispatched twice");
    next = S.g0 + S.n;
  }
  CHECK(g.nseg_tree + g.nseg_q1 + g.nseg_y == (int)g.hseg.size(), "segment counts do not add up");
  CHECK(next == g.ytiles, "Y blocks %d, tiles %d", next, g.ytiles);
  int64_t want = 0;
  for (const BatchProb& P : g.hp) {
    const int64_t rows = P.m - b, t = rows > 0 ? (rows + batched_y_tile_rows() - 1) / batched_y_tile_rows() : 0;
    want += t;
    for (int j = 0; j < t; ++j)
      CHECK(yblock.count(std::make_pair((int)(&P - g.hp.data()), j)) == 1, "Y tile missing");
  }
  CHECK(want == g.ytiles, "Y tiles %d, expected %lld", g.ytiles, (long long)want);
  for (int p = 0; p < np; ++p) {
    const Tree t(Ls[p]);
    const int nn = Ls[p] - 1;
    // local item numbering: node nid's items at (nn - 1 - nid)·ncb + cb,
    // leaf z's at (nn + z)·ncb + cb
    auto node_item = [&](int nid, int cb) { return (nn - 1 - nid) * ncb + cb; };
    for (int it = 0; it < q1_items_per_problem(Ls[p], b); ++it) {
      auto f = qblock.find(std::make_pair(p, it));
      CHECK(f != qblock.end(), "q1 item (%d, %d) never dispatched", p, it);
      if (f == qblock.end()) continue;
      const int cb = it % ncb;
      int parent;
      if (it < nn * ncb) {
        const int nid = nn - 1 - it / ncb, k = t.level_of(nid);
        parent = t.consumer(nid - t.node0[k], k);
      } else {
        parent = t.consumer(it / ncb - nn, -1);
      }
      if (parent >= 0)
        CHECK(at(qblock, p, node_item(parent, cb)) < f->second, "q1 item (%d, %d) waits on a later block", p, it);
    }
  }
}

}  // namespace

int main() {
  check_group({1}, 16);
  check_group({2, 1, 3}, 32);
  check_group({64, 7, 33, 1, 64, 2, 16}, 128);
  std::vector<int> mixed;
  for (int i = 0; i < 200; ++i) mixed.push_back(1 + (i * 37) % 65);
  check_group(mixed, 64);
  check_group(mixed, 48);
  if (fails) {
    std::printf("%d failures\n", fails);
    return 1;
  }
  std::printf("batch order ok\n");
  return 0;
}
