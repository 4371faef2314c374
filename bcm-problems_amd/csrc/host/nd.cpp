// Nested-dissection ordering and the binary-tree block partition of the
// parallel solver (CSMatrix::PartBinaryTreeND, ordmatr.cpp:1698-1805).
//
// The reference orders the symmetrized block pattern with METIS_NodeND
// (ordmatr.cpp:28-85) and then recovers the dissection tree from the ordered
// pattern: CSMatrix::PartBinaryTreeND(CTree&) (ordmatr.cpp:1204-1396) walks
// the binary tree, and at every node FindSeparator (smatrix.cpp:4355-4488)
// locates, in the node's contiguous index range, the split [part 1][part 2]
// [separator] that a nested-dissection order leaves behind. METIS is not part
// of this build, so nd_order() is this library's own dissection with the same
// layout (each part ordered recursively, its separator last): recursive
// bisection of the block graph by a vertex separator taken from the middle
// level of a breadth-first level structure rooted at a pseudo-peripheral
// block row (George's automatic nested dissection), the separator thinned to
// the rows that touch the far side, and parts below a leaf size ordered by
// reverse Cuthill-McKee. find_separator() and part_binary_tree_nd() follow
// the reference step by step, so the tree is the one the reference would
// build from the same ordering.
#include <algorithm>
#include <numeric>

#include "bcm_host.hpp"

namespace bcm {

namespace {

constexpr int kLeafRows = 64;  // parts up to this many block rows are not dissected further

struct Graph {
  const std::vector<int>* ia;
  const std::vector<int>* ja;
};

// Breadth-first levels of the subgraph marked in `in` (in[v] == tag) from
// `root`; returns the vertices level by level in `seq` and the level starts.
void bfs_levels(const Graph& g, const std::vector<int>& in, int tag, int root, std::vector<int>& mark, int stamp,
                std::vector<int>& seq, std::vector<int>& lev) {
  seq.clear();
  lev.clear();
  seq.push_back(root);
  mark[root] = stamp;
  size_t b = 0;
  while (b < seq.size()) {
    lev.push_back((int)b);
    const size_t e = seq.size();
    for (size_t k = b; k < e; ++k) {
      const int v = seq[k];
      for (int p = (*g.ia)[v]; p < (*g.ia)[v + 1]; ++p) {
        const int w = (*g.ja)[p];
        if (in[w] == tag && mark[w] != stamp) {
          mark[w] = stamp;
          seq.push_back(w);
        }
      }
    }
    b = e;
  }
  lev.push_back((int)seq.size());
}

class Dissector {
 public:
  Dissector(int n, const std::vector<int>& ia, const std::vector<int>& ja)
      : g_{&ia, &ja}, in_(n, -1), mark_(n, -1), deg_(n, 0) {
    for (int v = 0; v < n; ++v)
      for (int p = ia[v]; p < ia[v + 1]; ++p) deg_[v] += ja[p] != v ? 1 : 0;
  }

  // Appends the new-to-old sequence of `verts` (connected or not).
  void run(std::vector<int> verts, std::vector<int>& out) {
    if (verts.empty()) return;
    const int tag = next_tag_++;
    for (int v : verts) in_[v] = tag;
    // connected components, each dissected on its own (an empty separator
    // between them), the one holding the lowest row first; stamps only grow,
    // so a row reached by an earlier search of this call has a mark above `start`
    std::sort(verts.begin(), verts.end());
    const int start = stamp_;
    std::vector<std::vector<int>> comps;
    std::vector<int> seq, lev;
    for (int v : verts) {
      if (mark_[v] > start) continue;
      bfs_levels(g_, in_, tag, v, mark_, next_stamp(), seq, lev);
      comps.push_back(seq);
    }
    for (auto& c : comps) dissect_connected(std::move(c), out);
  }

 private:
  Graph g_;
  std::vector<int> in_, mark_, deg_;
  int next_tag_ = 0, stamp_ = 0;

  int next_stamp() { return ++stamp_; }

  // Pseudo-peripheral root (repeat BFS from a vertex of minimum degree in the
  // last level while the eccentricity grows) and its level structure.
  void rooted_levels(const std::vector<int>& verts, int tag, std::vector<int>& seq, std::vector<int>& lev) {
    int root = verts[0];
    for (int v : verts)
      if (deg_[v] < deg_[root]) root = v;
    bfs_levels(g_, in_, tag, root, mark_, next_stamp(), seq, lev);
    for (int it = 0; it < 8; ++it) {
      const int nl = (int)lev.size() - 1;
      int cand = seq[lev[nl - 1]];
      for (int k = lev[nl - 1]; k < lev[nl]; ++k)
        if (deg_[seq[k]] < deg_[cand]) cand = seq[k];
      std::vector<int> s2, l2;
      bfs_levels(g_, in_, tag, cand, mark_, next_stamp(), s2, l2);
      if (l2.size() <= lev.size()) break;
      seq.swap(s2);
      lev.swap(l2);
    }
  }

  void order_leaf(const std::vector<int>& verts, std::vector<int>& out) {
    // reverse Cuthill-McKee of the leaf's induced subgraph
    const int m = (int)verts.size();
    std::vector<int> xadj(m + 1, 0), adj;
    // local numbering by ascending global index (kept in mark_ as -2 - k)
    std::vector<int> sorted = verts;
    std::sort(sorted.begin(), sorted.end());
    const int tag = next_tag_++;
    for (int k = 0; k < m; ++k) {
      in_[sorted[k]] = tag;
      mark_[sorted[k]] = -2 - k;  // local index, encoded
    }
    for (int k = 0; k < m; ++k) {
      const int v = sorted[k];
      for (int p = (*g_.ia)[v]; p < (*g_.ia)[v + 1]; ++p) {
        const int w = (*g_.ja)[p];
        if (in_[w] == tag) adj.push_back(-2 - mark_[w]);
      }
      xadj[k + 1] = (int)adj.size();
    }
    std::vector<int> perm;
    genrcm(m, xadj, adj, perm);
    for (int k = 0; k < m; ++k) out.push_back(sorted[perm[k]]);
  }

  void dissect_connected(std::vector<int> verts, std::vector<int>& out) {
    if ((int)verts.size() <= kLeafRows) {
      order_leaf(verts, out);
      return;
    }
    const int tag = next_tag_++;
    for (int v : verts) in_[v] = tag;
    std::vector<int> seq, lev;
    rooted_levels(verts, tag, seq, lev);
    const int nl = (int)lev.size() - 1;
    if (nl < 3) {
      order_leaf(verts, out);
      return;
    }
    // separator level: the first whose preceding levels hold half the rows
    const int half = (int)verts.size() / 2;
    int m = 1;
    while (m < nl - 2 && lev[m + 1] <= half) ++m;
    std::vector<int> p1(seq.begin(), seq.begin() + lev[m]);
    std::vector<int> sep, p2(seq.begin() + lev[m + 1], seq.end());
    // thin: a separator row with no neighbour beyond the separator joins part 1
    const int far = next_tag_++;
    for (int v : p2) in_[v] = far;
    for (int k = lev[m]; k < lev[m + 1]; ++k) {
      const int v = seq[k];
      bool touches = false;
      for (int p = (*g_.ia)[v]; p < (*g_.ia)[v + 1] && !touches; ++p) touches = in_[(*g_.ja)[p]] == far;
      (touches ? sep : p1).push_back(v);
    }
    run(std::move(p1), out);
    run(std::move(p2), out);
    std::sort(sep.begin(), sep.end());
    for (int v : sep) out.push_back(v);
  }
};

}  // namespace

std::vector<int> nd_order(const BlockCsr& a) {
  std::vector<int> ias, jas;
  symm_pattern(a, ias, jas);
  const int n = a.nsup;
  std::vector<int> seq;
  seq.reserve(n);
  Dissector d(n, ias, jas);
  std::vector<int> all(n);
  std::iota(all.begin(), all.end(), 0);
  d.run(std::move(all), seq);
  if ((int)seq.size() != n) throw Error(kErrInternal, "nested dissection lost rows");
  std::vector<int> order(n, -1);
  for (int k = 0; k < n; ++k) order[seq[k]] = k;
  return order;
}

void find_separator(int n, const int* ia, const int* ja, int& ind1, int& ind2) {
  // smatrix.cpp:4355-4488, step by step
  ind1 = -1;
  ind2 = -1;
  int i1b = (1 * n) / 4, i1e = (3 * n) / 4;
  const int i2b = i1e + 1, i2e = n;
  if (i1b >= i1e || i2b >= i2e || i1e >= i2b) return;
  auto rows_reach_below = [&](int from, int to, int lim) {  // some row in [from, to) has a column < lim
    for (int i = from; i < to; ++i)
      for (int j = ia[i]; j < ia[i + 1]; ++j)
        if (ja[j] < lim) return true;
    return false;
  };
  if (rows_reach_below(i1e, i2b, i1b)) return;
  int i1en = i1e;  // the bound moves while the loop runs
  for (int i = i2b - 1; i >= i1en; --i)
    for (int j = ia[i]; j < ia[i + 1]; ++j)
      if (ja[j] < i1en - 1) i1en = ja[j] + 1;
  i1e = i1en;
  if (rows_reach_below(i1e, i2b, i1b)) return;
  int i1bn = i1b;
  for (int i = 0; i < i1bn; ++i)
    for (int j = ia[i]; j < ia[i + 1]; ++j)
      if (ja[j] < i2b && ja[j] > i1bn) i1bn = ja[j];
  i1b = i1bn;
  if (rows_reach_below(i1e, i2b, i1b)) return;
  while (i1b != i1e) {
    const int chk = i1e == i1b + 1 ? i1b : (i1b + i1e) / 2;
    const bool ok = !rows_reach_below(chk, i2b, chk);
    if (ok) {
      if (i1e != i1b + 1) i1b = chk;
      else i1e = chk;
    } else {
      i1e = chk;
    }
  }
  ind1 = i1b;
  ind2 = n;
  for (int i = 0; i < ind1; ++i)
    for (int j = ia[i]; j < ia[i + 1]; ++j)
      if (ja[j] > ind1 + 1 && ja[j] < ind2) ind2 = ja[j];
}

NdTree part_binary_tree_nd(const BlockCsr& ordered, int nproc) {
  if (nproc < 1 || (nproc & (nproc - 1))) throw Error(kErrArg, "PartBinaryTreeND: nproc must be a power of two");
  std::vector<int> ias, jas;
  symm_pattern(ordered, ias, jas);
  NdTree t;
  t.nproc = nproc;
  // node ranges of the binary tree, depth first (ordmatr.cpp:1204-1396): a
  // node splits its range into [child 0][child 1][its own separator]
  struct Node {
    int beg, end, cpu0, ncpu;
  };
  std::vector<Node> stack = {{0, ordered.nsup - 1, 0, nproc}};
  std::vector<std::pair<int, int>> blocks;  // (first row, cpu) of every block
  std::vector<int> lia, lja;
  while (!stack.empty()) {
    Node nd = stack.back();
    stack.pop_back();
    if (nd.ncpu == 1) {
      if (nd.end >= nd.beg) blocks.push_back({nd.beg, nd.cpu0});
      t.leaf_rows.push_back({nd.beg, nd.end + 1});
      continue;
    }
    // the node's submatrix of the symmetrized pattern, renumbered from 0
    const int nloc = nd.end - nd.beg + 1;
    lia.assign(std::max(nloc, 0) + 1, 0);
    lja.clear();
    for (int i = nd.beg; i <= nd.end; ++i) {
      for (int j = ias[i]; j < ias[i + 1]; ++j)
        if (jas[j] >= nd.beg && jas[j] <= nd.end) lja.push_back(jas[j] - nd.beg);
      lia[i - nd.beg + 1] = (int)lja.size();
    }
    int s1 = -1, s2 = -1;
    if (nloc > 0) find_separator(nloc, lia.data(), lja.data(), s1, s2);
    if (s1 == -1) s1 = s2 = 0;
    if (nd.end >= nd.beg + s2) blocks.push_back({nd.beg + s2, nd.cpu0});
    t.separators.push_back({nd.beg + s2, nd.end + 1});
    const int h = nd.ncpu / 2;
    stack.push_back({nd.beg + s1, nd.beg + s2 - 1, nd.cpu0 + h, h});
    stack.push_back({nd.beg, nd.beg + s1 - 1, nd.cpu0, h});
  }
  std::sort(blocks.begin(), blocks.end());
  t.blks.push_back(0);
  for (size_t k = 0; k < blocks.size(); ++k) {
    t.blks.push_back(k + 1 < blocks.size() ? blocks[k + 1].first : ordered.nsup);
    t.blk2cpu.push_back(blocks[k].second);
  }
  return t;
}

}  // namespace bcm
