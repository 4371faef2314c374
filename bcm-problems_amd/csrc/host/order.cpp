// Ordering and structural transforms of block-CSR matrices.
//
// Reference: ordmatr.cpp:28-227 (OrderPrfMtr), rcm.cpp (genrcm/fnroot/rcm,
// the SPARSPAK reverse Cuthill-McKee of George & Liu 1981), smatrix.cpp:1538
// (SymmMtr), smatrixrs.cpp:345 (OrdMtr), :456 (OrdMtrSymm), :1645-1803
// (CombineLU), :2758 (ExtendSparsity), :2955 (TranspMtr), svector.cpp:418
// (OrdVect).
#include "bcm_host.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>

namespace bcm {

void BlockCsr::set_blamx() {
  blamx = 0;
  for (int i = 0; i < nsup; ++i) blamx = std::max(blamx, bla(i));
}

void BlockCsr::validate(bool require_diag_first) const {
  if ((int)sprnd.size() != nsup + 1 || (int)ia.size() != nsup + 1)
    throw Error(kErrArg, "block matrix: sprnd/ia size mismatch");
  if ((int)ja.size() != nzja() || (int)bsa.size() != nzja())
    throw Error(kErrArg, "block matrix: ja/bsa size mismatch");
  int64_t off = 0;
  for (int i = 0; i < nsup; ++i) {
    if (ia[i + 1] < ia[i]) throw Error(kErrArg, "block matrix: ia not monotone");
    for (int j = ia[i]; j < ia[i + 1]; ++j) {
      int c = ja[j];
      if (c < 0 || c >= nsup) throw Error(kErrArg, "block matrix: column out of range");
      if (j > ia[i] && c <= ja[j - 1]) throw Error(kErrArg, "block matrix: columns not ascending");
      if (bsa[j] != off) throw Error(kErrArg, "block matrix: bsa not contiguous");
      off += (int64_t)bla(i) * bla(c);
    }
    if (require_diag_first && (ia[i] == ia[i + 1] || ja[ia[i]] != i))
      throw Error(kErrArg, "block matrix: diagonal block must lead every row");
  }
  if (off != nza()) throw Error(kErrArg, "block matrix: value count mismatch");
}

// ---------------------------------------------------------------------------
// Pattern helpers

static void transpose_pattern(int n, const std::vector<int>& ia, const std::vector<int>& ja,
                              std::vector<int>& iat, std::vector<int>& jat) {
  iat.assign(n + 1, 0);
  for (int c : ja) iat[c + 1]++;
  for (int i = 0; i < n; ++i) iat[i + 1] += iat[i];
  jat.resize(ja.size());
  std::vector<int> pos(iat.begin(), iat.end() - 1);
  for (int i = 0; i < n; ++i)
    for (int j = ia[i]; j < ia[i + 1]; ++j) jat[pos[ja[j]]++] = i;
}

void symm_pattern(const BlockCsr& a, std::vector<int>& ias, std::vector<int>& jas) {
  // Row i of the result is the ascending union of row i of A and row i of A^T
  // (SymmMtr merges the two sorted lists; the diagonal stays in).
  std::vector<int> iat, jat;
  transpose_pattern(a.nsup, a.ia, a.ja, iat, jat);
  ias.assign(a.nsup + 1, 0);
  jas.clear();
  jas.reserve(2 * a.ja.size() + a.nsup);
  for (int i = 0; i < a.nsup; ++i) {
    int p = a.ia[i], pe = a.ia[i + 1], q = iat[i], qe = iat[i + 1];
    while (p < pe || q < qe) {
      if (q >= qe || (p < pe && a.ja[p] < jat[q])) jas.push_back(a.ja[p++]);
      else if (p >= pe || jat[q] < a.ja[p]) jas.push_back(jat[q++]);
      else { jas.push_back(a.ja[p]); ++p; ++q; }
    }
    ias[i + 1] = (int)jas.size();
  }
}

// ---------------------------------------------------------------------------
// Reverse Cuthill-McKee (SPARSPAK GENRCM), 0-based. `avail[v] != 0` marks
// nodes not yet numbered.

namespace {

struct RcmWork {
  const int* xadj;
  const int* adj;
  std::vector<int> avail;
  std::vector<int> xls;
  std::vector<int> deg;
  std::vector<char> seen;
};

// Breadth-first level structure rooted at `root` over available nodes; the
// nodes are listed level by level in ls, level k occupying [xls[k], xls[k+1]).
void level_structure(RcmWork& w, int root, int* ls, int& nlvl) {
  w.avail[root] = 0;
  ls[0] = root;
  nlvl = 0;
  int end = 0, count = 1;
  while (true) {
    int begin = end;
    end = count;
    w.xls[nlvl++] = begin;
    for (int i = begin; i < end; ++i) {
      int v = ls[i];
      for (int e = w.xadj[v]; e < w.xadj[v + 1]; ++e) {
        int u = w.adj[e];
        if (w.avail[u]) { ls[count++] = u; w.avail[u] = 0; }
      }
    }
    if (count == end) break;
  }
  w.xls[nlvl] = end;
  for (int i = 0; i < count; ++i) w.avail[ls[i]] = 1;
}

// Pseudo-peripheral root: repeatedly restart from the lowest-degree node of
// the deepest level until the level count stops growing.
void pseudo_peripheral(RcmWork& w, int& root, int* ls, int& nlvl) {
  level_structure(w, root, ls, nlvl);
  const int size = w.xls[nlvl];
  if (nlvl == 1 || nlvl == size) return;
  while (true) {
    int best = size;
    int first = w.xls[nlvl - 1];
    root = ls[first];
    if (size - 1 > first) {
      for (int i = first; i < size; ++i) {
        int v = ls[i], d = 0;
        for (int e = w.xadj[v]; e < w.xadj[v + 1]; ++e) d += w.avail[w.adj[e]] > 0;
        if (d < best) { root = v; best = d; }
      }
    }
    int trial;
    level_structure(w, root, ls, trial);
    if (trial <= nlvl) return;
    nlvl = trial;
    if (nlvl >= size) return;
  }
}

// Numbers the component of `root` into perm[0..size) in reverse Cuthill-McKee order.
void number_component(RcmWork& w, int root, int* perm, int& size) {
  // Degrees count the still-available adjacency entries (the self loop included).
  std::vector<int> queue{root};
  w.seen[root] = 1;
  for (size_t h = 0; h < queue.size(); ++h) {
    int v = queue[h], d = 0;
    for (int e = w.xadj[v]; e < w.xadj[v + 1]; ++e) {
      int u = w.adj[e];
      if (!w.avail[u]) continue;
      ++d;
      if (!w.seen[u]) { w.seen[u] = 1; queue.push_back(u); }
    }
    w.deg[v] = d;
  }
  for (int v : queue) w.seen[v] = 0;
  size = (int)queue.size();
  perm[0] = root;
  w.avail[root] = 0;
  if (size <= 1) return;
  int end = 0, filled = 1;
  do {
    int begin = end;
    end = filled;
    for (int i = begin; i < end; ++i) {
      int v = perm[i];
      int first = filled;
      for (int e = w.xadj[v]; e < w.xadj[v + 1]; ++e) {
        int u = w.adj[e];
        if (w.avail[u]) { w.avail[u] = 0; perm[filled++] = u; }
      }
      // The reference's insertion sort (rcm.cpp:180-190) never moves an
      // element in front of the first newly numbered neighbour: that one keeps
      // its place and the later ones are stably sorted by degree among themselves.
      for (int k = first + 2; k < filled; ++k) {
        int u = perm[k];
        int l = k;
        while (l > first + 1 && w.deg[perm[l - 1]] > w.deg[u]) { perm[l] = perm[l - 1]; --l; }
        perm[l] = u;
      }
    }
  } while (filled > end);
  std::reverse(perm, perm + size);
}

}  // namespace

void genrcm(int n, const std::vector<int>& xadj, const std::vector<int>& adj, std::vector<int>& perm) {
  RcmWork w;
  w.xadj = xadj.data();
  w.adj = adj.data();
  w.avail.assign(n, 1);
  w.xls.assign(n + 1, 0);
  w.deg.assign(n, 0);
  w.seen.assign(n, 0);
  perm.assign(n, 0);
  int num = 0;
  for (int i = 0; i < n; ++i) {
    if (!w.avail[i]) continue;
    int root = i, nlvl = 0, size = 0;
    pseudo_peripheral(w, root, perm.data() + num, nlvl);
    number_component(w, root, perm.data() + num, size);
    num += size;
    if (num >= n) return;
  }
}

std::vector<int> rcm_order(const BlockCsr& a) {
  std::vector<int> ias, jas, perm;
  symm_pattern(a, ias, jas);
  genrcm(a.nsup, ias, jas, perm);
  std::vector<int> order(a.nsup);
  for (int k = 0; k < a.nsup; ++k) order[perm[k]] = k;
  return order;
}

// ---------------------------------------------------------------------------
// Structural transforms with values

BlockCsr ord_mtr(const BlockCsr& a, const std::vector<int>& order) {
  const int n = a.nsup;
  std::vector<int> iord(n);
  for (int i = 0; i < n; ++i) iord[order[i]] = i;
  BlockCsr r;
  r.nsup = n;
  r.sprnd.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) r.sprnd[order[i] + 1] = a.bla(i);
  for (int i = 0; i < n; ++i) r.sprnd[i + 1] += r.sprnd[i];
  r.ia.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) r.ia[order[i] + 1] = a.ia[i + 1] - a.ia[i];
  for (int i = 0; i < n; ++i) r.ia[i + 1] += r.ia[i];
  r.ja.resize(a.ja.size());
  r.bsa.resize(a.ja.size());
  r.a.resize(a.a.size());
  std::vector<std::pair<int, int>> row;
  int64_t nz = 0;
  for (int inew = 0; inew < n; ++inew) {
    int i = iord[inew];
    row.clear();
    for (int j = a.ia[i]; j < a.ia[i + 1]; ++j) row.emplace_back(order[a.ja[j]], j);
    std::sort(row.begin(), row.end());
    int bi = a.bla(i);
    int p = r.ia[inew];
    for (auto& e : row) {
      int bj = r.bla(e.first);
      r.ja[p] = e.first;
      r.bsa[p] = nz;
      std::memcpy(&r.a[nz], &a.a[a.bsa[e.second]], sizeof(double) * bi * bj);
      nz += (int64_t)bi * bj;
      ++p;
    }
  }
  r.set_blamx();
  return r;
}

BlockCsr transp_mtr(const BlockCsr& a) {
  const int n = a.nsup;
  BlockCsr t;
  t.nsup = n;
  t.sprnd = a.sprnd;
  transpose_pattern(n, a.ia, a.ja, t.ia, t.ja);
  t.bsa.resize(t.ja.size());
  int64_t nz = 0;
  for (int i = 0; i < n; ++i)
    for (int j = t.ia[i]; j < t.ia[i + 1]; ++j) {
      t.bsa[j] = nz;
      nz += (int64_t)t.bla(i) * t.bla(t.ja[j]);
    }
  t.a.resize(nz);
  std::vector<int> pos(t.ia.begin(), t.ia.end() - 1);
  for (int i = 0; i < n; ++i) {
    int bi = a.bla(i);
    for (int j = a.ia[i]; j < a.ia[i + 1]; ++j) {
      int c = a.ja[j], bj = a.bla(c);
      int k = pos[c]++;
      const double* src = &a.a[a.bsa[j]];
      double* dst = &t.a[t.bsa[k]];
      for (int r = 0; r < bi; ++r)
        for (int cc = 0; cc < bj; ++cc) dst[r * bj + cc] = src[cc * bi + r];
    }
  }
  t.set_blamx();
  return t;
}

BlockCsr symm_mtr(const BlockCsr& up) {
  BlockCsr t = transp_mtr(up);
  const int n = up.nsup;
  BlockCsr s;
  s.nsup = n;
  s.sprnd = up.sprnd;
  s.ia.assign(n + 1, 0);
  s.ja.reserve(up.ja.size() + t.ja.size());
  s.bsa.reserve(up.ja.size() + t.ja.size());
  s.a.reserve(up.a.size() + t.a.size());
  int64_t nz = 0;
  for (int i = 0; i < n; ++i) {
    int p = up.ia[i], pe = up.ia[i + 1], q = t.ia[i], qe = t.ia[i + 1];
    int bi = up.bla(i);
    auto push = [&](int c, const double* src) {
      s.ja.push_back(c);
      s.bsa.push_back(nz);
      int64_t sz = (int64_t)bi * up.bla(c);
      s.a.insert(s.a.end(), src, src + sz);
      nz += sz;
    };
    while (p < pe || q < qe) {
      if (q >= qe || (p < pe && up.ja[p] < t.ja[q])) { push(up.ja[p], &up.a[up.bsa[p]]); ++p; }
      else if (p >= pe || t.ja[q] < up.ja[p]) { push(t.ja[q], &t.a[t.bsa[q]]); ++q; }
      else { push(up.ja[p], &up.a[up.bsa[p]]); ++p; ++q; }
    }
    s.ia[i + 1] = (int)s.ja.size();
  }
  s.set_blamx();
  return s;
}

BlockCsr upper_part(const BlockCsr& a) {
  BlockCsr u;
  u.nsup = a.nsup;
  u.sprnd = a.sprnd;
  u.ia.assign(a.nsup + 1, 0);
  u.ja.reserve(a.ja.size());
  u.bsa.reserve(a.ja.size());
  u.a.reserve(a.a.size());
  int64_t nz = 0;
  for (int i = 0; i < a.nsup; ++i) {
    for (int j = a.ia[i]; j < a.ia[i + 1]; ++j) {
      int c = a.ja[j];
      if (c < i) continue;
      int64_t sz = (int64_t)a.bla(i) * a.bla(c);
      u.ja.push_back(c);
      u.bsa.push_back(nz);
      u.a.insert(u.a.end(), a.a.begin() + a.bsa[j], a.a.begin() + a.bsa[j] + sz);
      nz += sz;
    }
    u.ia[i + 1] = (int)u.ja.size();
  }
  u.set_blamx();
  return u;
}

// Same result as upper_part(ord_mtr(symm_mtr(up), order)) -- the
// reference's OrdMtrSymm -- in one pass: stored block (i, j) moves to
// (order[i], order[j]) when that stays in the upper triangle, otherwise its
// transpose moves to (order[j], order[i]).
BlockCsr ord_mtr_symm(const BlockCsr& up, const std::vector<int>& order) {
  const int n = up.nsup;
  BlockCsr r;
  r.nsup = n;
  r.sprnd.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) r.sprnd[order[i] + 1] = up.bla(i);
  for (int i = 0; i < n; ++i) r.sprnd[i + 1] += r.sprnd[i];
  struct Ent {
    int col;
    int src;
    bool tr;
  };
  std::vector<int> cnt(n + 1, 0);
  for (int i = 0; i < n; ++i)
    for (int j = up.ia[i]; j < up.ia[i + 1]; ++j) cnt[std::min(order[i], order[up.ja[j]]) + 1]++;
  r.ia.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) r.ia[i + 1] = r.ia[i] + cnt[i + 1];
  std::vector<Ent> ent(r.ia[n]);
  std::vector<int> pos(r.ia.begin(), r.ia.end() - 1);
  for (int i = 0; i < n; ++i)
    for (int j = up.ia[i]; j < up.ia[i + 1]; ++j) {
      const int p = order[i], q = order[up.ja[j]];
      if (p <= q) ent[pos[p]++] = Ent{q, j, false};
      else ent[pos[q]++] = Ent{p, j, true};
    }
  r.ja.resize(ent.size());
  r.bsa.resize(ent.size());
  int64_t nz = 0;
  for (int i = 0; i < n; ++i) {
    std::sort(ent.begin() + r.ia[i], ent.begin() + r.ia[i + 1],
              [](const Ent& a, const Ent& b) { return a.col < b.col; });
    for (int k = r.ia[i]; k < r.ia[i + 1]; ++k) {
      r.ja[k] = ent[k].col;
      r.bsa[k] = nz;
      nz += (int64_t)r.bla(i) * r.bla(ent[k].col);
    }
  }
  r.a.resize(nz);
  for (int i = 0; i < n; ++i)
    for (int k = r.ia[i]; k < r.ia[i + 1]; ++k) {
      const Ent& e = ent[k];
      const int bi = r.bla(i), bj = r.bla(e.col);
      const double* src = &up.a[up.bsa[e.src]];
      double* dst = &r.a[r.bsa[k]];
      if (!e.tr) {
        std::memcpy(dst, src, sizeof(double) * bi * bj);
      } else {
        // source block: bj x bi (column-major); dst = its transpose, bi x bj
        for (int rr = 0; rr < bj; ++rr)
          for (int cc = 0; cc < bi; ++cc) dst[rr * bi + cc] = src[cc * bj + rr];
      }
    }
  r.set_blamx();
  return r;
}

BlockCsr extend_sparsity(const BlockCsr& a, const BlockCsr& b) {
  if (a.nsup != b.nsup) throw Error(kErrArg, "extend_sparsity: block counts differ");
  BlockCsr r;
  r.nsup = a.nsup;
  r.sprnd = a.sprnd;
  r.ia.assign(a.nsup + 1, 0);
  int64_t nz = 0;
  for (int i = 0; i < a.nsup; ++i) {
    int p = a.ia[i], pe = a.ia[i + 1], q = b.ia[i], qe = b.ia[i + 1];
    int bi = a.bla(i);
    auto push = [&](int c, const double* src) {
      int64_t sz = (int64_t)bi * a.bla(c);
      r.ja.push_back(c);
      r.bsa.push_back(nz);
      if (src) r.a.insert(r.a.end(), src, src + sz);
      else r.a.insert(r.a.end(), sz, 0.0);
      nz += sz;
    };
    while (p < pe || q < qe) {
      if (q >= qe || (p < pe && a.ja[p] < b.ja[q])) { push(a.ja[p], &a.a[a.bsa[p]]); ++p; }
      else if (p >= pe || b.ja[q] < a.ja[p]) { push(b.ja[q], nullptr); ++q; }
      else { push(a.ja[p], &a.a[a.bsa[p]]); ++p; ++q; }
    }
    r.ia[i + 1] = (int)r.ja.size();
  }
  r.set_blamx();
  return r;
}

void combine_lu(const BlockCsr& a, BlockCsr& l, BlockCsr& u) {
  BlockCsr up = upper_part(a);
  BlockCsr lo;
  lo.nsup = a.nsup;
  lo.sprnd = a.sprnd;
  lo.ia.assign(a.nsup + 1, 0);
  int64_t nz = 0;
  for (int i = 0; i < a.nsup; ++i) {
    for (int j = a.ia[i]; j < a.ia[i + 1]; ++j) {
      int c = a.ja[j];
      if (c > i) continue;
      int64_t sz = (int64_t)a.bla(i) * a.bla(c);
      lo.ja.push_back(c);
      lo.bsa.push_back(nz);
      lo.a.insert(lo.a.end(), a.a.begin() + a.bsa[j], a.a.begin() + a.bsa[j] + sz);
      nz += sz;
    }
    lo.ia[i + 1] = (int)lo.ja.size();
  }
  BlockCsr lt = transp_mtr(lo);
  l = extend_sparsity(lt, up);
  u = extend_sparsity(up, lt);
}

void ord_vect(char type, const std::vector<int>& sp_old, const std::vector<int>& sp_new,
              const std::vector<int>& order, double* v, int n, int nrhs) {
  std::vector<double> tmp(n);
  const int nsup = (int)order.size();
  for (int r = 0; r < nrhs; ++r) {
    double* x = v + (int64_t)r * n;
    std::memcpy(tmp.data(), x, sizeof(double) * n);
    for (int io = 0; io < nsup; ++io) {
      int in = order[io];
      int b = sp_new[in + 1] - sp_new[in];
      if (type == 'D') std::memcpy(x + sp_new[in], tmp.data() + sp_old[io], sizeof(double) * b);
      else std::memcpy(x + sp_old[io], tmp.data() + sp_new[in], sizeof(double) * b);
    }
  }
}

}  // namespace bcm
