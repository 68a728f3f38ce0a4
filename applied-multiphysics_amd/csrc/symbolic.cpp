// Native symbolic analysis: fill-reducing ordering, elimination tree, column counts, relaxed supernodes and the
// dense-block layout of the supernodal factor. It produces exactly the integer structure of the reference's
// analyze() (packages/core/src/la/sparse.mjs:666-902, with amdOrder :297-592), so a factor computed here has the
// reference's Sp / Si / Lxp layout entry for entry. The numeric phase on the GPU consumes this structure.
#include "symbolic.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace amp {

namespace {

inline int32_t flip(int32_t i) { return -i - 2; }

// Approximate minimum degree on the quotient graph (Amestoy, Davis, Duff 1996), step for step as amdOrder in
// sparse.mjs:297-592 (the element absorption, hash and degree-list orders decide the permutation, so they are kept).
class AmdOrdering {
 public:
  AmdOrdering(int32_t n, const int32_t* rowptr, const int32_t* colidx, double denseFactor)
      : n_(n), rowptr_(rowptr), colidx_(colidx), denseFactor_(denseFactor) {}

  bool run(int32_t* perm);

 private:
  void clearMarks(int32_t lemax) {
    if (mark_ < 2 || int64_t(mark_) + lemax > 2000000000LL) {
      for (int32_t k = 0; k < n_; k++)
        if (w_[k] != 0) w_[k] = 1;
      mark_ = 2;
    }
  }
  void compact();
  void eliminate(int32_t k);
  void postorder(int32_t* perm, int32_t* count);

  int32_t n_;
  const int32_t *rowptr_, *colidx_;
  double denseFactor_;
  std::vector<int32_t> Cp_, Ci_, len_, nv_, next_, head_, elen_, degree_, w_, hhead_, last_;
  int64_t cnz_ = 0, nzmax_ = 0;
  int32_t mark_ = 0, lemax_ = 0, nel_ = 0, mindeg_ = 0;
};

bool AmdOrdering::run(int32_t* perm) {
  const int32_t n = n_;
  if (n == 0) return true;
  int64_t offdiag = 0;
  for (int32_t i = 0; i < n; i++)
    for (int32_t p = rowptr_[i]; p < rowptr_[i + 1]; p++) offdiag += colidx_[p] != i;
  cnz_ = offdiag;
  nzmax_ = offdiag + offdiag / 5 + 2 * int64_t(n);
  Cp_.assign(n + 1, 0);
  Ci_.assign(size_t(std::max<int64_t>(nzmax_, 1)), 0);
  int64_t q = 0;
  for (int32_t i = 0; i < n; i++) {
    Cp_[i] = int32_t(q);
    for (int32_t p = rowptr_[i]; p < rowptr_[i + 1]; p++)
      if (colidx_[p] != i) Ci_[q++] = colidx_[p];
  }
  Cp_[n] = int32_t(q);
  len_.assign(n + 1, 0);
  for (int32_t k = 0; k < n; k++) len_[k] = Cp_[k + 1] - Cp_[k];
  const double dense = std::min(double(n - 2), std::max(16.0, denseFactor_ * std::sqrt(double(n))));
  head_.assign(n + 1, -1);
  last_.assign(n + 1, -1);
  next_.assign(n + 1, -1);
  hhead_.assign(n + 1, -1);
  nv_.assign(n + 1, 1);
  w_.assign(n + 1, 1);
  elen_.assign(n + 1, 0);
  degree_ = len_;
  mark_ = 0;
  clearMarks(0);
  elen_[n] = -2;
  Cp_[n] = -1;
  w_[n] = 0;
  nv_[n] = 0;
  nel_ = 0;
  for (int32_t i = 0; i < n; i++) {
    const int32_t d = degree_[i];
    if (d == 0) {  // isolated node: eliminated first
      elen_[i] = -2;
      nel_++;
      Cp_[i] = -1;
      w_[i] = 0;
    } else if (double(d) > dense) {  // dense row: ordered last, under the placeholder node n
      nv_[i] = 0;
      elen_[i] = -1;
      nel_++;
      Cp_[i] = flip(n);
      nv_[n]++;
    } else {
      if (head_[d] != -1) last_[head_[d]] = i;
      next_[i] = head_[d];
      head_[d] = i;
    }
  }
  mindeg_ = 0;
  lemax_ = 0;
  while (nel_ < n) {
    int32_t k = -1;
    for (; mindeg_ < n && (k = head_[mindeg_]) == -1; mindeg_++) {
    }
    if (next_[k] != -1) last_[next_[k]] = -1;
    head_[mindeg_] = next_[k];
    eliminate(k);
  }
  int32_t count = 0;
  postorder(perm, &count);
  return count == n;
}

void AmdOrdering::compact() {
  for (int32_t j = 0; j < n_; j++) {
    const int32_t p = Cp_[j];
    if (p >= 0) {
      Cp_[j] = Ci_[p];
      Ci_[p] = flip(j);
    }
  }
  int64_t q = 0;
  for (int64_t p = 0; p < cnz_;) {
    const int32_t j = flip(Ci_[p++]);
    if (j >= 0) {
      Ci_[q] = Cp_[j];
      Cp_[j] = int32_t(q++);
      for (int32_t t = 0; t < len_[j] - 1; t++) Ci_[q++] = Ci_[p++];
    }
  }
  cnz_ = q;
}

void AmdOrdering::eliminate(int32_t k) {
  const int32_t n = n_;
  int32_t* Ci = Ci_.data();
  const int32_t elenk = elen_[k];
  int32_t nvk = nv_[k];
  nel_ += nvk;
  if (elenk > 0 && cnz_ + mindeg_ >= nzmax_) compact();
  // the new element Lk: union of the pivot's element lists and its own variables
  int32_t dk = 0;
  nv_[k] = -nvk;
  int64_t p = Cp_[k];
  const int64_t pk1 = elenk == 0 ? p : cnz_;
  int64_t pk2 = pk1;
  for (int32_t k1 = 1; k1 <= elenk + 1; k1++) {
    int32_t e, ln;
    int64_t pj;
    if (k1 > elenk) {
      e = k;
      pj = p;
      ln = len_[k] - elenk;
    } else {
      e = Ci[p++];
      pj = Cp_[e];
      ln = len_[e];
    }
    for (int32_t k2 = 1; k2 <= ln; k2++) {
      const int32_t i = Ci[pj++];
      const int32_t nvi = nv_[i];
      if (nvi <= 0) continue;
      dk += nvi;
      nv_[i] = -nvi;
      Ci[pk2++] = i;
      if (next_[i] != -1) last_[next_[i]] = last_[i];
      if (last_[i] != -1)
        next_[last_[i]] = next_[i];
      else
        head_[degree_[i]] = next_[i];
    }
    if (e != k) {
      Cp_[e] = flip(k);
      w_[e] = 0;
    }
  }
  if (elenk != 0) cnz_ = pk2;
  degree_[k] = dk;
  Cp_[k] = int32_t(pk1);
  len_[k] = int32_t(pk2 - pk1);
  elen_[k] = -2;
  // |Le \ Lk| for every element e adjacent to a variable of Lk
  clearMarks(lemax_);
  for (int64_t pk = pk1; pk < pk2; pk++) {
    const int32_t i = Ci[pk];
    const int32_t eln = elen_[i];
    if (eln <= 0) continue;
    const int32_t nvi = -nv_[i];
    const int32_t wnvi = mark_ - nvi;
    const int64_t qend = int64_t(Cp_[i]) + eln - 1;
    for (int64_t q = Cp_[i]; q <= qend; q++) {
      const int32_t e = Ci[q];
      if (w_[e] >= mark_)
        w_[e] -= nvi;
      else if (w_[e] != 0)
        w_[e] = degree_[e] + wnvi;
    }
  }
  // approximate external degrees, aggressive element absorption and the supervariable hash
  for (int64_t pk = pk1; pk < pk2; pk++) {
    const int32_t i = Ci[pk];
    const int64_t p1 = Cp_[i], p2 = p1 + elen_[i] - 1;
    int64_t pn = p1, h = 0;
    int32_t d = 0;
    for (int64_t q = p1; q <= p2; q++) {
      const int32_t e = Ci[q];
      if (w_[e] == 0) continue;
      const int32_t dext = w_[e] - mark_;
      if (dext > 0) {
        d += dext;
        Ci[pn++] = e;
        h += e;
      } else {
        Cp_[e] = flip(k);
        w_[e] = 0;
      }
    }
    elen_[i] = int32_t(pn - p1 + 1);
    const int64_t p3 = pn, p4 = p1 + len_[i];
    for (int64_t q = p2 + 1; q < p4; q++) {
      const int32_t j = Ci[q];
      const int32_t nvj = nv_[j];
      if (nvj <= 0) continue;
      d += nvj;
      Ci[pn++] = j;
      h += j;
    }
    if (d == 0) {  // mass elimination into the pivot
      Cp_[i] = flip(k);
      const int32_t nvi = -nv_[i];
      dk -= nvi;
      nvk += nvi;
      nel_ += nvi;
      nv_[i] = 0;
      elen_[i] = -1;
    } else {
      degree_[i] = std::min(degree_[i], d);
      Ci[pn] = Ci[p3];
      Ci[p3] = Ci[p1];
      Ci[p1] = k;
      len_[i] = int32_t(pn - p1 + 1);
      h = (h < 0 ? -h : h) % n;
      next_[i] = hhead_[h];
      hhead_[h] = i;
      last_[i] = int32_t(h);
    }
  }
  degree_[k] = dk;
  lemax_ = std::max(lemax_, dk);
  mark_ += lemax_;
  clearMarks(lemax_);
  // supervariables: variables with identical adjacency in one hash bucket merge
  for (int64_t pk = pk1; pk < pk2; pk++) {
    int32_t i = Ci[pk];
    if (nv_[i] >= 0) continue;
    const int32_t h = last_[i];
    i = hhead_[h];
    hhead_[h] = -1;
    for (; i != -1 && next_[i] != -1; i = next_[i], mark_++) {
      const int32_t ln = len_[i], eln = elen_[i];
      for (int64_t q = int64_t(Cp_[i]) + 1; q <= int64_t(Cp_[i]) + ln - 1; q++) w_[Ci[q]] = mark_;
      int32_t jlast = i;
      for (int32_t j = next_[i]; j != -1;) {
        bool same = len_[j] == ln && elen_[j] == eln;
        for (int64_t q = int64_t(Cp_[j]) + 1; same && q <= int64_t(Cp_[j]) + ln - 1; q++) same = w_[Ci[q]] == mark_;
        if (same) {
          Cp_[j] = flip(i);
          nv_[i] += nv_[j];
          nv_[j] = 0;
          elen_[j] = -1;
          j = next_[j];
          next_[jlast] = j;
        } else {
          jlast = j;
          j = next_[j];
        }
      }
    }
  }
  // finalize Lk: degree lists of the surviving variables
  int64_t q = pk1;
  for (int64_t pk = pk1; pk < pk2; pk++) {
    const int32_t i = Ci[pk];
    const int32_t nvi = -nv_[i];
    if (nvi <= 0) continue;
    nv_[i] = nvi;
    const int32_t d = std::min(degree_[i] + dk - nvi, n - nel_ - nvi);
    if (head_[d] != -1) last_[head_[d]] = i;
    next_[i] = head_[d];
    last_[i] = -1;
    head_[d] = i;
    mindeg_ = std::min(mindeg_, d);
    degree_[i] = d;
    Ci[q++] = i;
  }
  nv_[k] = nvk;
  if ((len_[k] = int32_t(q - pk1)) == 0) {
    Cp_[k] = -1;
    w_[k] = 0;
  }
  if (elenk != 0) cnz_ = q;
}

void AmdOrdering::postorder(int32_t* perm, int32_t* count) {
  const int32_t n = n_;
  std::vector<int32_t>& parent = Cp_;
  for (int32_t i = 0; i < n; i++) parent[i] = flip(parent[i]);
  std::vector<int32_t>& child = head_;
  std::vector<int32_t>& sibling = next_;
  std::fill(child.begin(), child.end(), -1);
  // non-principal variables hang under their representative; the reference leaves a variable whose parent is the
  // root marker alone (its sibling slot reads 0)
  for (int32_t j = n; j >= 0; j--) {
    if (nv_[j] > 0) continue;
    if (parent[j] == -1) {
      sibling[j] = 0;
      continue;
    }
    sibling[j] = child[parent[j]];
    child[parent[j]] = j;
  }
  for (int32_t e = n; e >= 0; e--) {
    if (nv_[e] <= 0 || parent[e] == -1) continue;
    sibling[e] = child[parent[e]];
    child[parent[e]] = e;
  }
  std::vector<int32_t>& stack = w_;
  int32_t o = 0;
  for (int32_t i = 0; i <= n; i++) {
    if (parent[i] != -1) continue;
    int32_t top = 0;
    stack[0] = i;
    while (top >= 0) {
      const int32_t node = stack[top];
      const int32_t c = child[node];
      if (c == -1) {
        top--;
        if (node < n) perm[o++] = node;
      } else {
        child[node] = sibling[c];
        stack[++top] = c;
      }
    }
  }
  *count = o;
}

// Depth-first postorder of a forest; children visited in increasing index (sparse.mjs:597-625).
std::vector<int32_t> forestPostorder(const std::vector<int32_t>& parent) {
  const int32_t n = int32_t(parent.size());
  std::vector<int32_t> child(n, -1), sibling(n, -1), stack(n), post(n);
  for (int32_t j = n - 1; j >= 0; j--) {
    if (parent[j] == -1) continue;
    sibling[j] = child[parent[j]];
    child[parent[j]] = j;
  }
  int32_t k = 0;
  for (int32_t r = 0; r < n; r++) {
    if (parent[r] != -1) continue;
    int32_t top = 0;
    stack[0] = r;
    while (top >= 0) {
      const int32_t node = stack[top];
      const int32_t c = child[node];
      if (c == -1) {
        top--;
        post[k++] = node;
      } else {
        child[node] = sibling[c];
        stack[++top] = c;
      }
    }
  }
  return post;
}

}  // namespace

std::string patternFingerprint(int32_t n, const int32_t* rowptr, const int32_t* colidx) {
  uint32_t h = 0x811c9dc5u;  // 32-bit FNV-1a over the four bytes of every rowptr then colidx entry
  auto mix = [&h](int32_t v) {
    const uint32_t u = uint32_t(v);
    for (int b = 0; b < 4; b++) {
      h ^= (u >> (8 * b)) & 0xffu;
      h *= 0x01000193u;
    }
  };
  for (int32_t i = 0; i <= n; i++) mix(rowptr[i]);
  const int32_t nnz = rowptr[n];
  for (int32_t p = 0; p < nnz; p++) mix(colidx[p]);
  char buf[64];
  std::snprintf(buf, sizeof buf, "%d:%d:%x", n, nnz, h);
  return buf;
}

// ndOrder (sparse.mjs:230-283). A range of more than `leaf` nodes is sorted stably by its coordinate of larger extent
// (x when the extents are equal) and split at half its length; a node of the first half with a neighbour in the second
// half is a separator node. The order of the range is the order of its first half without the separators, then that
// of its second half, then the separators (in sorted order); a range of at most `leaf` nodes keeps its order. Each
// range is rearranged in place into [first half without separators | second half | separators] and its first two
// parts are split in turn, so the array ends up holding the order itself.
void ndOrder(int32_t n, const int32_t* rowptr, const int32_t* colidx, const double* x, const double* y, int32_t leaf,
             int32_t* perm) {
  for (int32_t i = 0; i < n; i++) perm[i] = i;
  std::vector<int8_t> second(size_t(n), 0);
  std::vector<int32_t> tmp;
  std::vector<std::pair<int32_t, int32_t>> work{{0, n}};  // ranges [lo, hi) still to split
  while (!work.empty()) {
    const auto [lo, hi] = work.back();
    work.pop_back();
    if (hi - lo <= leaf) continue;
    double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
    for (int32_t i = lo; i < hi; i++) {
      const int32_t v = perm[i];
      x0 = std::min(x0, x[v]);
      x1 = std::max(x1, x[v]);
      y0 = std::min(y0, y[v]);
      y1 = std::max(y1, y[v]);
    }
    const double* key = x1 - x0 >= y1 - y0 ? x : y;
    std::stable_sort(perm + lo, perm + hi, [key](int32_t p, int32_t q) { return key[p] < key[q]; });
    const int32_t half = (hi - lo) >> 1;
    for (int32_t i = lo + half; i < hi; i++) second[perm[i]] = 1;
    // first half: non-separators stay in front (in order), separators go to tmp; then the second half moves up
    tmp.clear();
    int32_t a = lo;
    for (int32_t i = lo; i < lo + half; i++) {
      const int32_t v = perm[i];
      bool sep = false;
      for (int32_t p = rowptr[v]; p < rowptr[v + 1] && !sep; p++) sep = colidx[p] != v && second[colidx[p]];
      if (sep)
        tmp.push_back(v);
      else
        perm[a++] = v;
    }
    for (int32_t i = lo + half; i < hi; i++) second[perm[i]] = 0;
    const int32_t b = a + (hi - lo - half);
    std::memmove(perm + a, perm + lo + half, sizeof(int32_t) * size_t(hi - lo - half));
    std::copy(tmp.begin(), tmp.end(), perm + b);
    work.push_back({a, b});   // second half
    work.push_back({lo, a});  // first half without the separators
  }
}

bool amdOrder(int32_t n, const int32_t* rowptr, const int32_t* colidx, double denseFactor, int32_t* perm) {
  AmdOrdering amd(n, rowptr, colidx, denseFactor);
  return amd.run(perm);
}

bool analyzePattern(int32_t n, const int32_t* rowptr, const int32_t* colidx, const int32_t* perm0in, bool relax,
                    Symbolic& S) {
  S = Symbolic();
  S.n = n;
  S.patternKey = patternFingerprint(n, rowptr, colidx);
  S.nnzPattern = n > 0 ? rowptr[n] : 0;
  S.lowerP.assign(1, 0);
  S.superStart.assign(1, 0);
  S.Sp.assign(1, 0);
  S.Lxp.assign(1, 0);
  S.fill = 1;
  if (n == 0) return true;
  std::vector<int32_t> perm0(n);
  if (perm0in)
    std::copy(perm0in, perm0in + n, perm0.begin());
  else if (!amdOrder(n, rowptr, colidx, 10.0, perm0.data()))
    return false;

  // elimination tree of P A P^T (Liu, path compression) and its postorder folded into the permutation
  std::vector<int32_t> pinv(n), parent0(n, -1), ancestor(n, -1);
  for (int32_t i = 0; i < n; i++) pinv[perm0[i]] = i;
  for (int32_t k = 0; k < n; k++) {
    const int32_t v = perm0[k];
    for (int32_t p = rowptr[v]; p < rowptr[v + 1]; p++) {
      for (int32_t i = pinv[colidx[p]]; i != -1 && i < k;) {
        const int32_t up = ancestor[i];
        ancestor[i] = k;
        if (up == -1) parent0[i] = k;
        i = up;
      }
    }
  }
  const std::vector<int32_t> post = forestPostorder(parent0);
  std::vector<int32_t> postInv(n);
  S.perm.resize(n);
  S.parent.resize(n);
  for (int32_t k = 0; k < n; k++) {
    S.perm[k] = perm0[post[k]];
    postInv[post[k]] = k;
  }
  for (int32_t k = 0; k < n; k++) {
    const int32_t p = parent0[post[k]];
    S.parent[k] = p == -1 ? -1 : postInv[p];
  }
  S.pinv.resize(n);
  for (int32_t i = 0; i < n; i++) S.pinv[S.perm[i]] = i;
  const std::vector<int32_t>& parent = S.parent;

  // lower CSC of P A P^T, each entry remembering its CSR slot
  S.lowerP.assign(n + 1, 0);
  for (int32_t v = 0; v < n; v++) {
    const int32_t pv = S.pinv[v];
    for (int32_t p = rowptr[v]; p < rowptr[v + 1]; p++) S.lowerP[pv + 1] += S.pinv[colidx[p]] >= pv;
  }
  for (int32_t j = 0; j < n; j++) S.lowerP[j + 1] += S.lowerP[j];
  S.nnzA = S.lowerP[n];
  S.lowerI.resize(S.nnzA);
  S.lowerSrc.resize(S.nnzA);
  {
    std::vector<int32_t> fillPos(S.lowerP.begin(), S.lowerP.end() - 1);
    for (int32_t v = 0; v < n; v++) {
      const int32_t pv = S.pinv[v];
      for (int32_t p = rowptr[v]; p < rowptr[v + 1]; p++) {
        const int32_t pi = S.pinv[colidx[p]];
        if (pi < pv) continue;
        const int32_t q = fillPos[pv]++;
        S.lowerI[q] = pi;
        S.lowerSrc[q] = p;
      }
    }
  }

  // column counts by the Gilbert-Ng-Peyton skeleton algorithm on the postordered tree
  S.colCount.assign(n, 0);
  {
    std::vector<int32_t> first(n, -1), maxfirst(n, -1), prevleaf(n, -1), anc(n);
    for (int32_t k = 0; k < n; k++) {
      S.colCount[k] = first[k] == -1 ? 1 : 0;
      for (int32_t j = k; j != -1 && first[j] == -1; j = parent[j]) first[j] = k;
    }
    for (int32_t i = 0; i < n; i++) anc[i] = i;
    for (int32_t j = 0; j < n; j++) {
      if (parent[j] != -1) S.colCount[parent[j]]--;
      for (int32_t p = S.lowerP[j]; p < S.lowerP[j + 1]; p++) {
        const int32_t i = S.lowerI[p];
        if (i <= j || first[j] <= maxfirst[i]) continue;
        maxfirst[i] = first[j];
        const int32_t jprev = prevleaf[i];
        prevleaf[i] = j;
        S.colCount[j]++;
        if (jprev == -1) continue;
        int32_t root = jprev;
        while (root != anc[root]) root = anc[root];
        for (int32_t s = jprev; s != root;) {
          const int32_t up = anc[s];
          anc[s] = root;
          s = up;
        }
        S.colCount[root]--;
      }
      if (parent[j] != -1) anc[j] = parent[j];
    }
    for (int32_t j = 0; j < n; j++)
      if (parent[j] != -1) S.colCount[parent[j]] += S.colCount[j];
  }
  S.nnzL = 0;
  S.flops = 0;
  for (int32_t j = 0; j < n; j++) {
    S.nnzL += S.colCount[j];
    S.flops += double(S.colCount[j]) * double(S.colCount[j]);
  }

  // fundamental supernodes: a column joins its predecessor's supernode when it is the only child and the structure
  // shrinks by exactly one
  std::vector<int32_t> nchild(n, 0), fstart;
  for (int32_t j = 0; j < n; j++)
    if (parent[j] != -1) nchild[parent[j]]++;
  fstart.push_back(0);
  for (int32_t j = 1; j < n; j++)
    if (!(parent[j - 1] == j && S.colCount[j - 1] == S.colCount[j] + 1 && nchild[j] == 1)) fstart.push_back(j);
  fstart.push_back(n);
  const int32_t nsf = int32_t(fstart.size()) - 1;

  // relaxed amalgamation (CHOLMOD's default thresholds: 4 / any, 16 / 0.8, 48 / 0.1, else 0.05 zero fraction)
  std::vector<int32_t> starts;
  if (relax && nsf > 1) {
    std::vector<int32_t> colF(n), first(nsf), lastc(nsf), sparent(nsf);
    std::vector<double> rows(nsf), zeros(nsf, 0.0), entries(nsf);
    std::vector<char> alive(nsf, 1);
    for (int32_t s = 0; s < nsf; s++)
      for (int32_t j = fstart[s]; j < fstart[s + 1]; j++) colF[j] = s;
    for (int32_t s = 0; s < nsf; s++) {
      first[s] = fstart[s];
      lastc[s] = fstart[s + 1] - 1;
      rows[s] = S.colCount[fstart[s]];
      const double nc = fstart[s + 1] - fstart[s];
      entries[s] = nc * rows[s] - (nc * (nc - 1)) / 2;
      const int32_t pl = parent[lastc[s]];
      sparent[s] = pl == -1 ? -1 : colF[pl];
    }
    for (int32_t s = 0; s < nsf; s++) {
      const int32_t P = sparent[s];
      if (P == -1 || first[P] != lastc[s] + 1) continue;
      const double ncC = lastc[s] - first[s] + 1, ncP = lastc[P] - first[P] + 1;
      const double nc = ncC + ncP, m = ncC + rows[P];
      const double stored = nc * m - (nc * (nc - 1)) / 2;
      const double z = stored - entries[s] - entries[P] + zeros[s] + zeros[P];
      const double zf = z / stored;
      if (!(nc <= 4 || (nc <= 16 && zf < 0.8) || (nc <= 48 && zf < 0.1) || zf < 0.05)) continue;
      first[P] = first[s];
      rows[P] = m;
      zeros[P] = z;
      entries[P] += entries[s];
      alive[s] = 0;
    }
    for (int32_t s = 0; s < nsf; s++)
      if (alive[s]) starts.push_back(first[s]);
    starts.push_back(n);
  } else {
    starts = fstart;
  }
  S.superStart = starts;
  const int32_t nsuper = int32_t(starts.size()) - 1;
  S.nsuper = nsuper;
  S.colSuper.resize(n);
  for (int32_t s = 0; s < nsuper; s++)
    for (int32_t j = starts[s]; j < starts[s + 1]; j++) S.colSuper[j] = s;
  S.superParent.resize(nsuper);
  for (int32_t s = 0; s < nsuper; s++) {
    const int32_t pl = parent[starts[s + 1] - 1];
    S.superParent[s] = pl == -1 ? -1 : S.colSuper[pl];
  }
  // children linked in decreasing index (the multifrontal stack pops the last child first)
  S.childHead.assign(nsuper, -1);
  S.childNext.assign(nsuper, -1);
  for (int32_t s = 0; s < nsuper; s++) {
    const int32_t P = S.superParent[s];
    if (P == -1) continue;
    S.childNext[s] = S.childHead[P];
    S.childHead[P] = s;
  }
  // row structure: own columns, then the rows below, gathered from A's columns and the children, sorted
  S.Sp.assign(nsuper + 1, 0);
  S.Si.clear();
  std::vector<int32_t> marker(n, -1), below;
  S.maxFront = 0;
  for (int32_t s = 0; s < nsuper; s++) {
    const int32_t f = starts[s], l = starts[s + 1] - 1;
    below.clear();
    for (int32_t j = f; j <= l; j++)
      for (int32_t p = S.lowerP[j]; p < S.lowerP[j + 1]; p++) {
        const int32_t i = S.lowerI[p];
        if (i > l && marker[i] != s) {
          marker[i] = s;
          below.push_back(i);
        }
      }
    for (int32_t c = S.childHead[s]; c != -1; c = S.childNext[c]) {
      const int32_t ncc = starts[c + 1] - starts[c];
      for (int32_t r = S.Sp[c] + ncc; r < S.Sp[c + 1]; r++) {
        const int32_t i = S.Si[r];
        if (i > l && marker[i] != s) {
          marker[i] = s;
          below.push_back(i);
        }
      }
    }
    std::sort(below.begin(), below.end());
    for (int32_t j = f; j <= l; j++) S.Si.push_back(j);
    S.Si.insert(S.Si.end(), below.begin(), below.end());
    const int32_t m = l - f + 1 + int32_t(below.size());
    S.Sp[s + 1] = S.Sp[s] + m;
    S.maxFront = std::max(S.maxFront, m);
  }
  S.Lxp.assign(nsuper + 1, 0);
  for (int32_t s = 0; s < nsuper; s++)
    S.Lxp[s + 1] = S.Lxp[s] + int64_t(S.Sp[s + 1] - S.Sp[s]) * (starts[s + 1] - starts[s]);
  S.nnzLStored = S.Lxp[nsuper];
  S.fill = double(S.nnzL) / double(S.nnzA);
  return true;
}

}  // namespace amp
