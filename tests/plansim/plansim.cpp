// Test-only CPU executor of the device factor / solve plans (applied-multiphysics_amd/csrc/plan.h).
//
// It replays the exact launch and task sequence the GPU executes, with plain loops for every task kind, so that the
// plan (layout, offsets, task coverage, ordering) is checked against the oracle on the CPU. It is never linked into
// the shipped library.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <stdexcept>
#include <vector>

#include "../../applied-multiphysics_amd/csrc/plan.h"
#include "../../applied-multiphysics_amd/csrc/symbolic.h"

using namespace amp;

namespace {

struct Ctx {
  const Symbolic* S;
  const FactorPlan* P;
  std::vector<double> Lx, U, W, D;
  std::vector<uint8_t> pert;
  int ldlt = 0;
  double tau = 0;
  int32_t failCol = INT32_MAX;
};

double* buf(Ctx& c, int8_t b) { return b == kBufLx ? c.Lx.data() : b == kBufU ? c.U.data() : c.W.data(); }

// pivot + scaling + rank-1 update of an nb-column right-looking factorization of the m x m array F (ld)
bool pivotColumn(Ctx& c, double* F, int32_t ld, int32_t k, int32_t rows, int32_t col, double* dOut) {
  double d = F[k + int64_t(k) * ld];
  if (!c.ldlt) {
    if (!(d > 0)) c.failCol = std::min(c.failCol, col);
    const double sq = std::sqrt(d);
    F[k + int64_t(k) * ld] = sq;
    const double inv = 1 / sq;
    for (int32_t i = k + 1; i < rows; i++) F[i + int64_t(k) * ld] *= inv;
    return true;
  }
  if (!(std::fabs(d) >= c.tau)) {
    if (!std::isfinite(d))
      c.failCol = std::min(c.failCol, col);
    else {
      d = d < 0 ? -c.tau : c.tau;
      c.pert[col] = 1;
    }
  }
  *dOut = d;
  F[k + int64_t(k) * ld] = 1;
  const double inv = 1 / d;
  for (int32_t i = k + 1; i < rows; i++) F[i + int64_t(k) * ld] *= inv;
  return true;
}

// right-looking partial factorization of nc columns of a rows x rows lower array
void partialFactor(Ctx& c, double* F, int32_t ld, int32_t rows, int32_t nc, int32_t col0) {
  for (int32_t k = 0; k < nc; k++) {
    double d = 1;
    pivotColumn(c, F, ld, k, rows, col0 + k, &d);
    if (c.ldlt) c.D[col0 + k] = d;
    const double* Fk = F + int64_t(k) * ld;
    for (int32_t j = k + 1; j < rows; j++) {
      const double wj = c.ldlt ? Fk[j] * d : Fk[j];
      double* Fj = F + int64_t(j) * ld;
      for (int32_t i = j; i < rows; i++) Fj[i] -= Fk[i] * wj;
    }
  }
}

void runSmall(Ctx& c, int32_t s, const double* values) {
  const Symbolic& S = *c.S;
  const FactorPlan& P = *c.P;
  const int32_t f = S.superStart[s], nc = S.superStart[s + 1] - f, m = S.Sp[s + 1] - S.Sp[s], mu = m - nc;
  std::vector<double> F(size_t(m) * m, 0.0);
  for (int32_t q = S.lowerP[f]; q < S.lowerP[f + nc]; q++) F[P.amap[q]] = 0.0 + values[S.lowerSrc[q]];
  for (int32_t ch = S.childHead[s]; ch != -1; ch = S.childNext[ch]) {
    const int32_t ncc = S.superStart[ch + 1] - S.superStart[ch], muc = S.Sp[ch + 1] - S.Sp[ch] - ncc;
    const int32_t* rel = P.rel.data() + S.Sp[ch] + ncc;
    const double* Uc = c.U.data() + P.Uoff[ch];
    for (int32_t b = 0; b < muc; b++)
      for (int32_t a = b; a < muc; a++) F[rel[a] + int64_t(rel[b]) * m] += Uc[a + int64_t(b) * muc];
  }
  partialFactor(c, F.data(), m, m, nc, f);
  std::memcpy(c.Lx.data() + S.Lxp[s], F.data(), size_t(m) * nc * 8);
  if (mu > 0)
    for (int32_t b = 0; b < mu; b++)
      for (int32_t a = 0; a < mu; a++) c.U[P.Uoff[s] + a + int64_t(b) * mu] = F[nc + a + int64_t(nc + b) * m];
}

// the task's child column ranges, as the device does; every column a range lists must land in [j0, j1)
void runEa(Ctx& c, const EaTask& t) {
  const Symbolic& S = *c.S;
  const FactorPlan& P = *c.P;
  const int32_t s = t.s, nc = S.superStart[s + 1] - S.superStart[s], m = S.Sp[s + 1] - S.Sp[s], mu = m - nc;
  for (int32_t q = t.r0; q < t.r1; q++) {
    const EaRange& g = P.eaRanges[q];
    const int32_t ch = g.c;
    const int32_t ncc = S.superStart[ch + 1] - S.superStart[ch], muc = S.Sp[ch + 1] - S.Sp[ch] - ncc;
    const int32_t* rel = P.rel.data() + S.Sp[ch] + ncc;
    const double* Uc = c.U.data() + P.Uoff[ch];
    for (int32_t b = g.b0; b < g.b1; b++) {
      const int32_t j = rel[b];
      if (j < t.j0 || j >= t.j1) {
        c.failCol = -2;  // a range outside its work item: reported as a plan error
        continue;
      }
      for (int32_t a = b; a < muc; a++) {
        const int32_t i = rel[a];
        const double v = Uc[a + int64_t(b) * muc];
        if (j < nc)
          c.Lx[S.Lxp[s] + i + int64_t(j) * m] += v;
        else
          c.U[P.Uoff[s] + (i - nc) + int64_t(j - nc) * mu] += v;
      }
    }
  }
}

void runDiag(Ctx& c, const DiagTask& t) {
  double* F = c.Lx.data() + t.f_off;
  partialFactor(c, F, t.ld, t.nb, t.nb, t.col0);
  double* W = c.W.data() + t.w_off;
  const int32_t nb = t.nb;
  for (int32_t j = 0; j < nb; j++) {  // column j of the inverse of the lower factor
    for (int32_t i = 0; i < nb; i++) W[i + int64_t(j) * kTileEdge] = 0;
    for (int32_t i = j; i < nb; i++) {
      double s = i == j ? 1.0 : 0.0;
      for (int32_t k = j; k < i; k++) s -= F[i + int64_t(k) * t.ld] * W[k + int64_t(j) * kTileEdge];
      W[i + int64_t(j) * kTileEdge] = c.ldlt ? s : s / F[i + int64_t(i) * t.ld];
    }
  }
}

// A kPanel launch: every task reads the diagonal block and its panel rows as they were when the launch started (the
// device's writers of those inputs wait until every workgroup has read them), so the writes of all tasks are
// collected first and applied after. Per task: the block factored and inverted as runDiag does (kPanelPre: its W
// and pivots as the launch before stored them), the own rows L = A W^T (LDL^T: / d) as a mode-2 product, the column
// tiles' rows the same way, and the update of the own rows in the column tiles' columns, C -= L d Lc^T, as a mode-0
// product; kPanelNext: the updated first column tile's block factored and stored as the next step's owner would.
void runPanelLaunch(Ctx& c, const Launch& L) {
  const FactorPlan& P = *c.P;
  constexpr int32_t E = kTileEdge;
  std::vector<std::pair<int64_t, double>> lxw, ww, nextL;
  // counter contract: a front's tasks are consecutive, share a counter no other front of any launch uses, and number
  // ntask; one of them is the owner, none for a kPanelPre front, whose W must come from a kPanelNext task of an
  // earlier launch (the first row tile of a step with columns to update); else a plan error, reported as failCol -2
  static thread_local std::vector<int32_t> seen;
  static thread_local std::set<int64_t> ahead;  // W offsets stored by kPanelNext tasks
  if (L.first == 0 || seen.size() < size_t(P.panelCounters)) {
    seen.assign(size_t(P.panelCounters), 0);
    ahead.clear();
  }
  for (int64_t q = L.first; q < L.first + L.count;) {
    const PanelTask& t0 = P.panels[size_t(q)];
    const bool pre = t0.flags & kPanelPre;
    int64_t e = q;
    int32_t owners = 0;
    for (; e < L.first + L.count && P.panels[size_t(e)].ctr == t0.ctr; e++) {
      const PanelTask& t = P.panels[size_t(e)];
      owners += t.owner;
      if (bool(t.flags & kPanelPre) != pre) c.failCol = -2;
      if ((t.flags & kPanelNext) && (t.ro != t.nb || t.ucols <= 0 || t.nb != E || t.rows < std::min(E, t.ucols)))
        c.failCol = -2;
    }
    if (t0.ctr < 0 || t0.ctr >= P.panelCounters || seen[size_t(t0.ctr)]++ || e - q != t0.ntask ||
        owners != (pre ? 0 : 1) || t0.ntask > P.panelMaxTasks || (pre && (!ahead.count(t0.w_off) || t0.rows == 0)))
      c.failCol = -2;
    q = e;
  }
  for (int64_t q = L.first; q < L.first + L.count; q++) {
    const PanelTask& t = P.panels[size_t(q)];
    if (t.flags & kPanelNext) ahead.insert(t.w_off + int64_t(E) * E);
  }
  for (int64_t q = L.first; q < L.first + L.count; q++) {
    const PanelTask& t = P.panels[size_t(q)];
    const int32_t nb = t.nb;
    const int64_t ld = t.ld;
    const double* F0 = c.Lx.data() + t.f_off;
    std::vector<double> F(size_t(nb) * nb, 0.0), W(size_t(E) * E, 0.0);
    // the factor and inverse of an nb x nb lower block Fb (in place; its pivots recorded at columns col0 ..)
    auto factorBlock = [&c](std::vector<double>& Fb, std::vector<double>& Wb, int32_t n, int32_t col0) {
      partialFactor(c, Fb.data(), n, n, n, col0);
      for (int32_t j = 0; j < n; j++)
        for (int32_t i = j; i < n; i++) {
          double s = i == j ? 1.0 : 0.0;
          for (int32_t k = j; k < i; k++) s -= Fb[i + size_t(k) * n] * Wb[k + size_t(j) * E];
          Wb[i + size_t(j) * E] = c.ldlt ? s : s / Fb[i + size_t(i) * n];
        }
    };
    if (t.flags & kPanelPre) {
      std::copy(c.W.begin() + t.w_off, c.W.begin() + t.w_off + int64_t(E) * E, W.begin());
    } else {
      for (int32_t j = 0; j < nb; j++)
        for (int32_t i = j; i < nb; i++) F[i + size_t(j) * nb] = F0[i + j * ld];
      factorBlock(F, W, nb, t.col0);
    }
    if (t.owner) {
      for (int32_t j = 0; j < nb; j++)
        for (int32_t i = j; i < nb; i++) lxw.emplace_back(t.f_off + i + j * ld, F[i + size_t(j) * nb]);
      for (int32_t i = 0; i < E * E; i++) ww.emplace_back(t.w_off + i, W[size_t(i)]);
    }
    if (t.rows == 0) continue;
    // rows r0 .. r0 + rows - 1 of the panel (relative to the block's first row): L = A W^T (/ d)
    auto solveRows = [&](int32_t r0, int32_t rows) {
      std::vector<double> Lr(size_t(rows) * nb);
      for (int32_t j = 0; j < nb; j++)
        for (int32_t i = 0; i < rows; i++) {
          double s = 0;
          for (int32_t k = 0; k < nb; k++) s += F0[r0 + i + k * ld] * W[j + size_t(k) * E];
          if (c.ldlt) s /= c.D[t.col0 + j];
          Lr[i + size_t(j) * rows] = s;
        }
      return Lr;
    };
    const std::vector<double> Lo = solveRows(t.ro, t.rows);
    for (int32_t j = 0; j < nb; j++)
      for (int32_t i = 0; i < t.rows; i++) lxw.emplace_back(t.f_off + t.ro + i + j * ld, Lo[i + size_t(j) * t.rows]);
    const int32_t own = (t.ro - nb) / E;
    std::vector<double> Fn(size_t(E) * E, 0.0);  // (kPanelNext) the updated next block, its lower part
    for (int32_t ct = 0; ct * E < t.ucols && ct <= own; ct++) {
      const int32_t wc = std::min(E, t.ucols - E * ct);
      const std::vector<double> Lc = solveRows(nb + E * ct, wc);
      for (int32_t j = 0; j < wc; j++)
        for (int32_t i = 0; i < t.rows; i++) {
          double s = 0;
          for (int32_t k = 0; k < nb; k++) {
            double b = Lc[j + size_t(k) * wc];
            if (c.ldlt) b *= c.D[t.col0 + k];
            s += Lo[i + size_t(k) * t.rows] * b;
          }
          const int64_t d = t.f_off + t.ro + i + (nb + E * ct + j) * ld;
          lxw.emplace_back(d, c.Lx[size_t(d)] - s);
          if ((t.flags & kPanelNext) && ct == 0 && i >= j && i < wc) Fn[i + size_t(j) * wc] = c.Lx[size_t(d)] - s;
        }
      if ((t.flags & kPanelNext) && ct == 0) {
        // the next block (stored after the updates, so its factor replaces them)
        std::vector<double> Wn(size_t(E) * E, 0.0);
        factorBlock(Fn, Wn, wc, t.col0 + E);
        for (int32_t j = 0; j < wc; j++)
          for (int32_t i = j; i < wc; i++) nextL.emplace_back(t.f_off + E + i + (E + j) * ld, Fn[i + size_t(j) * wc]);
        for (int32_t i = 0; i < E * E; i++) ww.emplace_back(t.w_off + int64_t(E) * E + i, Wn[size_t(i)]);
      }
    }
  }
  lxw.insert(lxw.end(), nextL.begin(), nextL.end());
  for (const auto& e : lxw) c.Lx[size_t(e.first)] = e.second;
  for (const auto& e : ww) c.W[size_t(e.first)] = e.second;
}

void runGemm(Ctx& c, const GemmTask& t) {
  const double* A = buf(c, t.a_buf) + t.a_off;
  const double* B = buf(c, t.b_buf) + t.b_off;
  double* C = buf(c, t.c_buf) + t.c_off;
  std::vector<double> acc(size_t(t.rows) * t.cols, 0.0);
  for (int32_t j = 0; j < t.cols; j++)
    for (int32_t i = 0; i < t.rows; i++) {
      double s = 0;
      for (int32_t k = 0; k < t.k; k++) {
        double b = B[j + int64_t(k) * t.ldb];
        if (c.ldlt && t.mode != 2) b *= c.D[t.d_off + k];
        s += A[i + int64_t(k) * t.lda] * b;
      }
      if (c.ldlt && t.mode == 2) s /= c.D[t.d_off + j];
      acc[i + size_t(j) * t.rows] = s;
    }
  for (int32_t j = 0; j < t.cols; j++)
    for (int32_t i = 0; i < t.rows; i++) {
      double& dst = C[i + int64_t(j) * t.ldc];
      const double v = acc[i + size_t(j) * t.rows];
      dst = t.mode == 0 ? dst - v : t.mode == 1 ? -v : v;
    }
}

struct Handle {
  Symbolic S;
  FactorPlan F;
  SolvePlan SP;
  Ctx c;
  int32_t flowWidth = kFlowQuadsF32;  // task width of the flow launches ps_solve simulates
};

}  // namespace

extern "C" {

// prm (may be NULL for the defaults): smallMax, smallNarrow, wideLaunch, superPanel, flowWideMin, packMin, fusedPanel
void* ps_create_params(int32_t n, const int32_t* rowptr, const int32_t* colidx, const int32_t* perm, int32_t relax,
                       const int32_t* prm) {
  Handle* h = new Handle();
  if (!analyzePattern(n, rowptr, colidx, perm, relax != 0, h->S)) {
    delete h;
    return nullptr;
  }
  PlanParams pp;
  if (prm) {
    pp.smallMax = prm[0];
    pp.smallNarrow = prm[1];
    pp.wideLaunch = prm[2];
    pp.superPanel = prm[3];
    pp.flowWideMin = prm[4];
    pp.packMin = prm[5];
    pp.fusedPanel = prm[6];
    pp.fusedSolve = prm[7];
  }
  h->F = buildFactorPlan(h->S, pp);
  h->SP = buildSolvePlan(h->S, h->F, pp);
  return h;
}

void* ps_create(int32_t n, const int32_t* rowptr, const int32_t* colidx, const int32_t* perm, int32_t relax) {
  return ps_create_params(n, rowptr, colidx, perm, relax, nullptr);
}

// Assembly map of a synthetic analysis with one tall, wide tiled front: supernode 0 holds columns [0, nc) with rows
// 0 .. m - 1 (its below rows are the columns of supernode 1, the root, [nc, m)); A holds the diagonal and the entry
// (m - 1, nc - 1). out: [0] the plan's Lx index of that entry, [1] its position in the reference layout
// (Lxp[0] + (m - 1) + (nc - 1) m), [2] product tasks of the plan. Returns 0, or 1 when the plan refuses the shape.
int32_t ps_tall_front_map(int32_t nc, int32_t m, int32_t superPanel, int64_t* out) {
  Symbolic S;
  S.n = m;
  S.nsuper = 2;
  S.superStart = {0, nc, m};
  S.Sp = {0, m, m + (m - nc)};
  S.Si.resize(size_t(m) + (m - nc));
  for (int32_t i = 0; i < m; i++) S.Si[i] = i;
  for (int32_t i = nc; i < m; i++) S.Si[size_t(m) + (i - nc)] = i;
  S.superParent = {1, -1};
  S.childHead = {-1, 0};
  S.childNext = {-1, -1};
  S.colSuper.resize(m);
  for (int32_t j = 0; j < m; j++) S.colSuper[j] = j < nc ? 0 : 1;
  S.Lxp = {0, int64_t(m) * nc, int64_t(m) * nc + int64_t(m - nc) * (m - nc)};
  S.nnzLStored = S.Lxp[2];
  S.maxFront = m;
  S.lowerP.assign(size_t(m) + 1, 0);
  for (int32_t j = 0; j < m; j++) {
    S.lowerI.push_back(j);
    if (j == nc - 1) S.lowerI.push_back(m - 1);
    S.lowerP[j + 1] = int32_t(S.lowerI.size());
  }
  S.nnzA = int64_t(S.lowerI.size());
  S.lowerSrc.resize(S.lowerI.size());
  for (size_t q = 0; q < S.lowerSrc.size(); q++) S.lowerSrc[q] = int32_t(q);
  S.nnzPattern = S.nnzA;
  PlanParams pp;
  pp.superPanel = superPanel;
  FactorPlan F;
  try {
    F = buildFactorPlan(S, pp);
  } catch (const std::length_error&) {
    return 1;
  }
  out[0] = F.adst[S.lowerP[nc - 1] + 1];
  out[1] = S.Lxp[0] + (m - 1) + int64_t(nc - 1) * m;
  out[2] = int64_t(F.gemm.size());
  return 0;
}

void ps_free(void* p) { delete static_cast<Handle*>(p); }

// plan statistics: launches, gemm tasks, U entries, W entries
void ps_stats(void* p, int64_t* out) {
  Handle* h = static_cast<Handle*>(p);
  out[0] = int64_t(h->F.launches.size());
  out[1] = int64_t(h->F.gemm.size());
  out[2] = h->F.uEntries;
  out[3] = h->F.wEntries;
  out[4] = int64_t(h->SP.launches.size());
  out[5] = h->F.nlevels;
}

// factor: returns failCol (INT32_MAX when ok); Lx (nnzLStored), D (n)
int32_t ps_factor(void* p, const double* values, int32_t ldlt, double pivotTol, double* LxOut, double* DOut,
                  uint8_t* pertOut) {
  Handle* h = static_cast<Handle*>(p);
  const Symbolic& S = h->S;
  const FactorPlan& P = h->F;
  Ctx& c = h->c;
  c.S = &S;
  c.P = &P;
  c.Lx.assign(size_t(std::max<int64_t>(P.nnzLStored, 1)), std::nan(""));  // poison: every entry must be written
  c.U.assign(size_t(std::max<int64_t>(P.uEntries, 1)), std::nan(""));
  c.W.assign(size_t(std::max<int64_t>(P.wEntries, 1)), 0.0);
  c.D.assign(size_t(std::max(S.n, 1)), 0.0);
  c.pert.assign(size_t(std::max(S.n, 1)), 0);
  c.ldlt = ldlt;
  c.failCol = INT32_MAX;
  if (ldlt) {
    double amax = 0;
    for (int32_t j = 0; j < S.n; j++)
      for (int32_t q = S.lowerP[j]; q < S.lowerP[j + 1]; q++)
        if (S.lowerI[q] == j) amax = std::max(amax, std::fabs(values[S.lowerSrc[q]]));
    c.tau = pivotTol * (amax > 0 ? amax : 1);
  }
  for (const Launch& L : P.launches) {
    if (L.kind == kPanel) {
      runPanelLaunch(c, L);
      continue;
    }
    for (int64_t t = L.first; t < L.first + L.count; t++) {
      switch (L.kind) {
        case kZero:
          std::fill(c.Lx.begin() + P.zeros[t].off, c.Lx.begin() + P.zeros[t].off + P.zeros[t].len, 0.0);
          break;
        case kAssemble:
          if (P.adst[t] >= 0) c.Lx[P.adst[t]] = 0.0 + values[S.lowerSrc[t]];
          break;
        case kSmall:
          runSmall(c, P.smallFronts[t], values);
          break;
        case kExtendAdd:
          runEa(c, P.ea[t]);
          break;
        case kDiag:
          runDiag(c, P.diag[t]);
          break;
        case kGemm:
          runGemm(c, P.gemm[t]);
          break;
        case kSplit:  // bf16 operands of the bulk products: no effect on an FP64 replay
          break;
        case kPanel:  // run per launch above
          break;
      }
    }
  }
  std::memcpy(LxOut, c.Lx.data(), size_t(P.nnzLStored) * 8);
  if (DOut) std::memcpy(DOut, c.D.data(), size_t(S.n) * 8);
  if (pertOut) std::memcpy(pertOut, c.pert.data(), size_t(S.n));
  return c.failCol;
}

// Two-stream ordering contract of the factor plan. Launches run on the main stream or, with kSyncSide, on the side
// stream, in enqueue order per stream; kSyncWait makes a launch wait for the other stream's most recent kSyncSignal
// launch (as the device executor does). Every element of Lx, U, W and D that one launch writes must not be read or
// written by a launch of the other stream that is not ordered with it, nor by another task of the same launch.
// Returns the number of conflicting accesses (0 for a race-free plan). dropWaits ignores every kSyncWait (a plan
// with its waits removed must show conflicts, which pins the checker itself).
int64_t ps_check_streams(void* p, int32_t dropWaits) {
  Handle* h = static_cast<Handle*>(p);
  const Symbolic& S = h->S;
  const FactorPlan& P = h->F;
  const int32_t N = int32_t(P.launches.size());
  // vector clocks: vc[l][s] = sequence number (within stream s) of the last launch of stream s ordered before l
  std::vector<int32_t> vc0(N), vc1(N), st(N), seq(N);
  int32_t tail[2][2] = {{-1, -1}, {-1, -1}}, sig[2][2] = {{-1, -1}, {-1, -1}}, cnt[2] = {0, 0};
  bool haveSig[2] = {false, false};
  for (int32_t l = 0; l < N; l++) {
    const Launch& L = P.launches[l];
    const int s = (L.sync & kSyncSide) ? 1 : 0;
    int32_t v[2] = {tail[s][0], tail[s][1]};
    if (!dropWaits && (L.sync & kSyncWait) && haveSig[1 - s]) {
      v[0] = std::max(v[0], sig[1 - s][0]);
      v[1] = std::max(v[1], sig[1 - s][1]);
    }
    v[s] = cnt[s]++;
    vc0[l] = v[0];
    vc1[l] = v[1];
    st[l] = s;
    seq[l] = v[s];
    tail[s][0] = v[0];
    tail[s][1] = v[1];
    if (L.sync & kSyncSignal) {
      sig[s][0] = v[0];
      sig[s][1] = v[1];
      haveSig[s] = true;
    }
  }
  auto before = [&](int32_t a, int32_t b) {  // launch a complete before launch b starts (a != b)
    if (st[a] == st[b]) return a < b;
    return (st[a] == 0 ? vc0[b] : vc1[b]) >= seq[a];
  };
  // shadow state per element: last writer and last reader per stream (launch, task)
  struct Sh {
    std::vector<int32_t> wl, wt, rl[2], rt[2];
    void init(size_t n) {
      wl.assign(n, -1);
      wt.assign(n, -1);
      for (int s = 0; s < 2; s++) {
        rl[s].assign(n, -1);
        rt[s].assign(n, -1);
      }
    }
  };
  Sh sh[5];
  sh[0].init(size_t(std::max<int64_t>(P.nnzLStored, 1)));
  sh[1].init(size_t(std::max<int64_t>(P.uEntries, 1)));
  sh[2].init(size_t(std::max<int64_t>(P.wEntries, 1)));
  sh[3].init(size_t(std::max(S.n, 1)));
  sh[4].init(size_t(std::max<int64_t>(P.splitEntries, 1)));  // the split panel (one array stands for its four)
  int64_t bad = 0;
  int32_t curL = 0, curT = 0;
  bool readsOrdered = false;  // kPanel writes: the launch's own reads are ordered before them (launch counter)
  auto conflict = [&](int32_t ol, int32_t ot) {
    if (ol < 0) return false;
    if (ol == curL) return ot != curT;
    return !before(ol, curL);
  };
  auto acc = [&](int b, int64_t i, bool write) {
    Sh& x = sh[b];
    if (i < 0 || size_t(i) >= x.wl.size()) {
      bad++;
      return;
    }
    const int s = st[curL];
    if (conflict(x.wl[i], x.wt[i])) bad++;
    if (write) {
      for (int o = 0; o < 2; o++)
        if (!(readsOrdered && x.rl[o][i] == curL) && conflict(x.rl[o][i], x.rt[o][i])) bad++;
      x.wl[i] = curL;
      x.wt[i] = curT;
    } else {
      x.rl[s][i] = curL;
      x.rt[s][i] = curT;
    }
  };
  for (int32_t l = 0; l < N; l++) {
    const Launch& L = P.launches[l];
    curL = l;
    if (L.kind == kPanel) {
      // pass 0: every task's reads (the block, or kPanelPre its W and D; its A rows and the column tiles' A rows, its
      // C entries); pass 1: the writes (owner: the block, W, D; every task: its L rows and C entries; kPanelNext: the
      // next block, its W and D, before the task's wait, so not ordered after the launch's reads)
      for (int pass = 0; pass < 2; pass++) {
        readsOrdered = pass == 1;
        for (int64_t q = L.first; q < L.first + L.count; q++) {
          curT = int32_t(q - L.first);
          const PanelTask& t = P.panels[size_t(q)];
          const int32_t E = kTileEdge, nb = t.nb;
          const int64_t ld = t.ld;
          const bool wr = pass == 1, pre = t.flags & kPanelPre;
          if (!wr && pre) {
            for (int64_t i = 0; i < int64_t(E) * E; i++) acc(2, t.w_off + i, false);
            for (int32_t k = 0; k < nb; k++) acc(3, t.col0 + k, false);
          }
          if ((!wr && !pre) || (wr && t.owner))
            for (int32_t j = 0; j < nb; j++)
              for (int32_t i = j; i < nb; i++) acc(0, t.f_off + i + j * ld, wr);
          if (wr && t.owner) {
            for (int64_t i = 0; i < int64_t(E) * E; i++) acc(2, t.w_off + i, true);
            for (int32_t k = 0; k < nb; k++) acc(3, t.col0 + k, true);
          }
          if (wr && (t.flags & kPanelNext)) {
            readsOrdered = false;
            const int32_t nbn = std::min(E, t.ucols);
            for (int32_t j = 0; j < nbn; j++)
              for (int32_t i = j; i < nbn; i++) acc(0, t.f_off + E + i + (E + j) * ld, true);
            for (int64_t i = 0; i < int64_t(E) * E; i++) acc(2, t.w_off + int64_t(E) * E + i, true);
            for (int32_t k = 0; k < nbn; k++) acc(3, t.col0 + E + k, true);
            readsOrdered = true;
          }
          if (t.rows == 0) continue;
          for (int32_t j = 0; j < nb; j++)
            for (int32_t i = 0; i < t.rows; i++) acc(0, t.f_off + t.ro + i + j * ld, wr);
          const int32_t own = (t.ro - nb) / E;
          for (int32_t ct = 0; ct * E < t.ucols && ct <= own; ct++) {
            const int32_t wc = std::min(E, t.ucols - E * ct);
            if (!wr)
              for (int32_t j = 0; j < nb; j++)
                for (int32_t i = 0; i < wc; i++) acc(0, t.f_off + nb + E * ct + i + j * ld, false);
            for (int32_t j = 0; j < wc; j++)
              for (int32_t i = 0; i < t.rows; i++) acc(0, t.f_off + t.ro + i + (nb + E * ct + j) * ld, wr);
          }
        }
      }
      readsOrdered = false;
      continue;
    }
    for (int64_t t = L.first; t < L.first + L.count; t++) {
      curT = int32_t(t - L.first);
      switch (L.kind) {
        case kZero:
          for (int64_t i = 0; i < P.zeros[t].len; i++) acc(0, P.zeros[t].off + i, true);
          break;
        case kAssemble:
          if (P.adst[t] >= 0) acc(0, P.adst[t], true);
          break;
        case kSmall: {
          const int32_t s = P.smallFronts[t];
          const int32_t f = S.superStart[s], nc = S.superStart[s + 1] - f, m = S.Sp[s + 1] - S.Sp[s], mu = m - nc;
          for (int32_t ch = S.childHead[s]; ch != -1; ch = S.childNext[ch]) {
            const int32_t ncc = S.superStart[ch + 1] - S.superStart[ch], muc = S.Sp[ch + 1] - S.Sp[ch] - ncc;
            for (int32_t b = 0; b < muc; b++)
              for (int32_t a = b; a < muc; a++) acc(1, P.Uoff[ch] + a + int64_t(b) * muc, false);
          }
          for (int64_t i = 0; i < int64_t(m) * nc; i++) acc(0, S.Lxp[s] + i, true);
          for (int64_t i = 0; i < int64_t(mu) * mu; i++) acc(1, P.Uoff[s] + i, true);
          for (int32_t k = 0; k < nc; k++) acc(3, f + k, true);
          break;
        }
        case kExtendAdd: {
          const EaTask& e = P.ea[t];
          const int32_t s = e.s, nc = S.superStart[s + 1] - S.superStart[s], m = S.Sp[s + 1] - S.Sp[s], mu = m - nc;
          for (int32_t q = e.r0; q < e.r1; q++) {
            const int32_t ch = P.eaRanges[q].c;
            const int32_t ncc = S.superStart[ch + 1] - S.superStart[ch], muc = S.Sp[ch + 1] - S.Sp[ch] - ncc;
            const int32_t* rel = P.rel.data() + S.Sp[ch] + ncc;
            for (int32_t b = P.eaRanges[q].b0; b < P.eaRanges[q].b1; b++) {
              const int32_t j = rel[b];
              for (int32_t a = b; a < muc; a++) {
                acc(1, P.Uoff[ch] + a + int64_t(b) * muc, false);
                const int32_t i = rel[a];
                const int bb = j < nc ? 0 : 1;
                const int64_t d = j < nc ? S.Lxp[s] + i + int64_t(j) * m : P.Uoff[s] + (i - nc) + int64_t(j - nc) * mu;
                acc(bb, d, false);
                acc(bb, d, true);
              }
            }
          }
          break;
        }
        case kDiag: {
          const DiagTask& d = P.diag[t];
          for (int32_t j = 0; j < d.nb; j++)
            for (int32_t i = j; i < d.nb; i++) {
              acc(0, d.f_off + i + int64_t(j) * d.ld, false);
              acc(0, d.f_off + i + int64_t(j) * d.ld, true);
            }
          for (int64_t i = 0; i < int64_t(kTileEdge) * kTileEdge; i++) acc(2, d.w_off + i, true);
          for (int32_t k = 0; k < d.nb; k++) acc(3, d.col0 + k, true);
          break;
        }
        case kPanel:  // checked per launch above
          break;
        case kSplit: {
          const SplitTask& sp = P.splits[t];
          for (int32_t d = 0; d < sp.w; d++) {
            acc(3, sp.d_off + d, false);
            for (int32_t r = 0; r < sp.rows; r++) acc(0, sp.src + r + int64_t(d) * sp.m, false);
          }
          for (int32_t r = 0; r < sp.rows; r++)
            for (int32_t d = 0; d < sp.dp; d++)
              acc(4, sp.dst + int64_t(d / kSplitDepth) * sp.rows * kSplitDepth + int64_t(r) * kSplitDepth + d % kSplitDepth,
                  true);
          break;
        }
        case kGemm: {
          const GemmTask& g = P.gemm[t];
          for (int32_t k = 0; k < g.k; k++) {
            for (int32_t i = 0; i < g.rows; i++) acc(g.a_buf, g.a_off + i + int64_t(k) * g.lda, false);
            for (int32_t j = 0; j < g.cols; j++) acc(g.b_buf, g.b_off + j + int64_t(k) * g.ldb, false);
            if (g.mode != 2) acc(3, g.d_off + k, false);
          }
          if (g.mode == 2)
            for (int32_t j = 0; j < g.cols; j++) acc(3, g.d_off + j, false);
          const SplitRef& r = P.srefs[t];
          if (L.kind == kGemm && L.edge == kGemmEdge && (L.sync & kSyncSide) && r.ld > 0)
            for (int32_t k = 0; k < g.k; k++) {  // the bf16 kernel reads its operands from the split panel
              const int64_t c = int64_t(k / kSplitDepth) * r.ld + k % kSplitDepth;
              for (int32_t i = 0; i < g.rows; i++) acc(4, r.a + c + int64_t(i) * kSplitDepth, false);
              for (int32_t j = 0; j < g.cols; j++) acc(4, r.b + c + int64_t(j) * kSplitDepth, false);
            }
          for (int32_t j = 0; j < g.cols; j++)
            for (int32_t i = 0; i < g.rows; i++) {
              const int64_t d = g.c_off + i + int64_t(j) * g.ldc;
              if (g.mode == 0) acc(g.c_buf, d, false);
              acc(g.c_buf, d, true);
            }
          break;
        }
      }
    }
  }
  return bad;
}

// Task list of a flow launch for the kernels of FP32 factors (list == kFlowQuadsF32) or FP64 factors, and its blocks
// per task.
static int32_t flowRange(const SolveLaunch& L, int32_t list, int64_t& first, int64_t& count) {
  first = list == kFlowQuadsF32 ? L.first : L.first2;
  count = list == kFlowQuadsF32 ? L.count : L.count2;
  return list == kFlowQuadsF32 ? L.nq : L.nq2;
}

// Ordering contract of the tiled-supernode solve launches, for the task lists of both precisions (their workgroups wait on flags
// raised by earlier tasks of the same launch, so an order violation would be a deadlock on the device): returns the
// number of violations. Fused launches also wait on counters: every wait must name exactly the tasks whose stores the
// waiting task reads (the below tasks of a supernode's children for its gather tasks, its gather tasks for its flow
// tasks; backward, the chain tasks of its parent for its initialisation tasks, those for its chain tasks), all listed
// before it, and a counter is raised by the tasks of one launch only.
int64_t ps_check_flow(void* p) {
  Handle* h = static_cast<Handle*>(p);
  const Symbolic& S = h->S;
  const SolvePlan& P = h->SP;
  int64_t bad = 0;
  auto front = [&S](int32_t s, int32_t& nc, int32_t& m, int32_t& nb) {
    nc = S.superStart[s + 1] - S.superStart[s];
    m = S.Sp[s + 1] - S.Sp[s];
    nb = (nc + kTileEdge - 1) / kTileEdge;
  };
  for (const int32_t nq : {kFlowQuadsF32, kFlowQuadsF64}) {
    std::vector<int32_t> flag0(S.nsuper, -1);
    std::vector<int8_t> seenFwd(size_t(P.flowFlags), 0), seenBwd(size_t(P.flowFlags), 0);
    for (const SolveLaunch& L : P.launches) {
      if (!solveKindIsFlow(L.kind)) continue;
      const bool fwd = solveKindIsForward(L.kind), fused = L.kind == kSolveFwdFused || L.kind == kSolveBwdFused;
      std::vector<int8_t>& seen = fwd ? seenFwd : seenBwd;
      std::map<int32_t, std::vector<int8_t>> done;  // per supernode: blocks solved so far in this launch
      int64_t first, count;
      const int32_t width = flowRange(L, nq, first, count);
      bad += width != 1 && (width != nq || !fwd || fused);
      bad += fused && (L.first2 != L.first || L.count2 != L.count);
      for (int64_t t = first; t < first + count; t++) {
        const BlockTask& b = P.blocks[t];
        int32_t nc, m, nb;
        front(b.s, nc, m, nb);
        if (fused && ((fwd && b.writer < 0) || (!fwd && b.writer == 0))) continue;  // gather / initialisation task
        if (flag0[b.s] == -1) flag0[b.s] = b.flag0;
        if (flag0[b.s] != b.flag0 || b.flag0 < 0 || b.flag0 + nb > P.flowFlags) bad++;
        auto& d = done[b.s];
        d.resize(nb, 0);
        if (b.writer < 0 || b.writer > width) {
          bad++;
          continue;
        }
        if (!b.writer) {  // forward below task: every block, at most nq x 64 below rows
          bad += !fwd || b.blk != nb || b.r0 < nc || b.r1 > m || b.r1 <= b.r0 || b.r1 - b.r0 > width * kTileEdge;
          for (int32_t k = 0; k < nb; k++) bad += d[k] == 0;
          continue;
        }
        const int32_t lo = b.blk, hi = b.blk + b.writer;  // the task's blocks [lo, hi)
        if (lo < 0 || hi > nb || b.r0 != lo * kTileEdge || b.r1 != std::min(nc, hi * kTileEdge)) {
          bad++;
          continue;
        }
        if (fwd)
          for (int32_t k = 0; k < lo; k++) bad += d[k] == 0;
        else
          for (int32_t k = hi; k < nb; k++) bad += d[k] == 0;
        for (int32_t k = lo; k < hi; k++) {
          bad += d[k] != 0;
          d[k] = 1;
          seen[b.flag0 + k]++;
        }
      }
    }
    for (int64_t i = 0; i < P.flowFlags; i++) bad += (seenFwd[i] != 1) + (seenBwd[i] != 1);
    // every initialisation slot is written by exactly one task listed before the flow task that reads it (in an
    // earlier launch, or earlier in the same fused launch)
    std::vector<int64_t> written(size_t(P.initSlots), -1);
    int64_t seq = 0;  // position of a task in the order of all launches
    for (size_t li = 0; li < P.launches.size(); li++) {
      const SolveLaunch& L = P.launches[li];
      if (L.kind != kSolveBwdInit && L.kind != kSolveBwdFlow && L.kind != kSolveBwdFused) continue;
      int64_t first = L.first, count = L.count;
      if (L.kind != kSolveBwdInit) flowRange(L, nq, first, count);
      for (int64_t t = first; t < first + count; t++, seq++) {
        const BlockTask& b = P.blocks[t];
        if (L.kind == kSolveBwdInit || (L.kind == kSolveBwdFused && b.writer == 0)) {
          if (b.p0 < 0 || b.p0 >= P.initSlots || written[b.p0] != -1)
            bad++;
          else
            written[b.p0] = seq;
          continue;
        }
        const int32_t nc = S.superStart[b.s + 1] - S.superStart[b.s], mu = S.Sp[b.s + 1] - S.Sp[b.s] - nc;
        if (b.np != (mu + kInitRows - 1) / kInitRows) bad++;
        for (int32_t p = b.p0; p < b.p0 + b.writer * b.np; p++)
          bad += p < 0 || p >= P.initSlots || written[p] < 0 || written[p] >= seq;
      }
    }
    for (int64_t w : written) bad += w < 0;
  }
  // counters of the fused launches
  std::vector<int32_t> slotLaunch(size_t(P.doneSlots), -1);
  for (size_t li = 0; li < P.launches.size(); li++) {
    const SolveLaunch& L = P.launches[li];
    if (L.kind != kSolveFwdFused && L.kind != kSolveBwdFused) continue;
    const bool fwd = L.kind == kSolveFwdFused;
    // role of a task: 0 gather (forward) / initialisation (backward), 1 chain, 2 below (forward)
    auto role = [&](const BlockTask& b) { return fwd ? (b.writer < 0 ? 0 : b.writer > 0 ? 1 : 2) : (b.writer == 0 ? 0 : 1); };
    std::map<int32_t, std::vector<int64_t>> bySlot;     // counter -> tasks raising it
    std::map<std::pair<int32_t, int>, std::vector<int64_t>> byRole;  // (supernode, role) -> tasks
    for (int64_t t = L.first; t < L.first + L.count; t++) {
      const BlockTask& b = P.blocks[t];
      byRole[{b.s, role(b)}].push_back(t);
      if (b.dslot >= 0) {
        if (b.dslot >= P.doneSlots || (slotLaunch[b.dslot] != -1 && slotLaunch[b.dslot] != int32_t(li))) bad++;
        else slotLaunch[b.dslot] = int32_t(li);
        bySlot[b.dslot].push_back(t);
      }
    }
    // forward gather tasks of every supernode in the launch cover its rows in slices of kGatherRows, in order
    if (fwd)
      for (auto& kv : byRole) {
        if (kv.first.second != 0) continue;
        int32_t nc, m, nb;
        front(kv.first.first, nc, m, nb);
        bad += int64_t(kv.second.size()) != (m + kGatherRows - 1) / kGatherRows;
        for (size_t i = 0; i < kv.second.size(); i++) {
          const BlockTask& b = P.blocks[kv.second[i]];
          bad += b.r0 != int32_t(i) * kGatherRows || b.r1 != std::min(m, b.r0 + kGatherRows);
        }
      }
    for (int64_t t = L.first; t < L.first + L.count; t++) {
      const BlockTask& b = P.blocks[t];
      const int r = role(b);
      std::vector<int64_t> need;  // the tasks whose stores this task reads
      auto add = [&](int32_t s, int rl) {
        auto it = byRole.find({s, rl});
        if (it != byRole.end()) need.insert(need.end(), it->second.begin(), it->second.end());
      };
      if (fwd && r == 0) {
        for (int32_t c = S.childHead[b.s]; c != -1; c = S.childNext[c]) add(c, 2);
      } else if (fwd) {
        add(b.s, 0);
        bad += need.empty();  // every supernode of a forward fused launch is gathered in it
      } else if (r == 0) {
        if (S.superParent[b.s] >= 0) add(S.superParent[b.s], 1);
      } else {
        add(b.s, 0);
      }
      std::sort(need.begin(), need.end());
      std::vector<int64_t> got;
      if (b.wslot >= 0) {
        auto it = bySlot.find(b.wslot);
        if (it != bySlot.end()) got = it->second;
      }
      std::sort(got.begin(), got.end());
      bad += got != need;
      bad += b.wslot >= 0 ? b.wcount != int32_t(got.size()) : !need.empty();
      for (int64_t u : got) bad += u >= t;
    }
  }
  return bad;
}

// Solve launches of each kind (out[0 .. 8], SolveKind) and the counters of the fused launches (out[9]).
void ps_solve_launches(void* p, int64_t* out) {
  Handle* h = static_cast<Handle*>(p);
  for (int i = 0; i < 10; i++) out[i] = 0;
  for (const SolveLaunch& L : h->SP.launches) out[L.kind]++;
  out[9] = h->SP.doneSlots;
}

void ps_set_flow_width(void* p, int32_t nq) { static_cast<Handle*>(p)->flowWidth = nq; }

// Whole fronts in launches that pack several to a wave: out[0] / out[1] factor fronts packed by 8 / 16 lanes,
// out[2] / out[3] forward-solve fronts packed by 8 / 16, out[4] / out[5] the same for the backward solve. Returns the
// number of packed fronts taller than their segment (must be 0).
int64_t ps_packed(void* p, int64_t* out) {
  Handle* h = static_cast<Handle*>(p);
  const Symbolic& S = h->S;
  int64_t bad = 0;
  std::fill(out, out + 6, 0);
  auto rows = [&S](int32_t s) { return S.Sp[s + 1] - S.Sp[s]; };
  for (const Launch& L : h->F.launches) {
    if (L.kind != kSmall || L.threads != 0 || L.edge == 0) continue;
    out[L.edge == 8 ? 0 : 1] += L.count;
    for (int64_t i = L.first; i < L.first + L.count; i++) bad += rows(h->F.smallFronts[size_t(i)]) > L.edge;
  }
  for (const SolveLaunch& L : h->SP.launches) {
    if ((L.kind != kSolveSmallFwd && L.kind != kSolveSmallBwd) || L.seg == 0) continue;
    out[(L.kind == kSolveSmallFwd ? 2 : 4) + (L.seg == 8 ? 0 : 1)] += L.count;
    for (int64_t i = L.first; i < L.first + L.count; i++) bad += rows(h->SP.fronts[size_t(i)]) > L.seg;
  }
  return bad;
}

// solve q right-hand sides (one after another, original ordering) with the last factor
void ps_solve(void* p, const double* b, double* x, int32_t q) {
  Handle* h = static_cast<Handle*>(p);
  const Symbolic& S = h->S;
  const FactorPlan& F = h->F;
  const SolvePlan& P = h->SP;
  Ctx& c = h->c;
  const int32_t n = S.n;
  const double* Lx = c.Lx.data();
  std::vector<double> y(size_t(n) * q), fv(size_t(std::max<int64_t>(P.fvEntries, 1)) * q, std::nan(""));
  std::vector<double> pz(size_t(std::max<int64_t>(P.initSlots, 1)) * kTileEdge * q, std::nan(""));
  for (int32_t k = 0; k < q; k++)
    for (int32_t i = 0; i < n; i++) y[size_t(k) * n + i] = b[size_t(k) * n + S.perm[i]];
  bool scaled = false;
  for (const SolveLaunch& L : P.launches) {
    if (c.ldlt && !scaled && !solveKindIsForward(L.kind)) {
      for (int32_t k = 0; k < q; k++)
        for (int32_t j = 0; j < n; j++) y[size_t(k) * n + j] /= c.D[j];
      scaled = true;
    }
    int64_t first = L.first, count = L.count;
    if (solveKindIsFlow(L.kind)) flowRange(L, h->flowWidth, first, count);
    for (int64_t t = first; t < first + count; t++) {
      // a fused launch's task runs as the task of its kind (gather rows of one slice, initialisation, flow)
      SolveKind kind = L.kind;
      if (kind == kSolveFwdFused) kind = P.blocks[t].writer < 0 ? kSolveGather : kSolveFwdFlow;
      if (kind == kSolveBwdFused) kind = P.blocks[t].writer == 0 ? kSolveBwdInit : kSolveBwdFlow;
      const bool slice = L.kind == kSolveFwdFused && kind == kSolveGather;
      for (int32_t k = 0; k < q; k++) {
        double* yk = y.data() + size_t(k) * n;
        double* fk = fv.data() + size_t(k) * P.fvEntries;
        if (kind == kSolveBwdInit) {
          const BlockTask& bt = P.blocks[t];
          const int32_t s = bt.s, nc = S.superStart[s + 1] - S.superStart[s], m = S.Sp[s + 1] - S.Sp[s];
          const double* Ls = Lx + S.Lxp[s];
          for (int32_t j = 0; j < kTileEdge; j++) {
            const int32_t col = bt.blk * kTileEdge + j;
            double sum = 0;
            if (col < nc)
              for (int32_t r = bt.r0; r < bt.r1; r++) sum += Ls[nc + r + int64_t(col) * m] * yk[S.Si[S.Sp[s] + nc + r]];
            pz[(size_t(bt.p0) * q + k) * kTileEdge + j] = sum;
          }
        } else if (kind == kSolveBwdGather) {
          const int32_t s = P.fronts[t];
          const int32_t nc = S.superStart[s + 1] - S.superStart[s], m = S.Sp[s + 1] - S.Sp[s];
          for (int32_t r = nc; r < m; r++) fk[S.Sp[s] + r] = yk[S.Si[S.Sp[s] + r]];
        } else if (kind == kSolveSmallFwd || kind == kSolveGather) {
          const int32_t s = slice ? P.blocks[t].s : P.fronts[t];
          const int32_t f = S.superStart[s], nc = S.superStart[s + 1] - f, m = S.Sp[s + 1] - S.Sp[s];
          std::vector<double> v(m);
          for (int32_t i = 0; i < m; i++) v[i] = i < nc ? yk[f + i] : 0.0;
          for (int32_t ch = S.childHead[s]; ch != -1; ch = S.childNext[ch]) {
            const int32_t ncc = S.superStart[ch + 1] - S.superStart[ch], muc = S.Sp[ch + 1] - S.Sp[ch] - ncc;
            for (int32_t r = 0; r < muc; r++) v[F.rel[S.Sp[ch] + ncc + r]] += fk[S.Sp[ch] + ncc + r];
          }
          if (kind == kSolveSmallFwd) {
            const double* Ls = Lx + S.Lxp[s];
            for (int32_t col = 0; col < nc; col++) {
              const double xc = c.ldlt ? v[col] : v[col] / Ls[col + int64_t(col) * m];
              v[col] = xc;
              for (int32_t r = col + 1; r < m; r++) v[r] -= Ls[r + int64_t(col) * m] * xc;
            }
            for (int32_t i = 0; i < nc; i++) yk[f + i] = v[i];
          }
          const int32_t i0 = slice ? P.blocks[t].r0 : 0, i1 = slice ? P.blocks[t].r1 : m;
          for (int32_t i = i0; i < i1; i++) fk[S.Sp[s] + i] = v[i];
        } else if (kind == kSolveFwdFlow) {
          // tasks run in list order: every block a task consumes was solved by an earlier task or, in a chain task,
          // by the task itself before
          const BlockTask& bt = P.blocks[t];
          const int32_t s = bt.s, f = S.superStart[s], nc = S.superStart[s + 1] - f;
          const int32_t m = S.Sp[s + 1] - S.Sp[s];
          const double* Ls = Lx + S.Lxp[s];
          if (!bt.writer) {
            for (int32_t r = bt.r0; r < bt.r1; r++) {
              double sum = 0;
              for (int32_t col = 0; col < nc; col++) sum += Ls[r + int64_t(col) * m] * yk[f + col];
              fk[S.Sp[s] + r] -= sum;
            }
          }
          for (int32_t i = 0; i < bt.writer; i++) {
            const int32_t kb = bt.blk + i, r0 = kb * kTileEdge, r1 = std::min(nc, r0 + kTileEdge);
            std::vector<double> v(kTileEdge, 0.0);
            for (int32_t r = r0; r < r1; r++) {
              double sum = 0;
              for (int32_t col = 0; col < r0; col++) sum += Ls[r + int64_t(col) * m] * yk[f + col];
              v[r - r0] = fk[S.Sp[s] + r] - sum;
            }
            const double* W = c.W.data() + F.Woff[s] + int64_t(kb) * kTileEdge * kTileEdge;
            for (int32_t r = r0; r < r1; r++) {
              double sum = 0;
              for (int32_t j = 0; j < kTileEdge; j++) sum += W[(r - r0) + int64_t(j) * kTileEdge] * v[j];
              yk[f + r] = sum;
            }
          }
        } else if (kind == kSolveSmallBwd) {
          const int32_t s = P.fronts[t];
          const int32_t f = S.superStart[s], nc = S.superStart[s + 1] - f, m = S.Sp[s + 1] - S.Sp[s];
          const double* Ls = Lx + S.Lxp[s];
          std::vector<double> z(nc);
          for (int32_t i = 0; i < nc; i++) {
            double sum = yk[f + i];
            for (int32_t r = nc; r < m; r++) sum -= Ls[r + int64_t(i) * m] * yk[S.Si[S.Sp[s] + r]];
            z[i] = sum;
          }
          for (int32_t col = nc - 1; col >= 0; col--) {
            const double xc = c.ldlt ? z[col] : z[col] / Ls[col + int64_t(col) * m];
            z[col] = xc;
            for (int32_t i = 0; i < col; i++) z[i] -= Ls[col + int64_t(i) * m] * xc;
          }
          for (int32_t i = 0; i < nc; i++) yk[f + i] = z[i];
        } else if (kind == kSolveBwdFlow) {
          const BlockTask& bt = P.blocks[t];
          const int32_t s = bt.s, f = S.superStart[s], nc = S.superStart[s + 1] - f, m = S.Sp[s + 1] - S.Sp[s];
          const double* Ls = Lx + S.Lxp[s];
          for (int32_t i = bt.writer - 1; i >= 0; i--) {
            const int32_t kb = bt.blk + i, c0 = kb * kTileEdge, c1 = std::min(nc, c0 + kTileEdge);
            const double* W = c.W.data() + F.Woff[s] + int64_t(kb) * kTileEdge * kTileEdge;
            std::vector<double> z(kTileEdge, 0.0);
            for (int32_t col = c0; col < c1; col++) {
              double below = 0;  // the partial products of the initialisation tasks, in slot order
              for (int32_t pp = 0; pp < bt.np; pp++)
                below += pz[(size_t(bt.p0 + i * bt.np + pp) * q + k) * kTileEdge + (col - c0)];
              double sum = yk[f + col] - below;
              for (int32_t r = c1; r < nc; r++) sum -= Ls[r + int64_t(col) * m] * yk[f + r];
              z[col - c0] = sum;
            }
            for (int32_t j = 0; j < c1 - c0; j++) {
              double sum = 0;
              for (int32_t i2 = 0; i2 < kTileEdge; i2++) sum += W[i2 + int64_t(j) * kTileEdge] * z[i2];
              yk[f + c0 + j] = sum;
            }
          }
        }
      }
    }
  }
  for (int32_t k = 0; k < q; k++)
    for (int32_t i = 0; i < n; i++) x[size_t(k) * n + S.perm[i]] = y[size_t(k) * n + i];
}

}  // extern "C"
