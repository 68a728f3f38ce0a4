// Host construction of the factor and solve plans (see plan.h).
#include "plan.h"

#include <stdexcept>

#include <algorithm>
#include <functional>

namespace amp {

namespace {

inline int32_t frontRows(const Symbolic& S, int32_t s) { return S.Sp[s + 1] - S.Sp[s]; }
inline int32_t frontCols(const Symbolic& S, int32_t s) { return S.superStart[s + 1] - S.superStart[s]; }

}  // namespace

FactorPlan buildFactorPlan(const Symbolic& S, const PlanParams& prm) {
  const int32_t SPW = std::max(kTileEdge, prm.superPanel / kTileEdge * kTileEdge);  // super-panel width
  FactorPlan P;
  const int32_t n = S.n, ns = S.nsuper;
  P.n = n;
  P.nsuper = ns;
  P.nnzLStored = S.nnzLStored;
  P.nnzA = S.nnzA;
  P.nnzPattern = S.nnzPattern;
  P.flops = S.flops;
  P.level.assign(ns, 0);
  P.isBig.assign(ns, 0);
  P.Uoff.assign(ns, -1);
  P.Woff.assign(ns, -1);
  if (ns == 0) return P;

  // levels: children carry smaller indices than their parent (postorder)
  for (int32_t s = 0; s < ns; s++) {
    const int32_t p = S.superParent[s];
    if (p >= 0) P.level[p] = std::max(P.level[p], P.level[s] + 1);
  }
  P.nlevels = 1 + *std::max_element(P.level.begin(), P.level.end());

  int64_t uTotal = 0, wTotal = 0;
  for (int32_t s = 0; s < ns; s++) {
    const int32_t m = frontRows(S, s), nc = frontCols(S, s), mu = m - nc;
    P.isBig[s] = m > std::min(prm.smallMax, kSmallMax);
    // the product kernels address a tile's operands with 32-bit offsets: depth x leading dimension < 2^31
    if (P.isBig[s] && int64_t(m) * std::max(SPW, kGemmEdge) >= (int64_t(1) << 31))
      throw std::length_error("a tiled front is too tall for 32-bit tile offsets");
    if (mu > 0) {
      P.Uoff[s] = uTotal;
      uTotal += int64_t(mu) * mu;
    }
    if (P.isBig[s]) {
      P.Woff[s] = wTotal;
      wTotal += int64_t((nc + kTileEdge - 1) / kTileEdge) * kTileEdge * kTileEdge;
    }
  }
  P.uEntries = uTotal;
  P.wEntries = wTotal;

  // relative maps: position of each child below row in the parent's row list
  P.rel.assign(S.Si.size(), -1);
  {
    std::vector<int32_t> pos(n, -1);
    for (int32_t p = 0; p < ns; p++) {
      if (S.childHead[p] == -1) continue;
      for (int32_t i = S.Sp[p]; i < S.Sp[p + 1]; i++) pos[S.Si[i]] = i - S.Sp[p];
      for (int32_t c = S.childHead[p]; c != -1; c = S.childNext[c])
        for (int32_t r = S.Sp[c] + frontCols(S, c); r < S.Sp[c + 1]; r++) P.rel[r] = pos[S.Si[r]];
    }
  }
  // assembly maps of A's lower entries: in-block offsets (int32) for whole fronts, whose blocks hold at most
  // kSmallMax^2 entries; absolute Lx indices (int64) for tiled fronts, whose m x nc blocks can pass 2^31 entries
  P.amap.assign(S.nnzA, 0);
  P.adst.assign(S.nnzA, -1);
  {
    std::vector<int32_t> pos(n, -1);
    for (int32_t s = 0; s < ns; s++) {
      const int32_t f = S.superStart[s], nc = frontCols(S, s), m = frontRows(S, s);
      for (int32_t i = S.Sp[s]; i < S.Sp[s + 1]; i++) pos[S.Si[i]] = i - S.Sp[s];
      for (int32_t c = 0; c < nc; c++)
        for (int32_t q = S.lowerP[f + c]; q < S.lowerP[f + c + 1]; q++) {
          const int64_t off = int64_t(pos[S.lowerI[q]]) + int64_t(c) * m;
          if (P.isBig[s])
            P.adst[q] = S.Lxp[s] + off;
          else
            P.amap[q] = int32_t(off);
        }
    }
  }

  // extend-add work item: parent columns [j0, j1) of front s, with the column range of every child that lands there
  auto addEa = [&S, &P](int32_t s, int32_t j0, int32_t j1) {
    EaTask t{s, j0, j1, int32_t(P.eaRanges.size()), 0, 0};
    for (int32_t c = S.childHead[s]; c != -1; c = S.childNext[c]) {
      const int32_t ncc = frontCols(S, c);
      const int32_t* rel = P.rel.data() + S.Sp[c] + ncc;
      const int32_t muc = frontRows(S, c) - ncc;
      const int32_t b0 = int32_t(std::lower_bound(rel, rel + muc, j0) - rel);
      const int32_t b1 = int32_t(std::lower_bound(rel, rel + muc, j1) - rel);
      if (b1 > b0) P.eaRanges.push_back(EaRange{c, b0, b1, 0});
    }
    if (P.eaRanges.size() > size_t(INT32_MAX)) throw std::length_error("extend-add ranges exceed 2^31");
    t.r1 = int32_t(P.eaRanges.size());
    if (t.r1 > t.r0) P.ea.push_back(t);  // items with no contribution are left out
  };
  bool joinPending = false;  // the next main-stream launch must wait for the side stream
  auto push = [&P, &joinPending](LaunchKind kind, int32_t level, int64_t first, int64_t count, int32_t maxM = 0,
                                 int32_t threads = 0, int32_t edge = 0, int32_t sync = 0) {
    if (count <= 0) return false;
    if (joinPending && !(sync & kSyncSide)) {
      sync |= kSyncWait;
      joinPending = false;
    }
    P.launches.push_back(Launch{kind, level, first, count, maxM, threads, edge, sync});
    if (sync & kSyncSignal) P.signals++;
    return true;
  };
  // one product launch: its tiles at edge kGemmEdge when there are enough of them to fill the device, else at
  // kTileEdge; `add(edge, out)` appends the tiles of every front of the launch. Returns whether a launch was made.
  auto gemmLaunch = [&P, &push, &prm](int32_t level, const std::function<void(int32_t, std::vector<GemmTask>&)>& add,
                                int32_t sync = 0) {
    std::vector<GemmTask> tl;
    int32_t edge = kGemmEdge;
    add(edge, tl);
    if (int64_t(tl.size()) < prm.wideLaunch) {
      tl.clear();
      edge = kTileEdge;
      add(edge, tl);
    }
    const int64_t first = int64_t(P.gemm.size());
    P.gemm.insert(P.gemm.end(), tl.begin(), tl.end());
    return push(kGemm, level, first, int64_t(tl.size()), 0, 0, edge, sync);
  };

  // zero the L blocks of tiled fronts, then scatter A into them
  for (int32_t s = 0; s < ns; s++) {
    if (!P.isBig[s]) continue;
    const int64_t len = S.Lxp[s + 1] - S.Lxp[s];
    for (int64_t o = 0; o < len; o += kZeroChunk) P.zeros.push_back(ZeroTask{S.Lxp[s] + o, std::min<int64_t>(kZeroChunk, len - o)});
  }
  // on the side stream, overlapping the whole-front levels (which touch neither these blocks nor A's entries in
  // them); the first launch of the first level with tiled fronts waits for it
  push(kZero, -1, 0, int64_t(P.zeros.size()), 0, 0, 0, kSyncSide);
  push(kAssemble, -1, 0, S.nnzA, 0, 0, 0, kSyncSide | kSyncSignal);
  bool assemblyPending = true;

  std::vector<std::vector<int32_t>> byLevel(P.nlevels);
  for (int32_t s = 0; s < ns; s++) byLevel[P.level[s]].push_back(s);

  for (int32_t L = 0; L < P.nlevels; L++) {
    const std::vector<int32_t>& fronts = byLevel[L];
    // fronts factored in LDS, five size classes: up to 8 rows eight to a wave when the level has at least packMin of
    // them (else with the next class), up to 16 rows four to a wave when the level has at least packMin of them (else
    // one wave each), one wave per front up to kWaveFrontMax rows, then a workgroup of 64 threads up to smallNarrow
    // rows, else 256
    int64_t n8 = 0, n16 = 0;
    for (int32_t s : fronts)
      if (!P.isBig[s]) {
        n8 += frontRows(S, s) <= 8;
        n16 += frontRows(S, s) <= 16;
      }
    const bool pack8 = n8 >= prm.packMin, pack16 = n16 - (pack8 ? n8 : 0) >= prm.packMin;
    for (int cls = 0; cls < 5; cls++) {
      const int64_t first = int64_t(P.smallFronts.size());
      int32_t maxM = 0;
      for (int32_t s : fronts) {
        const int32_t m = frontRows(S, s);
        const int c = (m <= 8 && pack8) ? 0 : m <= 16 ? 1 : m <= kWaveFrontMax ? 2 : m <= prm.smallNarrow ? 3 : 4;
        if (P.isBig[s] || c != cls) continue;
        P.smallFronts.push_back(s);
        maxM = std::max(maxM, m);
      }
      const int32_t seg = cls == 0 ? 8 : (cls == 1 && pack16) ? 16 : 0;
      push(kSmall, L, first, int64_t(P.smallFronts.size()) - first, maxM, cls < 3 ? 0 : cls == 3 ? 64 : 256, seg);
    }
    std::vector<int32_t> big;
    for (int32_t s : fronts)
      if (P.isBig[s]) big.push_back(s);
    if (big.empty()) continue;
    if (assemblyPending) {
      joinPending = true;
      assemblyPending = false;
    }

    // children's contributions to the factor columns
    {
      const int64_t first = int64_t(P.ea.size());
      for (int32_t s : big) {
        if (S.childHead[s] == -1) continue;
        const int32_t nc = frontCols(S, s);
        for (int32_t j = 0; j < nc; j += kEaWidthF) addEa(s, j, std::min(nc, j + kEaWidthF));
      }
      push(kExtendAdd, L, first, int64_t(P.ea.size()) - first);
    }
    // panel factorization, all tiled fronts of the level in lockstep, right-looking in super-panels of SPW columns
    // columns: a super-panel is factored in 64-column steps whose trailing updates stay inside it (main stream);
    // then the next super-panel's columns are brought up to date from it (main stream, lookahead) while every later
    // column is updated from it on the side stream (bulk), concurrently with the next super-panel's factorization.
    // Ordering: the bulk update of super-panel s waits for s to be factored (signal after its last step); the
    // lookahead update from s + 1 waits for the bulk update of s (the side stream's last signal at that point), which
    // also wrote the columns of s + 2; the update-matrix extend-add after the level's last super-panel joins the side
    // stream (the bulk launches also accumulate the Schur complement, super-panel by super-panel).
    int32_t maxNc = 0;
    for (int32_t s : big) maxNc = std::max(maxNc, frontCols(S, s));
    bool sideUsed = false;
    // C[i0.., j0..] -= L[i0.., d0 : d0 + depth] L[j0.., d0 : d0 + depth]^T over columns [j0from, j0to) and rows
    // [j0, m) of front s (lower tiles only), as tiles of the given edge
    auto lowerUpdate = [&S](std::vector<GemmTask>& out, int32_t edge, int32_t s, int32_t j0from, int32_t j0to,
                            int32_t d0, int32_t depth) {
      const int32_t m = frontRows(S, s), f = S.superStart[s];
      const int64_t base = S.Lxp[s];
      for (int32_t j0 = j0from; j0 < j0to; j0 += edge)
        for (int32_t i0 = j0; i0 < m; i0 += edge) {
          GemmTask t{};
          t.a_off = base + i0 + int64_t(d0) * m;
          t.b_off = base + j0 + int64_t(d0) * m;
          t.c_off = base + i0 + int64_t(j0) * m;
          t.lda = t.ldb = t.ldc = m;
          t.rows = std::min(edge, m - i0);
          t.cols = std::min(edge, j0to - j0);
          t.k = depth;
          t.d_off = f + d0;
          t.a_buf = t.b_buf = t.c_buf = kBufLx;
          t.mode = 0;
          out.push_back(t);
        }
    };
    // U[i0.., j0..] (op)= -L21[., d0 : d0 + depth] D L21[., d0 : d0 + depth]^T over the lower tiles of the mu x mu
    // update matrix of front s: written by the first super-panel (mode 1), accumulated by the later ones (mode 0)
    auto schurUpdate = [&S, &P](std::vector<GemmTask>& out, int32_t edge, int32_t s, int32_t d0, int32_t depth) {
      const int32_t nc = frontCols(S, s), m = frontRows(S, s), mu = m - nc, f = S.superStart[s];
      if (mu == 0 || depth <= 0) return;
      const int64_t base = S.Lxp[s];
      for (int32_t j0 = 0; j0 < mu; j0 += edge)
        for (int32_t i0 = j0; i0 < mu; i0 += edge) {
          GemmTask t{};
          t.a_off = base + nc + i0 + int64_t(d0) * m;
          t.b_off = base + nc + j0 + int64_t(d0) * m;
          t.c_off = P.Uoff[s] + i0 + int64_t(j0) * mu;
          t.lda = t.ldb = m;
          t.ldc = mu;
          t.rows = std::min(edge, mu - i0);
          t.cols = std::min(edge, mu - j0);
          t.k = depth;
          t.d_off = f + d0;
          t.a_buf = t.b_buf = kBufLx;
          t.c_buf = kBufU;
          t.mode = d0 == 0 ? 1 : 0;
          out.push_back(t);
        }
    };
    // In front of a bulk launch of kGemmEdge tiles (the last launch so far), the split of its panel for the bf16
    // products: rows [r0, m) of columns [c0, c0 + w) of every front the launch updates, r0 = c2 when it updates
    // later columns, else nc (the Schur complement's rows only). The split launch takes over the bulk launch's wait
    // (it reads the factored super-panel); the bulk launch follows it on the side stream. The bulk tasks' split
    // operands are their A and B rows in the split panel.
    auto addSplit = [&](int32_t L, int32_t c0, int32_t c2, const std::vector<int32_t>& fronts) {
      Launch bulk = P.launches.back();
      P.launches.pop_back();
      std::vector<int64_t> dst(size_t(ns), -1);
      std::vector<int32_t> row0(size_t(ns), 0);
      const int64_t first = int64_t(P.splits.size());
      int64_t total = 0;
      int32_t maxRows = 0, maxDp = 0;
      for (int32_t s : fronts) {
        const int32_t nc = frontCols(S, s), m = frontRows(S, s);
        if (nc <= c0 || (nc <= c2 && m == nc)) continue;
        const int32_t r0 = nc > c2 ? c2 : nc, w = std::min(SPW, nc - c0);
        const int32_t dp = (w + kSplitDepth - 1) / kSplitDepth * kSplitDepth;
        P.splits.push_back(SplitTask{S.Lxp[s] + r0 + int64_t(c0) * m, total, m, m - r0, w, dp, S.superStart[s] + c0, 0});
        dst[s] = total;
        row0[s] = r0;
        total += int64_t(m - r0) * dp;
        maxRows = std::max(maxRows, m - r0);
        maxDp = std::max(maxDp, dp);
      }
      P.splitEntries = std::max(P.splitEntries, total);
      P.srefs.resize(P.gemm.size(), SplitRef{0, 0, 0, 0});
      for (int64_t i = bulk.first; i < bulk.first + bulk.count; i++) {
        const GemmTask& g = P.gemm[size_t(i)];
        const int32_t s = int32_t(std::upper_bound(S.Lxp.begin(), S.Lxp.end(), g.a_off) - S.Lxp.begin()) - 1;
        const int32_t m = frontRows(S, s);
        const int64_t ra = (g.a_off - S.Lxp[s]) % m, rb = (g.b_off - S.Lxp[s]) % m;
        if (dst[s] < 0 || ra < row0[s] || rb < row0[s]) throw std::logic_error("bulk task outside its split panel");
        const int32_t cp = (m - row0[s]) * kSplitDepth;  // chunk pitch of the front's split panel
        P.srefs[size_t(i)] = SplitRef{dst[s] + (ra - row0[s]) * kSplitDepth, dst[s] + (rb - row0[s]) * kSplitDepth, cp, 0};
      }
      P.launches.push_back(Launch{kSplit, L, first, int64_t(P.splits.size()) - first, maxRows, 0, maxDp,
                                  kSyncSide | (bulk.sync & kSyncWait)});
      bulk.sync &= ~kSyncWait;
      P.launches.push_back(bulk);
    };
    for (int32_t c0 = 0; c0 < maxNc; c0 += SPW) {
      for (int32_t k0 = c0; k0 < std::min(maxNc, c0 + SPW); k0 += kTileEdge) {
        if (prm.fusedPanel && SPW <= kSuperPanel) {
          // one launch: every row tile of the panel (or the diagonal block alone) is a task
          const int64_t pFirst = int64_t(P.panels.size());
          for (int32_t s : big) {
            const int32_t nc = frontCols(S, s), m = frontRows(S, s);
            if (nc <= k0) continue;
            const int32_t nb = std::min(kTileEdge, nc - k0);
            const int32_t ucols = std::max(0, std::min(nc, c0 + SPW) - (k0 + nb));
            const int32_t ntask = std::max(1, (m - k0 - nb + kTileEdge - 1) / kTileEdge);
            // a later step of the super-panel: its block was factored by the previous step's first row tile (which
            // exists: the previous step had these rows below it and this block's columns to update)
            const bool pre = k0 > c0;
            if (pre && m - k0 == nb) continue;  // factored already, and no rows below it: nothing left to do
            PanelTask t{S.Lxp[s] + k0 + int64_t(k0) * m,
                        P.Woff[s] + int64_t(k0 / kTileEdge) * kTileEdge * kTileEdge, m, nb, S.superStart[s] + k0,
                        0, 0, ucols, P.panelCounters++, ntask, pre ? 0 : 1, pre ? kPanelPre : 0};
            P.panelMaxTasks = std::max(P.panelMaxTasks, ntask);
            if (m - k0 == nb) {
              P.panels.push_back(t);
              continue;
            }
            for (int32_t ro = nb; ro < m - k0; ro += kTileEdge) {
              t.ro = ro;
              t.rows = std::min(kTileEdge, m - k0 - ro);
              t.owner = !pre && ro == nb;
              t.flags = (pre ? kPanelPre : 0) | (ro == nb && ucols > 0 ? kPanelNext : 0);
              P.panels.push_back(t);
            }
          }
          push(kPanel, L, pFirst, int64_t(P.panels.size()) - pFirst);
          continue;
        }
        const int64_t dFirst = int64_t(P.diag.size());
        for (int32_t s : big) {
          const int32_t nc = frontCols(S, s), m = frontRows(S, s);
          if (nc <= k0) continue;
          const int32_t nb = std::min(kTileEdge, nc - k0);
          P.diag.push_back(DiagTask{S.Lxp[s] + k0 + int64_t(k0) * m,
                                    P.Woff[s] + int64_t(k0 / kTileEdge) * kTileEdge * kTileEdge, m, nb,
                                    S.superStart[s] + k0, 0});
        }
        push(kDiag, L, dFirst, int64_t(P.diag.size()) - dFirst);
        // L21 rows of this panel: A21 L11^-T (Cholesky) or A21 L11^-T D^-1 (LDL^T), in place
        const int64_t tFirst = int64_t(P.gemm.size());
        for (int32_t s : big) {
          const int32_t nc = frontCols(S, s), m = frontRows(S, s), f = S.superStart[s];
          if (nc <= k0) continue;
          const int32_t nb = std::min(kTileEdge, nc - k0);
          const int64_t base = S.Lxp[s];
          const int64_t woff = P.Woff[s] + int64_t(k0 / kTileEdge) * kTileEdge * kTileEdge;
          for (int32_t r0 = k0 + nb; r0 < m; r0 += kTileEdge) {
            GemmTask t{};
            t.a_off = base + r0 + int64_t(k0) * m;
            t.b_off = woff;
            t.c_off = t.a_off;
            t.lda = m;
            t.ldb = kTileEdge;
            t.ldc = m;
            t.rows = std::min(kTileEdge, m - r0);
            t.cols = nb;
            t.k = nb;
            t.d_off = f + k0;
            t.a_buf = kBufLx;
            t.b_buf = kBufW;
            t.c_buf = kBufLx;
            t.mode = 2;
            P.gemm.push_back(t);
          }
        }
        push(kGemm, L, tFirst, int64_t(P.gemm.size()) - tFirst, 0, 0, kTileEdge);
        // trailing update of the remaining columns of this super-panel
        gemmLaunch(L, [&](int32_t edge, std::vector<GemmTask>& out) {
          for (int32_t s : big) {
            const int32_t nc = frontCols(S, s);
            if (nc <= k0) continue;
            const int32_t nb = std::min(kTileEdge, nc - k0);
            lowerUpdate(out, edge, s, k0 + nb, std::min(nc, c0 + SPW), k0, nb);
          }
        });
      }
      // signal: super-panel c0 factored (its last step is the last main-stream launch so far)
      P.launches.back().sync |= kSyncSignal;
      P.signals++;
      const int32_t c1 = c0 + SPW, c2 = c1 + SPW;
      // lookahead: the next super-panel's columns, after the previous bulk update (which also wrote them)
      gemmLaunch(L, [&](int32_t edge, std::vector<GemmTask>& out) {
        for (int32_t s : big) {
          const int32_t nc = frontCols(S, s);
          if (nc > c1) lowerUpdate(out, edge, s, c1, std::min(nc, c2), c0, SPW);
        }
      }, sideUsed ? kSyncWait : 0);
      // bulk, on the side stream once the super-panel is factored: every later column, and the super-panel's share of
      // the Schur complement U = -L21 D L21^T (so that only the last super-panel's share follows the panel chain)
      if (gemmLaunch(L, [&](int32_t edge, std::vector<GemmTask>& out) {
            for (int32_t s : big) {
              const int32_t nc = frontCols(S, s);
              if (nc > c2) lowerUpdate(out, edge, s, c2, nc, c0, SPW);
              if (nc > c0) schurUpdate(out, edge, s, c0, std::min(SPW, nc - c0));
            }
          }, kSyncSide | kSyncWait | kSyncSignal)) {
        sideUsed = true;
        if (P.launches.back().edge == kGemmEdge) addSplit(L, c0, c2, big);
      }
    }
    // the next main-stream launch (the update-matrix extend-add, or the next level) joins the side stream: also when
    // the level's last super-panel had no bulk launch (no front with a Schur share or columns left there), since an
    // earlier bulk launch may still be writing update matrices that nothing on the main stream has waited for
    if (sideUsed) joinPending = true;
    // children's contributions to the update matrix
    {
      const int64_t first = int64_t(P.ea.size());
      for (int32_t s : big) {
        if (S.childHead[s] == -1) continue;
        const int32_t nc = frontCols(S, s), m = frontRows(S, s);
        for (int32_t j = nc; j < m; j += kEaWidth) addEa(s, j, std::min(m, j + kEaWidth));
      }
      push(kExtendAdd, L, first, int64_t(P.ea.size()) - first);
    }
  }
  for (const GemmTask& t : P.gemm) {
    if (t.mode == 2) {
      P.gemmFlops += double(t.rows) * t.cols * t.k;  // triangular solve: rows x nb^2
      continue;
    }
    // symmetric outputs: a tile whose corner sits on the diagonal only needs its lower part
    const bool onDiag = (t.mode == 1 && t.a_off == t.b_off) || (t.mode == 0 && t.a_off == t.b_off);
    double entries = double(t.rows) * t.cols;
    if (onDiag) {
      entries = 0;
      for (int32_t j = 0; j < t.cols; j++) entries += std::max(0, t.rows - j);
    }
    P.gemmFlops += 2.0 * entries * t.k;
  }
  for (size_t l = 1; l < P.launches.size(); l++) {
    const Launch& L = P.launches[l];
    const Launch& Sl = P.launches[l - 1];
    if (L.kind != kGemm || Sl.kind != kSplit) continue;
    P.bulkLaunches++;
    for (int64_t i = Sl.first; i < Sl.first + Sl.count; i++)
      P.bulkBytes += 4.0 * double(P.splits[size_t(i)].rows) * P.splits[size_t(i)].w;
    for (int64_t i = L.first; i < L.first + L.count; i++) {
      const GemmTask& t = P.gemm[size_t(i)];
      double entries = double(t.rows) * t.cols;
      if (t.a_off == t.b_off) {  // a tile on the diagonal: its lower part
        entries = 0;
        for (int32_t j = 0; j < t.cols; j++) entries += std::max(0, t.rows - j);
      }
      P.bulkBytes += entries * (t.mode == 0 ? 8.0 : 4.0);
    }
  }
  P.srefs.resize(P.gemm.size(), SplitRef{0, 0, 0, 0});
  return P;
}

SolvePlan buildSolvePlan(const Symbolic& S, const FactorPlan& F, const PlanParams& prm) {
  SolvePlan P;
  const int32_t ns = S.nsuper;
  P.fvEntries = ns ? S.Sp[ns] : 0;
  if (ns == 0) return P;
  std::vector<std::vector<int32_t>> byLevel(F.nlevels);
  for (int32_t s = 0; s < ns; s++) byLevel[F.level[s]].push_back(s);
  std::vector<int32_t> flag0(ns, -1);
  for (int32_t s = 0; s < ns; s++)
    if (F.isBig[s]) {
      flag0[s] = int32_t(P.flowFlags);
      P.flowFlags += (frontCols(S, s) + kTileEdge - 1) / kTileEdge;
    }
  auto push = [&P](SolveKind kind, int32_t level, int64_t first, int64_t count, int32_t maxM = 0) {
    if (count <= 0) return;
    P.launches.push_back(SolveLaunch{kind, level, first, count, maxM});
  };
  auto nblocks = [&S](int32_t s) { return (frontCols(S, s) + kTileEdge - 1) / kTileEdge; };
  // the whole-front supernodes of a level: those with m <= kSolveWaveMax (one wave each) and the rest, one launch each
  // (packed several to a wave: up to 8 rows when the level has at least packMin of them, then up to 16 rows when it
  // has at least packMin of those left)
  auto pushSmall = [&](SolveKind kind, int32_t L) {
    int64_t n8 = 0, n16 = 0;
    for (int32_t s : byLevel[L])
      if (!F.isBig[s]) {
        n8 += frontRows(S, s) <= 8;
        n16 += frontRows(S, s) <= 16;
      }
    const bool pack8 = n8 >= prm.packMin, pack16 = n16 - (pack8 ? n8 : 0) >= prm.packMin;
    for (int cls = 0; cls < 4; cls++) {
      const int64_t first = int64_t(P.fronts.size());
      int32_t maxM = 0;
      for (int32_t s : byLevel[L]) {
        const int32_t m = frontRows(S, s);
        const int c = (m <= 8 && pack8) ? 0 : (m <= 16 && pack16) ? 1 : m <= kSolveWaveMax ? 2 : 3;
        if (F.isBig[s] || c != cls) continue;
        P.fronts.push_back(s);
        maxM = std::max(maxM, m);
      }
      push(kind, L, first, int64_t(P.fronts.size()) - first, maxM);
      if (cls < 2 && int64_t(P.fronts.size()) > first) P.launches.back().seg = cls == 0 ? 8 : 16;
    }
  };
  // a flow launch with its task lists for FP32 and FP64 factors: `emit(nq)` appends the tasks of nq blocks to P.blocks;
  // a forward launch takes wide tasks when it has at least prm.flowWideMin of them
  auto pushFlow = [&](SolveKind kind, int32_t L, auto emit) {
    auto list = [&](int32_t wide, int64_t& first, int64_t& count, int32_t& nq) {
      first = int64_t(P.blocks.size());
      nq = 1;
      if (kind == kSolveFwdFlow) {
        emit(wide);
        if (int64_t(P.blocks.size()) - first >= prm.flowWideMin) {
          nq = wide;
          count = int64_t(P.blocks.size()) - first;
          return;
        }
        P.blocks.resize(size_t(first));
      }
      emit(1);
      count = int64_t(P.blocks.size()) - first;
    };
    SolveLaunch sl{kind, L, 0, 0, 0};
    list(kFlowQuadsF32, sl.first, sl.count, sl.nq);
    list(kFlowQuadsF64, sl.first2, sl.count2, sl.nq2);
    if (sl.count > 0) P.launches.push_back(sl);
  };
  std::vector<std::vector<int32_t>> bigOf(F.nlevels);
  for (int32_t L = 0; L < F.nlevels; L++)
    for (int32_t s : byLevel[L])
      if (F.isBig[s]) bigOf[L].push_back(s);
  // fused runs: a level joins one when its flow launches would take one block per task in both precisions (fewer than
  // flowWideMin tasks of kFlowQuadsF64 blocks, hence of kFlowQuadsF32) and its fronts have at most
  // kFusedSolveMaxBlocks blocks; every level of a run after the first holds
  // tiled supernodes only, so no whole-front launch has to run between two of its levels (the first level's whole
  // fronts run before the forward launch and after the backward one). runEnd[L]: last level of the run starting at L
  std::vector<int32_t> runEnd(F.nlevels, -1), runStart(F.nlevels, -1);
  if (prm.fusedSolve & 3) {
    std::vector<int8_t> narrow(F.nlevels, 0), pure(F.nlevels, 0);
    for (int32_t L = 0; L < F.nlevels; L++) {
      int64_t wide = 0;
      int32_t maxBlocks = 0;
      for (int32_t s : bigOf[L]) {
        maxBlocks = std::max(maxBlocks, nblocks(s));
        const int32_t mu = frontRows(S, s) - frontCols(S, s);
        wide += (nblocks(s) + kFlowQuadsF64 - 1) / kFlowQuadsF64 + (mu + kFlowQuadsF64 * kTileEdge - 1) / (kFlowQuadsF64 * kTileEdge);
      }
      narrow[L] = !bigOf[L].empty() && wide < prm.flowWideMin && maxBlocks <= kFusedSolveMaxBlocks;
      pure[L] = bigOf[L].size() == byLevel[L].size();
    }
    for (int32_t L = 0; L < F.nlevels;) {
      int32_t e = L;
      if (narrow[L])
        while (e + 1 < F.nlevels && narrow[e + 1] && pure[e + 1]) e++;
      if (e > L) {
        runEnd[L] = e;
        runStart[e] = L;
      }
      L = e + 1;
    }
  }
  auto parentOf = [&S](int32_t s) { return S.superParent[s]; };
  // a fused launch over `tasks` (the same list serves both precisions: one block per task)
  auto pushFused = [&](SolveKind kind, int32_t level, int64_t first) {
    SolveLaunch sl{kind, level, first, int64_t(P.blocks.size()) - first, 0};
    sl.first2 = sl.first;
    sl.count2 = sl.count;
    if (sl.count > 0) P.launches.push_back(sl);
  };
  // forward, leaves first
  for (int32_t L = 0; L < F.nlevels; L++) {
    const std::vector<int32_t>& big = bigOf[L];
    pushSmall(kSolveSmallFwd, L);
    if (big.empty()) continue;
    if (runEnd[L] >= 0 && (prm.fusedSolve & 1)) {
      // levels L .. runEnd[L] in one launch; per supernode two counters: below tasks of its children done (slot) and
      // its gather tasks done (slot + 1)
      const int32_t Le = runEnd[L];
      std::vector<int32_t> slot(ns, -1);
      for (int32_t l = L; l <= Le; l++)
        for (int32_t s : bigOf[l]) {
          slot[s] = int32_t(P.doneSlots);
          P.doneSlots += 2;
        }
      auto belowTasks = [&](int32_t s) { return (frontRows(S, s) - frontCols(S, s) + kTileEdge - 1) / kTileEdge; };
      const int64_t first = int64_t(P.blocks.size());
      for (int32_t l = L; l <= Le; l++) {
        for (int32_t s : bigOf[l]) {
          int32_t fromChildren = 0;
          for (int32_t c = S.childHead[s]; c != -1; c = S.childNext[c])
            if (slot[c] >= 0) fromChildren += belowTasks(c);
          const int32_t m = frontRows(S, s);
          for (int32_t r = 0; r < m; r += kGatherRows) {
            BlockTask t{s, r / kGatherRows, r, std::min(m, r + kGatherRows), -1, flag0[s], 0, 0};
            t.wslot = fromChildren > 0 ? slot[s] : -1;
            t.wcount = fromChildren;
            t.dslot = slot[s] + 1;
            P.blocks.push_back(t);
          }
        }
        int32_t maxBlk = 0;
        for (int32_t s : bigOf[l]) maxBlk = std::max(maxBlk, nblocks(s));
        auto flowTask = [&](BlockTask t) {
          t.wslot = slot[t.s] + 1;
          t.wcount = (frontRows(S, t.s) + kGatherRows - 1) / kGatherRows;
          P.blocks.push_back(t);
          return &P.blocks.back();
        };
        for (int32_t t = 0; t < maxBlk; t++)
          for (int32_t s : bigOf[l])
            if (t < nblocks(s))
              flowTask(BlockTask{s, t, t * kTileEdge, std::min(frontCols(S, s), (t + 1) * kTileEdge), 1, flag0[s], 0, 0});
        for (int32_t s : bigOf[l]) {
          const int32_t nc = frontCols(S, s), m = frontRows(S, s), par = parentOf(s);
          for (int32_t r = nc; r < m; r += kTileEdge)
            flowTask(BlockTask{s, nblocks(s), r, std::min(m, r + kTileEdge), 0, flag0[s], 0, 0})->dslot =
                par >= 0 && slot[par] >= 0 ? slot[par] : -1;
        }
      }
      pushFused(kSolveFwdFused, L, first);
      L = Le;
      continue;
    }
    int64_t first = int64_t(P.fronts.size());
    P.fronts.insert(P.fronts.end(), big.begin(), big.end());
    int32_t maxM = 0;
    for (int32_t s : big) maxM = std::max(maxM, frontRows(S, s));
    push(kSolveGather, L, first, int64_t(big.size()), maxM);
    // every task depends only on tasks listed before it: the chain tasks by block index across the level's
    // supernodes, then the below tasks (which consume every block of their supernode)
    int32_t maxBlk = 0;
    for (int32_t s : big) maxBlk = std::max(maxBlk, nblocks(s));
    pushFlow(kSolveFwdFlow, L, [&](int32_t nq) {
      for (int32_t t = 0; t < maxBlk; t += nq)
        for (int32_t s : big) {
          const int32_t nc = frontCols(S, s), nb = nblocks(s);
          if (t < nb)
            P.blocks.push_back(BlockTask{s, t, t * kTileEdge, std::min(nc, (t + nq) * kTileEdge),
                                         std::min(nq, nb - t), flag0[s], 0, 0});
        }
      for (int32_t s : big) {
        const int32_t nc = frontCols(S, s), m = frontRows(S, s);
        for (int32_t r = nc; r < m; r += nq * kTileEdge)
          P.blocks.push_back(BlockTask{s, nblocks(s), r, std::min(m, r + nq * kTileEdge), 0, flag0[s], 0, 0});
      }
    });
  }
  // backward, root first; blocks of each supernode from the last one down
  // partial products -L21[chunk, block]^T y[below chunk], one slot per (block, row chunk); y[below] is read through
  // the row structure, so no gather precedes them
  std::vector<int32_t> p0(ns, 0), np(ns, 0);
  auto initTasks = [&](int32_t s, int32_t wslot, int32_t wcount, int32_t dslot) {
    const int32_t nc = frontCols(S, s), mu = frontRows(S, s) - nc;
    np[s] = (mu + kInitRows - 1) / kInitRows;
    p0[s] = int32_t(P.initSlots);
    for (int32_t b = 0; b < nblocks(s); b++)
      for (int32_t c = 0; c < np[s]; c++) {
        BlockTask t{s, b, c * kInitRows, std::min(mu, (c + 1) * kInitRows), 0, flag0[s], int32_t(P.initSlots++), 1};
        t.wslot = wslot;
        t.wcount = wcount;
        t.dslot = dslot;
        P.blocks.push_back(t);
      }
  };
  for (int32_t L = F.nlevels - 1; L >= 0; L--) {
    const std::vector<int32_t>& big = bigOf[L];
    if (runStart[L] >= 0 && (prm.fusedSolve & 2)) {
      // levels L down to runStart[L] in one launch; per supernode two counters: its initialisation tasks done (slot)
      // and its chain tasks done (slot + 1)
      const int32_t Ls = runStart[L];
      std::vector<int32_t> slot(ns, -1);
      for (int32_t l = L; l >= Ls; l--)
        for (int32_t s : bigOf[l]) {
          slot[s] = int32_t(P.doneSlots);
          P.doneSlots += 2;
        }
      const int64_t first = int64_t(P.blocks.size());
      for (int32_t l = L; l >= Ls; l--) {
        for (int32_t s : bigOf[l]) {
          const int32_t par = parentOf(s);
          const bool waits = par >= 0 && slot[par] >= 0;
          initTasks(s, waits ? slot[par] + 1 : -1, waits ? nblocks(par) : 0, slot[s]);
        }
        int32_t maxBlk = 0;
        for (int32_t s : bigOf[l]) maxBlk = std::max(maxBlk, nblocks(s));
        for (int32_t i = 0; i < maxBlk; i++)
          for (int32_t s : bigOf[l]) {
            const int32_t nc = frontCols(S, s), nb = nblocks(s);
            if (i >= nb) continue;
            const int32_t lo = nb - 1 - i;
            BlockTask t{s, lo, lo * kTileEdge, std::min(nc, (lo + 1) * kTileEdge), 1, flag0[s], p0[s] + lo * np[s],
                        np[s]};
            t.wslot = np[s] > 0 ? slot[s] : -1;
            t.wcount = nb * np[s];
            t.dslot = slot[s] + 1;
            P.blocks.push_back(t);
          }
      }
      pushFused(kSolveBwdFused, L, first);
      pushSmall(kSolveSmallBwd, Ls);  // they read the solutions of the run's supernodes
      L = Ls;
      continue;
    }
    pushSmall(kSolveSmallBwd, L);
    if (big.empty()) continue;
    const int64_t bf = int64_t(P.blocks.size());
    for (int32_t s : big) initTasks(s, -1, 0, -1);
    push(kSolveBwdInit, L, bf, int64_t(P.blocks.size()) - bf);
    int32_t maxBlk = 0;
    for (int32_t s : big) maxBlk = std::max(maxBlk, nblocks(s));
    pushFlow(kSolveBwdFlow, L, [&](int32_t nq) {
      for (int32_t i = 0; i < maxBlk; i += nq)
        for (int32_t s : big) {
          const int32_t nc = frontCols(S, s), nb = nblocks(s);
          if (i >= nb) continue;
          const int32_t hi = nb - i, lo = std::max(0, hi - nq);  // blocks [lo, hi)
          P.blocks.push_back(BlockTask{s, lo, lo * kTileEdge, std::min(nc, hi * kTileEdge), hi - lo, flag0[s],
                                       p0[s] + lo * np[s], np[s]});
        }
    });
  }
  return P;
}

}  // namespace amp
