// Host-side symbolic structure of a supernodal factor (the SymbolicAnalysis of packages/core/src/la/types.d.mts:13-20).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace amp {

struct Symbolic {
  int32_t n = 0, nsuper = 0, maxFront = 0;
  int64_t nnzL = 0, nnzA = 0, nnzLStored = 0, nnzPattern = 0;
  double flops = 0, fill = 1;
  std::vector<int32_t> perm, pinv, parent, colCount;
  std::vector<int32_t> lowerP, lowerI, lowerSrc;
  std::vector<int32_t> superStart, colSuper, superParent, childHead, childNext;
  std::vector<int32_t> Sp, Si;
  std::vector<int64_t> Lxp;
  std::string patternKey;
};

std::string patternFingerprint(int32_t n, const int32_t* rowptr, const int32_t* colidx);
bool amdOrder(int32_t n, const int32_t* rowptr, const int32_t* colidx, double denseFactor, int32_t* perm);
// Nested dissection by recursive coordinate bisection (ndOrder, sparse.mjs:230-283): perm new -> old. leaf >= 1.
void ndOrder(int32_t n, const int32_t* rowptr, const int32_t* colidx, const double* x, const double* y, int32_t leaf,
             int32_t* perm);
bool analyzePattern(int32_t n, const int32_t* rowptr, const int32_t* colidx, const int32_t* perm0, bool relax,
                    Symbolic& out);

}  // namespace amp
