#include <vector>

#include "registry.hpp"

namespace bosss {

static const std::vector<KernelSet>& all_sets() {
  static std::vector<KernelSet> sets = [] {
    std::vector<KernelSet> v;
    register_sip(v);
    register_advdiff(v);
    register_euler(v);
    return v;
  }();
  return sets;
}

const KernelSet* find_kernels(int kind, int D, int N, int K1, int precision) {
  for (const KernelSet& k : all_sets())
    if (k.kind == kind && k.D == D && k.N == N && k.K1 == K1) {
      if (precision == 1 && !k.k32[KernelSet::RES]) return nullptr;
      return &k;
    }
  return nullptr;
}

int KernelSet::prepare() const {
  for (int i = 0; i < NVARIANT; i++) {
    if (k64[i] && hipFuncSetAttribute((const void*)k64[i], hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds64) != hipSuccess) return 1;
    if (k32[i] && hipFuncSetAttribute((const void*)k32[i], hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds32) != hipSuccess) return 2;
    if (kt64[i] && hipFuncSetAttribute((const void*)kt64[i], hipFuncAttributeMaxDynamicSharedMemorySize, (int)ktlds64[i]) != hipSuccess) return 3;
    if (kt32[i] && hipFuncSetAttribute((const void*)kt32[i], hipFuncAttributeMaxDynamicSharedMemorySize, (int)ktlds32[i]) != hipSuccess) return 4;
  }
  return 0;
}

}  // namespace bosss
