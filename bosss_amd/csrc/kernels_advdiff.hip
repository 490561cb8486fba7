// Upwind advection + SIP viscosity instantiations: 3D p=2 (3 pts), p=3 (4 pts); 2D p=2.
#include <vector>

#include "registry.hpp"

namespace bosss {
void register_advdiff(std::vector<KernelSet>& v) {
  v.push_back(make_kernel_set<FormAdvDiff, 2, 6, 3, false>(1));
  v.push_back(make_kernel_set<FormAdvDiff, 3, 10, 3, false>(1));
  v.push_back(make_kernel_set<FormAdvDiff, 3, 20, 4, true>(1));
}
}  // namespace bosss
