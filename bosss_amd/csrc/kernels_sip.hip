// SIPLaplace instantiations: 2D p=1..3, 3D p=1..3 (Gauss points p+1, order 2p).
#include <vector>

#include "registry.hpp"

namespace bosss {
void register_sip(std::vector<KernelSet>& v) {
  v.push_back(make_kernel_set<FormSIP, 2, 3, 2, false>(0));
  v.push_back(make_kernel_set<FormSIP, 2, 6, 3, false>(0));
  v.push_back(make_kernel_set<FormSIP, 2, 10, 4, false>(0));
  v.push_back(make_kernel_set<FormSIP, 3, 4, 2, false>(0));
  v.push_back(make_kernel_set<FormSIP, 3, 10, 3, false>(0));
  v.push_back(make_kernel_set<FormSIP, 3, 20, 4, false>(0));
}
}  // namespace bosss
