// Euler (Rusanov and HLLC) instantiations: 3D p=2 (4 pts), p=3 (5 pts); 2D p=2 (4 pts).
#include <vector>

#include "registry.hpp"

namespace bosss {
// The fused residual + VJP kernels of 3D p = 3 are compiled in their own translation unit,
// kernels_euler_fused.hip (built in parallel with this one).
extern template __global__ void kt_vjp_euler<FormEuler<3, false>, 3, 3, 5, double, true>(TArgs<double>);
extern template __global__ void kt_vjp_euler<FormEuler<3, true>, 3, 3, 5, double, true>(TArgs<double>);

void register_euler(std::vector<KernelSet>& v) {
  v.push_back(make_kernel_set<FormEuler<2, false>, 2, 6, 4, false>(2));
  v.push_back(make_kernel_set<FormEuler<2, true>, 2, 6, 4, false>(3));
  v.push_back(make_kernel_set<FormEuler<3, false>, 3, 10, 4, true>(2));
  v.push_back(make_kernel_set<FormEuler<3, true>, 3, 10, 4, true>(3));
  v.push_back(make_kernel_set<FormEuler<3, false>, 3, 20, 5, true>(2));
  v.push_back(make_kernel_set<FormEuler<3, true>, 3, 20, 5, true>(3));
}
}  // namespace bosss
