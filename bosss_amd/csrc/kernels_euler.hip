// Euler (Rusanov and HLLC) instantiations: 3D p=2 (4 pts), p=3 (5 pts); 2D p=2 (4 pts).
#include <vector>

#include "registry.hpp"

namespace bosss {
// The 3D p = 3 tensor kernels are compiled in their own translation units: the FP64 fused
// residual + VJP in kernels_euler_fused.hip, the others in kernels_euler_p3.hip (built in
// parallel with this one, each with its own scheduling flags).
extern template __global__ void kt_vjp_euler<FormEuler<3, false>, 3, 3, 5, double, true>(TArgs<double>);
extern template __global__ void kt_vjp_euler<FormEuler<3, true>, 3, 3, 5, double, true>(TArgs<double>);
#define BOSSS_EULER_P3_EXTERN(H)                                                                           \
  extern template __global__ void kt_forward<FormEuler<3, H>, 3, 3, 5, double, MODE_RES>(TArgs<double>);  \
  extern template __global__ void kt_forward<FormEuler<3, H>, 3, 3, 5, float, MODE_RES>(TArgs<float>);    \
  extern template __global__ void kt_jvp_euler<FormEuler<3, H>, 3, 3, 5, double>(TArgs<double>);          \
  extern template __global__ void kt_jvp_euler<FormEuler<3, H>, 3, 3, 5, float>(TArgs<float>);            \
  extern template __global__ void kt_vjp_euler<FormEuler<3, H>, 3, 3, 5, double, false>(TArgs<double>);   \
  extern template __global__ void kt_vjp_euler<FormEuler<3, H>, 3, 3, 5, float, false>(TArgs<float>);     \
  extern template __global__ void kt_vjp_euler<FormEuler<3, H>, 3, 3, 5, float, true>(TArgs<float>);
BOSSS_EULER_P3_EXTERN(false)
BOSSS_EULER_P3_EXTERN(true)
#undef BOSSS_EULER_P3_EXTERN

void register_euler(std::vector<KernelSet>& v) {
  v.push_back(make_kernel_set<FormEuler<2, false>, 2, 6, 4, false>(2));
  v.push_back(make_kernel_set<FormEuler<2, true>, 2, 6, 4, false>(3));
  v.push_back(make_kernel_set<FormEuler<3, false>, 3, 10, 4, true>(2));
  v.push_back(make_kernel_set<FormEuler<3, true>, 3, 10, 4, true>(3));
  v.push_back(make_kernel_set<FormEuler<3, false>, 3, 20, 5, true>(2));
  v.push_back(make_kernel_set<FormEuler<3, true>, 3, 20, 5, true>(3));
}
}  // namespace bosss
