// Euler (Rusanov and HLLC) 3D p = 3 kernels other than the fused residual + VJP: residual, JVP
// and VJP in FP64 and FP32, and the FP32 fused residual + VJP, compiled in their own translation
// unit (own scheduling flags; built in parallel with the others).
#include <vector>

#include "registry.hpp"

namespace bosss {
#define BOSSS_EULER_P3(H)                                                                                   \
  template __global__ void kt_forward<FormEuler<3, H>, 3, 3, 5, double, MODE_RES>(TArgs<double>);          \
  template __global__ void kt_forward<FormEuler<3, H>, 3, 3, 5, float, MODE_RES>(TArgs<float>);            \
  template __global__ void kt_jvp_euler<FormEuler<3, H>, 3, 3, 5, double>(TArgs<double>);                  \
  template __global__ void kt_jvp_euler<FormEuler<3, H>, 3, 3, 5, float>(TArgs<float>);                    \
  template __global__ void kt_vjp_euler<FormEuler<3, H>, 3, 3, 5, double, false>(TArgs<double>);           \
  template __global__ void kt_vjp_euler<FormEuler<3, H>, 3, 3, 5, float, false>(TArgs<float>);             \
  template __global__ void kt_vjp_euler<FormEuler<3, H>, 3, 3, 5, float, true>(TArgs<float>);
BOSSS_EULER_P3(false)
BOSSS_EULER_P3(true)
#undef BOSSS_EULER_P3
}  // namespace bosss
