// Fused residual + VJP kernels of the 3D p = 3 Euler forms (the 64^3 workload), compiled in
// their own translation unit so they build in parallel with the other Euler kernels.
#include <vector>

#include "registry.hpp"

namespace bosss {
template __global__ void kt_vjp_euler<FormEuler<3, false>, 3, 3, 5, double, true>(TArgs<double>);
template __global__ void kt_vjp_euler<FormEuler<3, true>, 3, 3, 5, double, true>(TArgs<double>);
}  // namespace bosss
