// Fused residual + VJP kernels of the 3D p = 3 Euler forms (the 64^3 headline workload),
// compiled on their own so the Makefile can give them the AMDGPU "max-memory-clause"
// scheduling strategy: it issues the global coefficient loads of a contraction as one
// clause instead of one load per MFMA (measured 5.39 -> 5.12 ms); the other kernels keep
// the default strategy, under which they are faster.
#include <vector>

#include "registry.hpp"

namespace bosss {
template __global__ void kt_vjp_euler<FormEuler<3, false>, 3, 3, 5, double, true>(TArgs<double>);
template __global__ void kt_vjp_euler<FormEuler<3, true>, 3, 3, 5, double, true>(TArgs<double>);
}  // namespace bosss
