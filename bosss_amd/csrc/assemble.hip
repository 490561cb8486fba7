// Operator-matrix assembly and block-sparse apply.
//
// Assembly (EvaluatorLinear.ComputeMatrix, DifferentialOperator.cs:1721-1838, and the
// coloured FDJacobianBuilder, DifferentialOperator.cs:1905-2880): the cells are
// distance-2 coloured on the host, so that every cell's stencil (itself and its
// face neighbours) holds at most one cell of each colour.  For colour k and local
// coefficient q the probe X = sum_{cells c of colour k} e_{c,q} goes through the
// exact JVP kernel; in the result, each cell's block row picks up exactly one
// column q of one of its blocks, which k_extract writes into place.  For the linear
// forms (SIP, upwind) the Jacobian is the operator matrix.
//
// Storage ("block-ELL", cell-centric): blocks[j][s][r][c], row-major Nt x Nt blocks,
// s = 0 the diagonal block, s = 1 + f the block coupling cell j to its neighbour
// across local face f (cols[j][s] = column cell, -1 for a boundary face: zero block).
// This is the BlockMsrMatrix block pattern of a DG operator (BlockMsrMatrix.cs:186,
// AccBlock :1388) laid out for one-pass streaming.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace bosss {

// X[c][q] = 1 for the cells of colour k, 0 elsewhere (X holds J*Nt entries)
__global__ void k_probe(int J, int Nt, const int32_t* __restrict__ color, int k, int q, double* __restrict__ X) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)J * Nt) return;
  int c = (int)(i / Nt), e = (int)(i - (size_t)c * Nt);
  X[i] = (e == q && color[c] == k) ? 1.0 : 0.0;
}

// blocks[j][s][:, q] <- alpha * Y[j][:] + beta * old, for the slot s whose column cell has colour k
__global__ void k_extract(int J, int Nt, int S, const int32_t* __restrict__ cols, const int32_t* __restrict__ color, int k, int q,
                          const double* __restrict__ Y, double alpha, double beta, double* __restrict__ blocks) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)J * Nt) return;
  int j = (int)(i / Nt), r = (int)(i - (size_t)j * Nt);
  for (int s = 0; s < S; s++) {
    int c = cols[(size_t)j * S + s];
    if (c >= 0 && color[c] == k) {
      double* b = blocks + (((size_t)j * S + s) * Nt + r) * Nt + q;
      *b = alpha * Y[i] + (beta != 0.0 ? beta * *b : 0.0);
      return;   // at most one slot per colour (distance-2 colouring)
    }
  }
}

// y_j <- alpha * sum_s B[j][s] x_{cols[j][s]} + beta * y_j.  One wavefront per cell, lane per row.
__global__ void __launch_bounds__(64) k_bell_spmv(int J, int Nt, int S, const int32_t* __restrict__ cols,
                                                  const double* __restrict__ blocks, double alpha, const double* __restrict__ x,
                                                  double beta, double* __restrict__ y) {
  const int j = blockIdx.x;
  if (j >= J) return;
  for (int r = threadIdx.x; r < Nt; r += 64) {
    double acc = 0.0;
    for (int s = 0; s < S; s++) {
      int c = cols[(size_t)j * S + s];
      if (c < 0) continue;
      const double* b = blocks + (((size_t)j * S + s) * Nt + r) * Nt;
      const double* xc = x + (size_t)c * Nt;
      for (int m = 0; m < Nt; m++) acc += b[m] * xc[m];
    }
    double* yj = y + (size_t)j * Nt + r;
    *yj = alpha * acc + (beta != 0.0 ? beta * *yj : 0.0);
  }
}

void launch_probe(int J, int Nt, const int32_t* color, int k, int q, double* X, hipStream_t s) {
  size_t n = (size_t)J * Nt;
  hipLaunchKernelGGL(k_probe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, J, Nt, color, k, q, X);
}
void launch_extract(int J, int Nt, int S, const int32_t* cols, const int32_t* color, int k, int q, const double* Y, double alpha,
                    double beta, double* blocks, hipStream_t s) {
  size_t n = (size_t)J * Nt;
  hipLaunchKernelGGL(k_extract, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, J, Nt, S, cols, color, k, q, Y, alpha, beta, blocks);
}
void launch_bell_spmv(int J, int Nt, int S, const int32_t* cols, const double* blocks, double alpha, const double* x, double beta,
                      double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_bell_spmv, dim3(J), dim3(64), 0, s, J, Nt, S, cols, blocks, alpha, x, beta, y);
}

}  // namespace bosss
