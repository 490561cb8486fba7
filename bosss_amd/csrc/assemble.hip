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

#include <algorithm>
#include <cstdint>
#include <vector>

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
    double* b = blocks + (((size_t)j * S + s) * Nt + r) * Nt + q;
    if (c < 0) {
      // no block here (a boundary face, or a neighbour an earlier slot already holds): the
      // slot becomes beta * old, written once, at the first colour
      if (k == 0) *b = beta != 0.0 ? beta * *b : 0.0;
    } else if (color[c] == k) {
      *b = alpha * Y[i] + (beta != 0.0 ? beta * *b : 0.0);   // at most one slot per colour
    }
  }
}

// y_j <- alpha * sum_s B[j][s] x_{cols[j][s]} + beta * y_j.  One wavefront per cell, lane per row.
// General fallback (any S, blocks wider than 128); launch_bell_spmv prefers the streaming
// kernel below, which reads the matrix about four times faster.
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
// Streaming block-ELL apply, one wavefront per cell row, four per workgroup.  The S blocks of a
// cell row are read as flat row-major arrays: element e = 64 i + lane of a panel of R whole rows
// (R * Nt <= 64 NL), so every load instruction reads 512 consecutive bytes.  The x entries of the
// S column cells are staged in LDS first; a lane accumulates sum_s B_s[e] x_s[e % Nt] for each of
// its NL elements, and the R row sums of the panel are reduced through LDS.  Loads past the panel
// are clamped to its last element and their products are never read.
template <int S, int NL, int XMAX>
__global__ void __launch_bounds__(256) k_bell_spmv_flat(int J, int Nt, const int32_t* __restrict__ cols,
                                                        const double* __restrict__ blocks, double alpha,
                                                        const double* __restrict__ x, double beta, double* __restrict__ y) {
  __shared__ double xs[4][S][XMAX];
  __shared__ double pr[4][NL * 64];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;   // wave-uniform: scalar index loads
  const int j0 = blockIdx.x * 4 + w;
  const bool valid = j0 < J;
  const int j = valid ? j0 : J - 1;
#pragma unroll
  for (int s = 0; s < S; s++) {
    const int k = cols[(size_t)j * S + s];
    const double* xk = x + (size_t)(k < 0 ? j : k) * Nt;
#pragma unroll
    for (int q0 = 0; q0 < XMAX; q0 += 64) {
      const int q = q0 + lane;
      if (q < XMAX) {
        const double t = xk[q < Nt ? q : Nt - 1];
        xs[w][s][q] = k >= 0 ? t : 0.0;
      }
    }
  }
  __syncthreads();
  const int NN = Nt * Nt;
  const int R = min(Nt, (64 * NL) / Nt);
  const double* bj = blocks + (size_t)j * S * NN;
  for (int r0 = 0; r0 < Nt; r0 += R) {
    const int rows = min(R, Nt - r0), pe = rows * Nt;
    const double* bp = bj + (size_t)r0 * Nt;
    double acc[NL];
    int col[NL];
#pragma unroll
    for (int i = 0; i < NL; i++) {
      const int e = 64 * i + lane;
      acc[i] = 0.0;
      col[i] = (e < pe ? e : pe - 1) % Nt;
    }
#pragma unroll
    for (int s = 0; s < S; s++) {
#pragma unroll
      for (int i = 0; i < NL; i++) {
        const int e = 64 * i + lane;
        acc[i] += bp[(size_t)s * NN + (e < pe ? e : pe - 1)] * xs[w][s][col[i]];
      }
    }
#pragma unroll
    for (int i = 0; i < NL; i++) pr[w][64 * i + lane] = acc[i];
    __syncthreads();
    if (valid && lane < rows) {
      double a = 0.0;
      for (int cc = 0; cc < Nt; cc++) a += pr[w][lane * Nt + cc];
      double* yj = y + (size_t)j * Nt + r0 + lane;
      *yj = alpha * a + (beta != 0.0 ? beta * *yj : 0.0);
    }
    __syncthreads();
  }
}

// Panel sizes (NL: 64-element loads per lane and panel) compiled for the apply.  Every choice
// computes each row sum in the same order (columns in order, slots in order within a column), so
// the result does not depend on it; only the number of panels per row and the registers per
// lane do.  Nt <= 32 stages x in 32-entry rows, Nt <= 128 in 128-entry rows.
static const int kBellNL32[] = {1, 2, 4, 7, 8, 12, 16};
static const int kBellNL128[] = {4, 8, 16};

// The fixed rule: the smallest compiled NL whose panel holds a whole block (Nt <= 32), else 16.
int bell_default_nl(int Nt) {
  if (Nt > 128) return 0;
  if (Nt > 32) return 16;
  const int nl = (Nt * Nt + 63) / 64;
  for (int c : kBellNL32)
    if (c >= nl) return c;
  return 16;
}

// Candidates an autotuner may time: the compiled sizes up to the fixed rule's (larger ones only
// add padding lanes).
void bell_candidates(int Nt, std::vector<int>& out) {
  out.clear();
  const int def = bell_default_nl(Nt);
  if (def == 0) return;
  if (Nt <= 32) {
    for (int c : kBellNL32)
      if (c <= def) out.push_back(c);
  } else {
    for (int c : kBellNL128) out.push_back(c);
  }
}

template <int S>
static bool launch_bell_flat(int J, int Nt, int nl, const int32_t* cols, const double* blocks, double alpha, const double* x,
                             double beta, double* y, hipStream_t st) {
  const dim3 g((unsigned)((J + 3) / 4)), b(256);
#define BELL_FLAT(NL, XMAX) hipLaunchKernelGGL((k_bell_spmv_flat<S, NL, XMAX>), g, b, 0, st, J, Nt, cols, blocks, alpha, x, beta, y)
  if (Nt <= 32) {
    switch (nl) {
      case 1: BELL_FLAT(1, 32); break;
      case 2: BELL_FLAT(2, 32); break;
      case 4: BELL_FLAT(4, 32); break;
      case 7: BELL_FLAT(7, 32); break;
      case 8: BELL_FLAT(8, 32); break;
      case 12: BELL_FLAT(12, 32); break;
      case 16: BELL_FLAT(16, 32); break;
      default: return false;
    }
  } else if (Nt <= 128) {
    switch (nl) {
      case 4: BELL_FLAT(4, 128); break;
      case 8: BELL_FLAT(8, 128); break;
      case 16: BELL_FLAT(16, 128); break;
      default: return false;
    }
  } else {
    return false;
  }
#undef BELL_FLAT
  return true;
}

// y = alpha A x + beta y; nl = 0 takes the fixed rule, another value must be one of
// bell_candidates(Nt) (anything else falls back to the fixed rule).
void launch_bell_spmv(int J, int Nt, int S, const int32_t* cols, const double* blocks, double alpha, const double* x, double beta,
                      double* y, hipStream_t s, int nl) {
  std::vector<int> cand;
  bell_candidates(Nt, cand);
  if (std::find(cand.begin(), cand.end(), nl) == cand.end()) nl = bell_default_nl(Nt);
  if (S == 7 && launch_bell_flat<7>(J, Nt, nl, cols, blocks, alpha, x, beta, y, s)) return;
  if (S == 5 && launch_bell_flat<5>(J, Nt, nl, cols, blocks, alpha, x, beta, y, s)) return;
  hipLaunchKernelGGL(k_bell_spmv, dim3(J), dim3(64), 0, s, J, Nt, S, cols, blocks, alpha, x, beta, y);
}

}  // namespace bosss
