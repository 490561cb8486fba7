// Two-species XDG Poisson operator (agglomeration threshold 0) on precomputed cut-cell
// quadrature items: matrix assembly, uneven block-sparse apply, and the block-Jacobi
// preconditioned CG solver (FP64, and FP32 with FP64 iterative refinement), captured in
// a HIP graph.
//
// Forms (restated, see the oracle for the line-by-line reading):
//   XLaplace_Bulk   XLaplace.cs:50-200  (SIPLaplace with Nu = species mu, ipLaplace.cs:183-249)
//   XLaplace_Interface (mode SIP)  XLaplace.cs:205-373
//   RHSSource       XdgPoisson3Main.cs:328-358
// The FP32 solve runs CG on the diagonal-block-equilibrated matrix Linv M Linv^T
// (D_j = L_j L_j^T), which keeps the FP32 operator positive definite where the
// unagglomerated cut cells make M itself too ill-conditioned for FP32.
// Layout: cell j holds nspec[j] species blocks of N coefficients (species spec0[j] first,
// XDGBasis.cs:92-113); the matrix is block-ELL with uneven blocks: row j, slot s (0 diagonal,
// 1+f neighbour across local face f) is the row-major n_j x n_col block at boff[j*S+s].
//
// One wavefront per quadrature item.  The basis (orthonormal Legendre products, the
// reference's mode order given by exponents[N][D]) and its physical gradients are staged
// for 64 nodes at a time in LDS; lanes own (test, trial) mode pairs.  Volume and interface
// items own their sub-blocks, and so does a face item for its two off-diagonal sub-blocks.  The
// diagonal sub-blocks (and boundary right-hand sides) that all faces of a cell touch are staged
// per face item and added by one workgroup per cell in face-item order, so the assembled
// matrix is the same bits on every run (no atomics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bosss_hip.h"
#include "devbuf.hpp"

namespace bosss {
namespace {

constexpr int XN = 20;          // max modes (p <= 3 in 3D)
constexpr int XD = 3;
constexpr int XCHUNK = 32;      // nodes staged per pass
constexpr int XPAIRS = (XN * XN + 63) / 64;

thread_local std::string g_xdg_err;   // last error of this host thread
int xfail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_xdg_err = buf;
  return code;
}
#define XHIP(x)                                                                     \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) return xfail(BOSSS_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

struct XArgs {
  int D, N, P, J, S;
  const int32_t* nspec;
  const int32_t* spec0;
  const int32_t* expo;      // [N][D]
  const double* invJ;       // [J][D][D]
  const double* scaling;    // [J]
  const double* lscale;     // [J][2]
  const int32_t* cols;      // [J][S]
  const int64_t* boff;      // [J][S]
  const int64_t* dof;       // [J+1]
  double penalty_base, pdeg, mu[2], src[2];
};

// normalised Legendre values/derivatives L_k(x) = sqrt((2k+1)/2) P_k(x), k <= P
__device__ inline void legendre(double x, double (&L)[4], double (&dL)[4]) {
  // all degrees up to 3 (the kernels' maximum) so every index below is a compile-time constant
  double p0 = 1.0, p1 = x, d0 = 0.0, d1 = 1.0;
  L[0] = p0; dL[0] = d0;
  L[1] = p1; dL[1] = d1;
#pragma unroll
  for (int k = 1; k < 3; k++) {
    double p2 = ((2 * k + 1) * x * p1 - k * p0) / (k + 1);
    double d2 = d0 + (2 * k + 1) * p1;
    p0 = p1; p1 = p2; d0 = d1; d1 = d2;
    L[k + 1] = p2; dL[k + 1] = d2;
  }
#pragma unroll
  for (int k = 0; k < 4; k++) { double s = sqrt((2 * k + 1) * 0.5); L[k] *= s; dL[k] *= s; }
}

// table[e] for e in 0..3 by selects (a run-time index into a private array would put the
// array in scratch memory)
__device__ inline double pick4(const double (&t)[4], int e) { return e == 0 ? t[0] : (e == 1 ? t[1] : (e == 2 ? t[2] : t[3])); }

// Per-dimension factors of mode n at the staged point: Lv[d] = L_{e_d}(xi_d), dLv[d] its
// derivative, with Lv = 1 for the dimensions beyond D.
__device__ inline void mode_factors(const XArgs& a, int n, const double (&L)[XD][4], const double (&dL)[XD][4], double (&Lv)[XD],
                                    double (&dLv)[XD]) {
  const int32_t* e = a.expo + n * a.D;
#pragma unroll
  for (int d = 0; d < XD; d++) {
    const int ed = d < a.D ? e[d] : 0;
    Lv[d] = d < a.D ? pick4(L[d], ed) : 1.0;
    dLv[d] = pick4(dL[d], ed);
  }
}

// physical basis values v[n] and gradients g[n][d] of cell j at reference point xi
__device__ inline void phys_basis(const XArgs& a, int j, const double* xi, double* v, double (*g)[XD]) {
  double L[XD][4], dL[XD][4];
#pragma unroll
  for (int d = 0; d < XD; d++) legendre(d < a.D ? xi[d] : 0.0, L[d], dL[d]);
  const double sc = a.scaling[j];
  const double* iJ = a.invJ + (size_t)j * a.D * a.D;
  for (int n = 0; n < a.N; n++) {
    double Lv[XD], dLv[XD];
    mode_factors(a, n, L, dL, Lv, dLv);
    double val = 1.0, gr[XD];
#pragma unroll
    for (int d = 0; d < XD; d++) val *= Lv[d];
#pragma unroll
    for (int k = 0; k < XD; k++) {
      double t = dLv[k];
#pragma unroll
      for (int d = 0; d < XD; d++) if (d != k) t *= Lv[d];
      gr[k] = t;
    }
    v[n] = sc * val;
    for (int d = 0; d < a.D; d++) {
      double s = 0.0;
      for (int k = 0; k < a.D; k++) s += gr[k] * iJ[k * a.D + d];
      g[n][d] = sc * s;
    }
  }
}

// Values and normal derivatives grad(phi_n) . nrm of the physical basis at one node (what the
// face and interface forms contract; the full gradient is not needed there).
__device__ inline void phys_basis_n(const XArgs& a, int j, const double* xi, const double* nrm, double* v, double* gn) {
  double L[XD][4], dL[XD][4];
#pragma unroll
  for (int d = 0; d < XD; d++) legendre(d < a.D ? xi[d] : 0.0, L[d], dL[d]);
  const double sc = a.scaling[j];
  const double* iJ = a.invJ + (size_t)j * a.D * a.D;
  double in_[XD];   // invJ . nrm: d(phi)/dn = sum_k d(phi)/dxi_k (invJ . nrm)_k
#pragma unroll
  for (int k = 0; k < XD; k++) {
    double t = 0.0;
    if (k < a.D)
      for (int d = 0; d < a.D; d++) t += iJ[k * a.D + d] * nrm[d];
    in_[k] = t;
  }
  for (int n = 0; n < a.N; n++) {
    double Lv[XD], dLv[XD];
    mode_factors(a, n, L, dL, Lv, dLv);
    double val = 1.0, s = 0.0;
#pragma unroll
    for (int d = 0; d < XD; d++) val *= Lv[d];
#pragma unroll
    for (int k = 0; k < XD; k++) {
      double t = dLv[k];
#pragma unroll
      for (int d = 0; d < XD; d++) if (d != k) t *= Lv[d];
      s += t * in_[k];
    }
    v[n] = sc * val;
    gn[n] = sc * s;
  }
}

__device__ inline int spec_off(const XArgs& a, int j, int sp) { return a.spec0[j] == sp ? 0 : a.N; }
__device__ inline int nrow(const XArgs& a, int j) { return a.nspec[j] * a.N; }
__device__ inline double* block_at(const XArgs& a, double* blocks, int j, int slot, int r, int c) {
  const int col = a.cols[j * a.S + slot];
  return blocks + a.boff[(size_t)j * a.S + slot] + (int64_t)r * nrow(a, col) + c;
}

// staged basis of up to XCHUNK nodes: values [XCHUNK][XN], gradients [XCHUNK][XN][XD]
struct Stage {
  double v[XCHUNK][XN];
  double g[XCHUNK][XN][XD];
};
// staged values and normal derivatives of up to XCHUNK nodes (faces, interface)
struct NStage {
  double v[XCHUNK][XN];
  double gn[XCHUNK][XN];
};

__global__ void __launch_bounds__(64) k_xdg_volume(XArgs a, const int32_t* __restrict__ cell, const int32_t* __restrict__ spec,
                                                   const int64_t* __restrict__ off, const double* __restrict__ nodes,
                                                   const double* __restrict__ w, double* __restrict__ blocks, double* __restrict__ affine) {
  __shared__ Stage st;
  __shared__ double sw[XCHUNK];
  const int it = blockIdx.x, lane = threadIdx.x;
  const int j = cell[it], sp = spec[it], o = spec_off(a, j, sp);
  const int N = a.N, NP = N * N;
  const double mu = a.mu[sp];
  double acc[XPAIRS];
  for (int i = 0; i < XPAIRS; i++) acc[i] = 0.0;
  double aff = 0.0;
  for (int64_t q0 = off[it]; q0 < off[it + 1]; q0 += XCHUNK) {
    const int nq = (int)std::min<int64_t>(XCHUNK, off[it + 1] - q0);
    if (lane < nq) {
      phys_basis(a, j, nodes + (q0 + lane) * a.D, st.v[lane], st.g[lane]);
      sw[lane] = w[q0 + lane];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < XPAIRS; i++) {
      const int pr = lane + 64 * i;
      if (pr < NP) {
        const int r = pr / N, c = pr % N;
        double s = 0.0;
        for (int q = 0; q < nq; q++) {
          double dot = 0.0;
          for (int d = 0; d < a.D; d++) dot += st.g[q][c][d] * st.g[q][r][d];
          s -= sw[q] * mu * dot;
        }
        acc[i] += s;
      }
    }
    if (lane < N)
      for (int q = 0; q < nq; q++) aff -= sw[q] * a.src[sp] * st.v[q][lane];
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < XPAIRS; i++) {
    const int pr = lane + 64 * i;
    if (pr < NP) *block_at(a, blocks, j, 0, o + pr / N, o + pr % N) += acc[i];
  }
  if (lane < N) affine[a.dof[j] + o + lane] += aff;
}

__global__ void __launch_bounds__(64) k_xdg_face(XArgs a, const int32_t* __restrict__ cells, const int32_t* __restrict__ fslot,
                                                 const int32_t* __restrict__ spec, const int64_t* __restrict__ off,
                                                 const double* __restrict__ nin, const double* __restrict__ nout,
                                                 const double* __restrict__ w, const double* __restrict__ normal,
                                                 const double* __restrict__ gD, double* __restrict__ blocks,
                                                 double* __restrict__ stage, double* __restrict__ astage) {
  __shared__ NStage si, so;
  __shared__ double sw[XCHUNK], sg[XCHUNK];
  const int it = blockIdx.x, lane = threadIdx.x;
  const int ji = cells[2 * it], jo = cells[2 * it + 1], sp = spec[it];
  const int N = a.N, NP = N * N, D = a.D;
  const double mu = a.mu[sp];
  double nrm[XD];
  for (int d = 0; d < D; d++) nrm[d] = normal[(size_t)it * D + d];
  const bool inner = jo >= 0;
  const double eta = inner ? a.penalty_base * fmax(1.0 / a.lscale[2 * ji + sp], 1.0 / a.lscale[2 * jo + sp]) * mu * a.pdeg : 0.0;
  const double pn = 2.0 * a.penalty_base * a.pdeg / a.lscale[2 * ji + sp];
  double acc[XPAIRS][4];
#pragma unroll
  for (int i = 0; i < XPAIRS; i++)
#pragma unroll
    for (int k = 0; k < 4; k++) acc[i][k] = 0.0;
  double aff = 0.0;
  for (int64_t q0 = off[it]; q0 < off[it + 1]; q0 += XCHUNK) {
    const int nq = (int)std::min<int64_t>(XCHUNK, off[it + 1] - q0);
    if (lane < nq) {
      phys_basis_n(a, ji, nin + (q0 + lane) * D, nrm, si.v[lane], si.gn[lane]);
      if (inner) phys_basis_n(a, jo, nout + (q0 + lane) * D, nrm, so.v[lane], so.gn[lane]);
      sw[lane] = w[q0 + lane];
      sg[lane] = gD ? gD[q0 + lane] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < XPAIRS; i++) {
      const int pr = lane + 64 * i;
      if (pr < NP) {
        const int r = pr / N, c = pr % N;
        for (int q = 0; q < nq; q++) {
          const double ww = sw[q];
          const double gri = si.gn[q][r], gci = si.gn[q][c];
          if (inner) {
            // (test side ts, trial side us): vA - vB = s_t phi_r, uA - uB = s_u phi_c
            const double vr[2] = {si.v[q][r], so.v[q][r]}, vc[2] = {si.v[q][c], so.v[q][c]};
            const double gr[2] = {gri, so.gn[q][r]}, gc[2] = {gci, so.gn[q][c]};
#pragma unroll
            for (int ts = 0; ts < 2; ts++)
#pragma unroll
              for (int us = 0; us < 2; us++) {
                const double st_ = ts ? -1.0 : 1.0, su = us ? -1.0 : 1.0;
                double f = 0.5 * mu * gc[us] * st_ * vr[ts] + 0.5 * mu * gr[ts] * su * vc[us] - eta * su * st_ * vc[us] * vr[ts];
                acc[i][ts * 2 + us] += ww * f;
              }
          } else {
            const double vr = si.v[q][r], vc = si.v[q][c];
            acc[i][0] += ww * (mu * gci * vr + mu * gri * vc - mu * vc * vr * pn);
          }
        }
      }
    }
    if (!inner && lane < N)
      for (int q = 0; q < nq; q++) aff += sw[q] * (-mu * si.gn[q][lane] * sg[q] + mu * sg[q] * si.v[q][lane] * pn);
    __syncthreads();
  }
  const int oi = spec_off(a, ji, sp), oo = inner ? spec_off(a, jo, sp) : 0;
  double* st = stage + (size_t)it * 2 * NP;
#pragma unroll
  for (int i = 0; i < XPAIRS; i++) {
    const int pr = lane + 64 * i;
    if (pr < NP) {
      const int r = pr / N, c = pr % N;
      st[pr] = acc[i][0];
      if (inner) {
        // the off-diagonal sub-blocks belong to this item alone (one item per face and species)
        *block_at(a, blocks, ji, fslot[2 * it], oi + r, oo + c) = acc[i][1];
        *block_at(a, blocks, jo, fslot[2 * it + 1], oo + r, oi + c) = acc[i][2];
        st[NP + pr] = acc[i][3];
      }
    }
  }
  if (!inner && lane < N) astage[(size_t)it * N + lane] = aff;
}

// Adds the staged diagonal contributions of every face item touching cell j (side 0: the item's
// IN cell, side 1: its OUT cell) in item order, and the boundary items' right-hand sides.
__global__ void __launch_bounds__(256) k_xdg_face_gather(XArgs a, const int32_t* __restrict__ goff, const int32_t* __restrict__ gent,
                                                         const int32_t* __restrict__ cells, const int32_t* __restrict__ spec,
                                                         const double* __restrict__ stage, const double* __restrict__ astage,
                                                         double* __restrict__ blocks, double* __restrict__ affine) {
  const int j = blockIdx.x, N = a.N, NP = N * N;
  for (int e = goff[j]; e < goff[j + 1]; e++) {
    const int it = gent[e] >> 1, side = gent[e] & 1;
    const int o = spec_off(a, j, spec[it]);
    const double* st = stage + ((size_t)it * 2 + side) * NP;
    for (int pr = threadIdx.x; pr < NP; pr += blockDim.x) *block_at(a, blocks, j, 0, o + pr / N, o + pr % N) += st[pr];
    if (cells[2 * it + 1] < 0 && (int)threadIdx.x < N) affine[a.dof[j] + o + threadIdx.x] += astage[(size_t)it * N + threadIdx.x];
  }
}

__global__ void __launch_bounds__(64) k_xdg_interface(XArgs a, const int32_t* __restrict__ cell, const int64_t* __restrict__ off,
                                                      const double* __restrict__ nodes, const double* __restrict__ w,
                                                      const double* __restrict__ normal, double* __restrict__ blocks) {
  __shared__ NStage st;
  __shared__ double sw[XCHUNK];
  const int it = blockIdx.x, lane = threadIdx.x;
  const int j = cell[it];
  if (a.nspec[j] != 2) return;
  const int N = a.N, NP = N * N, D = a.D;
  const double muA = a.mu[0], muB = a.mu[1];
  const double pmu = fmax(fabs(muA), fabs(muB)) * (muA < 0 ? -1.0 : 1.0);
  const double eta = a.penalty_base * fmax(1.0 / a.lscale[2 * j], 1.0 / a.lscale[2 * j + 1]) * pmu * a.pdeg;
  const double muS[2] = {muA, muB};
  double acc[XPAIRS][4];
#pragma unroll
  for (int i = 0; i < XPAIRS; i++)
#pragma unroll
    for (int k = 0; k < 4; k++) acc[i][k] = 0.0;
  for (int64_t q0 = off[it]; q0 < off[it + 1]; q0 += XCHUNK) {
    const int nq = (int)std::min<int64_t>(XCHUNK, off[it + 1] - q0);
    if (lane < nq) {
      phys_basis_n(a, j, nodes + (q0 + lane) * D, normal + (q0 + lane) * D, st.v[lane], st.gn[lane]);
      sw[lane] = w[q0 + lane];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < XPAIRS; i++) {
      const int pr = lane + 64 * i;
      if (pr < NP) {
        const int r = pr / N, c = pr % N;
        for (int q = 0; q < nq; q++) {
          const double gr = st.gn[q][r], gc = st.gn[q][c];
          const double vr = st.v[q][r], vc = st.v[q][c], ww = sw[q];
#pragma unroll
          for (int ts = 0; ts < 2; ts++)
#pragma unroll
            for (int us = 0; us < 2; us++) {
              const double st_ = ts ? -1.0 : 1.0, su = us ? -1.0 : 1.0;
              double f = muS[us] * 0.5 * gc * st_ * vr + muS[ts] * 0.5 * gr * su * vc - eta * su * st_ * vc * vr;
              acc[i][ts * 2 + us] += ww * f;
            }
        }
      }
    }
    __syncthreads();
  }
  const int o[2] = {spec_off(a, j, 0), spec_off(a, j, 1)};
#pragma unroll
  for (int i = 0; i < XPAIRS; i++) {
    const int pr = lane + 64 * i;
    if (pr < NP) {
      const int r = pr / N, c = pr % N;
#pragma unroll
      for (int ts = 0; ts < 2; ts++)
#pragma unroll
        for (int us = 0; us < 2; us++) *block_at(a, blocks, j, 0, o[ts] + r, o[us] + c) += acc[i][ts * 2 + us];
    }
  }
}

template <typename T>
__global__ void k_fill(int64_t n, T* __restrict__ x, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) x[i] = v;
}

// y <- alpha A x + beta y, one thread per matrix row (rowcell: the cell of each row)
template <typename Real>
__global__ void __launch_bounds__(256) k_xdg_spmv(int64_t ndof, int S, const int32_t* __restrict__ rowcell, const int32_t* __restrict__ nrows,
                                                  const int32_t* __restrict__ cols, const int64_t* __restrict__ boff,
                                                  const int64_t* __restrict__ dof, const Real* __restrict__ A, double alpha,
                                                  const Real* __restrict__ x, double beta, Real* __restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ndof) return;
  const int j = rowcell[t];
  const int r = (int)(t - dof[j]);
  double acc = 0.0;
  for (int s = 0; s < S; s++) {
    const int k = cols[j * S + s];
    if (k < 0) continue;
    const int m = nrows[k];
    const Real* row = A + boff[(size_t)j * S + s] + (int64_t)r * m;
    const Real* xk = x + dof[k];
    for (int c = 0; c < m; c++) acc += (double)row[c] * (double)xk[c];
  }
  y[t] = (Real)(alpha * acc + (beta != 0.0 ? beta * (double)y[t] : 0.0));
}

// Sum of v over the 64 lanes of a wavefront (fixed butterfly order: deterministic).
__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Apply on the solver's column-major copy (element (r, c) of block (j, s) at boff + c n_j + r).
// The blocks of cell row j follow each other in slot order, so together they are one n_j x M_j
// column-major panel, M_j = sum of the column cells' sizes.  One wavefront per cell row, four
// per workgroup.  Lane l < G n_j (G = 64 / n_j column groups) reads panel entries l + G n_j t,
// i.e. row l % n_j of panel columns l / n_j + G t: every load instruction reads G n_j consecutive
// entries.  The column cells' x entries are staged in LDS in panel-column order and the G
// partial sums of a row are added through LDS in a fixed order.  With DOT, the workgroup's
// sum of y_i x_i over its rows (p.q of CG, q as stored) goes to partial[2 b] (partial[2 b + 1] = 0).
constexpr int XPANEL = 7 * 2 * XN;   // panel columns: at most 7 slots of at most 2 N entries
template <typename Real, bool DOT, int T>
__global__ void __launch_bounds__(256 * T) k_xdg_spmv_panel(int J, int S, const int32_t* __restrict__ nrows, const int32_t* __restrict__ mcol,
                                                            const int64_t* __restrict__ xoff, const int64_t* __restrict__ boff,
                                                            const int64_t* __restrict__ dof, const Real* __restrict__ A, double alpha,
                                                            const Real* __restrict__ x, double beta, Real* __restrict__ y,
                                                            double* __restrict__ partial, int ngroups) {
  __shared__ double xs[4 * T][XPANEL];
  __shared__ double red[4 * T][64];
  __shared__ double wsum[4 * T];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;   // wave-uniform: scalar index loads
  const int j0 = blockIdx.x * 4 * T + w;
  const bool valid = j0 < J;
  const int j = valid ? j0 : J - 1;
  const int n = nrows[j];
  int M = 0;
  for (int s = 0; s < S; s++) {
    const int m = mcol[(size_t)j * S + s];
    if (lane < m) xs[w][M + lane] = (double)x[xoff[(size_t)j * S + s] + lane];
    M += m;
  }
  __syncthreads();
  const int G = 64 / n, GN = G * n, L = n * M;
  const Real* pa = A + boff[(size_t)j * S];
  double acc = 0.0;
  if (lane < GN) {
    int e = lane, c = lane / n;
    for (; e + 3 * GN < L; e += 4 * GN, c += 4 * G) {
      const double a0 = (double)pa[e], a1 = (double)pa[e + GN], a2 = (double)pa[e + 2 * GN], a3 = (double)pa[e + 3 * GN];
      acc += a0 * xs[w][c];
      acc += a1 * xs[w][c + G];
      acc += a2 * xs[w][c + 2 * G];
      acc += a3 * xs[w][c + 3 * G];
    }
    for (; e < L; e += GN, c += G) acc += (double)pa[e] * xs[w][c];
  }
  red[w][lane] = acc;
  __syncthreads();
  double t = 0.0;
  if (lane < n) {
    double a = 0.0;
    for (int g = 0; g < G; g++) a += red[w][g * n + lane];
    Real* yp = y + dof[j] + lane;
    const Real yi = (Real)(alpha * a + (beta != 0.0 ? beta * (double)*yp : 0.0));
    if (valid) *yp = yi;
    t = valid ? (double)yi * xs[w][lane] : 0.0;   // slot 0 is the diagonal: xs[w][0..n) = own x
  }
  if (DOT) {
    // one partial per group of four cell rows, whatever T: the same sums in the same order
    t = wave_sum(t);
    if (lane == 0) wsum[w] = t;
    __syncthreads();
    const int rg = blockIdx.x * T + (w >> 2);
    if (lane == 0 && (w & 3) == 0 && rg < ngroups) {
      partial[2 * rg] = ((wsum[w] + wsum[w + 1]) + wsum[w + 2]) + wsum[w + 3];
      partial[2 * rg + 1] = 0.0;
    }
  }
}

// row-major blocks -> column-major copy, one workgroup per block
__global__ void k_blocks_to_cm(int S, const int32_t* __restrict__ nrows, const int32_t* __restrict__ cols,
                               const int64_t* __restrict__ boff, const double* __restrict__ A, double* __restrict__ At) {
  const int j = blockIdx.x / S, s = blockIdx.x % S;
  const int k = cols[j * S + s];
  if (k < 0) return;
  const int n = nrows[j], m = nrows[k];
  const int64_t o = boff[(size_t)j * S + s];
  for (int i = threadIdx.x; i < n * m; i += blockDim.x) {
    const int r = i / m, c = i % m;
    At[o + (int64_t)c * n + r] = A[o + i];
  }
}

// gather the diagonal blocks into a compact batch (offsets doff) for the block LU
__global__ void k_gather_diag(int J, int S, const int32_t* __restrict__ nrows, const int64_t* __restrict__ boff,
                              const int64_t* __restrict__ doff, const double* __restrict__ A, double* __restrict__ Dg) {
  const int j = blockIdx.x;
  const int n = nrows[j];
  for (int i = threadIdx.x; i < n * n; i += blockDim.x) Dg[doff[j] + i] = A[boff[(size_t)j * S] + i];
}

// In-place transpose of the n_j x n_j blocks at doff[j] (the LU's row-major inverses -> the
// column-major form the fused CG update reads), one workgroup per block.
__global__ void __launch_bounds__(64) k_transpose_blocks(int J, const int32_t* __restrict__ nrows, const int64_t* __restrict__ doff,
                                                         double* __restrict__ Dg) {
  __shared__ double tile[2 * XN * 2 * XN];
  const int j = blockIdx.x, n = nrows[j];
  double* b = Dg + doff[j];
  for (int i = threadIdx.x; i < n * n; i += 64) tile[i] = b[i];
  __syncthreads();
  for (int i = threadIdx.x; i < n * n; i += 64) b[i] = tile[(i % n) * n + i / n];
}

// Diagonal-block equilibration for the FP32 solve (the reference's SymPart_DiagBlockEquilib
// pre-preconditioning, MultigridOperator.Mode): D_j = L_j L_j^T (Cholesky), the FP32 CG runs
// on Linv M Linv^T, whose diagonal blocks are the identity.
// L <- inv(chol(D)) in place (lower triangular, upper part zero); info[j] = first bad pivot.
__global__ void __launch_bounds__(64) k_chol_inv(int J, const int32_t* __restrict__ nrows, const int64_t* __restrict__ doff,
                                                 double* __restrict__ Dg, int32_t* __restrict__ info) {
  __shared__ double L[40 * 41], X[40 * 41];
  const int j = blockIdx.x, lane = threadIdx.x;
  const int n = nrows[j], ld = 41;
  if (n > 40) { if (lane == 0) info[j] = -1; return; }
  double* g = Dg + doff[j];
  for (int i = lane; i < n * n; i += 64) L[(i / n) * ld + (i % n)] = g[i];
  __syncthreads();
  int bad = 0;
  for (int c = 0; c < n; c++) {
    double d = L[c * ld + c];
    for (int k = 0; k < c; k++) d -= L[c * ld + k] * L[c * ld + k];
    if (!(d > 0.0)) { bad = c + 1; break; }
    const double dc = sqrt(d);
    __syncthreads();
    for (int i = lane; i < n; i += 64) {
      if (i > c) {
        double v = L[i * ld + c];
        for (int k = 0; k < c; k++) v -= L[i * ld + k] * L[c * ld + k];
        L[i * ld + c] = v / dc;
      }
    }
    __syncthreads();
    if (lane == 0) L[c * ld + c] = dc;
    __syncthreads();
  }
  if (bad) { if (lane == 0) info[j] = bad; return; }
  // X = L^-1: lane c owns column c
  for (int c = lane; c < n; c += 64)
    for (int i = 0; i < n; i++) {
      double v = (i == c) ? 1.0 : 0.0;
      if (i < c) { X[i * ld + c] = 0.0; continue; }
      for (int k = c; k < i; k++) v -= L[i * ld + k] * X[k * ld + c];
      X[i * ld + c] = v / L[i * ld + i];
    }
  __syncthreads();
  for (int i = lane; i < n * n; i += 64) g[i] = X[(i / n) * ld + (i % n)];
  if (lane == 0) info[j] = 0;
}

// A32(j, s) = (float)(Linv_j A(j, s) Linv_k^T), column-major, one workgroup per block
__global__ void __launch_bounds__(64) k_scale_blocks(int J, int S, const int32_t* __restrict__ nrows, const int32_t* __restrict__ cols,
                                                     const int64_t* __restrict__ boff, const int64_t* __restrict__ doff,
                                                     const double* __restrict__ Linv, const double* __restrict__ A,
                                                     float* __restrict__ A32) {
  __shared__ double Lj[40 * 40], Lk[40 * 40], B[40 * 40], T[40 * 40];
  const int j = blockIdx.x / S, s = blockIdx.x % S, lane = threadIdx.x;
  const int k = cols[j * S + s];
  if (k < 0) return;
  const int n = nrows[j], m = nrows[k];
  const int64_t o = boff[(size_t)j * S + s];
  for (int i = lane; i < n * n; i += 64) Lj[i] = Linv[doff[j] + i];
  for (int i = lane; i < m * m; i += 64) Lk[i] = Linv[doff[k] + i];
  for (int i = lane; i < n * m; i += 64) B[i] = A[o + i];
  __syncthreads();
  for (int i = lane; i < n * m; i += 64) {          // T = B Lk^T
    const int r = i / m, c = i % m;
    double v = 0.0;
    for (int l = 0; l <= c; l++) v += B[r * m + l] * Lk[c * m + l];
    T[i] = v;
  }
  __syncthreads();
  for (int i = lane; i < n * m; i += 64) {          // Lj T
    const int r = i / m, c = i % m;
    double v = 0.0;
    for (int l = 0; l <= r; l++) v += Lj[r * n + l] * T[l * m + c];
    A32[o + (int64_t)c * n + r] = (float)v;      // column-major solver copy
  }
}

// out_j = scale * (Linv_j or Linv_j^T) in_j  (+ out_j if accumulate)
template <typename In, typename Out>
__global__ void __launch_bounds__(64) k_lower_apply(int J, const int32_t* __restrict__ nrows, const int64_t* __restrict__ doff,
                                                    const int64_t* __restrict__ dof, const double* __restrict__ Linv, int transpose,
                                                    double scale, const In* __restrict__ in, Out* __restrict__ out, int accumulate) {
  const int j = blockIdx.x, i = threadIdx.x;
  const int n = nrows[j];
  if (i >= n) return;
  const double* L = Linv + doff[j];
  const In* v = in + dof[j];
  double acc = 0.0;
  if (!transpose) { for (int c = 0; c <= i; c++) acc += L[i * n + c] * (double)v[c]; }
  else { for (int c = i; c < n; c++) acc += L[c * n + i] * (double)v[c]; }
  Out* o = out + dof[j] + i;
  *o = (Out)((accumulate ? (double)*o : 0.0) + scale * acc);
}


// ------------------------------------------------------------ reductions
constexpr int RED_BLOCKS = 256, RED_THREADS = 256;

__device__ inline double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) t += sh[i];
  return t;
}

// partial[2*b + {0,1}] = sums over this block of x.y and x.x (second only if want2)
template <typename Real>
__global__ void __launch_bounds__(RED_THREADS) k_dot_partial(int64_t n, const Real* __restrict__ x, const Real* __restrict__ y,
                                                             int want2, double* __restrict__ partial) {
  __shared__ double sh[RED_THREADS / 64];
  double s1 = 0.0, s2 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double xi = (double)x[i];
    s1 += xi * (double)y[i];
    if (want2) s2 += xi * xi;
  }
  double t1 = block_sum(s1, sh);
  __syncthreads();
  double t2 = want2 ? block_sum(s2, sh) : 0.0;
  if (threadIdx.x == 0) { partial[2 * blockIdx.x] = t1; partial[2 * blockIdx.x + 1] = t2; }
}

// scal[slot1] = sum of partial sums 0, scal[slot2] = sum of partial sums 1 (slot2 < 0: skip);
// with cdst >= 0, scal[cdst] = scal[csrc] first (rz_old <- rz_new at the start of a CG step).
// Any block size that is a multiple of 64 up to 1024; fixed summation order.
__global__ void __launch_bounds__(1024) k_dot_final(int nb, const double* __restrict__ partial, double* __restrict__ scal,
                                                    int slot1, int slot2, int cdst, int csrc) {
  __shared__ double sh[16];
  const int T = blockDim.x;
  double a1[4] = {0.0, 0.0, 0.0, 0.0}, a2[4] = {0.0, 0.0, 0.0, 0.0};
  int i = threadIdx.x;
  for (; i + 3 * T < nb; i += 4 * T) {
#pragma unroll
    for (int u = 0; u < 4; u++) { a1[u] += partial[2 * (i + u * T)]; a2[u] += partial[2 * (i + u * T) + 1]; }
  }
  for (; i < nb; i += T) { a1[0] += partial[2 * i]; a2[0] += partial[2 * i + 1]; }
  const double s1 = (a1[0] + a1[1]) + (a1[2] + a1[3]), s2 = (a2[0] + a2[1]) + (a2[2] + a2[3]);
  double t1 = block_sum(s1, sh);
  __syncthreads();
  double t2 = block_sum(s2, sh);
  if (threadIdx.x == 0) {
    if (cdst >= 0) scal[cdst] = scal[csrc];
    scal[slot1] = t1;
    if (slot2 >= 0) scal[slot2] = t2;
  }
}

// PCG scalar slots
enum { S_RZ = 0, S_PQ = 1, S_RZN = 2, S_RR = 3, S_BB = 4, S_NSLOT = 8 };

// Fused CG step on one cell row per wavefront (4 T per workgroup): with STEP, x += alpha p and
// r -= alpha q (alpha = rz / pq, 0 if pq == 0); then z = Dinv r with the column-major inverse
// diagonal block (lane l < G n reads entries l + G n t, as the panel apply), or z = r when IDENT
// (the FP32 solve runs on the equilibrated matrix, whose diagonal blocks are the identity); the
// workgroup's sums of r.z and r.r go to partial[2 b], partial[2 b + 1].
template <typename Real, bool IDENT, int T>
__global__ void __launch_bounds__(256 * T) k_pcg_update(int J, int step, const int32_t* __restrict__ nrows, const int64_t* __restrict__ dof,
                                                        const int64_t* __restrict__ doff, const double* __restrict__ scal,
                                                        const Real* __restrict__ Dinv, const Real* __restrict__ p, const Real* __restrict__ q,
                                                        Real* __restrict__ x, Real* __restrict__ r, Real* __restrict__ z,
                                                        double* __restrict__ partial, int ngroups) {
  __shared__ double rs[4 * T][64];
  __shared__ double red[4 * T][64];
  __shared__ double wsum[4 * T][2];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;   // wave-uniform: scalar index loads
  const int j0 = blockIdx.x * 4 * T + w;
  const bool valid = j0 < J;
  const int j = valid ? j0 : J - 1;
  const int n = nrows[j];
  const int64_t o = dof[j];
  double ri = 0.0;
  if (lane < n) {
    if (step) {
      const double pq = scal[S_PQ];
      const double al = (pq != 0.0) ? scal[S_RZ] / pq : 0.0;
      const Real xi = (Real)((double)x[o + lane] + al * (double)p[o + lane]);
      const Real rr = (Real)((double)r[o + lane] - al * (double)q[o + lane]);
      if (valid) { x[o + lane] = xi; r[o + lane] = rr; }
      ri = (double)rr;
    } else {
      ri = (double)r[o + lane];
    }
  }
  double zi = ri;
  if (!IDENT) {
    rs[w][lane] = ri;
    __syncthreads();
    const int G = 64 / n, GN = G * n, L = n * n;
    const Real* pd = Dinv + doff[j];
    double acc = 0.0;
    if (lane < GN)
      for (int e = lane, c = lane / n; e < L; e += GN, c += G) acc += (double)pd[e] * rs[w][c];
    red[w][lane] = acc;
    __syncthreads();
    zi = 0.0;
    if (lane < n)
      for (int g = 0; g < G; g++) zi += red[w][g * n + lane];
  }
  double t1 = 0.0, t2 = 0.0;
  if (lane < n && valid) {
    const Real zr = (Real)zi;
    z[o + lane] = zr;
    t1 = ri * (double)zr;
    t2 = ri * ri;
  }
  t1 = wave_sum(t1);
  t2 = wave_sum(t2);
  if (lane == 0) { wsum[w][0] = t1; wsum[w][1] = t2; }
  __syncthreads();
  const int rg = blockIdx.x * T + (w >> 2);
  if (lane == 0 && (w & 3) == 0 && rg < ngroups) {
    partial[2 * rg] = ((wsum[w][0] + wsum[w + 1][0]) + wsum[w + 2][0]) + wsum[w + 3][0];
    partial[2 * rg + 1] = ((wsum[w][1] + wsum[w + 1][1]) + wsum[w + 2][1]) + wsum[w + 3][1];
  }
}
// p = z + beta p with beta = rz_new / rz (0 if rz == 0)
template <typename Real>
__global__ void k_pcg_p(int64_t n, const double* __restrict__ scal, const Real* __restrict__ z, Real* __restrict__ p) {
  const double rz = scal[S_RZ];
  const double be = (rz != 0.0) ? scal[S_RZN] / rz : 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (Real)((double)z[i] + be * (double)p[i]);
}
// y = (Real)(s * x)   (FP64 residual -> scaled FP32 right-hand side, and back)
template <typename A, typename B>
__global__ void k_scale_copy(int64_t n, double s, const A* __restrict__ x, B* __restrict__ y, int accumulate) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (B)((accumulate ? (double)y[i] : 0.0) + s * (double)x[i]);
}

}  // namespace
}  // namespace bosss

using namespace bosss;

struct bosss_xdg {
  bosss_ctx* ctx = nullptr;
  int device = 0;
  hipStream_t stream = nullptr;
  XArgs a{};
  int nvol = 0, nface = 0, nifc = 0;
  int64_t ndof = 0, nentries = 0, ndiag = 0;
  std::vector<void*> allocs;
  int32_t *d_nspec, *d_spec0, *d_expo, *d_cols, *d_nrows, *d_rowcell, *d_mcol;
  int64_t *d_boff, *d_dof, *d_doff, *d_xoff;
  double *d_invJ, *d_scal, *d_ls;
  int32_t *v_cell, *v_spec, *f_cells, *f_slot, *f_spec, *i_cell;
  int64_t *v_off, *f_off, *i_off;
  double *v_nodes, *v_w, *f_nin, *f_nout, *f_w, *f_normal, *f_gD, *i_nodes, *i_w, *i_normal;
  // face items touching each cell (CSR, entries item * 2 + side, in item order) and the staged
  // diagonal sub-blocks [nface][2][N N] and boundary right-hand sides [nface][N] of the items
  int32_t *g_off = nullptr, *g_ent = nullptr;
  double *f_stage = nullptr, *f_astage = nullptr;
  // solver workspace (allocated on first solve)
  bool ws = false;
  double *Dinv = nullptr, *Lchol = nullptr, *At = nullptr, *scal = nullptr, *partial = nullptr;
  int32_t* info = nullptr;
  double *r64 = nullptr, *z64 = nullptr, *p64 = nullptr, *q64 = nullptr;
  float *A32 = nullptr, *b32 = nullptr, *r32 = nullptr, *z32 = nullptr, *p32 = nullptr, *q32 = nullptr, *x32 = nullptr;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  const void* graph_key[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  int graph_k[2] = {0, 0};
  int graph_tile[2] = {0, 0};
  // tile shape of the solver's cell-row kernels (T groups of four rows per workgroup), per
  // precision: 0 = tune at the next solve; tile_ms = timed ms per apply + update of each candidate
  int tile[2] = {0, 0};
  float tile_ms[2][3] = {{-1.f, -1.f, -1.f}, {-1.f, -1.f, -1.f}};
  ~bosss_xdg() {
    for (int i = 0; i < 2; i++) if (graph[i]) (void)hipGraphExecDestroy(graph[i]);
    for (void* p : allocs) (void)hipFree(p);
  }
};

namespace bosss {
// the context's device and internal stream (bosss_hip.hip)
int ctx_device_stream(bosss_ctx* ctx, int* device, hipStream_t* stream);
}

namespace {

template <typename T>
int upload(bosss_xdg* x, T** dst, const T* src, size_t n) {
  *dst = nullptr;
  if (n == 0) return BOSSS_OK;
  if (!src) return xfail(BOSSS_ERR_ARG, "missing array (%zu elements)", n);
  XHIP(hipMalloc((void**)dst, n * sizeof(T)));
  x->allocs.push_back(*dst);
  XHIP(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return BOSSS_OK;
}
template <typename T>
int alloc(bosss_xdg* x, T** dst, size_t n) {
  XHIP(hipMalloc((void**)dst, std::max<size_t>(n, 1) * sizeof(T)));
  x->allocs.push_back(*dst);
  return BOSSS_OK;
}
#define XCHK(x)                    \
  do {                             \
    int rc_ = (x);                 \
    if (rc_ != BOSSS_OK) return rc_; \
  } while (0)

inline unsigned grid_for(int64_t n, int threads = 256) {
  int64_t b = (n + threads - 1) / threads;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 4096));
}

template <typename Real>
void launch_dot(bosss_xdg* X, int64_t n, const Real* x, const Real* y, int want2, int slot1, int slot2, hipStream_t s) {
  hipLaunchKernelGGL(k_dot_partial<Real>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, n, x, y, want2, X->partial);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(RED_THREADS), 0, s, RED_BLOCKS, X->partial, X->scal, slot1, slot2, -1, 0);
}

template <typename Real>
void launch_spmv(bosss_xdg* X, const Real* A, double alpha, const Real* x, double beta, Real* y, hipStream_t s) {
  hipLaunchKernelGGL(k_xdg_spmv<Real>, dim3((unsigned)((X->ndof + 255) / 256)), dim3(256), 0, s, X->ndof, X->a.S, X->d_rowcell,
                     X->d_nrows, X->d_cols, X->d_boff, X->d_dof, A, alpha, x, beta, y);
}

// groups of four cell rows: one dot-product partial each, whatever the tile shape
inline int row_groups(const bosss_xdg* X) { return (X->a.J + 3) / 4; }

// tile shapes of the cell-row kernels: T groups of four rows (4 T wavefronts) per workgroup
constexpr int XDG_TILES[3] = {1, 2, 4};
inline int tile_of(const bosss_xdg* X, int prec) { return X->tile[prec] > 0 ? X->tile[prec] : 1; }

template <typename Real, bool DOT, int T>
void spmv_panel_T(bosss_xdg* X, const Real* A, double alpha, const Real* x, double beta, Real* y, hipStream_t s) {
  const int ng = row_groups(X);
  hipLaunchKernelGGL((k_xdg_spmv_panel<Real, DOT, T>), dim3((unsigned)((ng + T - 1) / T)), dim3(256 * T), 0, s, X->a.J, X->a.S, X->d_nrows,
                     X->d_mcol, X->d_xoff, X->d_boff, X->d_dof, A, alpha, x, beta, y, X->partial, ng);
}
template <typename Real, bool IDENT, int T>
void update_T(bosss_xdg* X, int step, const Real* Dinv, const Real* p, const Real* q, Real* x, Real* r, Real* z, hipStream_t s) {
  const int ng = row_groups(X);
  hipLaunchKernelGGL((k_pcg_update<Real, IDENT, T>), dim3((unsigned)((ng + T - 1) / T)), dim3(256 * T), 0, s, X->a.J, step, X->d_nrows,
                     X->d_dof, X->d_doff, X->scal, Dinv, p, q, x, r, z, X->partial, ng);
}

// y = alpha A x + beta y on the column-major copy; with dot_slot >= 0 also scal[dot_slot] = x.y
// (and first scal[cdst] = scal[csrc] when cdst >= 0)
template <typename Real>
void launch_spmv_cm(bosss_xdg* X, const Real* A, double alpha, const Real* x, double beta, Real* y, hipStream_t s, int dot_slot = -1,
                    int cdst = -1, int csrc = 0) {
  const int T = tile_of(X, sizeof(Real) == 4);
  if (dot_slot < 0) {
    if (T == 4) spmv_panel_T<Real, false, 4>(X, A, alpha, x, beta, y, s);
    else if (T == 2) spmv_panel_T<Real, false, 2>(X, A, alpha, x, beta, y, s);
    else spmv_panel_T<Real, false, 1>(X, A, alpha, x, beta, y, s);
  } else {
    if (T == 4) spmv_panel_T<Real, true, 4>(X, A, alpha, x, beta, y, s);
    else if (T == 2) spmv_panel_T<Real, true, 2>(X, A, alpha, x, beta, y, s);
    else spmv_panel_T<Real, true, 1>(X, A, alpha, x, beta, y, s);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(1024), 0, s, row_groups(X), X->partial, X->scal, dot_slot, -1, cdst, csrc);
  }
}

// Fused x/r update (step != 0), z = Dinv r (Dinv == nullptr: identity blocks), then
// scal[S_RZN] = r.z, scal[S_RR] = r.r.
template <typename Real>
void launch_update(bosss_xdg* X, int step, const Real* Dinv, const Real* p, const Real* q, Real* x, Real* r, Real* z, hipStream_t s) {
  const int T = tile_of(X, sizeof(Real) == 4);
  if (Dinv) {
    if (T == 4) update_T<Real, false, 4>(X, step, Dinv, p, q, x, r, z, s);
    else if (T == 2) update_T<Real, false, 2>(X, step, Dinv, p, q, x, r, z, s);
    else update_T<Real, false, 1>(X, step, Dinv, p, q, x, r, z, s);
  } else {
    if (T == 4) update_T<Real, true, 4>(X, step, Dinv, p, q, x, r, z, s);
    else if (T == 2) update_T<Real, true, 2>(X, step, Dinv, p, q, x, r, z, s);
    else update_T<Real, true, 1>(X, step, Dinv, p, q, x, r, z, s);
  }
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(1024), 0, s, row_groups(X), X->partial, X->scal, (int)S_RZN, (int)S_RR, -1, 0);
}

// Autotuner of the tile shape (once per precision, not under stream capture): times five rounds of
// the iteration's apply (with its p.q partials) and update (step 0: z = Dinv r and its partials,
// x and r untouched) for each candidate on the solver's own matrix and keeps the fastest.  Every
// candidate forms the same partial sums per group of four rows in the same order, so the CG
// iterates do not depend on the choice.  The buffers it runs on are set again by pcg_start.
template <typename Real>
int tune_tiles(bosss_xdg* X, const Real* A, const Real* Dinv, Real* p, Real* q, Real* r, Real* z, hipStream_t s) {
  const int prec = sizeof(Real) == 4;
  if (X->tile[prec] > 0) return BOSSS_OK;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  XHIP(hipStreamIsCapturing(s, &cs));
  if (cs != hipStreamCaptureStatusNone) return BOSSS_OK;     // keep the default shape for this solve
  hipEvent_t e0 = nullptr, e1 = nullptr;
  XHIP(hipEventCreate(&e0));
  hipError_t e = hipEventCreate(&e1);
  float best = 1e30f;
  int bestT = 1;
  for (int c = 0; c < 3 && e == hipSuccess; c++) {
    X->tile[prec] = XDG_TILES[c];
    launch_spmv_cm<Real>(X, A, 1.0, p, 0.0, q, s, S_PQ, S_RZ, S_RZN);
    launch_update<Real>(X, 0, Dinv, p, q, nullptr, r, z, s);
    e = hipEventRecord(e0, s);
    for (int k = 0; k < 5 && e == hipSuccess; k++) {
      launch_spmv_cm<Real>(X, A, 1.0, p, 0.0, q, s, S_PQ, S_RZ, S_RZN);
      launch_update<Real>(X, 0, Dinv, p, q, nullptr, r, z, s);
    }
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess) {
      X->tile_ms[prec][c] = ms / 5;
      if (ms < best) { best = ms; bestT = XDG_TILES[c]; }
    }
  }
  (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (e != hipSuccess) { X->tile[prec] = 0; return xfail(BOSSS_ERR_HIP, "tile autotuner: %s", hipGetErrorString(e)); }
  X->tile[prec] = bestT;
  return BOSSS_OK;
}

// One PCG iteration (Real = double or float), all scalars on the device; A column-major, Dinv
// the column-major inverse diagonal blocks (nullptr: identity).  Slot S_RZN holds r.z of the
// previous step; the p.q reduction moves it to S_RZ before alpha = rz / pq is used.
template <typename Real>
void pcg_iteration(bosss_xdg* X, const Real* A, const Real* Dinv, Real* x, Real* r, Real* z, Real* p, Real* q, hipStream_t s) {
  const int64_t n = X->ndof;
  launch_spmv_cm<Real>(X, A, 1.0, p, 0.0, q, s, S_PQ, S_RZ, S_RZN);
  launch_update<Real>(X, 1, Dinv, p, q, x, r, z, s);
  hipLaunchKernelGGL(k_pcg_p<Real>, dim3(grid_for(n)), dim3(256), 0, s, n, X->scal, z, p);
}

// r = b - A x, z = Dinv r, p = z, scal[S_RZN] = r.z, rr = r.r, bb = b.b
template <typename Real>
void pcg_start(bosss_xdg* X, const Real* A, const Real* Dinv, const Real* b, Real* x, Real* r, Real* z, Real* p, hipStream_t s) {
  const int64_t n = X->ndof;
  hipLaunchKernelGGL((k_scale_copy<Real, Real>), dim3(grid_for(n)), dim3(256), 0, s, n, 1.0, b, r, 0);
  launch_spmv_cm<Real>(X, A, -1.0, x, 1.0, r, s);
  launch_update<Real>(X, 0, Dinv, p, p, x, r, z, s);
  hipLaunchKernelGGL((k_scale_copy<Real, Real>), dim3(grid_for(n)), dim3(256), 0, s, n, 1.0, z, p, 0);
  launch_dot<Real>(X, n, b, b, 0, S_BB, -1, s);
}

// Run PCG until sqrt(rr/bb) <= tol or maxit; K iterations per check (one graph launch).
template <typename Real>
int pcg_solve(bosss_xdg* X, int gslot, const Real* A, const Real* Dinv, const Real* b, Real* x, Real* r, Real* z, Real* p, Real* q,
              double tol, int maxit, int K, int use_graph, bool stop_on_stall, int* iters, double* relres) {
  hipStream_t s = X->stream;
  if (int rc = tune_tiles<Real>(X, A, Dinv, p, q, r, z, s)) return rc;
  pcg_start<Real>(X, A, Dinv, b, x, r, z, p, s);
  double h[S_NSLOT];
  XHIP(hipMemcpyAsync(h, X->scal, sizeof h, hipMemcpyDeviceToHost, s));
  XHIP(hipStreamSynchronize(s));
  double bb = h[S_BB], rel = bb > 0 ? std::sqrt(h[S_RR] / bb) : 0.0;
  int it = 0;
  if (use_graph) {
    // the graph bakes in the matrix and solution pointers: rebuild when they change
    const void* key0 = (const void*)A;
    const void* key1 = (const void*)x;
    if (!X->graph[gslot] || X->graph_key[gslot][0] != key0 || X->graph_key[gslot][1] != key1 || X->graph_k[gslot] != K ||
        X->graph_tile[gslot] != tile_of(X, sizeof(Real) == 4)) {
      if (X->graph[gslot]) { XHIP(hipGraphExecDestroy(X->graph[gslot])); X->graph[gslot] = nullptr; }
      hipGraph_t g;
      XHIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      for (int k = 0; k < K; k++) pcg_iteration<Real>(X, A, Dinv, x, r, z, p, q, s);
      XHIP(hipStreamEndCapture(s, &g));
      hipError_t e = hipGraphInstantiate(&X->graph[gslot], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return xfail(BOSSS_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
      X->graph_key[gslot][0] = key0;
      X->graph_key[gslot][1] = key1;
      X->graph_k[gslot] = K;
      X->graph_tile[gslot] = tile_of(X, sizeof(Real) == 4);
    }
  }
  // The recurrence residual drifts from b - A x in finite precision: on convergence the true
  // residual is recomputed (pcg_start) and CG restarted from the current iterate if it misses.
  // An inner (FP32) solve that stops improving -- the attainable accuracy of the working
  // precision -- ends early: no 10 % gain on the best residual over `stall` consecutive checks
  // (>= 512 iterations; the CG residual norm is not monotone).
  const int stall = std::max(8, 512 / K);
  bool stalled = false;
  for (int restart = 0; restart < 8 && !stalled; restart++) {
    double best = rel;
    int since = 0;
    while (rel > tol && it < maxit) {
      if (use_graph) {
        XHIP(hipGraphLaunch(X->graph[gslot], s));
      } else {
        for (int k = 0; k < K; k++) pcg_iteration<Real>(X, A, Dinv, x, r, z, p, q, s);
      }
      it += K;
      XHIP(hipMemcpyAsync(h, X->scal, sizeof h, hipMemcpyDeviceToHost, s));
      XHIP(hipStreamSynchronize(s));
      rel = std::sqrt(std::max(h[S_RR], 0.0) / bb);
      if (!std::isfinite(rel)) return xfail(BOSSS_ERR_NONFINITE, "PCG residual is not finite after %d iterations", it);
      if (rel < 0.9 * best) { best = rel; since = 0; }
      else if (stop_on_stall && ++since >= stall) { stalled = true; break; }
    }
    if (it == 0) break;
    pcg_start<Real>(X, A, Dinv, b, x, r, z, p, s);        // true residual r = b - A x
    XHIP(hipMemcpyAsync(h, X->scal, sizeof h, hipMemcpyDeviceToHost, s));
    XHIP(hipStreamSynchronize(s));
    rel = std::sqrt(std::max(h[S_RR], 0.0) / bb);
    if (rel <= tol || it >= maxit) break;
  }
  *iters = it;
  *relres = rel;
  return BOSSS_OK;
}

}  // namespace

extern "C" {

const char* bosss_hip_xdg_last_error(void) { return g_xdg_err.c_str(); }

int bosss_hip_xdg_create(bosss_ctx* ctx, const bosss_xdg_desc* d, bosss_xdg** out) {
  if (!ctx || !d || !out) return xfail(BOSSS_ERR_ARG, "NULL argument");
  if (d->D < 2 || d->D > 3) return xfail(BOSSS_ERR_ARG, "D must be 2 or 3");
  if (d->N < 1 || d->N > XN || d->degree < 0 || d->degree > 3) return xfail(BOSSS_ERR_UNSUPPORTED, "XDG kernels cover p <= 3 (N <= %d)", XN);
  if (d->S != 1 + 2 * d->D) return xfail(BOSSS_ERR_ARG, "S must be 1 + 2D");
  if (d->J <= 0 || d->nvol < 0 || d->nface < 0 || d->nifc < 0) return xfail(BOSSS_ERR_ARG, "bad sizes J=%d nvol=%d nface=%d nifc=%d", d->J, d->nvol, d->nface, d->nifc);
  if (!d->mode_exponents || !d->nspec || !d->spec0 || !d->invJ || !d->scaling || !d->lscale || !d->cols || !d->boff)
    return xfail(BOSSS_ERR_ARG, "XDG descriptor has NULL arrays");
  if ((d->nvol && (!d->vol_cell || !d->vol_spec || !d->vol_off || !d->vol_nodes || !d->vol_w)) ||
      (d->nface && (!d->face_cells || !d->face_slot || !d->face_spec || !d->face_off || !d->face_nodes_in || !d->face_nodes_out ||
                    !d->face_w || !d->face_normal)) ||   // face_gD may be NULL: homogeneous Dirichlet data
      (d->nifc && (!d->ifc_cell || !d->ifc_off || !d->ifc_nodes || !d->ifc_w || !d->ifc_normal)))
    return xfail(BOSSS_ERR_ARG, "XDG descriptor has NULL quadrature-item arrays");
  for (int n = 0; n < d->N * d->D; n++)
    if (d->mode_exponents[n] < 0 || d->mode_exponents[n] > d->degree) return xfail(BOSSS_ERR_ARG, "mode exponent out of range");
  for (int j = 0; j < d->J; j++)
    if (d->nspec[j] < 1 || d->nspec[j] > 2 || d->spec0[j] < 0 || d->spec0[j] > 1) return xfail(BOSSS_ERR_ARG, "bad species layout at cell %d", j);
  int dev = 0;
  hipStream_t st = nullptr;
  int rc = ctx_device_stream(ctx, &dev, &st);
  if (rc != BOSSS_OK) return rc;
  XHIP(hipSetDevice(dev));
  bosss_xdg* x = new bosss_xdg();
  x->ctx = ctx; x->device = dev; x->stream = st;
  const int J = d->J, D = d->D, S = d->S, N = d->N;
  x->nvol = d->nvol; x->nface = d->nface; x->nifc = d->nifc;
  // host-side derived layout
  std::vector<int32_t> nrows(J);
  std::vector<int64_t> dof(J + 1, 0), doff(J + 1, 0);
  for (int j = 0; j < J; j++) { nrows[j] = d->nspec[j] * N; dof[j + 1] = dof[j] + nrows[j]; doff[j + 1] = doff[j] + (int64_t)nrows[j] * nrows[j]; }
  int64_t ent = 0;
  for (int j = 0; j < J; j++)
    for (int s = 0; s < S; s++) {
      int k = d->cols[j * S + s];
      if (k < -1 || k >= J || (s == 0 && k != j)) { delete x; return xfail(BOSSS_ERR_ARG, "bad column cell at (%d, %d)", j, s); }
      if (k >= 0) {
        if (d->boff[(size_t)j * S + s] < 0) { delete x; return xfail(BOSSS_ERR_ARG, "bad block offset at (%d, %d)", j, s); }
        ent = std::max<int64_t>(ent, d->boff[(size_t)j * S + s] + (int64_t)nrows[j] * nrows[k]);
      }
    }
  // the solver streams each cell row's blocks as one panel: they must follow each other in
  // slot order (the layout block_layout builds; rows themselves may be anywhere)
  for (int j = 0; j < J; j++) {
    int64_t next = d->boff[(size_t)j * S];
    for (int s = 0; s < S; s++) {
      const int k = d->cols[j * S + s];
      if (k < 0) continue;
      if (d->boff[(size_t)j * S + s] != next) {
        delete x;
        return xfail(BOSSS_ERR_ARG, "cell row %d: blocks must follow each other in slot order (slot %d)", j, s);
      }
      next += (int64_t)nrows[j] * nrows[k];
    }
  }
  x->ndof = dof[J]; x->nentries = ent; x->ndiag = doff[J];
  // bounds checks of the quadrature items against the layout
  auto bad_off = [](const int64_t* o, int n) { if (n && o[0] != 0) return true; for (int i = 0; i < n; i++) if (o[i + 1] < o[i]) return true; return false; };
  if (bad_off(d->vol_off, d->nvol) || bad_off(d->face_off, d->nface) || bad_off(d->ifc_off, d->nifc)) {
    delete x; return xfail(BOSSS_ERR_ARG, "node offsets must start at 0 and increase");
  }
  for (int i = 0; i < d->nvol; i++)
    if (d->vol_cell[i] < 0 || d->vol_cell[i] >= J || d->vol_spec[i] < 0 || d->vol_spec[i] > 1 ||
        !(d->spec0[d->vol_cell[i]] == d->vol_spec[i] || d->nspec[d->vol_cell[i]] == 2)) { delete x; return xfail(BOSSS_ERR_ARG, "volume item %d: cell/species not in the layout", i); }
  for (int i = 0; i < d->nface; i++) {
    int ji = d->face_cells[2 * i], jo = d->face_cells[2 * i + 1], sp = d->face_spec[i];
    auto ok = [&](int j) { return j >= 0 && j < J && (d->spec0[j] == sp || d->nspec[j] == 2); };
    if (sp < 0 || sp > 1 || !ok(ji) || (jo >= 0 && !ok(jo))) { delete x; return xfail(BOSSS_ERR_ARG, "face item %d: cell/species not in the layout", i); }
    if (jo >= 0) {
      int s1 = d->face_slot[2 * i], s2 = d->face_slot[2 * i + 1];
      if (s1 < 1 || s1 >= S || s2 < 1 || s2 >= S || d->cols[ji * S + s1] != jo || d->cols[jo * S + s2] != ji) {
        delete x; return xfail(BOSSS_ERR_ARG, "face item %d: slots do not match the block columns", i);
      }
    }
  }
  {
    // each (cell, slot, species) pair of off-diagonal sub-blocks is written by one face item only
    std::vector<int32_t> seen((size_t)J * S * 2, -1);
    for (int i = 0; i < d->nface; i++) {
      const int ji = d->face_cells[2 * i], jo = d->face_cells[2 * i + 1];
      if (jo < 0) continue;
      int32_t& w = seen[((size_t)ji * S + d->face_slot[2 * i]) * 2 + d->face_spec[i]];
      if (w >= 0) { delete x; return xfail(BOSSS_ERR_ARG, "face items %d and %d: same face and species", w, i); }
      w = i;
    }
  }
  for (int i = 0; i < d->nifc; i++)
    if (d->ifc_cell[i] < 0 || d->ifc_cell[i] >= J) { delete x; return xfail(BOSSS_ERR_ARG, "interface item %d: bad cell", i); }
  const size_t KV = d->nvol ? (size_t)d->vol_off[d->nvol] : 0, KF = d->nface ? (size_t)d->face_off[d->nface] : 0,
               KI = d->nifc ? (size_t)d->ifc_off[d->nifc] : 0;
  rc = BOSSS_OK;
#define UP(dst, src, n) if (rc == BOSSS_OK) rc = upload(x, &x->dst, src, n)
  UP(d_nspec, d->nspec, J); UP(d_spec0, d->spec0, J); UP(d_expo, d->mode_exponents, (size_t)N * D);
  UP(d_cols, d->cols, (size_t)J * S); UP(d_boff, d->boff, (size_t)J * S);
  std::vector<int32_t> rowcell(dof[J]);
  for (int j = 0; j < J; j++)
    for (int64_t t = dof[j]; t < dof[j + 1]; t++) rowcell[t] = j;
  UP(d_nrows, nrows.data(), J); UP(d_dof, dof.data(), J + 1); UP(d_doff, doff.data(), J + 1);
  UP(d_rowcell, rowcell.data(), rowcell.size());
  // per (cell row, slot): size and offset of the column cell's x entries (0 / 0 for no column)
  std::vector<int32_t> mcol((size_t)J * S, 0);
  std::vector<int64_t> xoff((size_t)J * S, 0);
  for (size_t i = 0; i < (size_t)J * S; i++)
    if (d->cols[i] >= 0) { mcol[i] = nrows[d->cols[i]]; xoff[i] = dof[d->cols[i]]; }
  UP(d_mcol, mcol.data(), mcol.size()); UP(d_xoff, xoff.data(), xoff.size());
  UP(d_invJ, d->invJ, (size_t)J * D * D); UP(d_scal, d->scaling, J); UP(d_ls, d->lscale, (size_t)J * 2);
  UP(v_cell, d->vol_cell, d->nvol); UP(v_spec, d->vol_spec, d->nvol); UP(v_off, d->vol_off, d->nvol + 1);
  UP(v_nodes, d->vol_nodes, KV * D); UP(v_w, d->vol_w, KV);
  UP(f_cells, d->face_cells, (size_t)2 * d->nface); UP(f_slot, d->face_slot, (size_t)2 * d->nface); UP(f_spec, d->face_spec, d->nface);
  UP(f_off, d->face_off, d->nface + 1); UP(f_nin, d->face_nodes_in, KF * D); UP(f_nout, d->face_nodes_out, KF * D);
  UP(f_w, d->face_w, KF); UP(f_normal, d->face_normal, (size_t)d->nface * D);
  if (d->face_gD) { UP(f_gD, d->face_gD, KF); } else { x->f_gD = nullptr; }
  UP(i_cell, d->ifc_cell, d->nifc); UP(i_off, d->ifc_off, d->nifc + 1); UP(i_nodes, d->ifc_nodes, KI * D);
  UP(i_w, d->ifc_w, KI); UP(i_normal, d->ifc_normal, KI * D);
  {
    std::vector<int32_t> goff(J + 1, 0), gent;
    for (int i = 0; i < d->nface; i++) {
      goff[d->face_cells[2 * i] + 1]++;
      if (d->face_cells[2 * i + 1] >= 0) goff[d->face_cells[2 * i + 1] + 1]++;
    }
    for (int j = 0; j < J; j++) goff[j + 1] += goff[j];
    gent.resize(goff[J]);
    std::vector<int32_t> fill(goff.begin(), goff.end() - 1);
    for (int i = 0; i < d->nface; i++) {
      gent[fill[d->face_cells[2 * i]]++] = 2 * i;
      if (d->face_cells[2 * i + 1] >= 0) gent[fill[d->face_cells[2 * i + 1]]++] = 2 * i + 1;
    }
    UP(g_off, goff.data(), goff.size());
    UP(g_ent, gent.data(), gent.size());
    if (rc == BOSSS_OK && d->nface) rc = alloc(x, &x->f_stage, (size_t)d->nface * 2 * N * N);
    if (rc == BOSSS_OK && d->nface) rc = alloc(x, &x->f_astage, (size_t)d->nface * N);
  }
#undef UP
  if (rc != BOSSS_OK) { delete x; return rc; }
  XArgs& a = x->a;
  a.D = D; a.N = N; a.P = d->degree; a.J = J; a.S = S;
  a.nspec = x->d_nspec; a.spec0 = x->d_spec0; a.expo = x->d_expo; a.invJ = x->d_invJ; a.scaling = x->d_scal; a.lscale = x->d_ls;
  a.cols = x->d_cols; a.boff = x->d_boff; a.dof = x->d_dof;
  a.penalty_base = d->penalty_base;
  const double p = d->degree;
  a.pdeg = std::max((p + 1) * (p + D) / D, (p + 1.0) * (p + 1.0));
  a.mu[0] = d->mu[0]; a.mu[1] = d->mu[1]; a.src[0] = d->source[0]; a.src[1] = d->source[1];
  *out = x;
  return BOSSS_OK;
}

int bosss_hip_xdg_destroy(bosss_xdg* x) {
  if (x) { (void)hipSetDevice(x->device); delete x; }
  return BOSSS_OK;
}

int bosss_hip_xdg_info(bosss_xdg* x, int64_t* ndof, int64_t* nentries) {
  if (!x) return xfail(BOSSS_ERR_ARG, "NULL operator");
  if (ndof) *ndof = x->ndof;
  if (nentries) *nentries = x->nentries;
  return BOSSS_OK;
}

int bosss_hip_xdg_assemble_dev(bosss_xdg* x, double* blocks, double* affine, void* stream) {
  if (!x || !blocks || !affine) return xfail(BOSSS_ERR_ARG, "NULL argument");
  XHIP(hipSetDevice(x->device));
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(x->nentries)), dim3(256), 0, s, x->nentries, blocks, 0.0);
  hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(x->ndof)), dim3(256), 0, s, x->ndof, affine, 0.0);
  if (x->nvol)
    hipLaunchKernelGGL(k_xdg_volume, dim3(x->nvol), dim3(64), 0, s, x->a, x->v_cell, x->v_spec, x->v_off, x->v_nodes, x->v_w, blocks, affine);
  if (x->nface) {
    hipLaunchKernelGGL(k_xdg_face, dim3(x->nface), dim3(64), 0, s, x->a, x->f_cells, x->f_slot, x->f_spec, x->f_off, x->f_nin, x->f_nout,
                       x->f_w, x->f_normal, x->f_gD, blocks, x->f_stage, x->f_astage);
    hipLaunchKernelGGL(k_xdg_face_gather, dim3(x->a.J), dim3(256), 0, s, x->a, x->g_off, x->g_ent, x->f_cells, x->f_spec, x->f_stage,
                       x->f_astage, blocks, affine);
  }
  if (x->nifc)
    hipLaunchKernelGGL(k_xdg_interface, dim3(x->nifc), dim3(64), 0, s, x->a, x->i_cell, x->i_off, x->i_nodes, x->i_w, x->i_normal, blocks);
  XHIP(hipGetLastError());
  return BOSSS_OK;
}

int bosss_hip_xdg_spmv_dev(bosss_xdg* x, const double* blocks, double alpha, const double* v, double beta, double* y, void* stream) {
  if (!x || !blocks || !v || !y) return xfail(BOSSS_ERR_ARG, "NULL argument");
  XHIP(hipSetDevice(x->device));
  launch_spmv<double>(x, blocks, alpha, v, beta, y, (hipStream_t)stream);
  XHIP(hipGetLastError());
  return BOSSS_OK;
}

int bosss_hip_xdg_solve_dev(bosss_xdg* x, const double* blocks, const double* b, double* sol, double tol, int32_t maxit,
                            int32_t precision, int32_t use_graph, int32_t check_every, int32_t* iterations, int32_t* outer,
                            double* relres, void* stream) {
  if (!x || !blocks || !b || !sol || !iterations || !outer || !relres) return xfail(BOSSS_ERR_ARG, "NULL argument");
  if (precision != BOSSS_FP64 && precision != BOSSS_FP32) return xfail(BOSSS_ERR_ARG, "precision must be BOSSS_FP64 or BOSSS_FP32");
  if (check_every < 1 || check_every > 256) return xfail(BOSSS_ERR_ARG, "check_every must be 1..256");
  if (!(tol > 0)) return xfail(BOSSS_ERR_ARG, "tol must be positive");
  XHIP(hipSetDevice(x->device));
  hipStream_t s = x->stream;
  {
    // order after the caller's stream (blocks, b and sol may have been written there)
    hipEvent_t ev;
    XHIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    hipError_t e1 = hipEventRecord(ev, (hipStream_t)stream);
    hipError_t e2 = (e1 == hipSuccess) ? hipStreamWaitEvent(s, ev, 0) : e1;
    (void)hipEventDestroy(ev);
    if (e2 != hipSuccess) return xfail(BOSSS_ERR_HIP, "stream ordering: %s", hipGetErrorString(e2));
  }
  const int64_t n = x->ndof;
  if (!x->ws) {
    XCHK(alloc(x, &x->Dinv, x->ndiag)); XCHK(alloc(x, &x->scal, S_NSLOT)); XCHK(alloc(x, &x->partial, 2 * std::max<int64_t>(RED_BLOCKS, (x->a.J + 3) / 4)));
    XCHK(alloc(x, &x->info, x->a.J));
    XCHK(alloc(x, &x->r64, n)); XCHK(alloc(x, &x->z64, n)); XCHK(alloc(x, &x->p64, n)); XCHK(alloc(x, &x->q64, n));
    XCHK(alloc(x, &x->At, x->nentries));
    x->ws = true;
  }
  // column-major copy of the matrix for the solver's applies
  hipLaunchKernelGGL(k_blocks_to_cm, dim3(x->a.J * x->a.S), dim3(256), 0, s, x->a.S, x->d_nrows, x->d_cols, x->d_boff, blocks, x->At);
  if (precision == BOSSS_FP32 && !x->A32) {
    XCHK(alloc(x, &x->A32, x->nentries));
    XCHK(alloc(x, &x->b32, n)); XCHK(alloc(x, &x->r32, n)); XCHK(alloc(x, &x->z32, n)); XCHK(alloc(x, &x->p32, n)); XCHK(alloc(x, &x->q32, n)); XCHK(alloc(x, &x->x32, n));
  }
  // block-Jacobi: inverse diagonal blocks (batched uneven LU + DGETRI)
  hipLaunchKernelGGL(k_gather_diag, dim3(x->a.J), dim3(256), 0, s, x->a.J, x->a.S, x->d_nrows, x->d_boff, x->d_doff, blocks, x->Dinv);
  int nmax = 2 * x->a.N;
  int rc = bosss_hip_block_lu_dev(x->ctx, x->a.J, nmax, x->d_nrows, x->d_doff, x->Dinv, nullptr, nullptr, x->info, 1, (void*)s);
  if (rc != BOSSS_OK) return xfail(rc, "block LU: %s", bosss_hip_block_last_error());
  {
    std::vector<int32_t> info(x->a.J);
    XHIP(hipMemcpyAsync(info.data(), x->info, info.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    XHIP(hipStreamSynchronize(s));
    for (int j = 0; j < x->a.J; j++)
      if (info[j] != 0) return info[j] > 0 ? j + 1 : xfail(BOSSS_ERR_ARG, "block %d larger than %d", j, nmax);
  }
  if (precision == BOSSS_FP64) {
    hipLaunchKernelGGL(k_transpose_blocks, dim3(x->a.J), dim3(64), 0, s, x->a.J, x->d_nrows, x->d_doff, x->Dinv);
    *outer = 1;
    return pcg_solve<double>(x, 0, x->At, x->Dinv, b, sol, x->r64, x->z64, x->p64, x->q64, tol, maxit, check_every, use_graph, false,
                             iterations, relres);
  }
  // FP32 CG on the equilibrated matrix Linv M Linv^T inside FP64 iterative refinement:
  // r = b - M x (FP64), solve (Linv M Linv^T) y = Linv r / |r| in FP32, x += |r| Linv^T y.
  if (!x->Lchol) XCHK(alloc(x, &x->Lchol, x->ndiag));
  hipLaunchKernelGGL(k_gather_diag, dim3(x->a.J), dim3(256), 0, s, x->a.J, x->a.S, x->d_nrows, x->d_boff, x->d_doff, blocks, x->Lchol);
  hipLaunchKernelGGL(k_chol_inv, dim3(x->a.J), dim3(64), 0, s, x->a.J, x->d_nrows, x->d_doff, x->Lchol, x->info);
  {
    std::vector<int32_t> info(x->a.J);
    XHIP(hipMemcpyAsync(info.data(), x->info, info.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    XHIP(hipStreamSynchronize(s));
    for (int j = 0; j < x->a.J; j++)
      if (info[j] != 0) return xfail(BOSSS_ERR_ARG, "diagonal block %d is not symmetric positive definite (pivot %d)", j, info[j]);
  }
  hipLaunchKernelGGL(k_scale_blocks, dim3(x->a.J * x->a.S), dim3(64), 0, s, x->a.J, x->a.S, x->d_nrows, x->d_cols, x->d_boff, x->d_doff,
                     x->Lchol, blocks, x->A32);
  double h[S_NSLOT];
  launch_dot<double>(x, n, b, b, 0, S_BB, -1, s);
  XHIP(hipMemcpyAsync(h, x->scal, sizeof h, hipMemcpyDeviceToHost, s));
  XHIP(hipStreamSynchronize(s));
  const double bb = h[S_BB];
  int total = 0, k = 0;
  double rel = 0.0;
  const double inner_tol = 1e-4;
  const int max_outer = 40;
  // Every exit passes through the residual check below, which synchronises the stream: the last
  // correction of sol is complete and relres describes the returned sol.
  for (k = 0;; k++) {
    hipLaunchKernelGGL((k_scale_copy<double, double>), dim3(grid_for(n)), dim3(256), 0, s, n, 1.0, b, x->r64, 0);
    launch_spmv_cm<double>(x, x->At, -1.0, sol, 1.0, x->r64, s);
    launch_dot<double>(x, n, x->r64, x->r64, 0, S_RR, -1, s);
    XHIP(hipMemcpyAsync(h, x->scal, sizeof h, hipMemcpyDeviceToHost, s));
    XHIP(hipStreamSynchronize(s));
    rel = bb > 0 ? std::sqrt(h[S_RR] / bb) : 0.0;
    if (!std::isfinite(rel)) return xfail(BOSSS_ERR_NONFINITE, "refinement residual is not finite");
    if (rel <= tol || total >= maxit || k >= max_outer) break;
    const double rn = std::sqrt(h[S_RR]);
    hipLaunchKernelGGL((k_lower_apply<double, float>), dim3(x->a.J), dim3(64), 0, s, x->a.J, x->d_nrows, x->d_doff, x->d_dof, x->Lchol, 0,
                       1.0 / rn, x->r64, x->b32, 0);
    hipLaunchKernelGGL(k_fill<float>, dim3(grid_for(n)), dim3(256), 0, s, n, x->x32, 0.0f);
    int it = 0;
    double r_in = 0.0;
    rc = pcg_solve<float>(x, 1, x->A32, nullptr, x->b32, x->x32, x->r32, x->z32, x->p32, x->q32, inner_tol,
                          std::max(1, maxit - total), check_every, use_graph, true, &it, &r_in);
    if (rc != BOSSS_OK) return rc;
    total += it;
    hipLaunchKernelGGL((k_lower_apply<float, double>), dim3(x->a.J), dim3(64), 0, s, x->a.J, x->d_nrows, x->d_doff, x->d_dof, x->Lchol, 1,
                       rn, x->x32, sol, 1);
  }
  *iterations = total;
  *outer = k;
  *relres = rel;
  XHIP(hipGetLastError());
  return BOSSS_OK;
}

int bosss_hip_xdg_set_tile(bosss_xdg* x, int32_t precision, int32_t t) {
  if (!x) return xfail(BOSSS_ERR_ARG, "NULL operator");
  if (precision != BOSSS_FP64 && precision != BOSSS_FP32) return xfail(BOSSS_ERR_ARG, "precision must be BOSSS_FP64 or BOSSS_FP32");
  if (t != 0 && t != 1 && t != 2 && t != 4) return xfail(BOSSS_ERR_ARG, "tile must be 0 (tune), 1, 2 or 4 groups of four rows");
  const int prec = precision == BOSSS_FP32;
  x->tile[prec] = t;
  if (t == 0)
    for (int c = 0; c < 3; c++) x->tile_ms[prec][c] = -1.0f;
  return BOSSS_OK;
}

int bosss_hip_xdg_tiles(const bosss_xdg* x, int32_t precision, int32_t* chosen, float* ms3) {
  if (!x) return xfail(BOSSS_ERR_ARG, "NULL operator");
  if (precision != BOSSS_FP64 && precision != BOSSS_FP32) return xfail(BOSSS_ERR_ARG, "precision must be BOSSS_FP64 or BOSSS_FP32");
  const int prec = precision == BOSSS_FP32;
  if (chosen) *chosen = x->tile[prec];
  if (ms3)
    for (int c = 0; c < 3; c++) ms3[c] = x->tile_ms[prec][c];
  return BOSSS_OK;
}

// ---- host-pointer forms: host arrays copied to the device and back on the operator's stream
int bosss_hip_xdg_assemble(bosss_xdg* x, double* blocks, double* affine) {
  if (!x || !blocks || !affine) return xfail(BOSSS_ERR_ARG, "NULL argument");
  XHIP(hipSetDevice(x->device));
  DevArrays d(x->stream);
  double* dB = d.out<double>(x->nentries);
  double* da = d.out<double>(x->ndof);
  if (d.error() != hipSuccess) return xfail(BOSSS_ERR_HIP, "xdg_assemble staging: %s", hipGetErrorString(d.error()));
  if (int rc = bosss_hip_xdg_assemble_dev(x, dB, da, x->stream)) return rc;
  d.back(blocks, (const double*)dB, x->nentries);
  d.back(affine, (const double*)da, x->ndof);
  if (hipError_t e = d.finish()) return xfail(BOSSS_ERR_HIP, "xdg_assemble: %s", hipGetErrorString(e));
  return BOSSS_OK;
}

int bosss_hip_xdg_spmv(bosss_xdg* x, const double* blocks, double alpha, const double* v, double beta, double* y) {
  if (!x || !blocks || !v || !y) return xfail(BOSSS_ERR_ARG, "NULL argument");
  XHIP(hipSetDevice(x->device));
  DevArrays d(x->stream);
  const double* dB = d.in(blocks, x->nentries);
  const double* dv = d.in(v, x->ndof);
  double* dy = d.in((const double*)y, x->ndof);
  if (d.error() != hipSuccess) return xfail(BOSSS_ERR_HIP, "xdg_spmv staging: %s", hipGetErrorString(d.error()));
  if (int rc = bosss_hip_xdg_spmv_dev(x, dB, alpha, dv, beta, dy, x->stream)) return rc;
  d.back(y, (const double*)dy, x->ndof);
  if (hipError_t e = d.finish()) return xfail(BOSSS_ERR_HIP, "xdg_spmv: %s", hipGetErrorString(e));
  return BOSSS_OK;
}

int bosss_hip_xdg_solve(bosss_xdg* x, const double* blocks, const double* b, double* sol, double tol, int32_t maxit, int32_t precision,
                        int32_t* iterations, int32_t* outer, double* relres) {
  if (!x || !blocks || !b || !sol || !iterations || !outer || !relres) return xfail(BOSSS_ERR_ARG, "NULL argument");
  XHIP(hipSetDevice(x->device));
  DevArrays d(x->stream);
  const double* dB = d.in(blocks, x->nentries);
  const double* db = d.in(b, x->ndof);
  double* ds = d.in((const double*)sol, x->ndof);
  if (d.error() != hipSuccess) return xfail(BOSSS_ERR_HIP, "xdg_solve staging: %s", hipGetErrorString(d.error()));
  int rc = bosss_hip_xdg_solve_dev(x, dB, db, ds, tol, maxit, precision, 1, 16, iterations, outer, relres, x->stream);
  if (rc < 0) return rc;
  d.back(sol, (const double*)ds, x->ndof);
  if (hipError_t e = d.finish()) return xfail(BOSSS_ERR_HIP, "xdg_solve: %s", hipGetErrorString(e));
  return rc;
}

}  // extern "C"
