// Cell-centric DG operator kernels, tensor-table variant (hypercube cells, tensor Gauss rules).
//
// Same algorithm and data flow as dg_kernels.hpp (one wavefront = 16 cells = the
// 16 MFMA columns; volume + all 2D faces per cell; transposed MFMA contractions),
// with the register and LDS budget rearranged so the FP64 Euler p=3 kernels do
// not spill:
//
//  * the reference tables phi, dphi, traces are never stored: every A operand is
//    the product of D one-dimensional factors L_a(x_q) (or L'_a, or the endpoint
//    values at -1/+1) read from a ~0.5 KiB table in LDS.  The host verifies at
//    operator creation that the uploaded tables are exactly this tensor product
//    (to 1e-13), which they are for the reference's hypercube elements;
//  * the LDS that the tables no longer need holds each wavefront's own 16 cells'
//    coefficients (U, and W for the adjoint), staged with one coalesced copy and
//    then read as B operands;
//  * the neighbour's B operands are loaded just in time from global memory;
//  * the Euler JVP and VJP (kt_jvp_euler, kt_vjp_euler at the end of this file) park
//    the 4-nodes-per-lane MFMA outputs that are read more than once, or across register-
//    heavy flux code, in a per-lane LDS buffer and run the pointwise flux node by node;
//    values used once (the volume W gradients of the VJP, the volume state and tangent
//    of the JVP) are read straight from the MFMA results.  They read their own cell's
//    operands from global memory (L2) rather than LDS; the fused residual + VJP and the
//    FP32 JVP keep them in registers for the whole launch.  The VJP merges the face
//    cotangent before the pointwise adjoint (fn = -fo, so only
//    sc_own*W_own.trace - sc_nb*W_nb.trace is needed) and uses the hand-written reverse
//    sweeps of euler_adjoint.hpp;
//  * workgroups are renumbered so that each XCD (own L2) walks one contiguous range
//    of cells, for the Euler JVP and VJP kernels on large planes plane-fastest (xcd_block).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "dg_kernels.hpp"
#include "euler_adjoint.hpp"
#include "forms.hpp"

namespace bosss {

constexpr int tdeg_count(int D, int p) { return D == 2 ? (p + 1) * (p + 2) / 2 : (p + 1) * (p + 2) * (p + 3) / 6; }

// LDS layout of the 1D tables: L[P+1][K2], DL[P+1][K2] (K2 = K1 + 2: Gauss nodes, then x=-1, x=+1),
// w1[K1], then the packed mode exponents code[N] = a | b << 4 | c << 8 (as Real-sized slots).
template <int D_, int P_, int K1_>
struct TShape {
  static constexpr int D = D_, P = P_, K1 = K1_, K2 = K1 + 2;
  static constexpr int N = tdeg_count(D, P);
  static constexpr int KV = ipow(K1, D), KE = ipow(K1, D - 1), NF = 2 * D;
  static constexpr int TV = (KV + 15) / 16, TE = (KE + 15) / 16;
  static constexpr int NK = (N + 3) / 4, MT = (N + 15) / 16;
  static constexpr int O_L = 0, O_DL = (P + 1) * K2, O_W = 2 * (P + 1) * K2, O_CODE = O_W + K1;
  static constexpr int TAB = O_CODE + N;     // elements (Real) of the 1D table block
  // Face trace table FT[f][node][m] = phi_m at node `node` of face f, stored after the 1D block
  // in global memory: FROWS rows per face (whole 16-node tiles; the padding rows node >= KE hold
  // node 0's values, so a trace evaluated at them is a valid state), FS Real per row (modes
  // m >= N and the columns FZERO, FZERO + 1 are zero).  FS = 4 NK + 2 keeps the LDS reads of
  // both contractions (lanes along nodes, or along modes) free of bank conflicts or 2-way.
  static constexpr int FROWS = TE * 16, FS = NK * 4 + 2, FZERO = NK * 4;
  static constexpr int FTAB = NF * FROWS * FS;
  // 1 + D face tables in this layout follow each other: phi_m, then d phi_m / d xi_e, e < D.
  static constexpr int NFTAB = 1 + D;
};

// Per-axis 1D indices of a node: volume node kv = (i*K1 + j)*K1 + k has x = k, y = j, z = i
// (Cube.cs:167-178; 2D: kv = i*K1 + j, x = j, y = i).  Face f node kf (Square rule
// kf = i*K1 + j: s0 = node j, s1 = node i; Line rule in 2D: s0 = node kf) is mapped by
// the face coordinate maps of Cube/Square.TransformFaceCoordinates; endpoint -1 -> K1, +1 -> K1+1.
template <int D, int K1>
__device__ __forceinline__ void vol_axes(int kv, int (&q)[3]) {
  if (D == 3) { q[0] = kv % K1; q[1] = (kv / K1) % K1; q[2] = kv / (K1 * K1); }
  else { q[0] = kv % K1; q[1] = kv / K1; q[2] = 0; }
}
template <int D, int K1>
__device__ __forceinline__ void face_axes(int f, int kf, int (&q)[3]) {
  const int lo = K1, hi = K1 + 1;
  if (D == 3) {
    int s0 = kf % K1, s1 = kf / K1;
    switch (f) {
      case 0: q[0] = lo; q[1] = s1; q[2] = s0; break;
      case 1: q[0] = hi; q[1] = s1; q[2] = s0; break;
      case 2: q[0] = s0; q[1] = hi; q[2] = s1; break;
      case 3: q[0] = s0; q[1] = lo; q[2] = s1; break;
      case 4: q[0] = s0; q[1] = s1; q[2] = hi; break;
      default: q[0] = s0; q[1] = s1; q[2] = lo; break;
    }
  } else {
    int s0 = kf;
    switch (f) {
      case 0: q[0] = lo; q[1] = s0; break;
      case 1: q[0] = hi; q[1] = s0; break;
      case 2: q[0] = s0; q[1] = hi; break;
      default: q[0] = s0; q[1] = lo; break;
    }
    q[2] = 0;
  }
}

// table value: deriv = -1 -> phi_n, deriv = e -> d phi_n / d xi_e
template <class TS, typename Real>
__device__ __forceinline__ Real tval(const Real* tab, const int (&q)[3], int code, int deriv) {
  const Real* L = tab + TS::O_L;
  const Real* DL = tab + TS::O_DL;
  int a = code & 15, b = (code >> 4) & 15;
  Real x = (deriv == 0 ? DL : L)[a * TS::K2 + q[0]];
  Real y = (deriv == 1 ? DL : L)[b * TS::K2 + q[1]];
  if (TS::D == 3) {
    int c = (code >> 8) & 15;
    Real z = (deriv == 2 ? DL : L)[c * TS::K2 + q[2]];
    return x * y * z;
  }
  return x * y;
}

// Mode code of mode n, clamped to the last mode for the padding rows n >= N (callers mask
// those rows by a factor 0): every LDS read is unconditional, so no branch separates a
// load from its use and the compiler can issue the loads of a whole contraction together.
template <class TS, typename Real>
__device__ __forceinline__ int tcode(const Real* tab, int n) {
  return (int)tab[TS::O_CODE + (n < TS::N ? n : TS::N - 1)];
}
template <class TS, typename Real>
__device__ __forceinline__ Real mode_mask(int n) {
  return n < TS::N ? Real(1) : Real(0);
}

// eval: acc[v] (nodes x cells) = table(node(col), mode) . B,   B operand from `Bsrc(v, kk)`
template <class TS, typename Real, int V, class BF>
__device__ __forceinline__ void t_eval(const Real* tab, const int (&q)[3], int deriv, int grp, BF Bsrc,
                                       typename MM<Real>::acc (&acc)[V]) {
#pragma unroll
  for (int v = 0; v < V; v++) acc[v] = acc_zero<Real>();
#pragma unroll
  for (int kk = 0; kk < TS::NK; kk++) {
    int code = tcode<TS>(tab, 4 * kk + grp);
    Real a = tval<TS>(tab, q, code, deriv) * mode_mask<TS, Real>(4 * kk + grp);
#pragma unroll
    for (int v = 0; v < V; v++) acc[v] = MM<Real>::mma(a, Bsrc(v, kk), acc[v]);
  }
}

// Test-contraction accumulators of V variables.  Full 16-mode tiles use
// v_mfma_f64_16x16x4f64 (4 rows per lane); when N = 16 m + 4 in FP64 (p = 3 in 3D: N = 20)
// the last four modes use v_mfma_f64_4x4x4f64 (4 blocks of 4x4x4) instead of a 16-row tile
// that would be three quarters padding.  Its lane layout (measured on the device): lane l
// supplies A = table(node k = l/16, mode 16 m + l%4) and B = G(node l/16, cell l%16) -- the
// same B operand as the 16x16x4 tile -- and receives mode 16 m + l/16 of cell l%16.
template <class TS, typename Real, int V>
struct TAcc {
  typedef typename MM<Real>::acc acc_t;
  static constexpr bool REM = std::is_same<Real, double>::value && TS::N > 16 && (TS::N % 16) == 4;
  static constexpr int MF = REM ? TS::N / 16 : TS::MT;   // 16-row tiles
  acc_t t[V][MF];
  Real rem[REM ? V : 1];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int v = 0; v < V; v++) {
#pragma unroll
      for (int mt = 0; mt < MF; mt++) t[v][mt] = acc_zero<Real>();
      rem[REM ? v : 0] = Real(0);
    }
  }
  // f(v, n, value) for every coefficient row n < N this lane holds
  template <class F>
  __device__ __forceinline__ void store(int grp, F f) const {
#pragma unroll
    for (int v = 0; v < V; v++) {
#pragma unroll
      for (int mt = 0; mt < MF; mt++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          int n = mt * 16 + MM<Real>::row(grp, r);
          if (n < TS::N) f(v, n, t[v][mt][r]);
        }
      if (REM) f(v, MF * 16 + grp, rem[REM ? v : 0]);
    }
  }
};

template <class TS, typename Real, int V>
__device__ __forceinline__ void t_test(const Real* tab, const int (&q)[3], int deriv, int col, const Real (&g)[V],
                                       TAcc<TS, Real, V>& acc) {
  typedef TAcc<TS, Real, V> A_;
#pragma unroll
  for (int mt = 0; mt < A_::MF; mt++) {
    int code = tcode<TS>(tab, mt * 16 + col);
    Real a = tval<TS>(tab, q, code, deriv) * mode_mask<TS, Real>(mt * 16 + col);
#pragma unroll
    for (int v = 0; v < V; v++) acc.t[v][mt] = MM<Real>::mma(a, g[v], acc.t[v][mt]);
  }
  if constexpr (A_::REM) {
    const int m = A_::MF * 16 + (col & 3);
    Real a = tval<TS>(tab, q, tcode<TS>(tab, m), deriv);
#pragma unroll
    for (int v = 0; v < V; v++) acc.rem[v] = __builtin_amdgcn_mfma_f64_4x4x4f64(a, g[v], acc.rem[v], 0, 0, 0);
  }
}

// Copy the first ntab face tables (global memory, after the 1D block) into LDS at ft.
template <class TS, typename Real>
__device__ __forceinline__ void ft_load(Real* ft, const Real* gtab, int ntab = 1) {
  for (int i = threadIdx.x; i < ntab * TS::FTAB; i += blockDim.x) ft[i] = gtab[TS::TAB + i];
}

// Trace contraction from the face trace table: acc[v] (nodes x cells) = FT(face, node, mode) . B.
// `row` points at FT(face, t * 16 + col, grp) of the lane; its A operands are row[4 kk], LDS reads
// at immediate offsets (the padding modes are zero in the table).
template <class TS, typename Real, int V, class BF>
__device__ __forceinline__ void t_eval_ft(const Real* row, BF Bsrc, typename MM<Real>::acc (&acc)[V]) {
#pragma unroll
  for (int v = 0; v < V; v++) acc[v] = acc_zero<Real>();
#pragma unroll
  for (int kk = 0; kk < TS::NK; kk++) {
    const Real a = row[4 * kk];
#pragma unroll
    for (int v = 0; v < V; v++) acc[v] = MM<Real>::mma(a, Bsrc(v, kk), acc[v]);
  }
}

// The lane's mode columns of the face trace table in the test contraction: mode mt * 16 + col of
// each 16-mode tile (padding modes >= N clamped to the zero column), and of the 4-mode remainder.
template <class TS, typename Real, int V>
struct FtCols {
  typedef TAcc<TS, Real, V> A_;
  int m[A_::MF];
  int rem;
  __device__ __forceinline__ explicit FtCols(int col) {
#pragma unroll
    for (int mt = 0; mt < A_::MF; mt++) m[mt] = mt * 16 + col < TS::N ? mt * 16 + col : TS::FZERO;
    rem = A_::MF * 16 + (col & 3);
  }
};

// Test contraction from the face trace table: `nrow` points at FT(face, node, 0) of the lane's
// node; same accumulators and MFMA shapes as t_test.
template <class TS, typename Real, int V>
__device__ __forceinline__ void t_test_ft(const Real* nrow, const FtCols<TS, Real, V>& mc, const Real (&g)[V],
                                          TAcc<TS, Real, V>& acc) {
  typedef TAcc<TS, Real, V> A_;
#pragma unroll
  for (int mt = 0; mt < A_::MF; mt++) {
    const Real a = nrow[mc.m[mt]];
#pragma unroll
    for (int v = 0; v < V; v++) acc.t[v][mt] = MM<Real>::mma(a, g[v], acc.t[v][mt]);
  }
  if constexpr (A_::REM) {
    const Real a = nrow[mc.rem];
#pragma unroll
    for (int v = 0; v < V; v++) acc.rem[v] = __builtin_amdgcn_mfma_f64_4x4x4f64(a, g[v], acc.rem[v], 0, 0, 0);
  }
}

template <class TS, typename Real>
__device__ __forceinline__ Real vol_weight(const Real* tab, const int (&q)[3]) {
  const Real* w = tab + TS::O_W;
  if (TS::D == 3) return (w[q[2]] * w[q[1]]) * w[q[0]];
  return w[q[1]] * w[q[0]];
}
template <class TS, typename Real>
__device__ __forceinline__ Real face_weight(const Real* tab, int kf) {
  const Real* w = tab + TS::O_W;
  if (TS::D == 3) return w[kf / TS::K1] * w[kf % TS::K1];
  return w[kf];
}

// Stage the wave's 16 cells of one array into LDS: dst[c * stride + i], i < V*N (coalesced copy).
template <typename Real, int NT>
__device__ __forceinline__ void stage16(Real* dst, int stride, const Real* __restrict__ src, int cell0, int J, int lane) {
  for (int i = lane; i < 16 * NT; i += 64) {
    int c = i / NT, e = i - c * NT;
    dst[c * stride + e] = (cell0 + c < J) ? src[(size_t)cell0 * NT + i] : Real(0);
  }
}

template <typename Real>
struct TArgs {
  int J;
  const double* detJ;
  const double* invJ;
  const double* scaling;
  const double* lenscale;
  CellFaces cf;
  const Real* tab;      // 1D tables (TShape layout) in global memory
  FormParams fp;
  double alpha, beta;
  const Real* U;
  const Real* X;        // tangent (JVP) or cotangent W (VJP)
  Real* R;
  Real* out;
  int nxcd;             // XCDs the workgroups are dealt over (1: no remapping)
  int plane_blocks;     // workgroups per plane of the slowest grid direction (0: range order)
  int own_only;         // JVP: neighbour tangents taken as zero (J(U)_jj X_j, the diagonal blocks)
};

// Workgroup -> cell-block mapping.  The dispatcher deals consecutive workgroups
// round-robin over the XCDs, each with its own L2; renumbering the blocks so that
// XCD x walks one contiguous range of cells keeps a cell's face neighbours (the
// adjacent cell blocks) in the same L2.  Bijective for any grid size.
//
// With plane_blocks = W > 0 and a range of whole planes (per = g W blocks), XCD x walks its
// g planes plane-fastest: its i-th workgroup takes block (i mod g) W + i / g of the range, so
// the blocks across the slowest direction (one plane apart, the x-neighbours of a Cartesian
// grid numbered x-slowest) run next to each other instead of W workgroups apart, and a cell's
// coefficients are still in L2 when its neighbour in the next plane reads them.  A bijection
// of the range onto itself: mixed-radix digits (i mod g, i / g) of i in [0, g W).
__device__ __forceinline__ int xcd_block(int nxcd, int plane_blocks) {
  const int b = blockIdx.x, nb = gridDim.x;
  if (nxcd <= 1) return b;
  const int per = nb / nxcd, rem = nb % nxcd, x = b % nxcd, i = b / nxcd;
  if (plane_blocks > 0 && rem == 0 && per >= plane_blocks && per % plane_blocks == 0) {
    const int g = per / plane_blocks;
    return x * per + (i % g) * plane_blocks + i / g;
  }
  return x < rem ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
}

template <class Form> struct HandAdjoint : std::false_type {};
template <int D, bool H> struct HandAdjoint<FormEuler<D, H>> : std::true_type {};
template <class Form> struct IsHLLC : std::false_type {};
template <int D> struct IsHLLC<FormEuler<D, true>> : std::true_type {};

constexpr int TWAVES = 4;

// Face tables kt_forward and kt_adjoint load: the trace table, and for forms whose faces use
// gradients (SIP, advection-diffusion) the D reference-derivative tables as well.
template <class Form, int D>
constexpr int form_face_tables() { return (Form::FACE_GRADU || Form::FACE_TESTG) ? 1 + D : 1; }

// Waves per SIMD the residual/JVP (kt_forward) and linear-form adjoint (kt_adjoint) kernels
// are compiled for.  Their LDS (1D tables, staged coefficients, face tables; at most about 90 KB
// per workgroup, for the FP64 p = 3 scalar forms at one wave) fits every occupancy below, so
// registers set it: FP32, and FP64 up to p = 2, fit two waves with little or no spilling and gain 13-28 % from the
// overlap.  The FP64 p = 3 kernels, and the FP64 fused residual + adjoint at p = 2, would
// spill 260-300 B per lane and run slower (measured), so they stay at one wave.
template <typename Real, int P>
constexpr int tensor_min_waves() { return (sizeof(Real) == 4 || P < 3) ? 2 : 1; }
template <typename Real, int P, bool FUSED>
constexpr int adjoint_min_waves() { return (sizeof(Real) == 4 || (P < 3 && !FUSED)) ? 2 : 1; }

// ----------------------------------------------------------------------------
// Forward: residual (MODE_RES) or JVP (MODE_JVP).
// ----------------------------------------------------------------------------
template <class Form, int D, int P, int K1, typename Real, int MODE>
__global__ void __launch_bounds__(64 * TWAVES, (tensor_min_waves<Real, P>())) kt_forward(TArgs<Real> a) {
  typedef TShape<D, P, K1> TS;
  typedef typename MM<Real>::acc acc_t;
  constexpr int V = Form::V, N = TS::N, NT = V * N, NF = TS::NF, MT = TS::MT;
  constexpr bool TAN = (MODE == MODE_JVP);
  constexpr int NARR = TAN ? 2 : 1;
  constexpr int STRIDE = NT | 1;
  typedef typename std::conditional<TAN, Dual<1, Real>, Real>::type T;

  extern __shared__ __align__(16) unsigned char smem_raw[];
  Real* tab = reinterpret_cast<Real*>(smem_raw);
  for (int i = threadIdx.x; i < TS::TAB; i += blockDim.x) tab[i] = a.tab[i];
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int col = lane & 15, grp = lane >> 4;
  const int cell0 = (xcd_block(a.nxcd, a.plane_blocks) * TWAVES + wave) * 16;
  Real* stU = tab + ((TS::TAB + 1) & ~1) + wave * (NARR * 16 * STRIDE);
  Real* stX = stU + 16 * STRIDE;
  // face tables: traces, and the reference-derivative traces where the faces use gradients
  Real* ft = tab + ((TS::TAB + 1) & ~1) + TWAVES * NARR * 16 * STRIDE;
  ft_load<TS>(ft, a.tab, form_face_tables<Form, D>());
  stage16<Real, NT>(stU, STRIDE, a.U, cell0, a.J, lane);
  if (TAN) stage16<Real, NT>(stX, STRIDE, a.X, cell0, a.J, lane);
  __syncthreads();
  const FtCols<TS, Real, V> ftc(col);

  const int cell = cell0 + col;
  const bool valid = cell < a.J;
  const int cj = valid ? cell : 0;
  const Real sc = Real(a.scaling[cj]);
  const Real detJ = Real(a.detJ[cj]);
  Real iJ[D][D];
#pragma unroll
  for (int e = 0; e < D; e++)
#pragma unroll
    for (int d = 0; d < D; d++) iJ[e][d] = Real(a.invJ[(size_t)cj * D * D + e * D + d]);

  TAcc<TS, Real, V> racc;
  racc.zero();

  const Real* ownU = stU + col * STRIDE;
  const Real* ownX = stX + col * STRIDE;
  auto Bu = [&](int v, int kk) { int m = 4 * kk + grp; return ownU[v * N + (m < N ? m : N - 1)] * ((m < N) ? sc : Real(0)); };
  auto Bx = [&](int v, int kk) { int m = 4 * kk + grp; return ownX[v * N + (m < N ? m : N - 1)] * ((m < N) ? sc : Real(0)); };

  // ------------------------------------------------------------------ volume
  for (int t = 0; t < TS::TV; t++) {
    int qe[3];
    const int nodeE = t * 16 + col;
    vol_axes<D, K1>(nodeE < TS::KV ? nodeE : 0, qe);
    acc_t u[V], du[V], gu[D][V], dgu[D][V];
    t_eval<TS, Real, V>(tab, qe, -1, grp, Bu, u);
    if (TAN) t_eval<TS, Real, V>(tab, qe, -1, grp, Bx, du);
    if (Form::VOL_GRADU) {
#pragma unroll
      for (int e = 0; e < D; e++) {
        t_eval<TS, Real, V>(tab, qe, e, grp, Bu, gu[e]);
        if (TAN) t_eval<TS, Real, V>(tab, qe, e, grp, Bx, dgu[e]);
      }
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      if (t * 16 + MM<Real>::row(0, r) >= TS::KV) continue;   // all 4 nodes of this step are padding
      const int node = t * 16 + MM<Real>::row(grp, r);
      const bool pad = node >= TS::KV;
      int q[3];
      vol_axes<D, K1>(pad ? 0 : node, q);
      const Real wq = pad ? Real(0) : vol_weight<TS>(tab, q) * detJ * sc;
      T Uq[V], GU[V][D], F[V][D];
#pragma unroll
      for (int v = 0; v < V; v++) {
        if constexpr (TAN) { Uq[v].v = u[v][r]; Uq[v].d[0] = du[v][r]; } else { Uq[v] = u[v][r]; }
#pragma unroll
        for (int d = 0; d < D; d++) {
          if (Form::VOL_GRADU) {
            Real gv = 0, gt = 0;
#pragma unroll
            for (int e = 0; e < D; e++) { gv += gu[e][v][r] * iJ[e][d]; if (TAN) gt += dgu[e][v][r] * iJ[e][d]; }
            if constexpr (TAN) { GU[v][d].v = gv; GU[v][d].d[0] = gt; } else { GU[v][d] = gv; }
          } else {
            GU[v][d] = Uq[v];
          }
        }
      }
      Form::template volume<D, T, Real>(a.fp, Uq, GU, F);
#pragma unroll
      for (int e = 0; e < D; e++) {
        Real g[V];
#pragma unroll
        for (int v = 0; v < V; v++) {
          Real s = 0;
#pragma unroll
          for (int d = 0; d < D; d++) {
            if constexpr (TAN) s += F[v][d].d[0] * iJ[e][d]; else s += F[v][d] * iJ[e][d];
          }
          g[v] = pad ? Real(0) : s * wq;
        }
        t_test<TS, Real, V>(tab, q, e, col, g, racc);
      }
    }
  }

  // ------------------------------------------------------------------- faces
  const Real pbd = Real(a.fp.penalty_base * a.fp.penalty_deg);
  const Real cj_own = Real(1.0 / a.lenscale[cj]);
  for (int f = 0; f < NF; f++) {
    const int nb_raw = a.cf.nb[(size_t)cj * NF + f];
    const int nb = valid ? nb_raw : -1;
    const bool has_nb = nb >= 0;
    const int cn = has_nb ? nb : cj;
    const int fo = a.cf.pair[f];
    const Real sg = Real(a.cf.sqrtg[(size_t)cj * NF + f]) * (valid ? Real(1) : Real(0));
    Real nrm[D];
#pragma unroll
    for (int d = 0; d < D; d++) nrm[d] = Real(a.cf.normal[((size_t)cj * NF + f) * D + d]);
    Real mu = pbd * cj_own;
    if (has_nb) { Real mo = pbd * Real(1.0 / a.lenscale[cn]); mu = (mo > mu) ? mo : mu; }
    const Real scn = Real(a.scaling[cn]);
    Real iJn[D][D];
#pragma unroll
    for (int e = 0; e < D; e++)
#pragma unroll
      for (int d = 0; d < D; d++) iJn[e][d] = Form::FACE_GRADU ? Real(a.invJ[(size_t)cn * D * D + e * D + d]) : Real(0);
    const Real* nbU = a.U + (size_t)cn * NT;
    const Real* nbX = TAN ? a.X + (size_t)cn * NT : nullptr;
    auto Bnu = [&](int v, int kk) { int m = 4 * kk + grp; return nbU[v * N + (m < N ? m : N - 1)] * ((has_nb && m < N) ? scn : Real(0)); };
    auto Bnx = [&](int v, int kk) { int m = 4 * kk + grp; return nbX[v * N + (m < N ? m : N - 1)] * ((has_nb && !a.own_only && m < N) ? scn : Real(0)); };

    for (int t = 0; t < TS::TE; t++) {
      // trace operands FT(f, node, mode) and FT(fo, node, mode) of the tile's nodes; the
      // derivative table of direction e follows at (1 + e) FTAB
      const Real* fto = ft + (f * TS::FROWS + t * 16 + col) * TS::FS + grp;
      const Real* ftn = ft + (fo * TS::FROWS + t * 16 + col) * TS::FS + grp;
      acc_t uo[V], un[V], duo[V], dun[V], guo[D][V], gun[D][V], dguo[D][V], dgun[D][V];
      t_eval_ft<TS, Real, V>(fto, Bu, uo);
      t_eval_ft<TS, Real, V>(ftn, Bnu, un);
      if (TAN) {
        t_eval_ft<TS, Real, V>(fto, Bx, duo);
        t_eval_ft<TS, Real, V>(ftn, Bnx, dun);
      }
      if (Form::FACE_GRADU) {
#pragma unroll
        for (int e = 0; e < D; e++) {
          t_eval_ft<TS, Real, V>(fto + (1 + e) * TS::FTAB, Bu, guo[e]);
          t_eval_ft<TS, Real, V>(ftn + (1 + e) * TS::FTAB, Bnu, gun[e]);
          if (TAN) {
            t_eval_ft<TS, Real, V>(fto + (1 + e) * TS::FTAB, Bx, dguo[e]);
            t_eval_ft<TS, Real, V>(ftn + (1 + e) * TS::FTAB, Bnx, dgun[e]);
          }
        }
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        if (t * 16 + MM<Real>::row(0, r) >= TS::KE) continue;   // all 4 nodes of this step are padding
        const int node = t * 16 + MM<Real>::row(grp, r);
        const bool pad = node >= TS::KE;
        const Real* nrow = ft + (f * TS::FROWS + node) * TS::FS;
        const Real wq = pad ? Real(0) : face_weight<TS>(tab, node) * sg * sc;
        T Uo[V], Un[V], Go[V][D], Gn[V][D];
#pragma unroll
        for (int v = 0; v < V; v++) {
          if constexpr (TAN) {
            Uo[v].v = uo[v][r]; Uo[v].d[0] = duo[v][r];
            Un[v].v = un[v][r]; Un[v].d[0] = dun[v][r];
          } else { Uo[v] = uo[v][r]; Un[v] = un[v][r]; }
          if (!has_nb) Un[v] = Uo[v];
#pragma unroll
          for (int d = 0; d < D; d++) {
            if (Form::FACE_GRADU) {
              Real go_ = 0, gn_ = 0, tgo = 0, tgn = 0;
#pragma unroll
              for (int e = 0; e < D; e++) {
                go_ += guo[e][v][r] * iJ[e][d];
                gn_ += gun[e][v][r] * iJn[e][d];
                if (TAN) { tgo += dguo[e][v][r] * iJ[e][d]; tgn += dgun[e][v][r] * iJn[e][d]; }
              }
              if constexpr (TAN) { Go[v][d].v = go_; Go[v][d].d[0] = tgo; Gn[v][d].v = gn_; Gn[v][d].d[0] = tgn; }
              else { Go[v][d] = go_; Gn[v][d] = gn_; }
              if (!has_nb) Gn[v][d] = Go[v][d];
            } else {
              Go[v][d] = Uo[v]; Gn[v][d] = Uo[v];
            }
          }
        }
        T fo_[V], go_[V][D], fn_[V], gn_[V][D];
        if (has_nb) {
          Form::template inner<D, T, Real>(a.fp, nrm, mu, Uo, Un, Go, Gn, fo_, go_, fn_, gn_);
        } else {
#pragma unroll
          for (int v = 0; v < V; v++) { fo_[v] = Uo[v]; for (int d = 0; d < D; d++) go_[v][d] = Uo[v]; }
          Form::template boundary<D, T, Real>(a.fp, nrm, mu, Uo, Go, fo_, go_);
        }
        Real g[V];
#pragma unroll
        for (int v = 0; v < V; v++) {
          Real s;
          if constexpr (TAN) s = fo_[v].d[0]; else s = fo_[v];
          g[v] = pad ? Real(0) : s * wq;
        }
        t_test_ft<TS, Real, V>(nrow, ftc, g, racc);
        if (Form::FACE_TESTG) {
#pragma unroll
          for (int e = 0; e < D; e++) {
#pragma unroll
            for (int v = 0; v < V; v++) {
              Real s = 0;
#pragma unroll
              for (int d = 0; d < D; d++) {
                if constexpr (TAN) s += go_[v][d].d[0] * iJ[e][d]; else s += go_[v][d] * iJ[e][d];
              }
              g[v] = pad ? Real(0) : s * wq;
            }
            t_test_ft<TS, Real, V>(nrow + (1 + e) * TS::FTAB, ftc, g, racc);
          }
        }
      }
    }
  }

  if (valid) {
    Real* Op = (TAN ? a.out : a.R) + (size_t)cell * NT;
    const Real al = Real(a.alpha), be = Real(a.beta);
    racc.store(grp, [&](int v, int n, Real val) {
      if (TAN) {
        Op[v * N + n] = val;
      } else {
        Real old = (be != Real(0)) ? be * Op[v * N + n] : Real(0);
        Op[v * N + n] = old + al * val;
      }
    });
  }
}

// ----------------------------------------------------------------------------
// Adjoint: out <- J(U)^T W  (FUSED: also R <- beta R + alpha Op(U)).
// ----------------------------------------------------------------------------
template <class Form, int D, int P, int K1, typename Real, bool FUSED>
__global__ void __launch_bounds__(64 * TWAVES, (adjoint_min_waves<Real, P, FUSED>())) kt_adjoint(TArgs<Real> a) {
  typedef TShape<D, P, K1> TS;
  typedef typename MM<Real>::acc acc_t;
  constexpr int V = Form::V, N = TS::N, NT = V * N, NF = TS::NF, MT = TS::MT;
  constexpr int STRIDE = NT | 1;
  static_assert(!HandAdjoint<Form>::value, "the Euler forms use kt_vjp_euler");
  constexpr bool GU_IN = Form::VOL_GRADU || Form::FACE_GRADU;
  constexpr int M = GU_IN ? V * (1 + D) : V;
  typedef Dual<M, Real> TD;

  extern __shared__ __align__(16) unsigned char smem_raw[];
  Real* tab = reinterpret_cast<Real*>(smem_raw);
  for (int i = threadIdx.x; i < TS::TAB; i += blockDim.x) tab[i] = a.tab[i];
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int col = lane & 15, grp = lane >> 4;
  const int cell0 = (xcd_block(a.nxcd, a.plane_blocks) * TWAVES + wave) * 16;
  Real* stU = tab + ((TS::TAB + 1) & ~1) + wave * (2 * 16 * STRIDE);
  Real* stW = stU + 16 * STRIDE;
  Real* ft = tab + ((TS::TAB + 1) & ~1) + TWAVES * 2 * 16 * STRIDE;   // face tables
  ft_load<TS>(ft, a.tab, form_face_tables<Form, D>());
  stage16<Real, NT>(stU, STRIDE, a.U, cell0, a.J, lane);
  stage16<Real, NT>(stW, STRIDE, a.X, cell0, a.J, lane);
  __syncthreads();
  const FtCols<TS, Real, V> ftc(col);

  const int cell = cell0 + col;
  const bool valid = cell < a.J;
  const int cj = valid ? cell : 0;
  const Real sc = Real(a.scaling[cj]);
  const Real detJ = Real(a.detJ[cj]);
  Real iJ[D][D];
#pragma unroll
  for (int e = 0; e < D; e++)
#pragma unroll
    for (int d = 0; d < D; d++) iJ[e][d] = Real(a.invJ[(size_t)cj * D * D + e * D + d]);

  TAcc<TS, Real, V> racc, aacc;
  racc.zero();
  aacc.zero();

  const Real* ownU = stU + col * STRIDE;
  const Real* ownW = stW + col * STRIDE;
  auto Bu = [&](int v, int kk) { int m = 4 * kk + grp; return ownU[v * N + (m < N ? m : N - 1)] * ((m < N) ? sc : Real(0)); };
  auto Bw = [&](int v, int kk) { int m = 4 * kk + grp; return ownW[v * N + (m < N ? m : N - 1)] * ((m < N) ? Real(1) : Real(0)); };

  // ------------------------------------------------------------------ volume
  for (int t = 0; t < TS::TV; t++) {
    int qe[3];
    const int nodeE = t * 16 + col;
    vol_axes<D, K1>(nodeE < TS::KV ? nodeE : 0, qe);
    acc_t u[V], gu[D][V], wg[D][V];
    t_eval<TS, Real, V>(tab, qe, -1, grp, Bu, u);
#pragma unroll
    for (int e = 0; e < D; e++) {
      if (Form::VOL_GRADU) t_eval<TS, Real, V>(tab, qe, e, grp, Bu, gu[e]);
      t_eval<TS, Real, V>(tab, qe, e, grp, Bw, wg[e]);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      if (t * 16 + MM<Real>::row(0, r) >= TS::KV) continue;   // all 4 nodes of this step are padding
      const int node = t * 16 + MM<Real>::row(grp, r);
      const bool pad = node >= TS::KV;
      int q[3];
      vol_axes<D, K1>(pad ? 0 : node, q);
      const Real wq = pad ? Real(0) : vol_weight<TS>(tab, q) * detJ * sc;
      // cotangent of the physical volume flux: Fbar[v][d] = wq * sum_e invJ[e][d] (grad_xi W)_e
      Real Fb[V][D];
#pragma unroll
      for (int v = 0; v < V; v++)
#pragma unroll
        for (int d = 0; d < D; d++) {
          Real s = 0;
#pragma unroll
          for (int e = 0; e < D; e++) s += iJ[e][d] * wg[e][v][r];
          Fb[v][d] = s * wq;
        }
      Real ub[V], gb[V][D];
      Real Fv[V][D];
      {
        TD Uq[V], GU[V][D], F[V][D];
#pragma unroll
        for (int v = 0; v < V; v++) {
          Uq[v] = dconst<M, Real>(u[v][r]);
          Uq[v].d[v] = Real(1);
#pragma unroll
          for (int d = 0; d < D; d++) {
            if (Form::VOL_GRADU) {
              Real gv = 0;
#pragma unroll
              for (int e = 0; e < D; e++) gv += gu[e][v][r] * iJ[e][d];
              GU[v][d] = dconst<M, Real>(gv);
              if (GU_IN) GU[v][d].d[(V + v * D + d) % M] = Real(1);
            } else {
              GU[v][d] = Uq[v];
            }
          }
        }
        Form::template volume<D, TD, Real>(a.fp, Uq, GU, F);
#pragma unroll
        for (int qq = 0; qq < V; qq++) { ub[qq] = 0; for (int dd = 0; dd < D; dd++) gb[qq][dd] = 0; }
#pragma unroll
        for (int v = 0; v < V; v++)
#pragma unroll
          for (int d = 0; d < D; d++) {
            Fv[v][d] = F[v][d].v;
#pragma unroll
            for (int qq = 0; qq < V; qq++) {
              ub[qq] += Fb[v][d] * F[v][d].d[qq];
              if (Form::VOL_GRADU) {
#pragma unroll
                for (int dd = 0; dd < D; dd++) gb[qq][dd] += Fb[v][d] * F[v][d].d[(V + qq * D + dd) % M];
              }
            }
          }
      }
      if (FUSED) {
#pragma unroll
        for (int e = 0; e < D; e++) {
          Real g[V];
#pragma unroll
          for (int v = 0; v < V; v++) {
            Real s = 0;
#pragma unroll
            for (int d = 0; d < D; d++) s += Fv[v][d] * iJ[e][d];
            g[v] = pad ? Real(0) : s * wq;
          }
          t_test<TS, Real, V>(tab, q, e, col, g, racc);
        }
      }
      Real bv[V];
#pragma unroll
      for (int v = 0; v < V; v++) bv[v] = pad ? Real(0) : ub[v] * sc;
      t_test<TS, Real, V>(tab, q, -1, col, bv, aacc);
      if (Form::VOL_GRADU) {
#pragma unroll
        for (int e = 0; e < D; e++) {
#pragma unroll
          for (int v = 0; v < V; v++) {
            Real s = 0;
#pragma unroll
            for (int d = 0; d < D; d++) s += gb[v][d] * iJ[e][d];
            bv[v] = pad ? Real(0) : s * sc;
          }
          t_test<TS, Real, V>(tab, q, e, col, bv, aacc);
        }
      }
    }
  }

  // ------------------------------------------------------------------- faces
  const Real pbd = Real(a.fp.penalty_base * a.fp.penalty_deg);
  const Real cj_own = Real(1.0 / a.lenscale[cj]);
  for (int f = 0; f < NF; f++) {
    const int nb_raw = a.cf.nb[(size_t)cj * NF + f];
    const int nb = valid ? nb_raw : -1;
    const bool has_nb = nb >= 0;
    const int cn = has_nb ? nb : cj;
    const int fo = a.cf.pair[f];
    const Real sg = Real(a.cf.sqrtg[(size_t)cj * NF + f]) * (valid ? Real(1) : Real(0));
    Real nrm[D];
#pragma unroll
    for (int d = 0; d < D; d++) nrm[d] = Real(a.cf.normal[((size_t)cj * NF + f) * D + d]);
    Real mu = pbd * cj_own;
    if (has_nb) { Real mo = pbd * Real(1.0 / a.lenscale[cn]); mu = (mo > mu) ? mo : mu; }
    const Real scn = Real(a.scaling[cn]);
    Real iJn[D][D];
#pragma unroll
    for (int e = 0; e < D; e++)
#pragma unroll
      for (int d = 0; d < D; d++) iJn[e][d] = Real(a.invJ[(size_t)cn * D * D + e * D + d]);
    const Real* nbU = a.U + (size_t)cn * NT;
    const Real* nbW = a.X + (size_t)cn * NT;
    auto Bnu = [&](int v, int kk) { int m = 4 * kk + grp; return nbU[v * N + (m < N ? m : N - 1)] * ((has_nb && m < N) ? scn : Real(0)); };
    auto Bnw = [&](int v, int kk) { int m = 4 * kk + grp; return nbW[v * N + (m < N ? m : N - 1)] * ((has_nb && m < N) ? Real(1) : Real(0)); };

    for (int t = 0; t < TS::TE; t++) {
      // trace operands from the face tables (derivative table of direction e at (1 + e) FTAB)
      const Real* fto = ft + (f * TS::FROWS + t * 16 + col) * TS::FS + grp;
      const Real* ftn = ft + (fo * TS::FROWS + t * 16 + col) * TS::FS + grp;
      acc_t uo[V], un[V];
      t_eval_ft<TS, Real, V>(fto, Bu, uo);
      t_eval_ft<TS, Real, V>(ftn, Bnu, un);
      acc_t guo[D][V], gun[D][V], wo[V], wn[V], gwo[D][V], gwn[D][V];
      t_eval_ft<TS, Real, V>(fto, Bw, wo);
      t_eval_ft<TS, Real, V>(ftn, Bnw, wn);
#pragma unroll
      for (int e = 0; e < D; e++) {
        if (Form::FACE_GRADU) {
          t_eval_ft<TS, Real, V>(fto + (1 + e) * TS::FTAB, Bu, guo[e]);
          t_eval_ft<TS, Real, V>(ftn + (1 + e) * TS::FTAB, Bnu, gun[e]);
        }
        if (Form::FACE_TESTG) {
          t_eval_ft<TS, Real, V>(fto + (1 + e) * TS::FTAB, Bw, gwo[e]);
          t_eval_ft<TS, Real, V>(ftn + (1 + e) * TS::FTAB, Bnw, gwn[e]);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        if (t * 16 + MM<Real>::row(0, r) >= TS::KE) continue;   // all 4 nodes of this step are padding
        const int node = t * 16 + MM<Real>::row(grp, r);
        const bool pad = node >= TS::KE;
        const Real* nrow = ft + (f * TS::FROWS + node) * TS::FS;
        const Real wsg = pad ? Real(0) : face_weight<TS>(tab, node) * sg;
        TD Uo[V], Un[V], Go[V][D], Gn[V][D];
#pragma unroll
        for (int v = 0; v < V; v++) {
          Uo[v] = dconst<M, Real>(uo[v][r]);
          Uo[v].d[v] = Real(1);
          Un[v] = has_nb ? dconst<M, Real>(un[v][r]) : Uo[v];
#pragma unroll
          for (int d = 0; d < D; d++) {
            if (Form::FACE_GRADU) {
              Real go_ = 0, gn_ = 0;
#pragma unroll
              for (int e = 0; e < D; e++) { go_ += guo[e][v][r] * iJ[e][d]; gn_ += gun[e][v][r] * iJn[e][d]; }
              Go[v][d] = dconst<M, Real>(go_);
              if (GU_IN) Go[v][d].d[(V + v * D + d) % M] = Real(1);
              Gn[v][d] = has_nb ? dconst<M, Real>(gn_) : Go[v][d];
            } else {
              Go[v][d] = Uo[v]; Gn[v][d] = Un[v];
            }
          }
        }
        TD fo_[V], go_[V][D], fn_[V], gn_[V][D];
        if (has_nb) {
          Form::template inner<D, TD, Real>(a.fp, nrm, mu, Uo, Un, Go, Gn, fo_, go_, fn_, gn_);
        } else {
#pragma unroll
          for (int v = 0; v < V; v++) {
            fo_[v] = Uo[v]; fn_[v] = dconst<M, Real>(Real(0));
            for (int d = 0; d < D; d++) { go_[v][d] = Uo[v]; gn_[v][d] = dconst<M, Real>(Real(0)); }
          }
          Form::template boundary<D, TD, Real>(a.fp, nrm, mu, Uo, Go, fo_, go_);
        }
        Real ub[V], gb[V][D];
#pragma unroll
        for (int qq = 0; qq < V; qq++) { ub[qq] = 0; for (int d = 0; d < D; d++) gb[qq][d] = 0; }
#pragma unroll
        for (int v = 0; v < V; v++) {
          Real fob = wsg * sc * wo[v][r];
          Real fnb = has_nb ? wsg * scn * wn[v][r] : Real(0);
#pragma unroll
          for (int qq = 0; qq < V; qq++) {
            ub[qq] += fob * fo_[v].d[qq] + fnb * fn_[v].d[qq];
            if (GU_IN && Form::FACE_GRADU) {
#pragma unroll
              for (int dd = 0; dd < D; dd++)
                gb[qq][dd] += fob * fo_[v].d[(V + qq * D + dd) % M] + fnb * fn_[v].d[(V + qq * D + dd) % M];
            }
          }
          if (Form::FACE_TESTG) {
#pragma unroll
            for (int d = 0; d < D; d++) {
              Real gob = 0, gnb = 0;
#pragma unroll
              for (int e = 0; e < D; e++) { gob += iJ[e][d] * gwo[e][v][r]; gnb += iJn[e][d] * gwn[e][v][r]; }
              gob *= wsg * sc;
              gnb = has_nb ? gnb * wsg * scn : Real(0);
#pragma unroll
              for (int qq = 0; qq < V; qq++) {
                ub[qq] += gob * go_[v][d].d[qq] + gnb * gn_[v][d].d[qq];
                if (GU_IN && Form::FACE_GRADU) {
#pragma unroll
                  for (int dd = 0; dd < D; dd++)
                    gb[qq][dd] += gob * go_[v][d].d[(V + qq * D + dd) % M] + gnb * gn_[v][d].d[(V + qq * D + dd) % M];
                }
              }
            }
          }
        }
        Real g[V];
        if (FUSED) {
          const Real wq = wsg * sc;
#pragma unroll
          for (int v = 0; v < V; v++) g[v] = pad ? Real(0) : fo_[v].v * wq;
          t_test_ft<TS, Real, V>(nrow, ftc, g, racc);
          if (Form::FACE_TESTG) {
#pragma unroll
            for (int e = 0; e < D; e++) {
#pragma unroll
              for (int v = 0; v < V; v++) {
                Real s = 0;
#pragma unroll
                for (int d = 0; d < D; d++) s += go_[v][d].v * iJ[e][d];
                g[v] = pad ? Real(0) : s * wq;
              }
              t_test_ft<TS, Real, V>(nrow + (1 + e) * TS::FTAB, ftc, g, racc);
            }
          }
        }
#pragma unroll
        for (int v = 0; v < V; v++) g[v] = pad ? Real(0) : ub[v] * sc;
        t_test_ft<TS, Real, V>(nrow, ftc, g, aacc);
        if (Form::FACE_GRADU) {
#pragma unroll
          for (int e = 0; e < D; e++) {
#pragma unroll
            for (int v = 0; v < V; v++) {
              Real s = 0;
#pragma unroll
              for (int d = 0; d < D; d++) s += gb[v][d] * iJ[e][d];
              g[v] = pad ? Real(0) : s * sc;
            }
            t_test_ft<TS, Real, V>(nrow + (1 + e) * TS::FTAB, ftc, g, aacc);
          }
        }
      }
    }
  }

  if (valid) {
    Real* Rp = FUSED ? a.R + (size_t)cell * NT : nullptr;
    Real* Op = a.out + (size_t)cell * NT;
    const Real al = Real(a.alpha), be = Real(a.beta);
    if (FUSED)
      racc.store(grp, [&](int v, int n, Real val) {
        Real old = (be != Real(0)) ? be * Rp[v * N + n] : Real(0);
        Rp[v * N + n] = old + al * val;
      });
    aacc.store(grp, [&](int v, int n, Real val) { Op[v * N + n] = val; });
  }
}

// LDS bytes of a launch: 1D tables + per-wave staging of the own coefficients
template <int D, int P, int K1, int V, typename Real>
constexpr size_t kt_lds_bytes(int narrays) {
  typedef TShape<D, P, K1> TS;
  return (size_t)(((TS::TAB + 1) & ~1) + TWAVES * narrays * 16 * ((V * TS::N) | 1)) * sizeof(Real);
}
// kt_forward and kt_adjoint: the same, followed by their face tables
template <class Form, int D, int P, int K1, typename Real>
constexpr size_t kt_forward_lds_bytes(int narrays) {
  return kt_lds_bytes<D, P, K1, Form::V, Real>(narrays) +
         (size_t)form_face_tables<Form, D>() * TShape<D, P, K1>::FTAB * sizeof(Real);
}

}  // namespace bosss

namespace bosss {

// ----------------------------------------------------------------------------
// Euler VJP with an LDS node buffer (tensor tables).
//
// Same contractions and pointwise adjoints as kt_adjoint's hand-written path, but
// the 4-nodes-per-lane MFMA outputs (states and cotangents of a 16-node tile) are
// parked in a per-lane slice of LDS right after each evaluation, and the nodes are
// then processed one at a time from there.  LDS is private per lane here (slot-major,
// lane-minor: conflict-free, no barriers), so it acts as an explicitly managed
// spill area at LDS speed; the own cell's coefficients are read just in time from
// global memory (L1/L2) instead of being staged.  This keeps the FP64 Euler VJP
// inside the register file.
// ----------------------------------------------------------------------------
// Waves per SIMD the FP32 Euler VJP kernels are compiled for.  In FP32 the per-lane state fits
// 256 registers and the LDS use is 79 KB per workgroup at p = 3 (node buffers 61 KB, face trace
// table 17 KB), so two workgroups share a CU and
// one wave's memory waits overlap the other's arithmetic (the fused p = 3 residual + VJP runs
// 1.4x faster than at one wave).  The non-fused p = 3 variants reach two waves without the
// hint and spill with it, so they are left to the compiler.
template <typename Real, int P, bool FUSED>
constexpr int vjp_euler_min_waves() { return (sizeof(Real) == 4 && (FUSED || P < 3)) ? 2 : 1; }

template <class Form, int D, int P, int K1, typename Real, bool FUSED>
__global__ void __launch_bounds__(64 * TWAVES, (vjp_euler_min_waves<Real, P, FUSED>())) kt_vjp_euler(TArgs<Real> a) {
  typedef TShape<D, P, K1> TS;
  typedef typename MM<Real>::acc acc_t;
  constexpr int V = Form::V, N = TS::N, NT = V * N, NF = TS::NF, MT = TS::MT;
  constexpr int NSLOT = 3 * V * 4;

  extern __shared__ __align__(16) unsigned char smem_raw[];
  Real* tab = reinterpret_cast<Real*>(smem_raw);
  Real* ft = tab + ((TS::TAB + 1) & ~1) + TWAVES * NSLOT * 64;
  for (int i = threadIdx.x; i < TS::TAB; i += blockDim.x) tab[i] = a.tab[i];
  ft_load<TS>(ft, a.tab);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int col = lane & 15, grp = lane >> 4;
  Real* buf = tab + ((TS::TAB + 1) & ~1) + wave * (NSLOT * 64);
  const FtCols<TS, Real, V> ftc(col);
  auto slot = [&](int arr, int v, int r) -> Real& { return buf[((arr * V + v) * 4 + r) * 64 + lane]; };

  const int cell = (xcd_block(a.nxcd, a.plane_blocks) * TWAVES + wave) * 16 + col;
  const bool valid = cell < a.J;
  const int cj = valid ? cell : 0;
  const Real sc = Real(a.scaling[cj]);
  const Real detJ = Real(a.detJ[cj]);
  Real iJ[D][D];
#pragma unroll
  for (int e = 0; e < D; e++)
#pragma unroll
    for (int d = 0; d < D; d++) iJ[e][d] = Real(a.invJ[(size_t)cj * D * D + e * D + d]);

  TAcc<TS, Real, V> racc, aacc;
  racc.zero();
  aacc.zero();

  const Real* ownU = a.U + (size_t)cj * NT;
  const Real* ownW = a.X + (size_t)cj * NT;
  // The fused kernel keeps the own cell's B operands (scaled U, W) in registers for the whole
  // launch: U is contracted 7 times and W 9 times.  The VJP-only kernel reads them just in time
  // from L2 (holding them there costs it more in register traffic than it saves).
  constexpr int PV = FUSED ? V : 1, PK = FUSED ? TS::NK : 1;
  Real pU[PV][PK], pW[PV][PK];
  if constexpr (FUSED) {
#pragma unroll
    for (int v = 0; v < V; v++)
#pragma unroll
      for (int kk = 0; kk < TS::NK; kk++) {
        const int m = 4 * kk + grp;
        const bool in = valid && m < N;
        pU[v][kk] = ownU[v * N + (m < N ? m : N - 1)] * (in ? sc : Real(0));
        pW[v][kk] = ownW[v * N + (m < N ? m : N - 1)] * (in ? Real(1) : Real(0));
      }
  }
  auto Bu = [&](int v, int kk) -> Real {
    if constexpr (FUSED) {
      return pU[v][kk];
    } else {
      const int m = 4 * kk + grp;
      return ownU[v * N + (m < N ? m : N - 1)] * ((valid && m < N) ? sc : Real(0));
    }
  };
  auto Bw = [&](int v, int kk) -> Real {
    if constexpr (FUSED) {
      return pW[v][kk];
    } else {
      const int m = 4 * kk + grp;
      return ownW[v * N + (m < N ? m : N - 1)] * ((valid && m < N) ? Real(1) : Real(0));
    }
  };
  auto park = [&](int arr, const acc_t (&acc)[V]) {
#pragma unroll
    for (int v = 0; v < V; v++)
#pragma unroll
      for (int r = 0; r < 4; r++) slot(arr, v, r) = acc[v][r];
  };

  // ------------------------------------------------------------------ volume
  for (int t = 0; t < TS::TV; t++) {
    int qe[3];
    const int nodeE = t * 16 + col;
    vol_axes<D, K1>(nodeE < TS::KV ? nodeE : 0, qe);
    {
      acc_t acc[V];
      t_eval<TS, Real, V>(tab, qe, -1, grp, Bu, acc);
      park(0, acc);
    }
    if (FUSED) {
      for (int r = 0; r < 4; r++) {
        if (t * 16 + MM<Real>::row(0, r) >= TS::KV) continue;   // all 4 nodes of this step are padding
        const int node = t * 16 + MM<Real>::row(grp, r);
        const bool pad = node >= TS::KV;
        int q[3];
        vol_axes<D, K1>(pad ? 0 : node, q);
        const Real wq = pad ? Real(0) : vol_weight<TS>(tab, q) * detJ * sc;
        Real Uq[V], GUd[V][D], Fv[V][D];
#pragma unroll
        for (int v = 0; v < V; v++) { Uq[v] = slot(0, v, r); for (int d = 0; d < D; d++) GUd[v][d] = Uq[v]; }
        Form::template volume<D, Real, Real>(a.fp, Uq, GUd, Fv);
#pragma unroll
        for (int e = 0; e < D; e++) {
          Real g[V];
#pragma unroll
          for (int v = 0; v < V; v++) {
            Real s = 0;
#pragma unroll
            for (int d = 0; d < D; d++) s += Fv[v][d] * iJ[e][d];
            g[v] = pad ? Real(0) : s * wq;
          }
          t_test<TS, Real, V>(tab, q, e, col, g, racc);
        }
      }
    }
    // Ubar accumulated per node in the buffer's third array
#pragma unroll
    for (int v = 0; v < V; v++)
#pragma unroll
      for (int r = 0; r < 4; r++) slot(2, v, r) = Real(0);
    for (int e = 0; e < D; e++) {
      // the W gradient along e is used once per node: read straight from the MFMA result
      acc_t gW[V];
      t_eval<TS, Real, V>(tab, qe, e, grp, Bw, gW);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        if (t * 16 + MM<Real>::row(0, r) >= TS::KV) continue;   // all 4 nodes of this step are padding
        const int node = t * 16 + MM<Real>::row(grp, r);
        const bool pad = node >= TS::KV;
        int q[3];
        vol_axes<D, K1>(pad ? 0 : node, q);
        const Real wq = pad ? Real(0) : vol_weight<TS>(tab, q) * detJ * sc;
        Real Uq[V], gw[V], ub[V], kd[D];
#pragma unroll
        for (int d = 0; d < D; d++) kd[d] = iJ[e][d] * wq;
#pragma unroll
        for (int v = 0; v < V; v++) {
          Uq[v] = slot(0, v, r);
          ub[v] = slot(2, v, r);
          gw[v] = gW[v][r];
        }
        EulerAdjoint<D, Real>::volume_vjp_dir(a.fp, Uq, kd, gw, ub);
#pragma unroll
        for (int v = 0; v < V; v++) slot(2, v, r) = ub[v];
      }
    }
    for (int r = 0; r < 4; r++) {
      if (t * 16 + MM<Real>::row(0, r) >= TS::KV) continue;   // all 4 nodes of this step are padding
      const int node = t * 16 + MM<Real>::row(grp, r);
      const bool pad = node >= TS::KV;
      int q[3];
      vol_axes<D, K1>(pad ? 0 : node, q);
      Real bv[V];
#pragma unroll
      for (int v = 0; v < V; v++) bv[v] = pad ? Real(0) : slot(2, v, r) * sc;
      t_test<TS, Real, V>(tab, q, -1, col, bv, aacc);
    }
  }

  // ------------------------------------------------------------------- faces
  for (int f = 0; f < NF; f++) {
    const int nb_raw = a.cf.nb[(size_t)cj * NF + f];
    const int nb = valid ? nb_raw : -1;
    const bool has_nb = nb >= 0;
    const int cn = has_nb ? nb : cj;
    const int fo = a.cf.pair[f];
    const Real sg = Real(a.cf.sqrtg[(size_t)cj * NF + f]) * (valid ? Real(1) : Real(0));
    Real nrm[D];
#pragma unroll
    for (int d = 0; d < D; d++) nrm[d] = Real(a.cf.normal[((size_t)cj * NF + f) * D + d]);
    const Real scn = Real(a.scaling[cn]);
    const Real* nbU = a.U + (size_t)cn * NT;
    const Real* nbW = a.X + (size_t)cn * NT;
    auto Bnu = [&](int v, int kk) { int m = 4 * kk + grp; return nbU[v * N + (m < N ? m : N - 1)] * ((has_nb && m < N) ? scn : Real(0)); };

    for (int t = 0; t < TS::TE; t++) {
      // trace operands FT(f, node, mode) and FT(fo, node, mode) of the tile's nodes
      const Real* fto = ft + (f * TS::FROWS + t * 16 + col) * TS::FS + grp;
      const Real* ftn = ft + (fo * TS::FROWS + t * 16 + col) * TS::FS + grp;
      {
        acc_t acc[V];
        t_eval_ft<TS, Real, V>(fto, Bu, acc);
        park(0, acc);
      }
      {
        acc_t acc[V];
        t_eval_ft<TS, Real, V>(ftn, Bnu, acc);
        park(1, acc);
      }
      {
        // merged face cotangent  sc W_own.trace - scn W_nb.trace  (fn = -fo)
        acc_t acc[V];
#pragma unroll
        for (int v = 0; v < V; v++) acc[v] = acc_zero<Real>();
#pragma unroll
        for (int kk = 0; kk < TS::NK; kk++) {
          int m = 4 * kk + grp;
          Real ao = fto[4 * kk];
          Real an = ftn[4 * kk];
#pragma unroll
          for (int v = 0; v < V; v++) {
            const int mc = m < N ? m : N - 1;
            Real wo_ = Bw(v, kk) * sc;
            Real wn_ = nbW[v * N + mc] * ((has_nb && m < N) ? scn : Real(0));
            acc[v] = MM<Real>::mma(ao, wo_, acc[v]);
            acc[v] = MM<Real>::mma(an, -wn_, acc[v]);
          }
        }
        park(2, acc);
      }
      for (int r = 0; r < 4; r++) {
        if (t * 16 + MM<Real>::row(0, r) >= TS::KE) continue;   // all 4 nodes of this step are padding
        const int node = t * 16 + MM<Real>::row(grp, r);
        const bool pad = node >= TS::KE;
        const Real* nrow = ft + (f * TS::FROWS + node) * TS::FS;
        const Real wsg = pad ? Real(0) : face_weight<TS>(tab, node) * sg;
        Real Uo[V], Un[V], gbar[V], fo_[V], ub[V];
#pragma unroll
        for (int v = 0; v < V; v++) {
          Uo[v] = slot(0, v, r);
          Un[v] = has_nb ? slot(1, v, r) : Uo[v];
          gbar[v] = slot(2, v, r) * wsg;
          ub[v] = 0;
        }
        EulerAdjoint<D, Real>::template inner_vjp<IsHLLC<Form>::value>(a.fp, nrm, Uo, Un, gbar, fo_, ub);
        Real g[V];
        if (FUSED) {
#pragma unroll
          for (int v = 0; v < V; v++) g[v] = pad ? Real(0) : fo_[v] * wsg * sc;
          t_test_ft<TS, Real, V>(nrow, ftc, g, racc);
        }
#pragma unroll
        for (int v = 0; v < V; v++) g[v] = pad ? Real(0) : ub[v] * sc;
        t_test_ft<TS, Real, V>(nrow, ftc, g, aacc);
      }
    }
  }

  if (valid) {
    Real* Rp = FUSED ? a.R + (size_t)cell * NT : nullptr;
    Real* Op = a.out + (size_t)cell * NT;
    const Real al = Real(a.alpha), be = Real(a.beta);
    if (FUSED)
      racc.store(grp, [&](int v, int n, Real val) {
        Real old = (be != Real(0)) ? be * Rp[v * N + n] : Real(0);
        Rp[v * N + n] = old + al * val;
      });
    aacc.store(grp, [&](int v, int n, Real val) { Op[v * N + n] = val; });
  }
}

// ----------------------------------------------------------------------------
// Euler JVP with the same LDS node buffer: own trace, neighbour trace and own
// tangent trace of a 16-node face tile are parked in LDS, the neighbour tangent
// trace stays in registers; the dual-number flux is then evaluated node by node.
// ----------------------------------------------------------------------------
template <class Form, int D, int P, int K1, typename Real>
__global__ void __launch_bounds__(64 * TWAVES) __attribute__((amdgpu_waves_per_eu(sizeof(Real) == 4 ? 2 : 1))) kt_jvp_euler(TArgs<Real> a) {
  typedef TShape<D, P, K1> TS;
  typedef typename MM<Real>::acc acc_t;
  typedef Dual<1, Real> T;
  constexpr int V = Form::V, N = TS::N, NT = V * N, NF = TS::NF, MT = TS::MT;
  constexpr int NSLOT = 3 * V * 4;

  extern __shared__ __align__(16) unsigned char smem_raw[];
  Real* tab = reinterpret_cast<Real*>(smem_raw);
  Real* ft = tab + ((TS::TAB + 1) & ~1) + TWAVES * NSLOT * 64;
  for (int i = threadIdx.x; i < TS::TAB; i += blockDim.x) tab[i] = a.tab[i];
  ft_load<TS>(ft, a.tab);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int col = lane & 15, grp = lane >> 4;
  Real* buf = tab + ((TS::TAB + 1) & ~1) + wave * (NSLOT * 64);
  const FtCols<TS, Real, V> ftc(col);
  auto slot = [&](int arr, int v, int r) -> Real& { return buf[((arr * V + v) * 4 + r) * 64 + lane]; };

  const int cell = (xcd_block(a.nxcd, a.plane_blocks) * TWAVES + wave) * 16 + col;
  const bool valid = cell < a.J;
  const int cj = valid ? cell : 0;
  const Real sc = Real(a.scaling[cj]);
  const Real detJ = Real(a.detJ[cj]);
  Real iJ[D][D];
#pragma unroll
  for (int e = 0; e < D; e++)
#pragma unroll
    for (int d = 0; d < D; d++) iJ[e][d] = Real(a.invJ[(size_t)cj * D * D + e * D + d]);

  TAcc<TS, Real, V> racc;
  racc.zero();

  const Real* ownU = a.U + (size_t)cj * NT;
  const Real* ownX = a.X + (size_t)cj * NT;
  // FP32 keeps the own cell's B operands (scaled U and tangent) in registers for the whole launch;
  // FP64 reads them just in time from L2 (its register file is full at one wavefront per SIMD and
  // holding them there costs more than the reloads).
  constexpr bool PRE = std::is_same<Real, float>::value;
  constexpr int PV = PRE ? V : 1, PK = PRE ? TS::NK : 1;
  Real pU[PV][PK], pX[PV][PK];
  if constexpr (PRE) {
#pragma unroll
    for (int v = 0; v < V; v++)
#pragma unroll
      for (int kk = 0; kk < TS::NK; kk++) {
        const int m = 4 * kk + grp;
        const bool in = valid && m < N;
        pU[v][kk] = ownU[v * N + (m < N ? m : N - 1)] * (in ? sc : Real(0));
        pX[v][kk] = ownX[v * N + (m < N ? m : N - 1)] * (in ? sc : Real(0));
      }
  }
  auto Bu = [&](int v, int kk) -> Real {
    if constexpr (PRE) {
      return pU[v][kk];
    } else {
      const int m = 4 * kk + grp;
      return ownU[v * N + (m < N ? m : N - 1)] * ((valid && m < N) ? sc : Real(0));
    }
  };
  auto Bx = [&](int v, int kk) -> Real {
    if constexpr (PRE) {
      return pX[v][kk];
    } else {
      const int m = 4 * kk + grp;
      return ownX[v * N + (m < N ? m : N - 1)] * ((valid && m < N) ? sc : Real(0));
    }
  };
  auto park = [&](int arr, const acc_t (&acc)[V]) {
#pragma unroll
    for (int v = 0; v < V; v++)
#pragma unroll
      for (int r = 0; r < 4; r++) slot(arr, v, r) = acc[v][r];
  };

  // ------------------------------------------------------------------ volume
  for (int t = 0; t < TS::TV; t++) {
    int qe[3];
    const int nodeE = t * 16 + col;
    vol_axes<D, K1>(nodeE < TS::KV ? nodeE : 0, qe);
    // state and tangent at the nodes are used once each: read straight from the MFMA results
    acc_t au[V], adu[V];
    t_eval<TS, Real, V>(tab, qe, -1, grp, Bu, au);
    t_eval<TS, Real, V>(tab, qe, -1, grp, Bx, adu);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      if (t * 16 + MM<Real>::row(0, r) >= TS::KV) continue;   // all 4 nodes of this step are padding
      const int node = t * 16 + MM<Real>::row(grp, r);
      const bool pad = node >= TS::KV;
      int q[3];
      vol_axes<D, K1>(pad ? 0 : node, q);
      const Real wq = pad ? Real(0) : vol_weight<TS>(tab, q) * detJ * sc;
      T Uq[V], GU[V][D], F[V][D];
#pragma unroll
      for (int v = 0; v < V; v++) {
        Uq[v].v = au[v][r];
        Uq[v].d[0] = adu[v][r];
#pragma unroll
        for (int d = 0; d < D; d++) GU[v][d] = Uq[v];
      }
      Form::template volume<D, T, Real>(a.fp, Uq, GU, F);
#pragma unroll
      for (int e = 0; e < D; e++) {
        Real g[V];
#pragma unroll
        for (int v = 0; v < V; v++) {
          Real s = 0;
#pragma unroll
          for (int d = 0; d < D; d++) s += F[v][d].d[0] * iJ[e][d];
          g[v] = pad ? Real(0) : s * wq;
        }
        t_test<TS, Real, V>(tab, q, e, col, g, racc);
      }
    }
  }

  // ------------------------------------------------------------------- faces
  for (int f = 0; f < NF; f++) {
    const int nb_raw = a.cf.nb[(size_t)cj * NF + f];
    const int nb = valid ? nb_raw : -1;
    const bool has_nb = nb >= 0;
    const int cn = has_nb ? nb : cj;
    const int fo = a.cf.pair[f];
    const Real sg = Real(a.cf.sqrtg[(size_t)cj * NF + f]) * (valid ? Real(1) : Real(0));
    Real nrm[D];
#pragma unroll
    for (int d = 0; d < D; d++) nrm[d] = Real(a.cf.normal[((size_t)cj * NF + f) * D + d]);
    const Real scn = Real(a.scaling[cn]);
    const Real* nbU = a.U + (size_t)cn * NT;
    const Real* nbX = a.X + (size_t)cn * NT;
    auto Bnu = [&](int v, int kk) { int m = 4 * kk + grp; return nbU[v * N + (m < N ? m : N - 1)] * ((has_nb && m < N) ? scn : Real(0)); };
    auto Bnx = [&](int v, int kk) { int m = 4 * kk + grp; return nbX[v * N + (m < N ? m : N - 1)] * ((has_nb && !a.own_only && m < N) ? scn : Real(0)); };

    for (int t = 0; t < TS::TE; t++) {
      // trace operands FT(f, node, mode) and FT(fo, node, mode) of the tile's nodes
      const Real* fto = ft + (f * TS::FROWS + t * 16 + col) * TS::FS + grp;
      const Real* ftn = ft + (fo * TS::FROWS + t * 16 + col) * TS::FS + grp;
      acc_t dun[V];
      {
        acc_t acc[V];
        t_eval_ft<TS, Real, V>(fto, Bu, acc);
        park(0, acc);
        t_eval_ft<TS, Real, V>(ftn, Bnu, acc);
        park(1, acc);
        t_eval_ft<TS, Real, V>(fto, Bx, acc);
        park(2, acc);
      }
      if (a.own_only) {   // neighbour tangents taken as zero (diagonal-block probes): no contraction
#pragma unroll
        for (int v = 0; v < V; v++) dun[v] = acc_zero<Real>();
      } else {
        t_eval_ft<TS, Real, V>(ftn, Bnx, dun);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        if (t * 16 + MM<Real>::row(0, r) >= TS::KE) continue;   // all 4 nodes of this step are padding
        const int node = t * 16 + MM<Real>::row(grp, r);
        const bool pad = node >= TS::KE;
        const Real* nrow = ft + (f * TS::FROWS + node) * TS::FS;
        const Real wq = pad ? Real(0) : face_weight<TS>(tab, node) * sg * sc;
        T Uo[V], Un[V], Go[V][D], Gn[V][D];
#pragma unroll
        for (int v = 0; v < V; v++) {
          Uo[v].v = slot(0, v, r); Uo[v].d[0] = slot(2, v, r);
          Un[v].v = slot(1, v, r); Un[v].d[0] = dun[v][r];
          if (!has_nb) Un[v] = Uo[v];
#pragma unroll
          for (int d = 0; d < D; d++) { Go[v][d] = Uo[v]; Gn[v][d] = Uo[v]; }
        }
        T fo_[V], go_[V][D], fn_[V], gn_[V][D];
        if (has_nb) {
          Form::template inner<D, T, Real>(a.fp, nrm, Real(0), Uo, Un, Go, Gn, fo_, go_, fn_, gn_);
        } else {
#pragma unroll
          for (int v = 0; v < V; v++) { fo_[v] = Uo[v]; for (int d = 0; d < D; d++) go_[v][d] = Uo[v]; }
          Form::template boundary<D, T, Real>(a.fp, nrm, Real(0), Uo, Go, fo_, go_);
        }
        Real g[V];
#pragma unroll
        for (int v = 0; v < V; v++) g[v] = pad ? Real(0) : fo_[v].d[0] * wq;
        t_test_ft<TS, Real, V>(nrow, ftc, g, racc);
      }
    }
  }

  if (valid) {
    Real* Op = a.out + (size_t)cell * NT;
    racc.store(grp, [&](int v, int n, Real val) { Op[v * N + n] = val; });
  }
}

// LDS bytes of the Euler VJP / JVP kernels: 1D tables, per-wave node buffers, face trace table
template <int D, int P, int K1, int V, typename Real>
constexpr size_t kt_vjp_euler_lds_bytes() {
  typedef TShape<D, P, K1> TS;
  return (size_t)(((TS::TAB + 1) & ~1) + TWAVES * 3 * V * 4 * 64 + TS::FTAB) * sizeof(Real);
}

}  // namespace bosss
