// Cell-centric DG operator kernels for gfx950 (MI355X).
//
// One wavefront owns 16 cells (the 16 columns of a 16x16 MFMA tile).  For each
// cell it evaluates the volume integral and the integrals over all 2D faces of
// the cell, so every cell's residual is written exactly once, without atomics,
// in a deterministic order.  Each interior face is therefore evaluated by both
// of its cells; that costs one extra trace evaluation per face and removes the
// scatter to two cells (NECQuadratureEdge.cs:326-410) that would otherwise need
// atomics or a face colouring with 2D passes over the output.
//
// All modal<->nodal contractions run on MFMA in the transposed form
//     nodes x cells   =  table[nodes x modes] * coeff^T[modes x cells]      (eval)
//     modes x cells  +=  table^T[modes x nodes] * G[nodes x cells]           (test)
// The MFMA output register r of lane (col, grp) holds node row(grp, r) of cell
// col, which is exactly the B operand the test contraction needs for k-step r;
// no data moves between the two contractions.  Tables (basis values and
// gradients at volume and face nodes) live in LDS and feed the A operand.
//
// Reference semantics (see forms.hpp for the pointwise forms):
//   u(k) = sum_n Scaling_j c_n phi_n(xi_k)                    DGField_Evaluation.cs:38-129
//   grad u = invJ^T grad_xi u                                 DGField_Evaluation.cs:241-314
//   R_j,n += detJ Scaling sum_k w_k sum_e (sum_d F_d invJ[e][d]) dphi_n,e   NECQuadratureVolume.cs:584-733
//   R_j,n += sqrtGram Scaling sum_k w_k (f psi_n + (invJ g) . dpsi_n)       NECQuadratureEdge.cs:1180-1420
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "forms.hpp"

namespace bosss {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <typename Real> struct MM;
template <> struct MM<double> {
  typedef f64x4 acc;
  static __device__ __forceinline__ acc mma(double a, double b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  // v_mfma_f64_16x16x4f64: lane (col, grp) holds output rows grp + 4r, r = 0..3
  static __device__ __forceinline__ int row(int grp, int r) { return grp + 4 * r; }
};
template <> struct MM<float> {
  typedef f32x4 acc;
  static __device__ __forceinline__ acc mma(float a, float b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  // v_mfma_f32_16x16x4f32: lane (col, grp) holds output rows 4 grp + r
  static __device__ __forceinline__ int row(int grp, int r) { return 4 * grp + r; }
};
template <typename Real> __device__ __forceinline__ typename MM<Real>::acc acc_zero() {
  typename MM<Real>::acc a;
  a[0] = a[1] = a[2] = a[3] = Real(0);
  return a;
}

constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }

template <int D_, int N_, int K1_, bool FACEGRAD = true>
struct Shape {
  static constexpr int D = D_, N = N_, K1 = K1_;
  static constexpr int KV = ipow(K1, D), KE = ipow(K1, D - 1), NF = 2 * D;
  static constexpr int TV = (KV + 15) / 16, TE = (KE + 15) / 16;
  static constexpr int KVP = TV * 16, KEP = TE * 16;
  static constexpr int NK = (N + 3) / 4, MT = (N + 15) / 16;
  // packed table layout (offsets in elements)
  static constexpr int O_VPHI = 0;                        // [KVP][N]
  static constexpr int O_VDPHI = O_VPHI + KVP * N;        // [D][KVP][N]
  static constexpr int O_VW = O_VDPHI + D * KVP * N;      // [KVP]
  static constexpr int O_TR = O_VW + KVP;                 // [NF][KEP][N]
  static constexpr int O_DTR = O_TR + NF * KEP * N;       // [NF][D][KEP][N] (only if the form needs face gradients)
  static constexpr int O_FW = O_DTR + (FACEGRAD ? NF * D * KEP * N : 0);   // [KEP]
  static constexpr int SIZE = O_FW + KEP;
};

// Per-cell face topology (built on the host from the edge list).
struct CellFaces {
  const int* nb;         // [J][NF]    neighbour across local face f, -1 on the boundary
  const double* sqrtg;   // [J][NF]
  const double* normal;  // [J][NF][D] outward from the cell
  int pair[6];           // trafo of the neighbour's side for local face f
};

template <typename Real>
struct KArgs {
  int J;
  const double* detJ;
  const double* invJ;
  const double* scaling;
  const double* lenscale;
  CellFaces cf;
  const Real* tables;
  FormParams fp;
  double alpha, beta;
  const Real* U;    // state
  const Real* dU;   // tangent (JVP) or cotangent W (VJP)
  Real* R;          // residual output (beta R + alpha Op(U))
  Real* out;        // JVP / VJP output
  int own_only;     // JVP: neighbour tangents taken as zero (J(U)_jj X_j, the diagonal blocks)
};

enum { MODE_RES = 0, MODE_JVP = 1, MODE_RES_JVP = 2 };

// Load the B operand for one cell column: c[v][kk] = scale * U[cell][v][4kk + grp]
template <typename Real, int V, int N, int NK>
__device__ __forceinline__ void load_coeffs(const Real* __restrict__ U, int cell, bool ok, Real scale, int grp, Real (&c)[V][NK]) {
  const Real* p = U + (size_t)cell * (V * N);
#pragma unroll
  for (int v = 0; v < V; v++)
#pragma unroll
    for (int kk = 0; kk < NK; kk++) {
      int m = 4 * kk + grp;
      c[v][kk] = (ok && m < N) ? p[v * N + m] * scale : Real(0);
    }
}

// acc[v] = tab[t*16 + col][:] . c[v][:]   (tab: [rows][N] in LDS)
template <typename Real, int V, int N, int NK>
__device__ __forceinline__ void eval_tile(const Real* tab, int t, int col, int grp, const Real (&c)[V][NK], typename MM<Real>::acc (&acc)[V]) {
#pragma unroll
  for (int v = 0; v < V; v++) acc[v] = acc_zero<Real>();
  const Real* rowp = tab + (t * 16 + col) * N;
#pragma unroll
  for (int kk = 0; kk < NK; kk++) {
    int m = 4 * kk + grp;
    Real a = (m < N) ? rowp[m] : Real(0);
#pragma unroll
    for (int v = 0; v < V; v++) acc[v] = MM<Real>::mma(a, c[v][kk], acc[v]);
  }
}

// racc[v][mt] += tab^T (column node) * g[v] for k-step r   (tab: [rows][N] in LDS)
template <typename Real, int V, int N, int MT>
__device__ __forceinline__ void test_step(const Real* tab, int node, int col, const Real (&g)[V], typename MM<Real>::acc (&racc)[V][MT]) {
#pragma unroll
  for (int mt = 0; mt < MT; mt++) {
    int n = mt * 16 + col;
    Real a = (n < N) ? tab[node * N + n] : Real(0);
#pragma unroll
    for (int v = 0; v < V; v++) racc[v][mt] = MM<Real>::mma(a, g[v], racc[v][mt]);
  }
}

template <typename Real, int N>
__device__ __forceinline__ void copy_tables(Real* lds, const Real* __restrict__ g, int size) {
  for (int i = threadIdx.x; i < size; i += blockDim.x) lds[i] = g[i];
  __syncthreads();
}

// ----------------------------------------------------------------------------
// Forward kernel: residual and/or JVP.
//   MODE_RES      R   <- beta R + alpha Op(U)
//   MODE_JVP      out <- J(U) dU
//   MODE_RES_JVP  both
// ----------------------------------------------------------------------------
template <class Form, int D, int N, int K1, typename Real, int MODE>
__global__ void __launch_bounds__(256) k_forward(KArgs<Real> a) {
  typedef Shape<D, N, K1, Form::FACE_GRADU || Form::FACE_TESTG> S;
  typedef typename MM<Real>::acc acc_t;
  constexpr int V = Form::V;
  constexpr int NK = S::NK, MT = S::MT, NF = S::NF;
  constexpr bool TAN = (MODE != MODE_RES);
  constexpr bool RES = (MODE != MODE_JVP);
  typedef typename std::conditional<TAN, Dual<1, Real>, Real>::type T;

  extern __shared__ __align__(16) unsigned char smem_raw[];
  Real* lds = reinterpret_cast<Real*>(smem_raw);
  copy_tables<Real, N>(lds, a.tables, S::SIZE);

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, grp = lane >> 4;
  const int cell = (blockIdx.x * (blockDim.x >> 6) + wave) * 16 + col;
  const bool valid = cell < a.J;
  const int cj = valid ? cell : 0;

  const Real sc = Real(a.scaling[cj]);
  const Real detJ = Real(a.detJ[cj]);
  Real iJ[D][D];
#pragma unroll
  for (int e = 0; e < D; e++)
#pragma unroll
    for (int d = 0; d < D; d++) iJ[e][d] = Real(a.invJ[(size_t)cj * D * D + e * D + d]);

  Real cown[V][NK], town[V][NK];
  load_coeffs<Real, V, N, NK>(a.U, cj, valid, sc, grp, cown);
  if (TAN) load_coeffs<Real, V, N, NK>(a.dU, cj, valid, sc, grp, town);

  acc_t racc[V][MT], tacc[V][MT];
#pragma unroll
  for (int v = 0; v < V; v++)
#pragma unroll
    for (int mt = 0; mt < MT; mt++) { racc[v][mt] = acc_zero<Real>(); tacc[v][mt] = acc_zero<Real>(); }

  const Real* vphi = lds + S::O_VPHI;
  const Real* vdphi = lds + S::O_VDPHI;
  const Real* vw = lds + S::O_VW;

  // ------------------------------------------------------------------ volume
  for (int t = 0; t < S::TV; t++) {
    acc_t u[V], du[V], gu[D][V], dgu[D][V];
    eval_tile<Real, V, N, NK>(vphi, t, col, grp, cown, u);
    if (TAN) eval_tile<Real, V, N, NK>(vphi, t, col, grp, town, du);
    if (Form::VOL_GRADU) {
#pragma unroll
      for (int e = 0; e < D; e++) {
        eval_tile<Real, V, N, NK>(vdphi + e * S::KVP * N, t, col, grp, cown, gu[e]);
        if (TAN) eval_tile<Real, V, N, NK>(vdphi + e * S::KVP * N, t, col, grp, town, dgu[e]);
      }
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int node = t * 16 + MM<Real>::row(grp, r);
      const bool pad = node >= S::KV;   // tile padding: zero weight, and the form may be NaN there
      const Real wq = vw[node] * detJ * sc;
      T Uq[V], GU[V][D], F[V][D];
#pragma unroll
      for (int v = 0; v < V; v++) {
        if constexpr (TAN) { Uq[v].v = u[v][r]; Uq[v].d[0] = du[v][r]; } else { Uq[v] = u[v][r]; }
#pragma unroll
        for (int d = 0; d < D; d++) {
          if (Form::VOL_GRADU) {
            Real gv = Real(0), gt = Real(0);
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
        Real gr[V], gt[V];
#pragma unroll
        for (int v = 0; v < V; v++) {
          Real sv = Real(0), st = Real(0);
#pragma unroll
          for (int d = 0; d < D; d++) {
            if constexpr (TAN) { sv += F[v][d].v * iJ[e][d]; st += F[v][d].d[0] * iJ[e][d]; } else { sv += F[v][d] * iJ[e][d]; }
          }
          gr[v] = pad ? Real(0) : sv * wq;
          gt[v] = pad ? Real(0) : st * wq;
        }
        if (RES) test_step<Real, V, N, MT>(vdphi + e * S::KVP * N, node, col, gr, racc);
        if (TAN) test_step<Real, V, N, MT>(vdphi + e * S::KVP * N, node, col, gt, tacc);
      }
    }
  }

  // ------------------------------------------------------------------- faces
  const Real* tr = lds + S::O_TR;
  const Real* dtr = lds + S::O_DTR;
  const Real* fw = lds + S::O_FW;
  const Real pbd = Real(a.fp.penalty_base * a.fp.penalty_deg);
  const Real cj_own = Real(1.0 / a.lenscale[cj]);

  for (int f = 0; f < NF; f++) {
    const int nb = valid ? a.cf.nb[(size_t)cj * NF + f] : -1;
    const bool has_nb = nb >= 0;
    const int cn = has_nb ? nb : cj;
    const int fo = a.cf.pair[f];
    const Real sg = Real(valid ? a.cf.sqrtg[(size_t)cj * NF + f] : 0.0);
    Real nrm[D];
#pragma unroll
    for (int d = 0; d < D; d++) nrm[d] = Real(a.cf.normal[((size_t)cj * NF + f) * D + d]);
    // SIPLaplace.GetPenalty: max over both cells; boundary: own cell only (x2 inside the form)
    Real mu = pbd * cj_own;
    if (has_nb) { Real mo = pbd * Real(1.0 / a.lenscale[cn]); mu = (mo > mu) ? mo : mu; }
    const Real scn = Real(a.scaling[cn]);
    Real iJn[D][D];
    if (Form::FACE_GRADU) {
#pragma unroll
      for (int e = 0; e < D; e++)
#pragma unroll
        for (int d = 0; d < D; d++) iJn[e][d] = Real(a.invJ[(size_t)cn * D * D + e * D + d]);
    }
    Real cnb[V][NK], tnb[V][NK];
    load_coeffs<Real, V, N, NK>(a.U, cn, has_nb, scn, grp, cnb);
    if (TAN) load_coeffs<Real, V, N, NK>(a.dU, cn, has_nb && !a.own_only, scn, grp, tnb);

    const Real* tro = tr + f * S::KEP * N;
    const Real* trn = tr + fo * S::KEP * N;
    const Real* dtro = dtr + f * D * S::KEP * N;
    const Real* dtrn = dtr + fo * D * S::KEP * N;

    for (int t = 0; t < S::TE; t++) {
      acc_t uo[V], un[V], duo[V], dun[V];
      acc_t guo[D][V], gun[D][V], dguo[D][V], dgun[D][V];
      eval_tile<Real, V, N, NK>(tro, t, col, grp, cown, uo);
      eval_tile<Real, V, N, NK>(trn, t, col, grp, cnb, un);
      if (TAN) {
        eval_tile<Real, V, N, NK>(tro, t, col, grp, town, duo);
        eval_tile<Real, V, N, NK>(trn, t, col, grp, tnb, dun);
      }
      if (Form::FACE_GRADU) {
#pragma unroll
        for (int e = 0; e < D; e++) {
          eval_tile<Real, V, N, NK>(dtro + e * S::KEP * N, t, col, grp, cown, guo[e]);
          eval_tile<Real, V, N, NK>(dtrn + e * S::KEP * N, t, col, grp, cnb, gun[e]);
          if (TAN) {
            eval_tile<Real, V, N, NK>(dtro + e * S::KEP * N, t, col, grp, town, dguo[e]);
            eval_tile<Real, V, N, NK>(dtrn + e * S::KEP * N, t, col, grp, tnb, dgun[e]);
          }
        }
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int node = t * 16 + MM<Real>::row(grp, r);
        const bool pad = node >= S::KE;
        const Real wq = fw[node] * sg * sc;
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
        Real gr[V], gt[V];
#pragma unroll
        for (int v = 0; v < V; v++) {
          if constexpr (TAN) { gr[v] = fo_[v].v * wq; gt[v] = fo_[v].d[0] * wq; } else { gr[v] = fo_[v] * wq; gt[v] = 0; }
          if (pad) { gr[v] = Real(0); gt[v] = Real(0); }
        }
        if (RES) test_step<Real, V, N, MT>(tro, node, col, gr, racc);
        if (TAN) test_step<Real, V, N, MT>(tro, node, col, gt, tacc);
        if (Form::FACE_TESTG) {
#pragma unroll
          for (int e = 0; e < D; e++) {
#pragma unroll
            for (int v = 0; v < V; v++) {
              Real sv = 0, st = 0;
#pragma unroll
              for (int d = 0; d < D; d++) {
                if constexpr (TAN) { sv += go_[v][d].v * iJ[e][d]; st += go_[v][d].d[0] * iJ[e][d]; } else { sv += go_[v][d] * iJ[e][d]; }
              }
              gr[v] = pad ? Real(0) : sv * wq;
              gt[v] = pad ? Real(0) : st * wq;
            }
            if (RES) test_step<Real, V, N, MT>(dtro + e * S::KEP * N, node, col, gr, racc);
            if (TAN) test_step<Real, V, N, MT>(dtro + e * S::KEP * N, node, col, gt, tacc);
          }
        }
      }
    }
  }

  // ------------------------------------------------------------------- store
  if (valid) {
    Real* Rp = RES ? a.R + (size_t)cell * (V * N) : nullptr;
    Real* Op = TAN ? a.out + (size_t)cell * (V * N) : nullptr;
    const Real al = Real(a.alpha), be = Real(a.beta);
#pragma unroll
    for (int v = 0; v < V; v++)
#pragma unroll
      for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          int n = mt * 16 + MM<Real>::row(grp, r);
          if (n < N) {
            if (RES) {
              Real old = (be != Real(0)) ? be * Rp[v * N + n] : Real(0);
              Rp[v * N + n] = old + al * racc[v][mt][r];
            }
            if (TAN) Op[v * N + n] = tacc[v][mt][r];
          }
        }
  }
}

// ----------------------------------------------------------------------------
// Adjoint kernel: out <- J(U)^T W   (and, if FUSED, R <- beta R + alpha Op(U)).
// Per node, the transposed flux Jacobian is applied with forward-mode duals
// seeded on the cell's own inputs (state and, where the form uses it, its
// gradient); the face's other half is evaluated by the neighbour's thread.
// ----------------------------------------------------------------------------
template <class Form, int D, int N, int K1, typename Real, bool FUSED>
__global__ void __launch_bounds__(256) k_adjoint(KArgs<Real> a) {
  typedef Shape<D, N, K1, Form::FACE_GRADU || Form::FACE_TESTG> S;
  typedef typename MM<Real>::acc acc_t;
  constexpr int V = Form::V;
  constexpr int NK = S::NK, MT = S::MT, NF = S::NF;
  constexpr bool GU_IN = Form::VOL_GRADU || Form::FACE_GRADU;
  constexpr int M = GU_IN ? V * (1 + D) : V;   // own-side inputs seeded
  typedef Dual<M, Real> T;

  extern __shared__ __align__(16) unsigned char smem_raw[];
  Real* lds = reinterpret_cast<Real*>(smem_raw);
  copy_tables<Real, N>(lds, a.tables, S::SIZE);

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, grp = lane >> 4;
  const int cell = (blockIdx.x * (blockDim.x >> 6) + wave) * 16 + col;
  const bool valid = cell < a.J;
  const int cj = valid ? cell : 0;

  const Real sc = Real(a.scaling[cj]);
  const Real detJ = Real(a.detJ[cj]);
  Real iJ[D][D];
#pragma unroll
  for (int e = 0; e < D; e++)
#pragma unroll
    for (int d = 0; d < D; d++) iJ[e][d] = Real(a.invJ[(size_t)cj * D * D + e * D + d]);

  Real cown[V][NK], wown[V][NK];
  load_coeffs<Real, V, N, NK>(a.U, cj, valid, sc, grp, cown);
  load_coeffs<Real, V, N, NK>(a.dU, cj, valid, Real(1), grp, wown);   // W is a cotangent: no scaling

  acc_t racc[V][MT], aacc[V][MT];
#pragma unroll
  for (int v = 0; v < V; v++)
#pragma unroll
    for (int mt = 0; mt < MT; mt++) { racc[v][mt] = acc_zero<Real>(); aacc[v][mt] = acc_zero<Real>(); }

  const Real* vphi = lds + S::O_VPHI;
  const Real* vdphi = lds + S::O_VDPHI;
  const Real* vw = lds + S::O_VW;

  // ------------------------------------------------------------------ volume
  for (int t = 0; t < S::TV; t++) {
    acc_t u[V], gu[D][V], wg[D][V];
    eval_tile<Real, V, N, NK>(vphi, t, col, grp, cown, u);
#pragma unroll
    for (int e = 0; e < D; e++) {
      if (Form::VOL_GRADU) eval_tile<Real, V, N, NK>(vdphi + e * S::KVP * N, t, col, grp, cown, gu[e]);
      eval_tile<Real, V, N, NK>(vdphi + e * S::KVP * N, t, col, grp, wown, wg[e]);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int node = t * 16 + MM<Real>::row(grp, r);
      const bool pad = node >= S::KV;
      const Real wq = vw[node] * detJ * sc;
      T Uq[V], GU[V][D], F[V][D];
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
      Form::template volume<D, T, Real>(a.fp, Uq, GU, F);
      // cotangent of the physical flux: Fbar[v][d] = wq * sum_e invJ[e][d] * (grad_xi W)_e
      Real ub[V], gb[V][D];
#pragma unroll
      for (int v = 0; v < V; v++) { ub[v] = 0; for (int d = 0; d < D; d++) gb[v][d] = 0; }
#pragma unroll
      for (int v = 0; v < V; v++)
#pragma unroll
        for (int d = 0; d < D; d++) {
          Real fb = 0;
#pragma unroll
          for (int e = 0; e < D; e++) fb += iJ[e][d] * wg[e][v][r];
          fb *= wq;
#pragma unroll
          for (int q = 0; q < V; q++) {
            ub[q] += fb * F[v][d].d[q];
            if (GU_IN && Form::VOL_GRADU) {
#pragma unroll
              for (int dd = 0; dd < D; dd++) gb[q][dd] += fb * F[v][d].d[(V + q * D + dd) % M];
            }
          }
        }
      if (FUSED) {
#pragma unroll
        for (int e = 0; e < D; e++) {
          Real gr[V];
#pragma unroll
          for (int v = 0; v < V; v++) {
            Real s = 0;
#pragma unroll
            for (int d = 0; d < D; d++) s += F[v][d].v * iJ[e][d];
            gr[v] = pad ? Real(0) : s * wq;
          }
          test_step<Real, V, N, MT>(vdphi + e * S::KVP * N, node, col, gr, racc);
        }
      }
      // pull back to coefficients: cbar += sc * (ub phi + sum_e (invJ gb)_e dphi_e)
      Real bv[V];
#pragma unroll
      for (int v = 0; v < V; v++) bv[v] = pad ? Real(0) : ub[v] * sc;
      test_step<Real, V, N, MT>(vphi, node, col, bv, aacc);
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
          test_step<Real, V, N, MT>(vdphi + e * S::KVP * N, node, col, bv, aacc);
        }
      }
    }
  }

  // ------------------------------------------------------------------- faces
  const Real* tr = lds + S::O_TR;
  const Real* dtr = lds + S::O_DTR;
  const Real* fw = lds + S::O_FW;
  const Real pbd = Real(a.fp.penalty_base * a.fp.penalty_deg);
  const Real cj_own = Real(1.0 / a.lenscale[cj]);

  for (int f = 0; f < NF; f++) {
    const int nb = valid ? a.cf.nb[(size_t)cj * NF + f] : -1;
    const bool has_nb = nb >= 0;
    const int cn = has_nb ? nb : cj;
    const int fo = a.cf.pair[f];
    const Real sg = Real(valid ? a.cf.sqrtg[(size_t)cj * NF + f] : 0.0);
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
    Real cnb[V][NK], wnb[V][NK];
    load_coeffs<Real, V, N, NK>(a.U, cn, has_nb, scn, grp, cnb);
    load_coeffs<Real, V, N, NK>(a.dU, cn, has_nb, Real(1), grp, wnb);

    const Real* tro = tr + f * S::KEP * N;
    const Real* trn = tr + fo * S::KEP * N;
    const Real* dtro = dtr + f * D * S::KEP * N;
    const Real* dtrn = dtr + fo * D * S::KEP * N;

    for (int t = 0; t < S::TE; t++) {
      acc_t uo[V], un[V], wo[V], wn[V];
      acc_t guo[D][V], gun[D][V], gwo[D][V], gwn[D][V];
      eval_tile<Real, V, N, NK>(tro, t, col, grp, cown, uo);
      eval_tile<Real, V, N, NK>(trn, t, col, grp, cnb, un);
      eval_tile<Real, V, N, NK>(tro, t, col, grp, wown, wo);
      eval_tile<Real, V, N, NK>(trn, t, col, grp, wnb, wn);
#pragma unroll
      for (int e = 0; e < D; e++) {
        if (Form::FACE_GRADU) {
          eval_tile<Real, V, N, NK>(dtro + e * S::KEP * N, t, col, grp, cown, guo[e]);
          eval_tile<Real, V, N, NK>(dtrn + e * S::KEP * N, t, col, grp, cnb, gun[e]);
        }
        if (Form::FACE_TESTG) {
          eval_tile<Real, V, N, NK>(dtro + e * S::KEP * N, t, col, grp, wown, gwo[e]);
          eval_tile<Real, V, N, NK>(dtrn + e * S::KEP * N, t, col, grp, wnb, gwn[e]);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int node = t * 16 + MM<Real>::row(grp, r);
        const bool pad = node >= S::KE;
        const Real wsg = fw[node] * sg;
        T Uo[V], Un[V], Go[V][D], Gn[V][D];
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
        T fo_[V], go_[V][D], fn_[V], gn_[V][D];
        if (has_nb) {
          Form::template inner<D, T, Real>(a.fp, nrm, mu, Uo, Un, Go, Gn, fo_, go_, fn_, gn_);
        } else {
#pragma unroll
          for (int v = 0; v < V; v++) {
            fo_[v] = Uo[v]; fn_[v] = dconst<M, Real>(Real(0));
            for (int d = 0; d < D; d++) { go_[v][d] = Uo[v]; gn_[v][d] = dconst<M, Real>(Real(0)); }
          }
          Form::template boundary<D, T, Real>(a.fp, nrm, mu, Uo, Go, fo_, go_);
        }
        // cotangents of the node coefficients
        Real ub[V], gb[V][D];
#pragma unroll
        for (int q = 0; q < V; q++) { ub[q] = 0; for (int d = 0; d < D; d++) gb[q][d] = 0; }
#pragma unroll
        for (int v = 0; v < V; v++) {
          Real fob = wsg * sc * wo[v][r];
          Real fnb = has_nb ? wsg * scn * wn[v][r] : Real(0);
#pragma unroll
          for (int q = 0; q < V; q++) {
            ub[q] += fob * fo_[v].d[q] + fnb * fn_[v].d[q];
            if (GU_IN && Form::FACE_GRADU) {
#pragma unroll
              for (int dd = 0; dd < D; dd++) gb[q][dd] += fob * fo_[v].d[(V + q * D + dd) % M] + fnb * fn_[v].d[(V + q * D + dd) % M];
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
              for (int q = 0; q < V; q++) {
                ub[q] += gob * go_[v][d].d[q] + gnb * gn_[v][d].d[q];
                if (GU_IN && Form::FACE_GRADU) {
#pragma unroll
                  for (int dd = 0; dd < D; dd++)
                    gb[q][dd] += gob * go_[v][d].d[(V + q * D + dd) % M] + gnb * gn_[v][d].d[(V + q * D + dd) % M];
                }
              }
            }
          }
        }
        if (FUSED) {
          const Real wq = wsg * sc;
          Real gr[V];
#pragma unroll
          for (int v = 0; v < V; v++) gr[v] = pad ? Real(0) : fo_[v].v * wq;
          test_step<Real, V, N, MT>(tro, node, col, gr, racc);
          if (Form::FACE_TESTG) {
#pragma unroll
            for (int e = 0; e < D; e++) {
#pragma unroll
              for (int v = 0; v < V; v++) {
                Real s = 0;
#pragma unroll
                for (int d = 0; d < D; d++) s += go_[v][d].v * iJ[e][d];
                gr[v] = pad ? Real(0) : s * wq;
              }
              test_step<Real, V, N, MT>(dtro + e * S::KEP * N, node, col, gr, racc);
            }
          }
        }
        Real bv[V];
#pragma unroll
        for (int v = 0; v < V; v++) bv[v] = pad ? Real(0) : ub[v] * sc;
        test_step<Real, V, N, MT>(tro, node, col, bv, aacc);
        if (Form::FACE_GRADU) {
#pragma unroll
          for (int e = 0; e < D; e++) {
#pragma unroll
            for (int v = 0; v < V; v++) {
              Real s = 0;
#pragma unroll
              for (int d = 0; d < D; d++) s += gb[v][d] * iJ[e][d];
              bv[v] = pad ? Real(0) : s * sc;
            }
            test_step<Real, V, N, MT>(dtro + e * S::KEP * N, node, col, bv, aacc);
          }
        }
      }
    }
  }

  if (valid) {
    Real* Rp = FUSED ? a.R + (size_t)cell * (V * N) : nullptr;
    Real* Op = a.out + (size_t)cell * (V * N);
    const Real al = Real(a.alpha), be = Real(a.beta);
#pragma unroll
    for (int v = 0; v < V; v++)
#pragma unroll
      for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          int n = mt * 16 + MM<Real>::row(grp, r);
          if (n < N) {
            if (FUSED) {
              Real old = (be != Real(0)) ? be * Rp[v * N + n] : Real(0);
              Rp[v * N + n] = old + al * racc[v][mt][r];
            }
            Op[v * N + n] = aacc[v][mt][r];
          }
        }
  }
}

}  // namespace bosss
