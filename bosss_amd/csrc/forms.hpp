// Pointwise equation components evaluated at quadrature nodes on the GPU.
//
// Each form reports, per node, the coefficients the quadrature contracts with the
// test functions (the "vectorised" view of NonlinFormVectorizer.cs:77-140, 370-430):
//   volume :  F[v][d] multiplies d(psi)/dx_d
//   faces  :  fo[v] multiplies psi_own,  go[v][d] multiplies d(psi_own)/dx_d,
//             fn[v] multiplies psi_nb,   gn[v][d] multiplies d(psi_nb)/dx_d
// where "own" is the cell a kernel thread works for and the normal points out of
// it.  Every inner-edge form here satisfies  c_out(A,B,n) == c_in(B,A,-n), so a
// face evaluated from either side gives the reference's IN/OUT split
// (NECQuadratureEdge.cs:1000, OUT flux = -IN flux).
//
// All functions are templates over the scalar type T (double, float, or the
// forward-mode Dual below) so the JVP/VJP kernels differentiate the same code.
#pragma once
#include <hip/hip_runtime.h>

namespace bosss {

struct FormParams {
  double penalty_base, nu, vel[3], gamma, mach;
  double penalty_deg;  // max((p+1)^2, (p+1)(p+D)/D), SIPLaplace.CoefficientUpdate (ipLaplace.cs:100-113)
};

// 1/x for normal, finite x.  On the device: the hardware reciprocal estimate refined by
// Newton-Raphson steps r <- r + r (1 - x r), each doubling the correct bits (FP64:
// v_rcp_f64 gives about 23 bits, two steps reach the last ulp; FP32: v_rcp_f32 is within an
// ulp, one step).  On the host: the quotient.
__host__ __device__ __forceinline__ double recip(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
#else
  return 1.0 / x;
#endif
}
__host__ __device__ __forceinline__ float recip(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  float r = __builtin_amdgcn_rcpf(x);
  return fmaf(r, fmaf(-x, r, 1.0f), r);
#else
  return 1.0f / x;
#endif
}

// ----------------------------------------------------------------- dual numbers
template <int M, typename R>
struct Dual {
  R v;
  R d[M];
};
template <int M, typename R> __host__ __device__ __forceinline__ Dual<M, R> dconst(R x) { Dual<M, R> r; r.v = x;
#pragma unroll
  for (int i = 0; i < M; i++) r.d[i] = R(0); return r; }

#define BOSSS_DUAL_BINOP(OP, VEXPR, DEXPR)                                                            \
  template <int M, typename R>                                                                       \
  __host__ __device__ __forceinline__ Dual<M, R> operator OP(const Dual<M, R>& a, const Dual<M, R>& b) {      \
    Dual<M, R> r; r.v = VEXPR;                                                                       \
    _Pragma("unroll") for (int i = 0; i < M; i++) r.d[i] = DEXPR;                                    \
    return r;                                                                                        \
  }
BOSSS_DUAL_BINOP(+, a.v + b.v, a.d[i] + b.d[i])
BOSSS_DUAL_BINOP(-, a.v - b.v, a.d[i] - b.d[i])
BOSSS_DUAL_BINOP(*, a.v * b.v, a.d[i] * b.v + a.v * b.d[i])
#undef BOSSS_DUAL_BINOP
template <int M, typename R> __host__ __device__ __forceinline__ Dual<M, R> operator/(const Dual<M, R>& a, const Dual<M, R>& b) {
  Dual<M, R> r; R ib = recip(b.v); r.v = a.v * ib;
#pragma unroll
  for (int i = 0; i < M; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
template <int M, typename R> __host__ __device__ __forceinline__ Dual<M, R> operator-(const Dual<M, R>& a) {
  Dual<M, R> r; r.v = -a.v;
#pragma unroll
  for (int i = 0; i < M; i++) r.d[i] = -a.d[i];
  return r;
}
template <int M, typename R> __host__ __device__ __forceinline__ Dual<M, R> operator*(const Dual<M, R>& a, R s) {
  Dual<M, R> r; r.v = a.v * s;
#pragma unroll
  for (int i = 0; i < M; i++) r.d[i] = a.d[i] * s;
  return r;
}
template <int M, typename R> __host__ __device__ __forceinline__ Dual<M, R> operator*(R s, const Dual<M, R>& a) { return a * s; }
template <int M, typename R> __host__ __device__ __forceinline__ Dual<M, R> operator+(const Dual<M, R>& a, R s) { Dual<M, R> r = a; r.v += s; return r; }
template <int M, typename R> __host__ __device__ __forceinline__ Dual<M, R> operator+(R s, const Dual<M, R>& a) { Dual<M, R> r = a; r.v += s; return r; }
template <int M, typename R> __host__ __device__ __forceinline__ Dual<M, R> operator-(const Dual<M, R>& a, R s) { Dual<M, R> r = a; r.v -= s; return r; }
template <int M, typename R> __host__ __device__ __forceinline__ Dual<M, R> operator-(R s, const Dual<M, R>& a) { Dual<M, R> r = -a; r.v += s; return r; }
template <int M, typename R> __host__ __device__ __forceinline__ Dual<M, R> operator/(const Dual<M, R>& a, R s) { return a * recip(s); }
template <int M, typename R> __host__ __device__ __forceinline__ Dual<M, R> recip(const Dual<M, R>& b) { return dconst<M, R>(R(1)) / b; }
template <int M, typename R> __host__ __device__ __forceinline__ Dual<M, R> operator/(R s, const Dual<M, R>& b) { return dconst<M, R>(s) / b; }
template <int M, typename R> __host__ __device__ __forceinline__ Dual<M, R> dsqrt(const Dual<M, R>& a) {
  Dual<M, R> r; r.v = sqrt(a.v); R f = R(0.5) * recip(r.v);
#pragma unroll
  for (int i = 0; i < M; i++) r.d[i] = a.d[i] * f;
  return r;
}
template <int M, typename R> __host__ __device__ __forceinline__ Dual<M, R> dabs(const Dual<M, R>& a) { return a.v < R(0) ? -a : a; }
template <int M, typename R> __host__ __device__ __forceinline__ R value(const Dual<M, R>& a) { return a.v; }

__host__ __device__ __forceinline__ double dsqrt(double a) { return sqrt(a); }
__host__ __device__ __forceinline__ float dsqrt(float a) { return sqrtf(a); }
__host__ __device__ __forceinline__ double dabs(double a) { return fabs(a); }
__host__ __device__ __forceinline__ float dabs(float a) { return fabsf(a); }
__host__ __device__ __forceinline__ double value(double a) { return a; }
__host__ __device__ __forceinline__ float value(float a) { return a; }

template <class T> struct Scalar { typedef T type; };
template <int M, typename R> struct Scalar<Dual<M, R>> { typedef R type; };

template <class T> struct Make { static __host__ __device__ __forceinline__ T c(typename Scalar<T>::type x) { return T(x); } };
template <int M, typename R> struct Make<Dual<M, R>> { static __host__ __device__ __forceinline__ Dual<M, R> c(R x) { return dconst<M, R>(x); } };
template <class T> __host__ __device__ __forceinline__ T tconst(typename Scalar<T>::type x) { return Make<T>::c(x); }

// Math.Max(a, b) keeps a on ties (b wins only when strictly greater).
template <class T> __host__ __device__ __forceinline__ T vmax(const T& a, const T& b) { return value(b) > value(a) ? b : a; }


// ------------------------------------------------------------------------- SIP
// SIPLaplace (ipLaplace.cs:183-249) with constant Nu and g_D = 0 on every boundary edge.
struct FormSIP {
  static constexpr int V = 1;
  static constexpr bool VOL_GRADU = true;
  static constexpr bool FACE_GRADU = true;
  static constexpr bool FACE_TESTG = true;
  static constexpr bool HAS_BOUNDARY = true;

  template <int D, class T, class R>
  __host__ __device__ static void volume(const FormParams& P, const T* U, const T (*GU)[D], T (*F)[D]) {
    (void)U;
#pragma unroll
    for (int d = 0; d < D; d++) F[0][d] = GU[0][d] * R(-P.nu);   // -Nu * grad(u) . grad(v)
  }
  template <int D, class T, class R>
  __host__ __device__ static void inner(const FormParams& P, const R* n, R mu, const T* Uo, const T* Un, const T (*Go)[D],
                               const T (*Gn)[D], T* fo, T (*go)[D], T* fn, T (*gn)[D]) {
    const R nu = R(P.nu);
    T jump = Uo[0] - Un[0];
    T flux = (Go[0][0] + Gn[0][0]) * (R(0.5) * nu * n[0]);
#pragma unroll
    for (int d = 1; d < D; d++) flux = flux + (Go[0][d] + Gn[0][d]) * (R(0.5) * nu * n[d]);
    T pen = jump * (mu * nu);
    fo[0] = flux - pen;   // consistency - penalty     (coefficient of v_A)
    fn[0] = pen - flux;   //                            (coefficient of v_B)
#pragma unroll
    for (int d = 0; d < D; d++) {
      go[0][d] = jump * (R(0.5) * nu * n[d]);  // symmetry term (coefficient of grad v_A)
      gn[0][d] = go[0][d];                     //                (coefficient of grad v_B)
    }
  }
  template <int D, class T, class R>
  __host__ __device__ static void boundary(const FormParams& P, const R* n, R mu, const T* Uo, const T (*Go)[D], T* fo, T (*go)[D]) {
    const R nu = R(P.nu);
    T flux = Go[0][0] * (nu * n[0]);
#pragma unroll
    for (int d = 1; d < D; d++) flux = flux + Go[0][d] * (nu * n[d]);
    fo[0] = flux - Uo[0] * (nu * R(2) * mu);   // Dirichlet: pnlty = 2*GetPenalty(jIn,-1)
#pragma unroll
    for (int d = 0; d < D; d++) go[0][d] = Uo[0] * (nu * n[d]);
  }
};

// -------------------------------------------------------- upwind + SIP viscosity
// ScalarTransportFlux3D (TransportFlux.cs:112-190, NonlinearFlux base) + SIPLaplace.
struct FormAdvDiff {
  static constexpr int V = 1;
  static constexpr bool VOL_GRADU = true;
  static constexpr bool FACE_GRADU = true;
  static constexpr bool FACE_TESTG = true;
  static constexpr bool HAS_BOUNDARY = true;

  template <int D, class T, class R>
  __host__ __device__ static void volume(const FormParams& P, const T* U, const T (*GU)[D], T (*F)[D]) {
#pragma unroll
    for (int d = 0; d < D; d++) F[0][d] = U[0] * R(-P.vel[d]) + GU[0][d] * R(-P.nu);
  }
  template <int D, class T, class R>
  __host__ __device__ static void inner(const FormParams& P, const R* n, R mu, const T* Uo, const T* Un, const T (*Go)[D],
                               const T (*Gn)[D], T* fo, T (*go)[D], T* fn, T (*gn)[D]) {
    FormSIP::inner<D, T, R>(P, n, mu, Uo, Un, Go, Gn, fo, go, fn, gn);
    R un = R(P.vel[0]) * n[0];
#pragma unroll
    for (int d = 1; d < D; d++) un += R(P.vel[d]) * n[d];
    T up = (un > R(0)) ? Uo[0] * un : Un[0] * un;
    fo[0] = fo[0] + up;
    fn[0] = fn[0] - up;
  }
  template <int D, class T, class R>
  __host__ __device__ static void boundary(const FormParams& P, const R* n, R mu, const T* Uo, const T (*Go)[D], T* fo, T (*go)[D]) {
    FormSIP::boundary<D, T, R>(P, n, mu, Uo, Go, fo, go);
    R un = R(P.vel[0]) * n[0];
#pragma unroll
    for (int d = 1; d < D; d++) un += R(P.vel[d]) * n[d];
    if (un >= R(0)) fo[0] = fo[0] + Uo[0] * un;   // outflow; inflow value is 0
  }
};

// ----------------------------------------------------------------------- Euler
// StateVector.cs / IdealGas.cs / Euler{Density,Momentum,Energy}Component.cs
template <int D, class T, class R>
struct EulerState {
  T rho, m[D], E, u[D], p, a, irho;
  __host__ __device__ __forceinline__ EulerState(const T* U, const FormParams& P) {
    rho = U[0];
    irho = recip(rho);
#pragma unroll
    for (int d = 0; d < D; d++) { m[d] = U[1 + d]; u[d] = m[d] * irho; }
    E = U[D + 1];
    T msq = m[0] * m[0];
#pragma unroll
    for (int d = 1; d < D; d++) msq = msq + m[d] * m[d];
    const R ms = R(P.gamma * P.mach * P.mach);
    p = (E - msq * irho * (R(0.5) * ms)) * R(P.gamma - 1.0);
    a = dsqrt(p * irho) * R(1.0 / P.mach);
  }
  __host__ __device__ __forceinline__ T un(const R* n) const {
    T r = u[0] * n[0];
#pragma unroll
    for (int d = 1; d < D; d++) r = r + u[d] * n[d];
    return r;
  }
  // F(U).n for all V = D+2 components
  __host__ __device__ __forceinline__ void flux_n(const R* n, const T& unn, R ims, T* f) const {
    f[0] = rho * unn;
#pragma unroll
    for (int d = 0; d < D; d++) f[1 + d] = m[d] * unn + p * (ims * n[d]);
    f[D + 1] = unn * (E + p);
  }
};

template <int D, bool HLLC>
struct FormEuler {
  static constexpr int V = D + 2;
  static constexpr bool VOL_GRADU = false;
  static constexpr bool FACE_GRADU = false;
  static constexpr bool FACE_TESTG = false;
  static constexpr bool HAS_BOUNDARY = false;

  template <int DD, class T, class R>
  __host__ __device__ static void volume(const FormParams& P, const T* U, const T (*GU)[DD], T (*F)[DD]) {
    (void)GU;
    EulerState<D, T, R> s(U, P);
    const R ims = R(1.0 / (P.gamma * P.mach * P.mach));
    // NonlinearFlux volume fluxes enter with -1 (NECQuadratureVolume.cs:452)
#pragma unroll
    for (int d = 0; d < D; d++) {
      F[0][d] = -s.m[d];
#pragma unroll
      for (int c = 0; c < D; c++) {
        T f = s.m[c] * s.u[d];
        if (c == d) f = f + s.p * ims;
        F[1 + c][d] = -f;
      }
      F[D + 1][d] = -(s.u[d] * (s.E + s.p));
    }
  }

  template <int DD, class T, class R>
  __host__ __device__ static void inner(const FormParams& P, const R* n, R mu, const T* Uo, const T* Un, const T (*Go)[DD],
                               const T (*Gn)[DD], T* fo, T (*go)[DD], T* fn, T (*gn)[DD]) {
    (void)mu; (void)Go; (void)Gn; (void)go; (void)gn;
    EulerState<D, T, R> si(Uo, P), so(Un, P);
    const R ms = R(P.gamma * P.mach * P.mach);
    const R ims = R(1.0) / ms;
    T uni = si.un(n), uno = so.un(n);
    T fi[V], fx[V];
    if (!HLLC) {
      // RusanovFlux: 0.5 (F_in.n + F_out.n - S+ (U_out - U_in)), S+ = max(|u.n|+a)
      T wi = dabs(uni) + si.a, wo = dabs(uno) + so.a;
      T pen = vmax(wi, wo);
      si.flux_n(n, uni, ims, fi);
      so.flux_n(n, uno, ims, fx);
      fo[0] = (fi[0] + fx[0] - pen * (Un[0] - Uo[0])) * R(0.5);
#pragma unroll
      for (int c = 1; c < V; c++) fo[c] = (fi[c] + fx[c] - pen * (Un[c] - Uo[c])) * R(0.5);
    } else {
      // EulerFlux.EstimateWaveSpeeds (EulerFlux.cs:216-251)
      T meanP = (si.p + so.p) * R(0.5);
      T meanR = (si.rho + so.rho) * R(0.5);
      T meanA = (si.a + so.a) * R(0.5);
      T jump = uno - uni;
      const R g = R(P.gamma);
      T pStar = meanP - jump * meanA * meanR * (R(0.5) * ms);
      pStar = vmax(tconst<T>(R(0)), pStar);
      T qIn = tconst<T>(R(1)), qOut = tconst<T>(R(1));
      const R cq = R(0.5) * (g + R(1)) / g;
      if (value(pStar) > value(si.p)) qIn = dsqrt((pStar * recip(si.p) - R(1)) * cq + R(1));
      if (value(pStar) > value(so.p)) qOut = dsqrt((pStar * recip(so.p) - R(1)) * cq + R(1));
      T wIn = uni - si.a * qIn;
      T wOut = uno + so.a * qOut;
      // HLLCFlux.InnerEdgeFlux (HLLCFlux.cs:48-86, Toro 10.70/10.71)
      T cIn = si.rho * (wIn - uni);
      T cOut = so.rho * (wOut - uno);
      T sm = (cOut * uno - cIn * uni + (si.p - so.p) * ims) * recip(cOut - cIn);
      int branch;
      if (value(wIn) > R(0)) branch = 0;
      else if (R(0) < value(sm)) branch = 1;
      else if (R(0) <= value(wOut)) branch = 2;
      else branch = 3;
      const EulerState<D, T, R>& s = (branch <= 1) ? si : so;
      const T& uk = (branch <= 1) ? uni : uno;
      s.flux_n(n, uk, ims, fo);
      if (branch == 1 || branch == 2) {
        const T& wk = (branch == 1) ? wIn : wOut;
        // HLLC{Density,Momentum,Energy}Flux.GetModifiedVariableValue
        T A = s.rho * (wk - uk);
        T iA = recip(A);
        T base = A * recip(wk - sm);
        T dsu = sm - uk;
        fo[0] = fo[0] + wk * (base - s.rho);
#pragma unroll
        for (int c = 0; c < D; c++) fo[1 + c] = fo[1 + c] + wk * (base * (s.u[c] + dsu * n[c]) - s.m[c]);
        T factor = sm * ms + s.p * iA;
        fo[D + 1] = fo[D + 1] + wk * (base * (s.E * s.irho + factor * dsu) - s.E);
      }
    }
#pragma unroll
    for (int c = 0; c < V; c++) fn[c] = -fo[c];
  }
  template <int DD, class T, class R>
  __host__ __device__ static void boundary(const FormParams&, const R*, R, const T*, const T (*)[DD], T* fo, T (*)[DD]) {
#pragma unroll
    for (int c = 0; c < V; c++) fo[c] = fo[c] * R(0);
  }
};

}  // namespace bosss
