// Hand-written reverse-mode (adjoint) derivatives of the pointwise Euler fluxes.
//
// For the cell-centric VJP each thread needs, per quadrature node, the
// transposed Jacobian of the flux with respect to its own cell's state applied
// to a cotangent.  Forward-mode duals seeded on the 5 state components carry
// 6x every intermediate; the reverse sweeps below carry one adjoint per
// intermediate, which is what keeps the VJP kernels out of register spills.
//
// Forward formulas (same as forms.hpp, same reference files):
//   StateVector.cs / IdealGas.cs:81-134   p = (g-1)(E - g Ma^2 |m|^2/(2 rho)),  a = sqrt(p/rho)/Ma
//   Euler{Density,Momentum,Energy}Component.Flux
//   RusanovFlux.cs:55-65, EulerFlux.cs:216-251 (wave speeds), HLLCFlux.cs:48-86 (+ HLLC*Flux.cs)
// Branch choices (|u.n|, max, HLLC regions) follow the forward values, i.e. the
// derivative of the active branch (a consistent subgradient at the kinks).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "forms.hpp"

namespace bosss {


template <int D, typename R>
struct EulerPrim {
  R rho, m[D], E, irho, u[D], msq, p, a, un;
  __host__ __device__ __forceinline__ void set(const R* U, const R* n, const FormParams& P) {
    rho = U[0];
    irho = recip(rho);
    msq = R(0);
#pragma unroll
    for (int d = 0; d < D; d++) { m[d] = U[1 + d]; u[d] = m[d] * irho; msq += m[d] * m[d]; }
    E = U[D + 1];
    const R ms = R(P.gamma * P.mach * P.mach);
    p = (E - msq * irho * (R(0.5) * ms)) * R(P.gamma - 1.0);
    a = sqrt(p * irho) * R(1.0 / P.mach);
    un = R(0);
#pragma unroll
    for (int d = 0; d < D; d++) un += u[d] * n[d];
  }
};

// Adjoints of the primitive quantities of one state, pulled back to (rho, m, E).
template <int D, typename R>
struct EulerPrimBar {
  R rho, m[D], E, u[D], p, a, un;
  __host__ __device__ __forceinline__ void zero() {
    rho = E = p = a = un = R(0);
#pragma unroll
    for (int d = 0; d < D; d++) { m[d] = R(0); u[d] = R(0); }
  }
  // accumulate into Ubar[0..D+1].  SOUND = false: the sound speed has no adjoint (the
  // volume flux does not use it), so neither a nor its derivative is formed.
  template <bool SOUND = true>
  __host__ __device__ __forceinline__ void pullback(const EulerPrim<D, R>& s, const R* n, const FormParams& P, R* Ubar) const {
    const R g1 = R(P.gamma - 1.0);
    const R ms = R(P.gamma * P.mach * P.mach);
    R ub[D];
#pragma unroll
    for (int d = 0; d < D; d++) ub[d] = u[d] + un * n[d];
    R pb = p, irb = R(0);
    if constexpr (SOUND) {
      // a = sqrt(q)/Ma, q = p irho
      R qb = a * R(0.5 / (P.mach * P.mach)) * recip(s.a);
      pb += qb * s.irho;
      irb += qb * s.p;
    }
    // p = g1 (E - msq irho ms/2)
    R Eb = E + g1 * pb;
    R msqb = -g1 * R(0.5) * ms * s.irho * pb;
    irb += -g1 * R(0.5) * ms * s.msq * pb;
    R mb[D];
#pragma unroll
    for (int d = 0; d < D; d++) { mb[d] = m[d] + ub[d] * s.irho + R(2) * s.m[d] * msqb; irb += ub[d] * s.m[d]; }
    Ubar[0] += rho - s.irho * s.irho * irb;
#pragma unroll
    for (int d = 0; d < D; d++) Ubar[1 + d] += mb[d];
    Ubar[D + 1] += Eb;
  }
};

// F(s).n and its reverse: f0 = rho un, f_{1+c} = m_c un + p ims n_c, f_{D+1} = un (E + p)
template <int D, typename R>
__host__ __device__ __forceinline__ void flux_n(const EulerPrim<D, R>& s, const R* n, R ims, R* f) {
  f[0] = s.rho * s.un;
#pragma unroll
  for (int c = 0; c < D; c++) f[1 + c] = s.m[c] * s.un + s.p * (ims * n[c]);
  f[D + 1] = s.un * (s.E + s.p);
}
template <int D, typename R>
__host__ __device__ __forceinline__ void flux_n_bar(const EulerPrim<D, R>& s, const R* n, R ims, const R* fb, EulerPrimBar<D, R>& sb) {
  sb.rho += fb[0] * s.un;
  sb.un += fb[0] * s.rho;
#pragma unroll
  for (int c = 0; c < D; c++) {
    sb.m[c] += fb[1 + c] * s.un;
    sb.un += fb[1 + c] * s.m[c];
    sb.p += fb[1 + c] * ims * n[c];
  }
  sb.un += fb[D + 1] * (s.E + s.p);
  sb.E += fb[D + 1] * s.un;
  sb.p += fb[D + 1] * s.un;
}

template <int D, typename R>
struct EulerAdjoint {
  static constexpr int V = D + 2;

  // Volume: F[c][d] = -flux_c,d(U)  (NonlinearFlux scaled by -1).  Given Fbar[c][d],
  // Ubar += (dF/dU)^T Fbar.
  __host__ __device__ static void volume_vjp(const FormParams& P, const R* U, const R (*Fbar)[D], R* Ubar) {
    R zero_n[D];
#pragma unroll
    for (int d = 0; d < D; d++) zero_n[d] = R(0);
    EulerPrim<D, R> s;
    s.set(U, zero_n, P);
    EulerPrimBar<D, R> sb;
    sb.zero();
    const R ims = R(1.0 / (P.gamma * P.mach * P.mach));
    // flux_{0,d} = m_d; flux_{1+c,d} = m_c u_d + delta_cd p ims; flux_{D+1,d} = u_d (E + p)
#pragma unroll
    for (int d = 0; d < D; d++) {
      R g0 = -Fbar[0][d];
      sb.m[d] += g0;
#pragma unroll
      for (int c = 0; c < D; c++) {
        R gc = -Fbar[1 + c][d];
        sb.m[c] += gc * s.u[d];
        sb.u[d] += gc * s.m[c];
        if (c == d) sb.p += gc * ims;
      }
      R ge = -Fbar[D + 1][d];
      sb.u[d] += ge * (s.E + s.p);
      sb.E += ge * s.u[d];
      sb.p += ge * s.u[d];
    }
    sb.template pullback<false>(s, zero_n, P, Ubar);
  }

  // Volume term of one reference direction: Fbar[c][d] = gW[c] k[d] (rank one: the W gradient
  // along the direction times a row of the inverse Jacobian, times the weight).  Then
  // sum_{c,d} Fbar[c][d] dF[c][d]/dU = -d/dU sum_c gW[c] (flux(U) k)_c, and flux(U) k is the
  // normal flux flux_n with n = k, so the adjoint is flux_n_bar along k: the same result as
  // volume_vjp for this Fbar without the D x D loop.
  __host__ __device__ static void volume_vjp_dir(const FormParams& P, const R* U, const R* k, const R* gW, R* Ubar) {
    EulerPrim<D, R> s;
    s.set(U, k, P);
    EulerPrimBar<D, R> sb;
    sb.zero();
    const R ims = R(1.0 / (P.gamma * P.mach * P.mach));
    R fb[V];
#pragma unroll
    for (int c = 0; c < V; c++) fb[c] = -gW[c];
    flux_n_bar(s, k, ims, fb, sb);
    sb.template pullback<false>(s, k, P, Ubar);
  }

  // Inner face, own side i, neighbour o, normal n out of i.  fo = Fhat(Ui, Uo, n).
  // Writes fo (forward value) and accumulates Ubar_i += (dfo/dUi)^T gbar.
  template <bool HLLC>
  __host__ __device__ static void inner_vjp(const FormParams& P, const R* n, const R* Ui, const R* Uo, const R* gbar, R* fo, R* Ubar) {
    EulerPrim<D, R> si, so;
    si.set(Ui, n, P);
    so.set(Uo, n, P);
    EulerPrimBar<D, R> bi, bo;
    bi.zero();
    bo.zero();
    const R ms = R(P.gamma * P.mach * P.mach);
    const R ims = R(1) / ms;
    if (!HLLC) {
      R fi[V], fx[V];
      flux_n(si, n, ims, fi);
      flux_n(so, n, ims, fx);
      R wi = fabs(si.un) + si.a, wo = fabs(so.un) + so.a;
      bool take_o = wo > wi;
      R pen = take_o ? wo : wi;
      R penb = R(0);
      R fib[V];
#pragma unroll
      for (int c = 0; c < V; c++) {
        R du = Uo[c] - Ui[c];
        fo[c] = (fi[c] + fx[c] - pen * du) * R(0.5);
        fib[c] = R(0.5) * gbar[c];
        penb += -R(0.5) * gbar[c] * du;
        Ubar[c] += R(0.5) * pen * gbar[c];
      }
      flux_n_bar(si, n, ims, fib, bi);
      if (!take_o) {
        bi.un += (si.un < R(0) ? -penb : penb);
        bi.a += penb;
      }
      bi.pullback(si, n, P, Ubar);
      return;
    }
    // ---------------- HLLC forward
    const R g = R(P.gamma);
    const R h = R(0.5) * ms;
    R meanP = (si.p + so.p) * R(0.5);
    R meanR = (si.rho + so.rho) * R(0.5);
    R meanA = (si.a + so.a) * R(0.5);
    R jump = so.un - si.un;
    R pStar0 = meanP - jump * meanA * meanR * h;
    bool pos = pStar0 > R(0);
    R pStar = pos ? pStar0 : R(0);
    const R cq = R(0.5) * (g + R(1)) / g;
    bool qi_on = pStar > si.p, qo_on = pStar > so.p;
    // shared reciprocals: every quotient below reuses one of ipi, ipo, iden, iA, iB, iqI, iqO
    const R ipi = recip(si.p), ipo = recip(so.p);
    R qI = qi_on ? sqrt((pStar * ipi - R(1)) * cq + R(1)) : R(1);
    R qO = qo_on ? sqrt((pStar * ipo - R(1)) * cq + R(1)) : R(1);
    R wIn = si.un - si.a * qI;
    R wOut = so.un + so.a * qO;
    R cIn = si.rho * (wIn - si.un);
    R cOut = so.rho * (wOut - so.un);
    R num = cOut * so.un - cIn * si.un + (si.p - so.p) * ims;
    R den = cOut - cIn;
    const R iden = recip(den);
    R sm = num * iden;
    int branch;
    if (wIn > R(0)) branch = 0;
    else if (R(0) < sm) branch = 1;
    else if (R(0) <= wOut) branch = 2;
    else branch = 3;
    if (branch == 0 || branch == 3) {
      const EulerPrim<D, R>& s = branch == 0 ? si : so;
      flux_n(s, n, ims, fo);
      if (branch == 0) {
        flux_n_bar(si, n, ims, gbar, bi);
        bi.pullback(si, n, P, Ubar);
      }
      return;   // branch 3 does not depend on the own state
    }
    const bool kin = (branch == 1);
    const EulerPrim<D, R>& s = kin ? si : so;
    EulerPrimBar<D, R>& sb = kin ? bi : bo;
    const R* UK = kin ? Ui : Uo;
    const R wk = kin ? wIn : wOut;
    R A = s.rho * (wk - s.un);          // also the z of HLLCFlux's energy factor
    R B = wk - sm;
    const R iA = recip(A), iB = recip(B);
    R base = A * iB;
    R dsu = sm - s.un;
    R factor = sm * ms + s.p * iA;
    R Er = s.E * s.irho;
    R us[V];
    us[0] = base;
#pragma unroll
    for (int c = 0; c < D; c++) us[1 + c] = base * (s.u[c] + dsu * n[c]);
    us[D + 1] = base * (Er + factor * dsu);
    flux_n(s, n, ims, fo);
#pragma unroll
    for (int c = 0; c < V; c++) fo[c] = fo[c] + wk * (us[c] - UK[c]);
    // ---------------- reverse
    R wkb = R(0), smb = R(0);
    flux_n_bar(s, n, ims, gbar, sb);
    R usb[V];
#pragma unroll
    for (int c = 0; c < V; c++) {
      wkb += gbar[c] * (us[c] - UK[c]);
      usb[c] = wk * gbar[c];
    }
    // direct U_K dependence (the -U_K term); only the own side is returned
    if (kin) {
#pragma unroll
      for (int c = 0; c < V; c++) Ubar[c] += -wk * gbar[c];
    }
    R baseb = usb[0], dsub = R(0);
#pragma unroll
    for (int c = 0; c < D; c++) {
      baseb += usb[1 + c] * (s.u[c] + dsu * n[c]);
      sb.u[c] += usb[1 + c] * base;
      dsub += usb[1 + c] * base * n[c];
    }
    {
      R tb = usb[D + 1] * base;
      baseb += usb[D + 1] * (Er + factor * dsu);
      sb.E += tb * s.irho;
      sb.rho += -tb * Er * s.irho;
      R factorb = tb * dsu;
      dsub += tb * factor;
      smb += factorb * ms;
      sb.p += factorb * iA;
      R zb = -factorb * s.p * iA * iA;
      sb.rho += zb * (wk - s.un);
      wkb += zb * s.rho;
      sb.un += -zb * s.rho;
    }
    smb += dsub;
    sb.un += -dsub;
    {
      R Ab = baseb * iB;
      R Bb = -baseb * base * iB;
      sb.rho += Ab * (wk - s.un);
      wkb += Ab * s.rho;
      sb.un += -Ab * s.rho;
      wkb += Bb;
      smb += -Bb;
    }
    R wInb = kin ? wkb : R(0), wOutb = kin ? R(0) : wkb;
    // sm = num / den
    R numb = smb * iden;
    R denb = -smb * sm * iden;
    R cOutb = numb * so.un + denb;
    R cInb = -numb * si.un - denb;
    bo.un += numb * cOut;
    bi.un += -numb * cIn;
    bi.p += numb * ims;
    bo.p += -numb * ims;
    // cIn = rho_i (wIn - un_i), cOut = rho_o (wOut - un_o)
    bi.rho += cInb * (wIn - si.un);
    wInb += cInb * si.rho;
    bi.un += -cInb * si.rho;
    bo.rho += cOutb * (wOut - so.un);
    wOutb += cOutb * so.rho;
    bo.un += -cOutb * so.rho;
    // wIn = un_i - a_i qI, wOut = un_o + a_o qO
    bi.un += wInb;
    bi.a += -wInb * qI;
    R qIb = -wInb * si.a;
    bo.un += wOutb;
    bo.a += wOutb * qO;
    R qOb = wOutb * so.a;
    R pStarb = R(0);
    if (qi_on) {
      R yb = qIb * R(0.5) * recip(qI);
      pStarb += yb * cq * ipi;
      bi.p += -yb * cq * pStar * ipi * ipi;
    }
    if (qo_on) {
      R yb = qOb * R(0.5) * recip(qO);
      pStarb += yb * cq * ipo;
      bo.p += -yb * cq * pStar * ipo * ipo;
    }
    R p0b = pos ? pStarb : R(0);
    R meanPb = p0b;
    R jumpb = -p0b * meanA * meanR * h;
    R meanAb = -p0b * jump * meanR * h;
    R meanRb = -p0b * jump * meanA * h;
    bi.p += R(0.5) * meanPb;
    bo.p += R(0.5) * meanPb;
    bi.rho += R(0.5) * meanRb;
    bo.rho += R(0.5) * meanRb;
    bi.a += R(0.5) * meanAb;
    bo.a += R(0.5) * meanAb;
    bo.un += jumpb;
    bi.un += -jumpb;
    bi.pullback(si, n, P, Ubar);
  }
};

}  // namespace bosss
