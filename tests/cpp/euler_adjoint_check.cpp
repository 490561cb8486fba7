// Host-side check of the hand-written Euler flux adjoints (euler_adjoint.hpp)
// against forward-mode duals through the forward flux code (forms.hpp).
// Prints the worst relative mismatch per case and the HLLC branches visited;
// exits non-zero on a mismatch above 1e-12.
#include <cmath>
#include <cstdio>
#include <random>

#include "../../bosss_amd/csrc/euler_adjoint.hpp"

using namespace bosss;
constexpr int D = 3, V = 5;
typedef Dual<V, double> T;

static void random_state(std::mt19937_64& g, double* U, double vmax, const FormParams& P) {
  std::uniform_real_distribution<double> r(0.8, 1.2), v(-vmax, vmax), p(0.9, 1.1);
  double rho = r(g), u[3] = {v(g), v(g), v(g)}, pr = p(g);
  U[0] = rho;
  for (int d = 0; d < 3; d++) U[1 + d] = rho * u[d];
  double ms = P.gamma * P.mach * P.mach;
  U[4] = pr / (P.gamma - 1) + 0.5 * ms * rho * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
}

int main() {
  FormParams P{};
  P.gamma = 1.4;
  P.mach = 1.0 / std::sqrt(1.4);
  std::mt19937_64 g(1);
  std::uniform_real_distribution<double> un(-1, 1), wd(-1, 1);
  double worst_vol = 0, worst_lf = 0, worst_hllc = 0;
  int branches[4] = {0, 0, 0, 0};
  for (int trial = 0; trial < 4000; trial++) {
    double vmax = (trial % 4 == 0) ? 2.0 : (trial % 4 == 1 ? 0.1 : 1.0);  // include supersonic states
    double Ui[V], Uo[V];
    random_state(g, Ui, vmax, P);
    random_state(g, Uo, vmax, P);
    double n[D] = {un(g), un(g), un(g)};
    double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int d = 0; d < D; d++) n[d] /= nn;
    // ---- volume
    double Fb[V][D];
    for (int c = 0; c < V; c++) for (int d = 0; d < D; d++) Fb[c][d] = wd(g);
    double Ub[V] = {0, 0, 0, 0, 0};
    EulerAdjoint<D, double>::volume_vjp(P, Ui, Fb, Ub);
    T Ud[V], GU[V][D], F[V][D];
    for (int c = 0; c < V; c++) { Ud[c] = dconst<V, double>(Ui[c]); Ud[c].d[c] = 1.0; for (int d = 0; d < D; d++) GU[c][d] = Ud[c]; }
    FormEuler<D, true>::volume<D, T, double>(P, Ud, GU, F);
    for (int q = 0; q < V; q++) {
      double ref = 0;
      for (int c = 0; c < V; c++) for (int d = 0; d < D; d++) ref += Fb[c][d] * F[c][d].d[q];
      worst_vol = std::fmax(worst_vol, std::fabs(ref - Ub[q]) / (std::fabs(ref) + 1e-3));
    }
    // ---- volume, one reference direction: Fbar = gW (x) k through volume_vjp_dir
    {
      double gW[V], k[D], Fk[V][D], Ub1[V] = {0, 0, 0, 0, 0}, Ub2[V] = {0, 0, 0, 0, 0};
      for (int c = 0; c < V; c++) gW[c] = wd(g);
      for (int d = 0; d < D; d++) k[d] = wd(g);
      for (int c = 0; c < V; c++) for (int d = 0; d < D; d++) Fk[c][d] = gW[c] * k[d];
      EulerAdjoint<D, double>::volume_vjp_dir(P, Ui, k, gW, Ub1);
      EulerAdjoint<D, double>::volume_vjp(P, Ui, Fk, Ub2);
      for (int q = 0; q < V; q++) {
        double ref = 0;
        for (int c = 0; c < V; c++) for (int d = 0; d < D; d++) ref += Fk[c][d] * F[c][d].d[q];
        worst_vol = std::fmax(worst_vol, std::fabs(ref - Ub1[q]) / (std::fabs(ref) + 1e-3));
        worst_vol = std::fmax(worst_vol, std::fabs(Ub2[q] - Ub1[q]) / (std::fabs(Ub2[q]) + 1e-3));
      }
    }
    // ---- faces
    double gb[V];
    for (int c = 0; c < V; c++) gb[c] = wd(g);
    for (int hllc = 0; hllc < 2; hllc++) {
      double fo[V], Ubi[V] = {0, 0, 0, 0, 0};
      if (hllc) EulerAdjoint<D, double>::inner_vjp<true>(P, n, Ui, Uo, gb, fo, Ubi);
      else EulerAdjoint<D, double>::inner_vjp<false>(P, n, Ui, Uo, gb, fo, Ubi);
      T Ti[V], To[V], Gi[V][D], Go[V][D], fd[V], gd[V][D], fnd[V], gnd[V][D];
      for (int c = 0; c < V; c++) {
        Ti[c] = dconst<V, double>(Ui[c]); Ti[c].d[c] = 1.0;
        To[c] = dconst<V, double>(Uo[c]);
        for (int d = 0; d < D; d++) { Gi[c][d] = Ti[c]; Go[c][d] = To[c]; }
      }
      if (hllc) FormEuler<D, true>::inner<D, T, double>(P, n, 0.0, Ti, To, Gi, Go, fd, gd, fnd, gnd);
      else FormEuler<D, false>::inner<D, T, double>(P, n, 0.0, Ti, To, Gi, Go, fd, gd, fnd, gnd);
      double& worst = hllc ? worst_hllc : worst_lf;
      for (int c = 0; c < V; c++) worst = std::fmax(worst, std::fabs(fd[c].v - fo[c]) / (std::fabs(fd[c].v) + 1e-3));
      for (int q = 0; q < V; q++) {
        double ref = 0;
        for (int c = 0; c < V; c++) ref += gb[c] * fd[c].d[q];
        worst = std::fmax(worst, std::fabs(ref - Ubi[q]) / (std::fabs(ref) + 1e-3));
      }
      if (hllc) {
        // which branch?  (recompute from the forward values)
        double uni = 0, uno = 0;
        EulerPrim<D, double> si, so; si.set(Ui, n, P); so.set(Uo, n, P);
        (void)uni; (void)uno;
        double ms = P.gamma * P.mach * P.mach;
        double meanP = 0.5 * (si.p + so.p), meanR = 0.5 * (si.rho + so.rho), meanA = 0.5 * (si.a + so.a);
        double pS = std::fmax(0.0, meanP - (so.un - si.un) * meanA * meanR * 0.5 * ms);
        double cq = 0.5 * (P.gamma + 1) / P.gamma;
        double qI = pS > si.p ? std::sqrt((pS / si.p - 1) * cq + 1) : 1, qO = pS > so.p ? std::sqrt((pS / so.p - 1) * cq + 1) : 1;
        double wIn = si.un - si.a * qI, wOut = so.un + so.a * qO;
        double cIn = si.rho * (wIn - si.un), cOut = so.rho * (wOut - so.un);
        double sm = (cOut * so.un - cIn * si.un + (si.p - so.p) / ms) / (cOut - cIn);
        branches[wIn > 0 ? 0 : (0 < sm ? 1 : (0 <= wOut ? 2 : 3))]++;
      }
    }
  }
  std::printf("volume %.3e  rusanov %.3e  hllc %.3e  (hllc branches %d %d %d %d)\n", worst_vol, worst_lf, worst_hllc, branches[0],
              branches[1], branches[2], branches[3]);
  bool ok = worst_vol < 1e-12 && worst_lf < 1e-12 && worst_hllc < 1e-12 && branches[0] && branches[1] && branches[2] && branches[3];
  return ok ? 0 : 1;
}
