// Table of compiled kernel instantiations, keyed by (form, D, N, 1D Gauss points).
#pragma once
#include <hip/hip_runtime.h>

#include "dg_kernels.hpp"
#include "dg_tensor.hpp"

namespace bosss {

struct KernelSet {
  enum { RES = 0, JVP = 1, VJP = 2, RES_VJP = 3, NVARIANT = 4 };
  static constexpr int BLOCK = 256;
  int kind, D, N, K1;
  int kvp, kep, table_size;
  bool facegrad;
  int o_vphi, o_vdphi, o_vw, o_tr, o_dtr, o_fw;
  size_t lds64, lds32;
  void (*k64[NVARIANT])(KArgs<double>);
  void (*k32[NVARIANT])(KArgs<float>);
  // tensor-table kernels (dg_tensor.hpp): 1D tables in LDS, own coefficients staged in LDS
  int P, ttab_size, t_o_l, t_o_dl, t_o_w, t_o_code, t_k2;
  int t_o_ft, t_frows, t_fs, t_ftab;   // face tables: offset (after the 1D block), rows per face, row stride, size of one
  void (*kt64[NVARIANT])(TArgs<double>);
  void (*kt32[NVARIANT])(TArgs<float>);
  size_t ktlds64[NVARIANT], ktlds32[NVARIANT];
  int prepare() const;
};

constexpr int degree_from_modes(int D, int N) {
  int p = 0;
  while (tdeg_count(D, p) < N) p++;
  return p;
}

const KernelSet* find_kernels(int kind, int D, int N, int K1, int precision);

// JVP, VJP and fused residual+VJP kernels: the Euler forms (hand-written adjoint)
// use the LDS node-buffer kernels, the forms with forward-mode duals the generic ones.
template <class Form, int D, int P, int K1, typename Real>
void set_tensor_linearisations(void (**kt)(TArgs<Real>), size_t* lds) {
  if constexpr (HandAdjoint<Form>::value) {
    kt[KernelSet::JVP] = kt_jvp_euler<Form, D, P, K1, Real>;
    kt[KernelSet::VJP] = kt_vjp_euler<Form, D, P, K1, Real, false>;
    kt[KernelSet::RES_VJP] = kt_vjp_euler<Form, D, P, K1, Real, true>;
    lds[KernelSet::JVP] = lds[KernelSet::VJP] = lds[KernelSet::RES_VJP] = kt_vjp_euler_lds_bytes<D, P, K1, Form::V, Real>();
  } else {
    kt[KernelSet::JVP] = kt_forward<Form, D, P, K1, Real, MODE_JVP>;
    kt[KernelSet::VJP] = kt_adjoint<Form, D, P, K1, Real, false>;
    kt[KernelSet::RES_VJP] = kt_adjoint<Form, D, P, K1, Real, true>;
    lds[KernelSet::JVP] = kt_forward_lds_bytes<Form, D, P, K1, Real>(2);
    lds[KernelSet::VJP] = lds[KernelSet::RES_VJP] = kt_forward_lds_bytes<Form, D, P, K1, Real>(2);
  }
}

template <class Form, int D, int N, int K1, bool WITH32>
KernelSet make_kernel_set(int kind) {
  typedef Shape<D, N, K1, Form::FACE_GRADU || Form::FACE_TESTG> S;
  KernelSet k;
  k.kind = kind; k.D = D; k.N = N; k.K1 = K1;
  k.facegrad = Form::FACE_GRADU || Form::FACE_TESTG;
  k.kvp = S::KVP; k.kep = S::KEP; k.table_size = S::SIZE;
  k.o_vphi = S::O_VPHI; k.o_vdphi = S::O_VDPHI; k.o_vw = S::O_VW; k.o_tr = S::O_TR; k.o_dtr = S::O_DTR; k.o_fw = S::O_FW;
  k.lds64 = (size_t)S::SIZE * sizeof(double);
  k.lds32 = (size_t)S::SIZE * sizeof(float);
  k.k64[KernelSet::RES] = k_forward<Form, D, N, K1, double, MODE_RES>;
  k.k64[KernelSet::JVP] = k_forward<Form, D, N, K1, double, MODE_JVP>;
  k.k64[KernelSet::VJP] = k_adjoint<Form, D, N, K1, double, false>;
  k.k64[KernelSet::RES_VJP] = k_adjoint<Form, D, N, K1, double, true>;
  for (int i = 0; i < KernelSet::NVARIANT; i++) { k.k32[i] = nullptr; k.kt32[i] = nullptr; k.ktlds32[i] = 0; }
  constexpr int P = degree_from_modes(D, N);
  static_assert(tdeg_count(D, P) == N, "N must be a full total-degree basis");
  typedef TShape<D, P, K1> TSh;
  k.P = P; k.ttab_size = TSh::TAB + TSh::NFTAB * TSh::FTAB; k.t_o_l = TSh::O_L; k.t_o_dl = TSh::O_DL; k.t_o_w = TSh::O_W; k.t_o_code = TSh::O_CODE;
  k.t_k2 = TSh::K2;
  k.t_o_ft = TSh::TAB; k.t_frows = TSh::FROWS; k.t_fs = TSh::FS; k.t_ftab = TSh::FTAB;
  k.kt64[KernelSet::RES] = kt_forward<Form, D, P, K1, double, MODE_RES>;
  k.ktlds64[KernelSet::RES] = kt_forward_lds_bytes<Form, D, P, K1, double>(1);
  set_tensor_linearisations<Form, D, P, K1, double>(k.kt64, k.ktlds64);
  if constexpr (WITH32) {
    k.kt32[KernelSet::RES] = kt_forward<Form, D, P, K1, float, MODE_RES>;
    k.ktlds32[KernelSet::RES] = kt_forward_lds_bytes<Form, D, P, K1, float>(1);
    set_tensor_linearisations<Form, D, P, K1, float>(k.kt32, k.ktlds32);
    k.k32[KernelSet::RES] = k_forward<Form, D, N, K1, float, MODE_RES>;
    k.k32[KernelSet::JVP] = k_forward<Form, D, N, K1, float, MODE_JVP>;
    k.k32[KernelSet::VJP] = k_adjoint<Form, D, N, K1, float, false>;
    k.k32[KernelSet::RES_VJP] = k_adjoint<Form, D, N, K1, float, true>;
  }
  return k;
}

// each translation unit kernels_<family>.hip provides one of these
void register_sip(std::vector<KernelSet>& out);
void register_advdiff(std::vector<KernelSet>& out);
void register_euler(std::vector<KernelSet>& out);

}  // namespace bosss
