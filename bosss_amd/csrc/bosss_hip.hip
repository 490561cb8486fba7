// C ABI of the gfx950 DG operator library (declarations: include/bosss_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <map>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/bosss_hip.h"
#include "devbuf.hpp"
#include "dg_kernels.hpp"
#include "registry.hpp"

using namespace bosss;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(x)                                                                                  \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess) return fail(BOSSS_ERR_HIP, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

template <class T>
int upload(T** dst, const T* src, size_t n) {
  HIPCHK(hipMalloc((void**)dst, n * sizeof(T) + 16));
  if (src) HIPCHK(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return BOSSS_OK;
}

}  // namespace

struct bosss_ctx {
  int device;
  hipStream_t stream;
  int num_cu;
  int nxcd;   // XCDs of the device (gfx950: 32 CUs each); BOSSS_HIP_XCDS overrides, 1 disables the block remapping
};

struct bosss_grid {
  bosss_ctx* ctx;
  int D, J, E, NF;
  bool has_boundary;
  std::vector<int> nb_host;  // [J][NF], for host-side colouring
  int plane;        // cell-index offset across the slowest grid direction (0: none found)
  double *detJ, *invJ, *scaling, *lenscale;
  int* nb;          // [J][NF]
  double* sqrtg;    // [J][NF]
  double* normal;   // [J][NF][D]
  int pair[6];
};

struct bosss_op {
  bosss_ctx* ctx;
  const bosss_grid* grid;
  int kind, D, N, K1, V, precision;
  const KernelSet* ks;
  void* tables;      // packed tables (Real) on the device
  void* ttab;        // 1D tensor tables (Real) on the device, if the uploaded tables are a tensor product
  bool tensor;
  FormParams fp;
  size_t ndof;
  // staging buffers for the host-pointer entry points
  double *hU, *hR, *hX, *hY;
  // matrix assembly: distance-2 colouring and block-ELL columns (built on first use)
  int ncolors;
  int32_t *d_color, *d_cols;
  std::vector<int32_t> cols_host;
  // tile shape of the block-ELL apply (panel size NL): tuned on the first device apply
  bool spmv_tuned = false;
  int spmv_nl = 0;
  std::vector<int> spmv_cand;
  std::vector<float> spmv_ms;
  // CheckForNanOrInf of every output (off by default): index of the first non-finite entry
  bool check_finite = false;
  unsigned long long* d_first = nullptr;
};

namespace bosss {
void launch_probe(int J, int Nt, const int32_t* color, int k, int q, double* X, hipStream_t s);
void launch_extract(int J, int Nt, int S, const int32_t* cols, const int32_t* color, int k, int q, const double* Y, double alpha,
                    double beta, double* blocks, hipStream_t s);
void launch_bell_spmv(int J, int Nt, int S, const int32_t* cols, const double* blocks, double alpha, const double* x, double beta,
                      double* y, hipStream_t s, int nl);
int bell_default_nl(int Nt);
void bell_candidates(int Nt, std::vector<int>& out);
}

namespace bosss {
int ctx_device_stream(bosss_ctx* c, int* device, hipStream_t* stream) {
  if (!c) return fail(BOSSS_ERR_ARG, "ctx is NULL");
  *device = c->device;
  *stream = c->stream;
  return BOSSS_OK;
}
}  // namespace bosss

extern "C" {

const char* bosss_hip_last_error(void) { return g_err.c_str(); }

int bosss_hip_ctx_create(int device, bosss_ctx** ctx) {
  if (!ctx) return fail(BOSSS_ERR_ARG, "ctx is NULL");
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(BOSSS_ERR_ARG, "device %d out of range (%d devices)", device, n);
  HIPCHK(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(BOSSS_ERR_UNSUPPORTED, "device %d is %s; this library is built for gfx950 only", device, prop.gcnArchName);
  bosss_ctx* c = new bosss_ctx();
  c->device = device;
  c->num_cu = prop.multiProcessorCount;
  c->nxcd = c->num_cu >= 64 ? c->num_cu / 32 : 1;
  if (const char* x = std::getenv("BOSSS_HIP_XCDS")) c->nxcd = std::max(1, std::atoi(x));
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete c; return fail(BOSSS_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e)); }
  *ctx = c;
  return BOSSS_OK;
}

int bosss_hip_ctx_destroy(bosss_ctx* ctx) {
  if (!ctx) return BOSSS_OK;
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return BOSSS_OK;
}

int bosss_hip_sync(bosss_ctx* ctx) {
  if (!ctx) return fail(BOSSS_ERR_ARG, "ctx is NULL");
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return BOSSS_OK;
}

void* bosss_hip_ctx_stream(bosss_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int bosss_hip_grid_upload(bosss_ctx* ctx, const bosss_grid_desc* g, bosss_grid** out) {
  if (!ctx || !g || !out) return fail(BOSSS_ERR_ARG, "NULL argument");
  const int D = g->D, J = g->J, E = g->E;
  if (D < 2 || D > 3 || J <= 0 || E < 0) return fail(BOSSS_ERR_ARG, "bad sizes D=%d J=%d E=%d", D, J, E);
  if (!g->detJ || !g->invJ || !g->scaling || !g->lenscale || !g->cells || !g->trafo || !g->normals || !g->sqrtgram)
    return fail(BOSSS_ERR_ARG, "grid descriptor has NULL arrays");
  const int NF = 2 * D;
  // Cell-centric face table: local face slot = trafo index of the cell's side.
  std::vector<int> nb((size_t)J * NF, -2);
  std::vector<double> sg((size_t)J * NF, 0.0), nrm((size_t)J * NF * D, 0.0);
  int pair[6] = {-1, -1, -1, -1, -1, -1};
  bool has_boundary = false;
  for (int e = 0; e < E; e++) {
    int ci = g->cells[2 * e], co = g->cells[2 * e + 1];
    int ti = g->trafo[2 * e], to = g->trafo[2 * e + 1];
    if (ci < 0 || ci >= J || co >= J || ti < 0 || ti >= NF) return fail(BOSSS_ERR_ARG, "edge %d: bad cell/trafo index", e);
    if (nb[(size_t)ci * NF + ti] != -2) return fail(BOSSS_ERR_UNSUPPORTED, "cell %d has two edges on face %d (non-conforming grid)", ci, ti);
    nb[(size_t)ci * NF + ti] = co;
    sg[(size_t)ci * NF + ti] = g->sqrtgram[e];
    for (int d = 0; d < D; d++) nrm[((size_t)ci * NF + ti) * D + d] = g->normals[(size_t)e * D + d];
    if (co >= 0) {
      if (to < 0 || to >= NF) return fail(BOSSS_ERR_ARG, "edge %d: bad OUT trafo", e);
      if (nb[(size_t)co * NF + to] != -2) return fail(BOSSS_ERR_UNSUPPORTED, "cell %d has two edges on face %d (non-conforming grid)", co, to);
      nb[(size_t)co * NF + to] = ci;
      sg[(size_t)co * NF + to] = g->sqrtgram[e];
      for (int d = 0; d < D; d++) nrm[((size_t)co * NF + to) * D + d] = -g->normals[(size_t)e * D + d];
      if (pair[ti] < 0) pair[ti] = to;
      if (pair[to] < 0) pair[to] = ti;
      if (pair[ti] != to || pair[to] != ti)
        return fail(BOSSS_ERR_UNSUPPORTED, "edge %d: face pairing %d<->%d differs from the rest of the grid", e, ti, to);
    } else {
      has_boundary = true;
    }
  }
  for (size_t i = 0; i < nb.size(); i++)
    if (nb[i] == -2) return fail(BOSSS_ERR_UNSUPPORTED, "cell %zu has no edge on local face %zu", i / NF, i % NF);
  for (int f = 0; f < NF; f++)
    if (pair[f] < 0) pair[f] = (f % 2 == 0) ? f + 1 : f - 1;  // only boundary edges on this face: pairing unused
  // Plane stride: the largest cell-index offset |OUT - IN| carried by at least half as many
  // interior edges as there are cells.  In a Cartesian numbering each direction's stride is
  // carried by all but one layer of edges and the periodic wrap-arounds by one layer only, so
  // this is the stride of the slowest direction (a workgroup-ordering hint, no effect on results).
  std::map<int, int> offsets;
  for (int e = 0; e < E; e++)
    if (g->cells[2 * e + 1] >= 0) offsets[std::abs(g->cells[2 * e + 1] - g->cells[2 * e])]++;
  int plane = 0;
  for (const auto& oc : offsets)
    if (2LL * oc.second >= J) plane = std::max(plane, oc.first);

  HIPCHK(hipSetDevice(ctx->device));
  bosss_grid* G = new bosss_grid();
  G->ctx = ctx; G->D = D; G->J = J; G->E = E; G->NF = NF; G->has_boundary = has_boundary;
  G->nb_host = nb;
  G->plane = plane;
  for (int f = 0; f < 6; f++) G->pair[f] = f < NF ? pair[f] : 0;
  int rc;
  if ((rc = upload(&G->detJ, g->detJ, J)) || (rc = upload(&G->invJ, g->invJ, (size_t)J * D * D)) ||
      (rc = upload(&G->scaling, g->scaling, J)) || (rc = upload(&G->lenscale, g->lenscale, J)) ||
      (rc = upload(&G->nb, nb.data(), nb.size())) || (rc = upload(&G->sqrtg, sg.data(), sg.size())) ||
      (rc = upload(&G->normal, nrm.data(), nrm.size()))) {
    bosss_hip_grid_destroy(G);
    return rc;
  }
  *out = G;
  return BOSSS_OK;
}

int bosss_hip_grid_destroy(bosss_grid* g) {
  if (!g) return BOSSS_OK;
  (void)hipFree(g->detJ); (void)hipFree(g->invJ); (void)hipFree(g->scaling); (void)hipFree(g->lenscale);
  (void)hipFree(g->nb); (void)hipFree(g->sqrtg); (void)hipFree(g->normal);
  delete g;
  return BOSSS_OK;
}

}  // extern "C"

static int form_vars(int kind, int D) { return (kind == BOSSS_FORM_EULER_LF || kind == BOSSS_FORM_EULER_HLLC) ? D + 2 : 1; }

template <typename Real>
static void pack_tables(const bosss_tables_desc* t, const KernelSet* ks, std::vector<Real>& out) {
  const int D = t->D, N = t->N, Kv = t->Kv, Ke = t->Ke, NF = 2 * D;
  const int KVP = ks->kvp, KEP = ks->kep;
  out.assign(ks->table_size, Real(0));
  for (int k = 0; k < Kv; k++)
    for (int n = 0; n < N; n++) {
      out[ks->o_vphi + k * N + n] = Real(t->phi[k * N + n]);
      for (int d = 0; d < D; d++) out[ks->o_vdphi + (d * KVP + k) * N + n] = Real(t->dphi[(k * N + n) * D + d]);
    }
  for (int k = 0; k < Kv; k++) out[ks->o_vw + k] = Real(t->vol_weights[k]);
  for (int f = 0; f < NF; f++)
    for (int k = 0; k < Ke; k++)
      for (int n = 0; n < N; n++) {
        out[ks->o_tr + (f * KEP + k) * N + n] = Real(t->trace[((size_t)f * Ke + k) * N + n]);
        if (ks->facegrad)
          for (int d = 0; d < D; d++)
            out[ks->o_dtr + ((f * D + d) * KEP + k) * N + n] = Real(t->dtrace[(((size_t)f * Ke + k) * N + n) * D + d]);
      }
  for (int k = 0; k < Ke; k++) out[ks->o_fw + k] = Real(t->face_weights[k]);
}

// ---- 1D factors of the hypercube tables (Gauss-Legendre rule, normalised Legendre polynomials)
static void gauss_legendre(int K, std::vector<double>& x, std::vector<double>& w) {
  x.assign(K, 0.0); w.assign(K, 0.0);
  for (int i = 0; i < K; i++) {
    double z = std::cos(M_PI * (i + 0.75) / (K + 0.5)), pp = 0;
    for (int it = 0; it < 100; it++) {
      double p0 = 1, p1 = 0;
      for (int n = 1; n <= K; n++) { double p2 = p1; p1 = p0; p0 = ((2 * n - 1) * z * p1 - (n - 1) * p2) / n; }
      pp = K * (z * p0 - p1) / (z * z - 1);
      double dz = p0 / pp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[K - 1 - i] = z;
    w[K - 1 - i] = 2.0 / ((1 - z * z) * pp * pp);
  }
}
static void legendre_normalised(int P, double x, double* L, double* DL) {
  double p0 = 1, p1 = x, d0 = 0, d1 = 1;
  for (int n = 0; n <= P; n++) {
    double pn, dn;
    if (n == 0) { pn = 1; dn = 0; }
    else if (n == 1) { pn = x; dn = 1; }
    else {
      pn = ((2 * n - 1) * x * p1 - (n - 1) * p0) / n;
      dn = d0 + (2 * n - 1) * p1;
      p0 = p1; p1 = pn; d0 = d1; d1 = dn;
    }
    double c = std::sqrt((2 * n + 1) / 2.0);
    L[n] = c * pn;
    DL[n] = c * dn;
  }
}

// Build the 1D table block of TShape and the face trace table after it, and check that the
// uploaded tables are their tensor product.
static bool build_tensor_tables(const bosss_tables_desc* t, const KernelSet* ks, int K1, std::vector<double>& tab, double* maxdiff) {
  const int D = t->D, P = ks->P, K2 = ks->t_k2, N = t->N;
  tab.assign(ks->ttab_size, 0.0);
  if (t->Ke > ks->t_frows || ks->t_ftab != 2 * D * ks->t_frows * ks->t_fs || ks->t_o_ft + (1 + D) * ks->t_ftab != ks->ttab_size)
    return false;
  std::vector<double> x, w;
  gauss_legendre(K1, x, w);
  std::vector<double> L(P + 1), DL(P + 1);
  for (int q = 0; q < K2; q++) {
    double xx = q < K1 ? x[q] : (q == K1 ? -1.0 : 1.0);
    legendre_normalised(P, xx, L.data(), DL.data());
    for (int a = 0; a <= P; a++) { tab[ks->t_o_l + a * K2 + q] = L[a]; tab[ks->t_o_dl + a * K2 + q] = DL[a]; }
  }
  for (int q = 0; q < K1; q++) tab[ks->t_o_w + q] = w[q];
  std::vector<int> ex;   // modes sorted by degree, then lexicographic in (e_x, e_y, e_z)
  for (int deg = 0; deg <= P; deg++)
    for (int a = 0; a <= deg; a++)
      for (int b = 0; b <= deg - a; b++) {
        int c = deg - a - b;
        if (D == 2 && c != 0) continue;
        if (D == 2) { ex.push_back(a | (b << 4)); }
        else ex.push_back(a | (b << 4) | (c << 8));
      }
  // for D == 2 the loop above emits (a, deg-a) once per a: same lexicographic order
  if ((int)ex.size() != N) return false;
  for (int n = 0; n < N; n++) tab[ks->t_o_code + n] = ex[n];
  auto val = [&](const int* q, int code, int deriv) {
    int e[3] = {code & 15, (code >> 4) & 15, (code >> 8) & 15};
    double r = 1;
    for (int d = 0; d < D; d++) {
      const double* T = &tab[(deriv == d ? ks->t_o_dl : ks->t_o_l)];
      r *= T[e[d] * K2 + q[d]];
    }
    return r;
  };
  double md = 0, scale = 1;
  auto cmp = [&](double a, double b) { md = std::max(md, std::fabs(a - b)); scale = std::max(scale, std::fabs(b)); };
  for (int k = 0; k < t->Kv; k++) {
    int q[3];
    if (D == 3) { q[0] = k % K1; q[1] = (k / K1) % K1; q[2] = k / (K1 * K1); } else { q[0] = k % K1; q[1] = k / K1; q[2] = 0; }
    double wk = D == 3 ? (w[q[2]] * w[q[1]]) * w[q[0]] : w[q[1]] * w[q[0]];
    cmp(wk, t->vol_weights[k]);
    for (int n = 0; n < N; n++) {
      cmp(val(q, ex[n], -1), t->phi[k * N + n]);
      for (int d = 0; d < D; d++) cmp(val(q, ex[n], d), t->dphi[(k * N + n) * D + d]);
    }
  }
  for (int f = 0; f < 2 * D; f++)
    for (int k = 0; k < t->Ke; k++) {
      int q[3];
      const int lo = K1, hi = K1 + 1;
      if (D == 3) {
        int s0 = k % K1, s1 = k / K1;
        int m[6][3] = {{lo, s1, s0}, {hi, s1, s0}, {s0, hi, s1}, {s0, lo, s1}, {s0, s1, hi}, {s0, s1, lo}};
        for (int d = 0; d < 3; d++) q[d] = m[f][d];
        if (f == 0) cmp(w[k / K1] * w[k % K1], t->face_weights[k]);
      } else {
        int m[4][2] = {{lo, k}, {hi, k}, {k, hi}, {k, lo}};
        q[0] = m[f][0]; q[1] = m[f][1]; q[2] = 0;
        if (f == 0) cmp(w[k], t->face_weights[k]);
      }
      for (int n = 0; n < N; n++) {
        // face table e (-1: phi, d: d phi / d xi_d), row (f, k); the padding rows Ke .. t_frows-1 repeat node 0
        for (int e = -1; e < D; e++) {
          const double v = val(q, ex[n], e);
          if (e < 0) cmp(v, t->trace[((size_t)f * t->Ke + k) * N + n]);
          else cmp(v, t->dtrace[(((size_t)f * t->Ke + k) * N + n) * D + e]);
          double* ftf = &tab[ks->t_o_ft + (size_t)(1 + e) * ks->t_ftab + (size_t)f * ks->t_frows * ks->t_fs];
          ftf[(size_t)k * ks->t_fs + n] = v;
          if (k == 0)
            for (int kp = t->Ke; kp < ks->t_frows; kp++) ftf[(size_t)kp * ks->t_fs + n] = v;
        }
      }
    }
  *maxdiff = md / scale;
  return md <= 1e-12 * scale;
}

extern "C" {

int bosss_hip_op_create(bosss_ctx* ctx, const bosss_grid* grid, const bosss_tables_desc* t, int kind,
                        const bosss_form_params* prm, int precision, bosss_op** out) {
  if (!ctx || !grid || !t || !prm || !out) return fail(BOSSS_ERR_ARG, "NULL argument");
  if (t->D != grid->D) return fail(BOSSS_ERR_ARG, "tables are %dD, grid is %dD", t->D, grid->D);
  if (kind < 0 || kind > 3) return fail(BOSSS_ERR_UNSUPPORTED, "form kind %d has no GPU kernel", kind);
  if (precision != BOSSS_FP64 && precision != BOSSS_FP32) return fail(BOSSS_ERR_ARG, "precision %d", precision);
  if (t->ntrafo != 2 * t->D) return fail(BOSSS_ERR_UNSUPPORTED, "expected %d face trafos, got %d", 2 * t->D, t->ntrafo);
  int K1 = (int)std::lround(std::pow((double)t->Kv, 1.0 / t->D));
  int Ke1 = 1; for (int d = 0; d < t->D - 1; d++) Ke1 *= K1;
  int Kv1 = Ke1 * K1;
  if (Kv1 != t->Kv || Ke1 != t->Ke) return fail(BOSSS_ERR_UNSUPPORTED, "volume/face rules must be tensor Gauss rules (Kv=%d, Ke=%d)", t->Kv, t->Ke);
  if ((kind == BOSSS_FORM_EULER_LF || kind == BOSSS_FORM_EULER_HLLC) && grid->has_boundary)
    return fail(BOSSS_ERR_UNSUPPORTED, "Euler kernels cover periodic grids only (boundary edges present)");
  const KernelSet* ks = find_kernels(kind, t->D, t->N, K1, precision);
  if (!ks) return fail(BOSSS_ERR_UNSUPPORTED, "no kernel for kind=%d D=%d N=%d K1=%d precision=%d", kind, t->D, t->N, K1, precision);
  if (ks->lds64 > 160 * 1024 || (precision == BOSSS_FP32 && ks->lds32 > 160 * 1024))
    return fail(BOSSS_ERR_UNSUPPORTED, "tables need %zu bytes of LDS (> 160 KiB)", precision == BOSSS_FP64 ? ks->lds64 : ks->lds32);
  if (!t->phi || !t->dphi || !t->vol_weights || !t->trace || !t->dtrace || !t->face_weights)
    return fail(BOSSS_ERR_ARG, "tables descriptor has NULL arrays");
  HIPCHK(hipSetDevice(ctx->device));
  bosss_op* op = new bosss_op();
  op->ctx = ctx; op->grid = grid; op->kind = kind; op->D = t->D; op->N = t->N; op->K1 = K1;
  op->V = form_vars(kind, t->D); op->precision = precision; op->ks = ks;
  op->ndof = (size_t)grid->J * op->V * op->N;
  op->fp.penalty_base = prm->penalty_base;
  op->fp.nu = prm->nu;
  for (int d = 0; d < 3; d++) op->fp.vel[d] = prm->velocity[d];
  op->fp.gamma = prm->gamma;
  op->fp.mach = prm->mach;
  {
    double p = t->degree, D = t->D;
    op->fp.penalty_deg = std::max((p + 1) * (p + D) / D, (p + 1.0) * (p + 1.0));
  }
  int rc;
  if (precision == BOSSS_FP64) {
    std::vector<double> packed;
    pack_tables<double>(t, ks, packed);
    rc = upload((double**)&op->tables, packed.data(), packed.size());
  } else {
    std::vector<float> packed;
    pack_tables<float>(t, ks, packed);
    rc = upload((float**)&op->tables, packed.data(), packed.size());
  }
  if (rc) { delete op; return rc; }
  {
    std::vector<double> tab;
    double md = 0;
    const char* force_generic = std::getenv("BOSSS_HIP_GENERIC_TABLES");
    op->tensor = ks->kt64[KernelSet::RES] && !(force_generic && force_generic[0] == '1') && build_tensor_tables(t, ks, K1, tab, &md);
    op->ttab = nullptr;
    if (op->tensor) {
      if (precision == BOSSS_FP64) {
        rc = upload((double**)&op->ttab, tab.data(), tab.size());
      } else {
        std::vector<float> tf(tab.begin(), tab.end());
        rc = upload((float**)&op->ttab, tf.data(), tf.size());
      }
      if (rc) { (void)hipFree(op->tables); delete op; return rc; }
    }
  }
  if ((rc = ks->prepare()) != 0) { (void)hipFree(op->tables); delete op; return fail(BOSSS_ERR_HIP, "kernel attribute setup failed (%d)", rc); }
  *out = op;
  return BOSSS_OK;
}

int bosss_hip_op_destroy(bosss_op* op) {
  if (!op) return BOSSS_OK;
  (void)hipFree(op->tables);
  if (op->ttab) (void)hipFree(op->ttab);
  if (op->hU) (void)hipFree(op->hU);
  if (op->hR) (void)hipFree(op->hR);
  if (op->hX) (void)hipFree(op->hX);
  if (op->hY) (void)hipFree(op->hY);
  if (op->d_color) (void)hipFree(op->d_color);
  if (op->d_cols) (void)hipFree(op->d_cols);
  if (op->d_first) (void)hipFree(op->d_first);
  delete op;
  return BOSSS_OK;
}

int bosss_hip_op_info(const bosss_op* op, int64_t* ndof, int32_t* nvars, int32_t* nmodes) {
  if (!op) return fail(BOSSS_ERR_ARG, "op is NULL");
  if (ndof) *ndof = (int64_t)op->ndof;
  if (nvars) *nvars = op->V;
  if (nmodes) *nmodes = op->N;
  return BOSSS_OK;
}

int bosss_hip_op_kernel_family(const bosss_op* op, int32_t* tensor_tables) {
  if (!op || !tensor_tables) return fail(BOSSS_ERR_ARG, "NULL argument");
  *tensor_tables = op->tensor ? 1 : 0;
  return BOSSS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- launching
template <typename Real>
static KArgs<Real> make_args(const bosss_op* op, double alpha, double beta, const void* U, const void* X, void* R, void* out,
                             int own_only = 0) {
  KArgs<Real> a;
  a.own_only = own_only;
  const bosss_grid* g = op->grid;
  a.J = g->J;
  a.detJ = g->detJ; a.invJ = g->invJ; a.scaling = g->scaling; a.lenscale = g->lenscale;
  a.cf.nb = g->nb; a.cf.sqrtg = g->sqrtg; a.cf.normal = g->normal;
  for (int f = 0; f < 6; f++) a.cf.pair[f] = g->pair[f];
  a.tables = (const Real*)op->tables;
  a.fp = op->fp;
  a.alpha = alpha; a.beta = beta;
  a.U = (const Real*)U; a.dU = (const Real*)X; a.R = (Real*)R; a.out = (Real*)out;
  return a;
}

// `stream` is a hipStream_t; NULL is HIP's null (default) stream, as everywhere in HIP.
template <typename Real>
static TArgs<Real> make_targs(const bosss_op* op, double alpha, double beta, const void* U, const void* X, void* R, void* out,
                              int own_only, int plane_blocks) {
  KArgs<Real> k = make_args<Real>(op, alpha, beta, U, X, R, out, own_only);
  TArgs<Real> a;
  a.own_only = own_only;
  a.J = k.J; a.detJ = k.detJ; a.invJ = k.invJ; a.scaling = k.scaling; a.lenscale = k.lenscale; a.cf = k.cf;
  a.tab = (const Real*)op->ttab;
  a.fp = k.fp; a.alpha = alpha; a.beta = beta;
  a.U = (const Real*)U; a.X = (const Real*)X; a.R = (Real*)R; a.out = (Real*)out;
  a.nxcd = op->ctx->nxcd;
  a.plane_blocks = plane_blocks;
  return a;
}

// own_only (JVP): the neighbour tangents are taken as zero, so the result is J(U)_jj X_j per cell.
static int launch(bosss_op* op, int which, double alpha, double beta, const void* U, const void* X, void* R, void* out, hipStream_t s,
                  int own_only = 0) {
  HIPCHK(hipSetDevice(op->ctx->device));
  if (op->tensor) {
    dim3 grid((op->grid->J + 16 * TWAVES - 1) / (16 * TWAVES)), block(64 * TWAVES);
    // Plane-fastest workgroup order (xcd_block) for the Euler JVP and VJP kernels when one
    // plane's operands (U and the tangent or cotangent) exceed half of an XCD's 4 MB L2: in the
    // range order they are evicted before the next plane's workgroups re-read them (64^3, p = 3:
    // HBM fetch of the fused kernel -23 %, of the JVP -37 %).  The residual kernel ran 2 % slower
    // with it and keeps the range order, as do smaller planes, whose re-reads already hit L2.
    const bool euler = op->kind == BOSSS_FORM_EULER_LF || op->kind == BOSSS_FORM_EULER_HLLC;
    const size_t plane_bytes = (size_t)op->grid->plane * 2 * op->V * op->N * (op->precision == BOSSS_FP64 ? 8 : 4);
    const int plane_blocks = (euler && which != KernelSet::RES && op->grid->plane % (16 * TWAVES) == 0 &&
                              plane_bytes > (size_t(2) << 20))
                                 ? op->grid->plane / (16 * TWAVES) : 0;
    if (op->precision == BOSSS_FP64) {
      void (*k)(TArgs<double>) = op->ks->kt64[which];
      if (!k) return fail(BOSSS_ERR_UNSUPPORTED, "tensor kernel variant %d not built", which);
      hipLaunchKernelGGL(k, grid, block, op->ks->ktlds64[which], s,
                         make_targs<double>(op, alpha, beta, U, X, R, out, own_only, plane_blocks));
    } else {
      void (*k)(TArgs<float>) = op->ks->kt32[which];
      if (!k) return fail(BOSSS_ERR_UNSUPPORTED, "tensor kernel variant %d not built", which);
      hipLaunchKernelGGL(k, grid, block, op->ks->ktlds32[which], s,
                         make_targs<float>(op, alpha, beta, U, X, R, out, own_only, plane_blocks));
    }
    HIPCHK(hipGetLastError());
    return BOSSS_OK;
  }
  const int cells_per_block = 16 * (KernelSet::BLOCK / 64);
  dim3 grid((op->grid->J + cells_per_block - 1) / cells_per_block), block(KernelSet::BLOCK);
  if (op->precision == BOSSS_FP64) {
    KArgs<double> a = make_args<double>(op, alpha, beta, U, X, R, out, own_only);
    void (*k)(KArgs<double>) = op->ks->k64[which];
    if (!k) return fail(BOSSS_ERR_UNSUPPORTED, "kernel variant %d not built", which);
    hipLaunchKernelGGL(k, grid, block, op->ks->lds64, s, a);
  } else {
    KArgs<float> a = make_args<float>(op, alpha, beta, U, X, R, out, own_only);
    void (*k)(KArgs<float>) = op->ks->k32[which];
    if (!k) return fail(BOSSS_ERR_UNSUPPORTED, "kernel variant %d not built", which);
    hipLaunchKernelGGL(k, grid, block, op->ks->lds32, s, a);
  }
  HIPCHK(hipGetLastError());
  return BOSSS_OK;
}

// CheckForNanOrInf (DifferentialOperator.cs:1602,1623,1658; there in DEBUG builds): the smallest
// index of a NaN or Inf entry of v, or ~0 if every entry is finite.
template <typename Real>
__global__ void k_first_nonfinite(int64_t n, const Real* __restrict__ v, unsigned long long* first) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    if (!isfinite(v[i])) atomicMin(first, (unsigned long long)i);
}
__global__ void k_set_u64(unsigned long long* p, unsigned long long v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// With the check enabled: scan each output of the call (nullptr entries skipped), wait for the
// stream and report the first non-finite entry.
static int check_outputs(bosss_op* op, hipStream_t s, const char* what, std::initializer_list<const void*> outs) {
  if (!op->check_finite) return BOSSS_OK;
  if (!op->d_first) HIPCHK(hipMalloc((void**)&op->d_first, sizeof(unsigned long long)));
  int k = 0;
  for (const void* v : outs) {
    if (v) {
      hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, s, op->d_first, ~0ull);
      const unsigned nb = (unsigned)std::min<size_t>(4096, (op->ndof + 255) / 256);
      if (op->precision == BOSSS_FP64)
        hipLaunchKernelGGL(k_first_nonfinite<double>, dim3(nb), dim3(256), 0, s, (int64_t)op->ndof, (const double*)v, op->d_first);
      else
        hipLaunchKernelGGL(k_first_nonfinite<float>, dim3(nb), dim3(256), 0, s, (int64_t)op->ndof, (const float*)v, op->d_first);
      HIPCHK(hipGetLastError());
      unsigned long long first = 0;
      HIPCHK(hipMemcpyAsync(&first, op->d_first, sizeof first, hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      if (first != ~0ull) {
        const size_t nt = (size_t)op->V * op->N, i = (size_t)first;
        return fail(BOSSS_ERR_NONFINITE, "%s: output %d entry %zu (cell %zu, variable %zu, mode %zu) is NaN or Inf", what, k, i,
                    i / nt, (i % nt) / op->N, i % op->N);
      }
    }
    k++;
  }
  return BOSSS_OK;
}

extern "C" {

int bosss_hip_op_set_check_finite(bosss_op* op, int32_t enable) {
  if (!op) return fail(BOSSS_ERR_ARG, "op is NULL");
  op->check_finite = enable != 0;
  return BOSSS_OK;
}

int bosss_hip_evaluate_dev(bosss_op* op, double alpha, double beta, const double* U, double* R, void* stream) {
  if (!op || !U || !R) return fail(BOSSS_ERR_ARG, "NULL argument");
  if (int rc = launch(op, KernelSet::RES, alpha, beta, U, nullptr, R, nullptr, (hipStream_t)stream)) return rc;
  return check_outputs(op, (hipStream_t)stream, "evaluate", {R});
}
int bosss_hip_jvp_dev(bosss_op* op, const double* U, const double* dU, double* JdU, void* stream) {
  if (!op || !U || !dU || !JdU) return fail(BOSSS_ERR_ARG, "NULL argument");
  if (int rc = launch(op, KernelSet::JVP, 1.0, 0.0, U, dU, nullptr, JdU, (hipStream_t)stream)) return rc;
  return check_outputs(op, (hipStream_t)stream, "jvp", {JdU});
}
int bosss_hip_vjp_dev(bosss_op* op, const double* U, const double* W, double* JtW, void* stream) {
  if (!op || !U || !W || !JtW) return fail(BOSSS_ERR_ARG, "NULL argument");
  if (int rc = launch(op, KernelSet::VJP, 1.0, 0.0, U, W, nullptr, JtW, (hipStream_t)stream)) return rc;
  return check_outputs(op, (hipStream_t)stream, "vjp", {JtW});
}
int bosss_hip_evaluate_vjp_dev(bosss_op* op, double alpha, double beta, const double* U, double* R, const double* W,
                               double* JtW, void* stream) {
  if (!op || !U || !R || !W || !JtW) return fail(BOSSS_ERR_ARG, "NULL argument");
  if (int rc = launch(op, KernelSet::RES_VJP, alpha, beta, U, W, R, JtW, (hipStream_t)stream)) return rc;
  return check_outputs(op, (hipStream_t)stream, "evaluate_vjp", {R, JtW});
}

static int ensure_staging(bosss_op* op) {
  if (op->hU) return BOSSS_OK;
  size_t bytes = op->ndof * (op->precision == BOSSS_FP64 ? 8 : 4) + 16;
  HIPCHK(hipMalloc((void**)&op->hU, bytes));
  HIPCHK(hipMalloc((void**)&op->hR, bytes));
  HIPCHK(hipMalloc((void**)&op->hX, bytes));
  HIPCHK(hipMalloc((void**)&op->hY, bytes));
  return BOSSS_OK;
}

// Host-pointer entry points: FP64 arrays in and out; an FP32 operator converts at the boundary.
static int to_device(bosss_op* op, void* dst, const double* src) {
  hipStream_t s = op->ctx->stream;
  if (op->precision == BOSSS_FP64) {
    HIPCHK(hipMemcpyAsync(dst, src, op->ndof * 8, hipMemcpyHostToDevice, s));
  } else {
    std::vector<float> tmp(op->ndof);
    for (size_t i = 0; i < op->ndof; i++) tmp[i] = (float)src[i];
    HIPCHK(hipMemcpyAsync(dst, tmp.data(), op->ndof * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  return BOSSS_OK;
}
static int to_host(bosss_op* op, double* dst, const void* src) {
  hipStream_t s = op->ctx->stream;
  if (op->precision == BOSSS_FP64) {
    HIPCHK(hipMemcpyAsync(dst, src, op->ndof * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  } else {
    std::vector<float> tmp(op->ndof);
    HIPCHK(hipMemcpyAsync(tmp.data(), src, op->ndof * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (size_t i = 0; i < op->ndof; i++) dst[i] = tmp[i];
  }
  return BOSSS_OK;
}

int bosss_hip_evaluate(bosss_op* op, double time, double alpha, double beta, const double* U, double* R) {
  (void)time;  // the supported forms do not depend on time
  if (!op || !U || !R) return fail(BOSSS_ERR_ARG, "NULL argument");
  int rc;
  HIPCHK(hipSetDevice(op->ctx->device));
  if ((rc = ensure_staging(op)) || (rc = to_device(op, op->hU, U))) return rc;
  if (beta != 0.0 && (rc = to_device(op, op->hR, R))) return rc;
  if ((rc = launch(op, KernelSet::RES, alpha, beta, op->hU, nullptr, op->hR, nullptr, op->ctx->stream)) ||
      (rc = check_outputs(op, op->ctx->stream, "evaluate", {op->hR})))
    return rc;
  return to_host(op, R, op->hR);
}

int bosss_hip_jvp(bosss_op* op, const double* U, const double* dU, double* JdU) {
  if (!op || !U || !dU || !JdU) return fail(BOSSS_ERR_ARG, "NULL argument");
  int rc;
  HIPCHK(hipSetDevice(op->ctx->device));
  if ((rc = ensure_staging(op)) || (rc = to_device(op, op->hU, U)) || (rc = to_device(op, op->hX, dU))) return rc;
  if ((rc = launch(op, KernelSet::JVP, 1.0, 0.0, op->hU, op->hX, nullptr, op->hY, op->ctx->stream)) ||
      (rc = check_outputs(op, op->ctx->stream, "jvp", {op->hY})))
    return rc;
  return to_host(op, JdU, op->hY);
}

int bosss_hip_vjp(bosss_op* op, const double* U, const double* W, double* JtW) {
  if (!op || !U || !W || !JtW) return fail(BOSSS_ERR_ARG, "NULL argument");
  int rc;
  HIPCHK(hipSetDevice(op->ctx->device));
  if ((rc = ensure_staging(op)) || (rc = to_device(op, op->hU, U)) || (rc = to_device(op, op->hX, W))) return rc;
  if ((rc = launch(op, KernelSet::VJP, 1.0, 0.0, op->hU, op->hX, nullptr, op->hY, op->ctx->stream)) ||
      (rc = check_outputs(op, op->ctx->stream, "vjp", {op->hY})))
    return rc;
  return to_host(op, JtW, op->hY);
}

}  // extern "C"

// ------------------------------------------------------------------ matrix assembly
static int ensure_matrix_setup(bosss_op* op) {
  if (op->d_cols) return BOSSS_OK;
  const bosss_grid* g = op->grid;
  const int J = g->J, NF = g->NF, S = 1 + NF;
  // block-ELL columns; a neighbour met twice in one row (periodic direction of 2 cells) keeps one slot
  op->cols_host.assign((size_t)J * S, -1);
  for (int j = 0; j < J; j++) {
    op->cols_host[(size_t)j * S] = j;
    for (int f = 0; f < NF; f++) {
      int c = g->nb_host[(size_t)j * NF + f];
      bool dup = (c == j);
      for (int s2 = 1; s2 <= f && !dup; s2++) dup = (op->cols_host[(size_t)j * S + s2] == c);
      op->cols_host[(size_t)j * S + 1 + f] = dup ? -1 : c;
    }
  }
  // greedy distance-2 colouring of the cell graph
  std::vector<int32_t> color(J, -1);
  std::vector<int> seen;
  int ncol = 0;
  for (int j = 0; j < J; j++) {
    seen.assign(ncol + 1, 0);
    auto mark = [&](int c) { if (c >= 0 && color[c] >= 0) seen[color[c]] = 1; };
    for (int f = 0; f < NF; f++) {
      int c = g->nb_host[(size_t)j * NF + f];
      mark(c);
      if (c >= 0)
        for (int f2 = 0; f2 < NF; f2++) mark(g->nb_host[(size_t)c * NF + f2]);
    }
    int k = 0;
    while (k < ncol && seen[k]) k++;
    color[j] = k;
    if (k == ncol) ncol++;
  }
  op->ncolors = ncol;
  int rc;
  if ((rc = upload(&op->d_color, color.data(), color.size())) || (rc = upload(&op->d_cols, op->cols_host.data(), op->cols_host.size())))
    return rc;
  return BOSSS_OK;
}

// Diagonal cell blocks by probing the own-cell JVP: X[c][e] = (e == q) in every cell, then
// blocks[j][r][q] = alpha * (J(U)_jj e_q)[r] + shift * (r == q)  (transposed: blocks[j][q][r], the
// probe's column written contiguously; the row-major form writes it with a stride of Nt).
template <typename Real>
__global__ void k_probe_all(int J, int Nt, int q, Real* __restrict__ X) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)J * Nt) return;
  X[i] = (int)(i % Nt) == q ? Real(1) : Real(0);
}
template <typename Real>
__global__ void k_extract_diag(int J, int Nt, int q, int transposed, const Real* __restrict__ Y, double alpha, double shift,
                               double* __restrict__ blocks) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)J * Nt) return;
  const size_t j = i / Nt;
  const int r = (int)(i - j * Nt);
  const size_t at = transposed ? (j * Nt + q) * Nt + r : (j * Nt + r) * Nt + q;
  blocks[at] = alpha * (double)Y[i] + (r == q ? shift : 0.0);
}

// Autotuner of the block-ELL apply's panel size: times every candidate on this operator's
// matrix and x (5 launches each after one warm-up, HIP events on the caller's stream, output
// into a scratch vector) and keeps the fastest.  Under stream capture it does not time (events
// and synchronisation are not allowed there) and the fixed rule is used until a later call.
static int tune_spmv(bosss_op* op, const double* blocks, const double* x, hipStream_t s) {
  const int J = op->grid->J, Nt = op->V * op->N, S = 1 + op->grid->NF;
  std::vector<int> cand;
  bell_candidates(Nt, cand);
  op->spmv_cand = cand;
  op->spmv_ms.assign(cand.size(), -1.0f);
  if (cand.size() <= 1) {
    op->spmv_nl = cand.empty() ? 0 : cand[0];
    op->spmv_tuned = true;
    return BOSSS_OK;
  }
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  HIPCHK(hipStreamIsCapturing(s, &cs));
  if (cs != hipStreamCaptureStatusNone) {
    op->spmv_nl = bell_default_nl(Nt);
    return BOSSS_OK;
  }
  double* ys = nullptr;
  HIPCHK(hipMalloc((void**)&ys, op->ndof * sizeof(double)));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  int best = bell_default_nl(Nt);
  float best_ms = 1e30f;
  for (size_t i = 0; i < cand.size() && e == hipSuccess; i++) {
    launch_bell_spmv(J, Nt, S, op->d_cols, blocks, 1.0, x, 0.0, ys, s, cand[i]);
    e = hipEventRecord(e0, s);
    for (int r = 0; r < 5 && e == hipSuccess; r++) launch_bell_spmv(J, Nt, S, op->d_cols, blocks, 1.0, x, 0.0, ys, s, cand[i]);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess) {
      op->spmv_ms[i] = ms / 5;
      if (ms < best_ms) { best_ms = ms; best = cand[i]; }
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  hipError_t ef = hipFree(ys);
  if (e != hipSuccess) return fail(BOSSS_ERR_HIP, "apply autotuner: %s", hipGetErrorString(e));
  if (ef != hipSuccess) return fail(BOSSS_ERR_HIP, "apply autotuner: %s", hipGetErrorString(ef));
  op->spmv_nl = best;
  op->spmv_tuned = true;
  return BOSSS_OK;
}

extern "C" {

int bosss_hip_op_matrix_layout(bosss_op* op, int32_t* blocks_per_row, int64_t* entries, int32_t* ncolors, int32_t* cols) {
  if (!op) return fail(BOSSS_ERR_ARG, "op is NULL");
  int rc;
  HIPCHK(hipSetDevice(op->ctx->device));
  if ((rc = ensure_matrix_setup(op))) return rc;
  const int S = 1 + op->grid->NF, Nt = op->V * op->N;
  if (blocks_per_row) *blocks_per_row = S;
  if (entries) *entries = (int64_t)op->grid->J * S * Nt * Nt;
  if (ncolors) *ncolors = op->ncolors;
  if (cols) std::copy(op->cols_host.begin(), op->cols_host.end(), cols);
  return BOSSS_OK;
}

int bosss_hip_op_assemble_dev(bosss_op* op, const double* U, double alpha, double beta, double* blocks, void* stream) {
  if (!op || !U || !blocks) return fail(BOSSS_ERR_ARG, "NULL argument");
  if (op->precision != BOSSS_FP64) return fail(BOSSS_ERR_UNSUPPORTED, "assembly is FP64 only");
  int rc;
  HIPCHK(hipSetDevice(op->ctx->device));
  if ((rc = ensure_matrix_setup(op)) || (rc = ensure_staging(op))) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int J = op->grid->J, Nt = op->V * op->N, S = 1 + op->grid->NF;
  for (int k = 0; k < op->ncolors; k++)
    for (int q = 0; q < Nt; q++) {
      launch_probe(J, Nt, op->d_color, k, q, op->hX, s);
      if ((rc = launch(op, KernelSet::JVP, 1.0, 0.0, U, op->hX, nullptr, op->hY, s))) return rc;
      launch_extract(J, Nt, S, op->d_cols, op->d_color, k, q, op->hY, alpha, beta, blocks, s);
    }
  HIPCHK(hipGetLastError());
  return BOSSS_OK;
}

int bosss_hip_op_diag_blocks_dev(bosss_op* op, const void* U, double alpha, double shift, int32_t transposed, double* blocks,
                                 void* stream) {
  if (!op || !U || !blocks) return fail(BOSSS_ERR_ARG, "NULL argument");
  int rc;
  HIPCHK(hipSetDevice(op->ctx->device));
  if ((rc = ensure_staging(op))) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int J = op->grid->J, Nt = op->V * op->N;
  const size_t n = (size_t)J * Nt;
  const unsigned nb = (unsigned)((n + 255) / 256);
  for (int q = 0; q < Nt; q++) {
    if (op->precision == BOSSS_FP64) hipLaunchKernelGGL(k_probe_all<double>, dim3(nb), dim3(256), 0, s, J, Nt, q, (double*)op->hX);
    else hipLaunchKernelGGL(k_probe_all<float>, dim3(nb), dim3(256), 0, s, J, Nt, q, (float*)op->hX);
    if ((rc = launch(op, KernelSet::JVP, 1.0, 0.0, U, op->hX, nullptr, op->hY, s, 1))) return rc;
    if (op->precision == BOSSS_FP64)
      hipLaunchKernelGGL(k_extract_diag<double>, dim3(nb), dim3(256), 0, s, J, Nt, q, (int)(transposed != 0), (const double*)op->hY,
                         alpha, shift, blocks);
    else
      hipLaunchKernelGGL(k_extract_diag<float>, dim3(nb), dim3(256), 0, s, J, Nt, q, (int)(transposed != 0), (const float*)op->hY, alpha,
                         shift, blocks);
  }
  HIPCHK(hipGetLastError());
  return BOSSS_OK;
}

int bosss_hip_bell_spmv_dev(bosss_op* op, const double* blocks, double alpha, const double* x, double beta, double* y, void* stream) {
  if (!op || !blocks || !x || !y) return fail(BOSSS_ERR_ARG, "NULL argument");
  int rc;
  HIPCHK(hipSetDevice(op->ctx->device));
  if ((rc = ensure_matrix_setup(op))) return rc;
  if (!op->spmv_tuned && (rc = tune_spmv(op, blocks, x, (hipStream_t)stream))) return rc;
  launch_bell_spmv(op->grid->J, op->V * op->N, 1 + op->grid->NF, op->d_cols, blocks, alpha, x, beta, y, (hipStream_t)stream,
                   op->spmv_nl);
  HIPCHK(hipGetLastError());
  return BOSSS_OK;
}

int bosss_hip_op_set_spmv_tile(bosss_op* op, int32_t nl) {
  if (!op) return fail(BOSSS_ERR_ARG, "op is NULL");
  std::vector<int> cand;
  bell_candidates(op->V * op->N, cand);
  if (nl == 0) {
    op->spmv_tuned = false;
    op->spmv_nl = 0;
    return BOSSS_OK;
  }
  if (std::find(cand.begin(), cand.end(), (int)nl) == cand.end())
    return fail(BOSSS_ERR_ARG, "panel size %d is not a candidate for %d x %d blocks", nl, op->V * op->N, op->V * op->N);
  op->spmv_nl = nl;
  op->spmv_tuned = true;
  return BOSSS_OK;
}

int bosss_hip_op_spmv_tiles(const bosss_op* op, int32_t* chosen, int32_t* ncand, int32_t* cand, float* ms, int32_t max) {
  if (!op) return fail(BOSSS_ERR_ARG, "op is NULL");
  std::vector<int> c;
  bell_candidates(op->V * op->N, c);
  if (chosen) *chosen = op->spmv_tuned ? op->spmv_nl : 0;
  if (ncand) *ncand = (int32_t)c.size();
  for (int i = 0; i < (int)c.size() && i < max; i++) {
    if (cand) cand[i] = c[i];
    if (ms) ms[i] = (i < (int)op->spmv_ms.size() && op->spmv_cand == c) ? op->spmv_ms[i] : -1.0f;
  }
  return BOSSS_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ device buffers and host-pointer matrix calls
extern "C" {

int bosss_hip_buffer_create(bosss_ctx* ctx, int64_t bytes, void** dptr) {
  if (!ctx || !dptr) return fail(BOSSS_ERR_ARG, "NULL argument");
  if (bytes < 0) return fail(BOSSS_ERR_ARG, "negative size %lld", (long long)bytes);
  HIPCHK(hipSetDevice(ctx->device));
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, bytes ? (size_t)bytes : 1));
  hipError_t e = hipMemsetAsync(p, 0, bytes ? (size_t)bytes : 1, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    (void)hipFree(p);
    return fail(BOSSS_ERR_HIP, "buffer_create: %s", hipGetErrorString(e));
  }
  *dptr = p;
  return BOSSS_OK;
}

int bosss_hip_buffer_destroy(bosss_ctx* ctx, void* dptr) {
  if (!ctx) return fail(BOSSS_ERR_ARG, "ctx is NULL");
  if (!dptr) return BOSSS_OK;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipFree(dptr));
  return BOSSS_OK;
}

int bosss_hip_buffer_upload(bosss_ctx* ctx, void* dptr, const void* host, int64_t bytes) {
  if (!ctx || !dptr || (!host && bytes)) return fail(BOSSS_ERR_ARG, "NULL argument");
  if (bytes < 0) return fail(BOSSS_ERR_ARG, "negative size %lld", (long long)bytes);
  if (!bytes) return BOSSS_OK;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipMemcpyAsync(dptr, host, (size_t)bytes, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return BOSSS_OK;
}

int bosss_hip_buffer_download(bosss_ctx* ctx, void* host, const void* dptr, int64_t bytes) {
  if (!ctx || !dptr || (!host && bytes)) return fail(BOSSS_ERR_ARG, "NULL argument");
  if (bytes < 0) return fail(BOSSS_ERR_ARG, "negative size %lld", (long long)bytes);
  if (!bytes) return BOSSS_OK;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipMemcpyAsync(host, dptr, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return BOSSS_OK;
}

int bosss_hip_op_assemble(bosss_op* op, const double* U, double alpha, double beta, double* blocks) {
  if (!op || !U || !blocks) return fail(BOSSS_ERR_ARG, "NULL argument");
  if (op->precision != BOSSS_FP64) return fail(BOSSS_ERR_UNSUPPORTED, "assembly is FP64 only");
  int rc;
  HIPCHK(hipSetDevice(op->ctx->device));
  if ((rc = ensure_matrix_setup(op))) return rc;
  const size_t entries = (size_t)op->grid->J * (1 + op->grid->NF) * op->V * op->N * op->V * op->N;
  DevArrays d(op->ctx->stream);
  const double* dU = d.in(U, op->ndof);
  double* dB = beta != 0.0 ? d.in((const double*)blocks, entries) : d.out<double>(entries);
  if (d.error() != hipSuccess) return fail(BOSSS_ERR_HIP, "op_assemble staging: %s", hipGetErrorString(d.error()));
  if ((rc = bosss_hip_op_assemble_dev(op, dU, alpha, beta, dB, op->ctx->stream))) return rc;
  d.back(blocks, (const double*)dB, entries);
  if (hipError_t e = d.finish()) return fail(BOSSS_ERR_HIP, "op_assemble: %s", hipGetErrorString(e));
  return BOSSS_OK;
}

int bosss_hip_bell_spmv(bosss_op* op, const double* blocks, double alpha, const double* x, double beta, double* y) {
  if (!op || !blocks || !x || !y) return fail(BOSSS_ERR_ARG, "NULL argument");
  int rc;
  HIPCHK(hipSetDevice(op->ctx->device));
  const size_t Nt = (size_t)op->V * op->N, entries = (size_t)op->grid->J * (1 + op->grid->NF) * Nt * Nt;
  const size_t n = (size_t)op->grid->J * Nt;   // the matrix acts on FP64 vectors whatever the operator's precision
  DevArrays d(op->ctx->stream);
  const double* dB = d.in(blocks, entries);
  const double* dx = d.in(x, n);
  double* dy = beta != 0.0 ? d.in((const double*)y, n) : d.out<double>(n);
  if (d.error() != hipSuccess) return fail(BOSSS_ERR_HIP, "bell_spmv staging: %s", hipGetErrorString(d.error()));
  if ((rc = bosss_hip_bell_spmv_dev(op, dB, alpha, dx, beta, dy, op->ctx->stream))) return rc;
  d.back(y, (const double*)dy, n);
  if (hipError_t e = d.finish()) return fail(BOSSS_ERR_HIP, "bell_spmv: %s", hipGetErrorString(e));
  return BOSSS_OK;
}

}  // extern "C"
