// Plain C++ host of the DG operator library: no torch, no Python, no HIP headers.
//
// It binds every symbol of include/bosss_hip.h by name at run time, the way BoSSS binds
// its native libraries (MPI.Wrappers/Utils/DynLibLoader.cs:365-410, LoadAllDelegates: one
// delegate field per entry point, resolved by name, a missing symbol is an error), and
// drives the library through host memory only -- the host-pointer entry points, or the
// device-buffer calls plus the *_dev entry points -- as a C# host holding managed arrays
// would (DGField_Evaluation.cs:78-110 pins them for one call).
//
// usage: abi_harness <library.so> bind      (bind every symbol and exit; needs no GPU)
//        abi_harness <library.so> <scenario> <input.rec> <output.rec>
//   scenario dg        evaluate / jvp / vjp through host pointers, the same residual through
//                      device buffers + evaluate_dev, alpha/beta accumulation, the NaN check
//   scenario blocks    block_lu (pivots, INFO), block_solve, block_lu with inverse + block_gemv
//   scenario assemble  op_assemble into a host block-ELL array, bell_spmv through host pointers
//   scenario xdg       xdg_create, xdg_assemble, xdg_spmv, xdg_solve (FP64 and FP32) on host arrays
// Records (input and output): u32 name length, name, u8 type ('d' f64, 'i' i32, 'l' i64),
// i64 count, count elements.
#include <dlfcn.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/bosss_hip.h"

// ---------------------------------------------------------------- by-name binding
#define BOSSS_API(X)                                                                                                        \
  X(last_error) X(buffer_create) X(buffer_destroy) X(buffer_upload) X(buffer_download) X(ctx_create) X(ctx_destroy) X(sync)   \
  X(ctx_stream) X(grid_upload) X(grid_destroy) X(op_create) X(op_destroy) X(op_info) X(op_kernel_family) X(evaluate) X(jvp) X(vjp)                \
  X(op_set_check_finite) X(evaluate_dev) X(jvp_dev) X(vjp_dev) X(evaluate_vjp_dev) X(op_matrix_layout) X(op_assemble_dev) X(op_diag_blocks_dev)    \
  X(bell_spmv_dev) X(op_assemble) X(bell_spmv) X(op_set_spmv_tile) X(op_spmv_tiles) X(block_lu_dev) X(block_lu_noalloc_dev) X(block_solve_dev)       \
  X(block_gemv_dev) X(block_lu) X(block_solve) X(block_gemv) X(block_last_error) X(xdg_create) X(xdg_destroy) X(xdg_info)     \
  X(xdg_assemble_dev) X(xdg_spmv_dev) X(xdg_solve_dev) X(xdg_assemble) X(xdg_spmv) X(xdg_solve) X(xdg_set_tile) X(xdg_tiles) X(xdg_last_error)            \
  X(gmres_workspace) X(gmres_arnoldi_dev) X(gmres_update_dev) X(block_gemv_uniform_dev) X(krylov_last_error) X(measure_peaks) X(measure_fp64_mix) X(peaks_last_error)

struct Api {
#define FIELD(n) decltype(&bosss_hip_##n) n = nullptr;
  BOSSS_API(FIELD)
#undef FIELD
};

static bool load_all(const char* path, Api& api) {
  void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!h) {
    std::fprintf(stderr, "dlopen %s: %s\n", path, dlerror());
    return false;
  }
  int missing = 0;
#define BIND(n)                                                                   \
  api.n = reinterpret_cast<decltype(api.n)>(dlsym(h, "bosss_hip_" #n));          \
  if (!api.n) { std::fprintf(stderr, "missing symbol bosss_hip_%s\n", #n); missing++; }
  BOSSS_API(BIND)
#undef BIND
  return missing == 0;
}

// ---------------------------------------------------------------- records
struct Rec {
  char type = 'd';
  int64_t n = 0;
  std::vector<char> bytes;
  template <class T>
  T* p() { return reinterpret_cast<T*>(bytes.data()); }
};
typedef std::map<std::string, Rec> Recs;

static size_t esize(char t) { return t == 'i' ? 4 : 8; }

static bool read_recs(const char* path, Recs& out) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  for (;;) {
    uint32_t len = 0;
    if (std::fread(&len, 4, 1, f) != 1) break;
    std::string name(len, '\0');
    Rec r;
    if (std::fread(&name[0], 1, len, f) != len || std::fread(&r.type, 1, 1, f) != 1 || std::fread(&r.n, 8, 1, f) != 1) {
      std::fclose(f);
      return false;
    }
    r.bytes.resize((size_t)r.n * esize(r.type) + 8);
    if (r.n && std::fread(r.bytes.data(), esize(r.type), (size_t)r.n, f) != (size_t)r.n) {
      std::fclose(f);
      return false;
    }
    out[name] = std::move(r);
  }
  std::fclose(f);
  return true;
}

struct Writer {
  FILE* f;
  explicit Writer(const char* path) : f(std::fopen(path, "wb")) {}
  ~Writer() { if (f) std::fclose(f); }
  void put(const std::string& name, char type, const void* data, int64_t n) {
    uint32_t len = (uint32_t)name.size();
    std::fwrite(&len, 4, 1, f);
    std::fwrite(name.data(), 1, len, f);
    std::fwrite(&type, 1, 1, f);
    std::fwrite(&n, 8, 1, f);
    if (n) std::fwrite(data, esize(type), (size_t)n, f);
  }
  void d(const std::string& name, const std::vector<double>& v) { put(name, 'd', v.data(), (int64_t)v.size()); }
  void i(const std::string& name, const std::vector<int32_t>& v) { put(name, 'i', v.data(), (int64_t)v.size()); }
  void scalar(const std::string& name, int32_t v) { put(name, 'i', &v, 1); }
  void scalar_d(const std::string& name, double v) { put(name, 'd', &v, 1); }
};

#define CHECK(call, errfn)                                                                  \
  do {                                                                                      \
    int rc_ = (call);                                                                       \
    if (rc_ < 0) {                                                                          \
      std::fprintf(stderr, "%s:%d %s -> %d: %s\n", __FILE__, __LINE__, #call, rc_, errfn()); \
      return 1;                                                                             \
    }                                                                                       \
  } while (0)

static Rec& need(Recs& in, const char* k) {
  auto it = in.find(k);
  if (it == in.end()) {
    std::fprintf(stderr, "input record %s missing\n", k);
    std::exit(2);
  }
  return it->second;
}
static int32_t ival(Recs& in, const char* k) { return need(in, k).p<int32_t>()[0]; }
static double dval(Recs& in, const char* k) { return need(in, k).p<double>()[0]; }

// ---------------------------------------------------------------- scenarios
static int scenario_dg(Api& A, Recs& in, Writer& out) {
  bosss_ctx* ctx = nullptr;
  CHECK(A.ctx_create(0, &ctx), A.last_error);
  bosss_grid_desc gd;
  gd.D = ival(in, "D"); gd.J = ival(in, "J"); gd.E = ival(in, "E");
  gd.detJ = need(in, "detJ").p<double>(); gd.invJ = need(in, "invJ").p<double>();
  gd.scaling = need(in, "scaling").p<double>(); gd.lenscale = need(in, "lenscale").p<double>();
  gd.cells = need(in, "cells").p<int32_t>(); gd.trafo = need(in, "trafo").p<int32_t>();
  gd.normals = need(in, "normals").p<double>(); gd.sqrtgram = need(in, "sqrtgram").p<double>();
  bosss_grid* grid = nullptr;
  CHECK(A.grid_upload(ctx, &gd, &grid), A.last_error);
  bosss_tables_desc td;
  td.D = gd.D; td.degree = ival(in, "degree"); td.N = ival(in, "N"); td.Kv = ival(in, "Kv"); td.Ke = ival(in, "Ke");
  td.ntrafo = ival(in, "ntrafo");
  td.vol_weights = need(in, "vw").p<double>(); td.phi = need(in, "phi").p<double>(); td.dphi = need(in, "dphi").p<double>();
  td.face_weights = need(in, "fw").p<double>(); td.trace = need(in, "trace").p<double>(); td.dtrace = need(in, "dtrace").p<double>();
  bosss_form_params fp;
  const double* prm = need(in, "params").p<double>();   // penalty, nu, velocity[3], gamma, mach
  fp.penalty_base = prm[0]; fp.nu = prm[1];
  for (int d = 0; d < 3; d++) fp.velocity[d] = prm[2 + d];
  fp.gamma = prm[5]; fp.mach = prm[6];
  bosss_op* op = nullptr;
  CHECK(A.op_create(ctx, grid, &td, ival(in, "kind"), &fp, BOSSS_FP64, &op), A.last_error);
  int64_t ndof = 0;
  int32_t nv = 0, nm = 0;
  CHECK(A.op_info(op, &ndof, &nv, &nm), A.last_error);
  const double* U = need(in, "U").p<double>();
  const double* dU = need(in, "dU").p<double>();
  const double* W = need(in, "W").p<double>();
  std::vector<double> R(ndof, 0.0), JdU(ndof), JtW(ndof), Rab(need(in, "R0").p<double>(), need(in, "R0").p<double>() + ndof);
  CHECK(A.evaluate(op, 0.0, 1.0, 0.0, U, R.data()), A.last_error);
  CHECK(A.jvp(op, U, dU, JdU.data()), A.last_error);
  CHECK(A.vjp(op, U, W, JtW.data()), A.last_error);
  CHECK(A.evaluate(op, 0.0, 2.0, 0.5, U, Rab.data()), A.last_error);   // Rab <- 0.5 R0 + 2 R(U)
  // the same residual and VJP through device buffers and the device-resident entry points
  void *dUb = nullptr, *dRb = nullptr, *dWb = nullptr, *dGb = nullptr;
  const int64_t bytes = ndof * 8;
  CHECK(A.buffer_create(ctx, bytes, &dUb), A.last_error);
  CHECK(A.buffer_create(ctx, bytes, &dRb), A.last_error);
  CHECK(A.buffer_create(ctx, bytes, &dWb), A.last_error);
  CHECK(A.buffer_create(ctx, bytes, &dGb), A.last_error);
  CHECK(A.buffer_upload(ctx, dUb, U, bytes), A.last_error);
  CHECK(A.buffer_upload(ctx, dWb, W, bytes), A.last_error);
  void* s = A.ctx_stream(ctx);
  CHECK(A.evaluate_vjp_dev(op, 1.0, 0.0, (const double*)dUb, (double*)dRb, (const double*)dWb, (double*)dGb, s), A.last_error);
  CHECK(A.sync(ctx), A.last_error);
  std::vector<double> Rdev(ndof), Gdev(ndof);
  CHECK(A.buffer_download(ctx, Rdev.data(), dRb, bytes), A.last_error);
  CHECK(A.buffer_download(ctx, Gdev.data(), dGb, bytes), A.last_error);
  // CheckForNanOrInf: a NaN in one coefficient must be reported, a finite state must pass
  CHECK(A.op_set_check_finite(op, 1), A.last_error);
  std::vector<double> Ubad(U, U + ndof), Rbad(ndof);
  const int64_t ibad = ndof / 3;
  Ubad[ibad] = std::nan("");
  int rc_ok = A.evaluate(op, 0.0, 1.0, 0.0, U, Rbad.data());
  int rc_nan = A.evaluate(op, 0.0, 1.0, 0.0, Ubad.data(), Rbad.data());
  std::string msg_nan = A.last_error();
  int rc_nan_dev = 0;
  CHECK(A.buffer_upload(ctx, dUb, Ubad.data(), bytes), A.last_error);
  rc_nan_dev = A.evaluate_dev(op, 1.0, 0.0, (const double*)dUb, (double*)dRb, s);
  CHECK(A.op_set_check_finite(op, 0), A.last_error);
  std::printf("nan check: finite rc %d, NaN rc %d (%s), NaN _dev rc %d\n", rc_ok, rc_nan, msg_nan.c_str(), rc_nan_dev);
  for (void* p : {dUb, dRb, dWb, dGb}) CHECK(A.buffer_destroy(ctx, p), A.last_error);
  out.d("R", R); out.d("JdU", JdU); out.d("JtW", JtW); out.d("Rab", Rab); out.d("Rdev", Rdev); out.d("Gdev", Gdev);
  out.scalar("rc_finite", rc_ok); out.scalar("rc_nan", rc_nan); out.scalar("rc_nan_dev", rc_nan_dev);
  CHECK(A.op_destroy(op), A.last_error);
  CHECK(A.grid_destroy(grid), A.last_error);
  CHECK(A.ctx_destroy(ctx), A.last_error);
  return 0;
}

static int scenario_blocks(Api& A, Recs& in, Writer& out) {
  bosss_ctx* ctx = nullptr;
  CHECK(A.ctx_create(0, &ctx), A.last_error);
  Rec& nr = need(in, "n");
  const int nblk = (int)nr.n;
  const int32_t* n = nr.p<int32_t>();
  const int64_t* off = need(in, "off").p<int64_t>();
  const int64_t* pivoff = need(in, "pivoff").p<int64_t>();
  const int64_t* xoff = need(in, "xoff").p<int64_t>();
  Rec& Ar = need(in, "A");
  Rec& br = need(in, "b");
  const int32_t nmax = ival(in, "nmax");
  std::vector<double> LU(Ar.p<double>(), Ar.p<double>() + Ar.n), x(br.p<double>(), br.p<double>() + br.n);
  std::vector<int32_t> ipiv(need(in, "npiv").p<int32_t>()[0], 0), info(nblk, -7);
  CHECK(A.block_lu(ctx, nblk, nmax, n, off, LU.data(), ipiv.data(), pivoff, info.data(), 0), A.block_last_error);
  CHECK(A.block_solve(ctx, nblk, n, off, LU.data(), ipiv.data(), pivoff, xoff, x.data()), A.block_last_error);
  out.d("LU", LU); out.i("ipiv", ipiv); out.i("info", info); out.d("x", x);
  if (nmax <= 100) {
    std::vector<double> inv(Ar.p<double>(), Ar.p<double>() + Ar.n), y(br.n, 1.0);
    std::vector<int32_t> info2(nblk, -7);
    CHECK(A.block_lu(ctx, nblk, nmax, n, off, inv.data(), nullptr, nullptr, info2.data(), 1), A.block_last_error);
    CHECK(A.block_gemv(ctx, nblk, n, off, inv.data(), xoff, 1.0, br.p<double>(), 0.0, y.data()), A.block_last_error);
    out.d("inv", inv); out.d("y", y); out.i("info_inv", info2);
  }
  CHECK(A.ctx_destroy(ctx), A.last_error);
  return 0;
}

static int scenario_assemble(Api& A, Recs& in, Writer& out) {
  // same grid/tables records as the dg scenario; a linear form
  bosss_ctx* ctx = nullptr;
  CHECK(A.ctx_create(0, &ctx), A.last_error);
  bosss_grid_desc gd;
  gd.D = ival(in, "D"); gd.J = ival(in, "J"); gd.E = ival(in, "E");
  gd.detJ = need(in, "detJ").p<double>(); gd.invJ = need(in, "invJ").p<double>();
  gd.scaling = need(in, "scaling").p<double>(); gd.lenscale = need(in, "lenscale").p<double>();
  gd.cells = need(in, "cells").p<int32_t>(); gd.trafo = need(in, "trafo").p<int32_t>();
  gd.normals = need(in, "normals").p<double>(); gd.sqrtgram = need(in, "sqrtgram").p<double>();
  bosss_grid* grid = nullptr;
  CHECK(A.grid_upload(ctx, &gd, &grid), A.last_error);
  bosss_tables_desc td;
  td.D = gd.D; td.degree = ival(in, "degree"); td.N = ival(in, "N"); td.Kv = ival(in, "Kv"); td.Ke = ival(in, "Ke");
  td.ntrafo = ival(in, "ntrafo");
  td.vol_weights = need(in, "vw").p<double>(); td.phi = need(in, "phi").p<double>(); td.dphi = need(in, "dphi").p<double>();
  td.face_weights = need(in, "fw").p<double>(); td.trace = need(in, "trace").p<double>(); td.dtrace = need(in, "dtrace").p<double>();
  bosss_form_params fp;
  const double* prm = need(in, "params").p<double>();
  fp.penalty_base = prm[0]; fp.nu = prm[1];
  for (int d = 0; d < 3; d++) fp.velocity[d] = prm[2 + d];
  fp.gamma = prm[5]; fp.mach = prm[6];
  bosss_op* op = nullptr;
  CHECK(A.op_create(ctx, grid, &td, ival(in, "kind"), &fp, BOSSS_FP64, &op), A.last_error);
  int32_t S = 0, ncol = 0;
  int64_t entries = 0, ndof = 0;
  CHECK(A.op_matrix_layout(op, &S, &entries, &ncol, nullptr), A.last_error);
  CHECK(A.op_info(op, &ndof, nullptr, nullptr), A.last_error);
  std::vector<int32_t> cols((size_t)gd.J * S);
  CHECK(A.op_matrix_layout(op, nullptr, nullptr, nullptr, cols.data()), A.last_error);
  std::vector<double> zero(ndof, 0.0), blocks(entries, 0.0), y(ndof, 0.0), b(ndof, 0.0);
  CHECK(A.op_assemble(op, zero.data(), 1.0, 0.0, blocks.data()), A.last_error);
  const double* x = need(in, "U").p<double>();
  CHECK(A.bell_spmv(op, blocks.data(), 1.0, x, 0.0, y.data()), A.last_error);
  CHECK(A.evaluate(op, 0.0, 1.0, 0.0, zero.data(), b.data()), A.last_error);   // affine part R(0)
  out.d("blocks", blocks); out.i("cols", cols); out.d("y", y); out.d("b", b); out.scalar("S", S);
  CHECK(A.op_destroy(op), A.last_error);
  CHECK(A.grid_destroy(grid), A.last_error);
  CHECK(A.ctx_destroy(ctx), A.last_error);
  return 0;
}

static int scenario_xdg(Api& A, Recs& in, Writer& out) {
  bosss_ctx* ctx = nullptr;
  CHECK(A.ctx_create(0, &ctx), A.last_error);
  bosss_xdg_desc d;
  std::memset(&d, 0, sizeof d);
  d.D = ival(in, "D"); d.degree = ival(in, "degree"); d.N = ival(in, "N"); d.J = ival(in, "J");
  d.mode_exponents = need(in, "ex").p<int32_t>(); d.nspec = need(in, "nspec").p<int32_t>(); d.spec0 = need(in, "spec0").p<int32_t>();
  d.invJ = need(in, "invJ").p<double>(); d.scaling = need(in, "scal").p<double>(); d.lscale = need(in, "ls").p<double>();
  d.penalty_base = dval(in, "penalty");
  for (int s = 0; s < 2; s++) { d.mu[s] = need(in, "mu").p<double>()[s]; d.source[s] = need(in, "source").p<double>()[s]; }
  d.S = ival(in, "S");
  d.cols = need(in, "cols").p<int32_t>(); d.boff = need(in, "boff").p<int64_t>();
  d.nvol = (int32_t)need(in, "vc").n;
  d.vol_cell = need(in, "vc").p<int32_t>(); d.vol_spec = need(in, "vs").p<int32_t>(); d.vol_off = need(in, "vo").p<int64_t>();
  d.vol_nodes = need(in, "vn").p<double>(); d.vol_w = need(in, "vw").p<double>();
  d.nface = (int32_t)need(in, "fsp").n;
  d.face_cells = need(in, "fc").p<int32_t>(); d.face_slot = need(in, "fsl").p<int32_t>(); d.face_spec = need(in, "fsp").p<int32_t>();
  d.face_off = need(in, "fo").p<int64_t>(); d.face_nodes_in = need(in, "fni").p<double>(); d.face_nodes_out = need(in, "fno").p<double>();
  d.face_w = need(in, "fw").p<double>(); d.face_normal = need(in, "fnm").p<double>();
  d.face_gD = in.count("gd") ? need(in, "gd").p<double>() : nullptr;
  d.nifc = (int32_t)need(in, "ic").n;
  d.ifc_cell = need(in, "ic").p<int32_t>(); d.ifc_off = need(in, "io").p<int64_t>(); d.ifc_nodes = need(in, "inn").p<double>();
  d.ifc_w = need(in, "iw").p<double>(); d.ifc_normal = need(in, "inm").p<double>();
  bosss_xdg* x = nullptr;
  CHECK(A.xdg_create(ctx, &d, &x), A.xdg_last_error);
  int64_t ndof = 0, nent = 0;
  CHECK(A.xdg_info(x, &ndof, &nent), A.xdg_last_error);
  std::vector<double> blocks(nent), affine(ndof), y(ndof, 0.0), rhs(ndof);
  CHECK(A.xdg_assemble(x, blocks.data(), affine.data()), A.xdg_last_error);
  const double* v = need(in, "v").p<double>();
  CHECK(A.xdg_spmv(x, blocks.data(), 1.0, v, 0.0, y.data()), A.xdg_last_error);
  for (int64_t i = 0; i < ndof; i++) rhs[i] = -affine[i];
  for (int prec : {BOSSS_FP64, BOSSS_FP32}) {
    std::vector<double> sol(ndof, 0.0);
    int32_t it = 0, outer = 0;
    double rel = 0;
    CHECK(A.xdg_solve(x, blocks.data(), rhs.data(), sol.data(), 1e-10, 20000, prec, &it, &outer, &rel), A.xdg_last_error);
    std::printf("xdg solve %s: %d iterations, %d outer, relres %.3e\n", prec == BOSSS_FP64 ? "fp64" : "fp32", it, outer, rel);
    const std::string tag = prec == BOSSS_FP64 ? "fp64" : "fp32";
    out.d("sol_" + tag, sol);
    out.scalar("it_" + tag, it);
    out.scalar_d("relres_" + tag, rel);
  }
  out.d("blocks", blocks); out.d("affine", affine); out.d("y", y);
  CHECK(A.xdg_destroy(x), A.xdg_last_error);
  CHECK(A.ctx_destroy(ctx), A.last_error);
  return 0;
}

int main(int argc, char** argv) {
  if (argc != 5 && !(argc == 3 && std::strcmp(argv[2], "bind") == 0)) {
    std::fprintf(stderr, "usage: %s <library.so> bind | <dg|blocks|assemble|xdg> <input.rec> <output.rec>\n", argv[0]);
    return 2;
  }
  Api api;
  if (!load_all(argv[1], api)) return 3;
  std::printf("bound every entry point of bosss_hip.h by name\n");
  if (argc == 3) return 0;
  Recs in;
  if (!read_recs(argv[3], in)) {
    std::fprintf(stderr, "cannot read %s\n", argv[3]);
    return 2;
  }
  Writer out(argv[4]);
  if (!out.f) return 2;
  const std::string sc = argv[2];
  int rc = sc == "dg" ? scenario_dg(api, in, out)
         : sc == "blocks" ? scenario_blocks(api, in, out)
         : sc == "assemble" ? scenario_assemble(api, in, out)
         : sc == "xdg" ? scenario_xdg(api, in, out)
         : 2;
  std::printf("scenario %s: rc %d\n", sc.c_str(), rc);
  return rc;
}
