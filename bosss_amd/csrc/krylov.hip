// Device-resident GMRES building blocks for the Newton-Krylov driver.
//
// The reference's matrix-free GMRES (Newton.MatrixFreeGMRES, Newton.cs:1128-1500) keeps the
// Krylov vectors, the Hessenberg matrix and the Givens rotations on the host.  Here the whole
// Arnoldi step runs on the GPU: classical Gram-Schmidt twice (one batched dot product per pass,
// as the reference's GMRES dots), the norm, the new basis vector, column k of the Hessenberg
// matrix, the previous rotations applied to it, rotation k and the rotated right-hand side g,
// whose entry k+1 is the residual estimate.  The host reads g only when it wants to test
// convergence, so a run of Arnoldi steps needs no synchronisation and can be captured into a
// graph (no allocation, no memset: the caller provides the workspace).  The final update
// x += Z y solves the k x k triangular system on the device too.
//
// Dot products are summed in a fixed order (per-thread strided partial sums, wavefront and
// workgroup trees, then the workgroup partials in index order), so results are bitwise
// reproducible for a given vector length.  Vectors are double or float; every dot product,
// norm and the Hessenberg data are double.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <initializer_list>
#include <string>

#include "../../include/bosss_hip.h"

namespace bosss {
int ctx_device_stream(bosss_ctx* ctx, int* device, hipStream_t* stream);
}

namespace {

constexpr int KR_THREADS = 256;
constexpr int KR_MAXBLK = 2048;     // workgroups of the streaming kernels (fewer for short vectors)
constexpr int KR_MAXM = 256;        // largest restart length
constexpr int KR_CHUNK = 16;        // basis vectors per pass of the dot kernel
constexpr int KR_FUSED = 32;        // longest basis whose second-pass dots ride on the first update (16 with 16-byte packs)

thread_local std::string g_kr_err;

int kr_blocks(int64_t n) {   // n: elements, or VEC-element packs
  int64_t b = (n + 4 * KR_THREADS - 1) / (4 * KR_THREADS);
  return (int)(b < 1 ? 1 : (b > KR_MAXBLK ? KR_MAXBLK : b));
}

// workspace layout (doubles): dot partials [KR_MAXM+1][KR_MAXBLK], norm partials [KR_MAXBLK],
// h [KR_MAXM+1], scalars [8] (0: 1/|w|)
constexpr int64_t WS_PART = 0;
constexpr int64_t WS_NORM = WS_PART + (int64_t)(KR_MAXM + 1) * KR_MAXBLK;
constexpr int64_t WS_H = WS_NORM + KR_MAXBLK;
constexpr int64_t WS_S = WS_H + KR_MAXM + 1;
constexpr int64_t WS_SIZE = WS_S + 8;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Workgroup sum in a fixed order (wavefront trees, then the wavefronts in index order); the
// result is valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int r = 0; r < KR_THREADS / 64; r++) s += red[r];
  return s;
}

// Stores the workgroup sums of acc[0..k1-i0) as part[i0 + q][blockIdx.x].
template <int C>
__device__ __forceinline__ void store_partials(const double (&acc)[C], int i0, int k1, double (*red)[C], double* part) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < C; q++) {
    const double s = wave_sum(acc[q]);
    if (lane == 0) red[wv][q] = s;
  }
  __syncthreads();
  if (threadIdx.x < C && i0 + (int)threadIdx.x < k1) {
    double s = 0.0;
    for (int r = 0; r < KR_THREADS / 64; r++) s += red[r][threadIdx.x];
    part[(int64_t)(i0 + threadIdx.x) * KR_MAXBLK + blockIdx.x] = s;
  }
  __syncthreads();
}

// VEC consecutive elements per thread (16-byte loads: 2 doubles or 4 floats, or 1 where the
// vector length, the row stride or a pointer does not allow it).  Per element every sum runs in
// the same index order as with VEC = 1; only the partial sums per thread group VEC elements.
template <typename Real, int VEC>
struct Pack {
  Real v[VEC];
};
template <typename Real, int VEC>
__device__ __forceinline__ Pack<Real, VEC> load_pack(const Real* __restrict__ p) {
  Pack<Real, VEC> r;
  if constexpr (VEC == 1) {
    r.v[0] = *p;
  } else {
    typedef Real vt __attribute__((ext_vector_type(VEC)));
    const vt x = *reinterpret_cast<const vt*>(p);
#pragma unroll
    for (int e = 0; e < VEC; e++) r.v[e] = x[e];
  }
  return r;
}
template <typename Real, int VEC>
__device__ __forceinline__ void store_pack(Real* __restrict__ p, const Pack<Real, VEC>& r) {
  if constexpr (VEC == 1) {
    *p = r.v[0];
  } else {
    typedef Real vt __attribute__((ext_vector_type(VEC)));
    vt x;
#pragma unroll
    for (int e = 0; e < VEC; e++) x[e] = r.v[e];
    *reinterpret_cast<vt*>(p) = x;
  }
}

// part[i][blk] = sum over the workgroup's share of t of V[i][t] w[t], i < k1
template <typename Real, int VEC>
__global__ void __launch_bounds__(KR_THREADS) k_mdot(int64_t n, int k1, const Real* __restrict__ V, int64_t ldv,
                                                     const Real* __restrict__ w, double* __restrict__ part) {
  __shared__ double red[KR_THREADS / 64][KR_CHUNK];
  const int64_t nv = n / VEC;
  for (int i0 = 0; i0 < k1; i0 += KR_CHUNK) {
    double acc[KR_CHUNK];
#pragma unroll
    for (int q = 0; q < KR_CHUNK; q++) acc[q] = 0.0;
    for (int64_t tv = (int64_t)blockIdx.x * KR_THREADS + threadIdx.x; tv < nv; tv += (int64_t)gridDim.x * KR_THREADS) {
      const Pack<Real, VEC> wt = load_pack<Real, VEC>(w + tv * VEC);
#pragma unroll
      for (int q = 0; q < KR_CHUNK; q++)
        if (i0 + q < k1) {
          const Pack<Real, VEC> vq = load_pack<Real, VEC>(V + (int64_t)(i0 + q) * ldv + tv * VEC);
#pragma unroll
          for (int e = 0; e < VEC; e++) acc[q] += (double)vq.v[e] * (double)wt.v[e];
        }
    }
    store_partials<KR_CHUNK>(acc, i0, k1, red, part);
  }
}

// h[i] = sum_blk part[i][blk]; H[i + k ldh] = h[i] (pass 0) or += h[i] (pass 1).  One workgroup per i.
__global__ void __launch_bounds__(KR_THREADS) k_hreduce(int nb, const double* __restrict__ part, double* __restrict__ h,
                                                        double* __restrict__ Hk, int accumulate) {
  __shared__ double red[KR_THREADS / 64];
  const int i = blockIdx.x;
  double s = 0.0;
#pragma unroll 8
  for (int b = threadIdx.x; b < nb; b += KR_THREADS) s += part[(int64_t)i * KR_MAXBLK + b];
  s = block_sum(s, red);
  if (threadIdx.x == 0) {
    h[i] = s;
    Hk[i] = accumulate ? Hk[i] + s : s;
  }
}

// s[e] = sum_{i<k1} V[i][t + e] h[i], in index order, with the loads of eight basis vectors issued together
template <typename Real, int VEC>
__device__ __forceinline__ void basis_combination(int k1, const Real* __restrict__ V, int64_t ldv, int64_t t, const double* hs,
                                                  double (&s)[VEC]) {
#pragma unroll
  for (int e = 0; e < VEC; e++) s[e] = 0.0;
  int i = 0;
  for (; i + 8 <= k1; i += 8) {
    Pack<Real, VEC> v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) v[q] = load_pack<Real, VEC>(V + (int64_t)(i + q) * ldv + t);
#pragma unroll
    for (int q = 0; q < 8; q++)
#pragma unroll
      for (int e = 0; e < VEC; e++) s[e] += (double)v[q].v[e] * hs[i + q];
  }
  for (; i < k1; i++) {
    const Pack<Real, VEC> v = load_pack<Real, VEC>(V + (int64_t)i * ldv + t);
#pragma unroll
    for (int e = 0; e < VEC; e++) s[e] += (double)v.v[e] * hs[i];
  }
}

// w[t] -= sum_{i<k1} V[i][t] h[i]; with norm: npart[blk] = workgroup's sum of the new w[t]^2
template <typename Real, int VEC>
__global__ void __launch_bounds__(KR_THREADS) k_mupdate(int64_t n, int k1, const Real* __restrict__ V, int64_t ldv,
                                                        const double* __restrict__ h, Real* __restrict__ w, double* __restrict__ npart) {
  __shared__ double hs[KR_MAXM + 1];
  __shared__ double red[KR_THREADS / 64];
  for (int i = threadIdx.x; i < k1; i += KR_THREADS) hs[i] = h[i];
  __syncthreads();
  const int64_t nv = n / VEC;
  double nrm = 0.0;
  for (int64_t tv = (int64_t)blockIdx.x * KR_THREADS + threadIdx.x; tv < nv; tv += (int64_t)gridDim.x * KR_THREADS) {
    double s[VEC];
    basis_combination<Real, VEC>(k1, V, ldv, tv * VEC, hs, s);
    Pack<Real, VEC> x = load_pack<Real, VEC>(w + tv * VEC);
#pragma unroll
    for (int e = 0; e < VEC; e++) {
      x.v[e] = (Real)((double)x.v[e] - s[e]);
      nrm += (double)x.v[e] * (double)x.v[e];
    }
    store_pack<Real, VEC>(w + tv * VEC, x);
  }
  if (npart) {
    nrm = block_sum(nrm, red);
    if (threadIdx.x == 0) npart[blockIdx.x] = nrm;
  }
}

// The first Gram-Schmidt update fused with the second pass's dot products: w[t] -= sum_i V[i][t] h[i],
// then part[i][blk] = workgroup's sum of V[i][t] w[t] over the updated w, with the basis values kept
// in registers between the two (k1 <= C).  Same sums, in the same order, as k_mupdate followed by
// k_mdot on the same grid.
template <typename Real, int C, int VEC>
__global__ void __launch_bounds__(KR_THREADS) k_mupdate_dot(int64_t n, int k1, const Real* __restrict__ V, int64_t ldv,
                                                            const double* __restrict__ h, Real* __restrict__ w,
                                                            double* __restrict__ part) {
  __shared__ double hs[C];
  __shared__ double red[KR_THREADS / 64][C];
  for (int i = threadIdx.x; i < C; i += KR_THREADS) hs[i] = i < k1 ? h[i] : 0.0;
  __syncthreads();
  const int64_t nv = n / VEC;
  double acc[C];
#pragma unroll
  for (int q = 0; q < C; q++) acc[q] = 0.0;
  for (int64_t tv = (int64_t)blockIdx.x * KR_THREADS + threadIdx.x; tv < nv; tv += (int64_t)gridDim.x * KR_THREADS) {
    Pack<Real, VEC> v[C];
#pragma unroll
    for (int q = 0; q < C; q++) {
      if (q < k1) v[q] = load_pack<Real, VEC>(V + (int64_t)q * ldv + tv * VEC);
      else {
#pragma unroll
        for (int e = 0; e < VEC; e++) v[q].v[e] = Real(0);
      }
    }
    Pack<Real, VEC> x = load_pack<Real, VEC>(w + tv * VEC);
    double xd[VEC];
#pragma unroll
    for (int e = 0; e < VEC; e++) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < C; q++)
        if (q < k1) s += (double)v[q].v[e] * hs[q];
      x.v[e] = (Real)((double)x.v[e] - s);
      xd[e] = (double)x.v[e];
    }
    store_pack<Real, VEC>(w + tv * VEC, x);
#pragma unroll
    for (int q = 0; q < C; q++)
#pragma unroll
      for (int e = 0; e < VEC; e++) acc[q] += (double)v[q].v[e] * xd[e];
  }
  store_partials<C>(acc, 0, k1, red, part);
}

// |w| from the norm partials, column k of H completed and rotated, rotation k formed, g updated,
// res[k] = |g[k+1]| (the residual estimate after step k), and 1/|w| (0 on breakdown) for the new
// basis vector.  One workgroup; thread 0 does the rotations.
__global__ void __launch_bounds__(KR_THREADS) k_finish(int nb, int k, int ldh, const double* __restrict__ npart, double* __restrict__ H,
                                                       double* __restrict__ cs, double* __restrict__ sn, double* __restrict__ g,
                                                       double* __restrict__ res, double* __restrict__ scal) {
  __shared__ double red[KR_THREADS / 64];
  double s = 0.0;
#pragma unroll 8
  for (int b = threadIdx.x; b < nb; b += KR_THREADS) s += npart[b];
  s = block_sum(s, red);
  if (threadIdx.x != 0) return;
  const double hn = sqrt(s);
  double* Hk = H + (int64_t)k * ldh;
  Hk[k + 1] = hn;
  scal[0] = hn > 0.0 ? 1.0 / hn : 0.0;
  for (int i = 0; i < k; i++) {
    const double a = Hk[i], b = Hk[i + 1];
    Hk[i] = cs[i] * a + sn[i] * b;
    Hk[i + 1] = -sn[i] * a + cs[i] * b;
  }
  const double a = Hk[k], b = Hk[k + 1], den = hypot(a, b);
  const double c = den == 0.0 ? 1.0 : a / den, sr = den == 0.0 ? 0.0 : b / den;
  cs[k] = c;
  sn[k] = sr;
  Hk[k] = c * a + sr * b;
  Hk[k + 1] = 0.0;
  g[k + 1] = -sr * g[k];
  g[k] = c * g[k];
  res[k] = fabs(g[k + 1]);   // later rotations change g[k+1]; the estimate of step k is kept here
}

template <typename Real, int VEC>
__global__ void __launch_bounds__(KR_THREADS) k_scale_to(int64_t n, const Real* __restrict__ w, const double* __restrict__ scal,
                                                         Real* __restrict__ out) {
  const double inv = scal[0];
  const int64_t nv = n / VEC;
  for (int64_t tv = (int64_t)blockIdx.x * KR_THREADS + threadIdx.x; tv < nv; tv += (int64_t)gridDim.x * KR_THREADS) {
    Pack<Real, VEC> x = load_pack<Real, VEC>(w + tv * VEC);
#pragma unroll
    for (int e = 0; e < VEC; e++) x.v[e] = (Real)((double)x.v[e] * inv);
    store_pack<Real, VEC>(out + tv * VEC, x);
  }
}

// y = R^-1 g[0..k1) with R the leading k1 x k1 upper triangle of H (back substitution)
__global__ void k_trisolve(int k1, int ldh, const double* __restrict__ H, const double* __restrict__ g, double* __restrict__ y) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int i = k1 - 1; i >= 0; i--) {
    double s = g[i];
    for (int j = i + 1; j < k1; j++) s -= H[(int64_t)j * ldh + i] * y[j];
    const double d = H[(int64_t)i * ldh + i];
    y[i] = d != 0.0 ? s / d : 0.0;
  }
}

// x[t] += sum_{i<k1} Z[i][t] y[i]
template <typename Real, int VEC>
__global__ void __launch_bounds__(KR_THREADS) k_xupdate(int64_t n, int k1, const Real* __restrict__ Z, int64_t ldz,
                                                        const double* __restrict__ y, Real* __restrict__ x) {
  __shared__ double ys[KR_MAXM];
  for (int i = threadIdx.x; i < k1; i += KR_THREADS) ys[i] = y[i];
  __syncthreads();
  const int64_t nv = n / VEC;
  for (int64_t tv = (int64_t)blockIdx.x * KR_THREADS + threadIdx.x; tv < nv; tv += (int64_t)gridDim.x * KR_THREADS) {
    double s[VEC];
    basis_combination<Real, VEC>(k1, Z, ldz, tv * VEC, ys, s);
    Pack<Real, VEC> xv = load_pack<Real, VEC>(x + tv * VEC);
#pragma unroll
    for (int e = 0; e < VEC; e++) xv.v[e] = (Real)((double)xv.v[e] + s[e]);
    store_pack<Real, VEC>(x + tv * VEC, xv);
  }
}

// y_b <- alpha A_b x_b + beta y_b for nblk uniform n x n blocks (n <= 128) stored transposed:
// AT[b][k][i] = A_b[i][k].  One workgroup per block, thread i owns row i; for each k the
// threads read one column of A_b, n consecutive entries.  x_b is staged in LDS.
template <typename Real>
__global__ void __launch_bounds__(128) k_bdiag_gemv_t(int nblk, int n, const Real* __restrict__ AT, double alpha,
                                                      const Real* __restrict__ x, double beta, Real* __restrict__ y) {
  __shared__ Real xs[128];
  const int b = blockIdx.x, i = threadIdx.x;
  if (b >= nblk) return;
  const int64_t xo = (int64_t)b * n;
  if (i < n) xs[i] = x[xo + i];
  __syncthreads();
  if (i >= n) return;
  const Real* A = AT + (int64_t)b * n * n + i;
  Real acc = Real(0);
  for (int k = 0; k < n; k++) acc += A[(int64_t)k * n] * xs[k];
  y[xo + i] = (Real)(alpha * (double)acc + (beta != 0.0 ? beta * (double)y[xo + i] : 0.0));
}

int kr_select(bosss_ctx* ctx) {
  if (!ctx) { g_kr_err = "ctx is NULL"; return BOSSS_ERR_ARG; }
  int dev = 0;
  hipStream_t st = nullptr;
  if (bosss::ctx_device_stream(ctx, &dev, &st) != BOSSS_OK) { g_kr_err = "bad context"; return BOSSS_ERR_ARG; }
  if (hipSetDevice(dev) != hipSuccess) { g_kr_err = "hipSetDevice failed"; return BOSSS_ERR_HIP; }
  return BOSSS_OK;
}

// 16-byte packs when the vector length, the row stride and every pointer allow them
template <typename Real>
bool kr_vec_ok(int64_t n, int64_t ld, std::initializer_list<const void*> ptrs) {
  constexpr int VEC = 16 / sizeof(Real);
  if (n % VEC || ld % VEC) return false;
  for (const void* p : ptrs)
    if (reinterpret_cast<uintptr_t>(p) % 16) return false;
  return true;
}

template <typename Real, int VEC>
void arnoldi_v(int64_t n, int k, int ldh, Real* V, int64_t ldv, Real* w, double* H, double* cs, double* sn, double* g, double* res,
               double* work, hipStream_t s) {
  const int nb = kr_blocks(n / VEC), k1 = k + 1;
  double* part = work + WS_PART;
  double* npart = work + WS_NORM;
  double* h = work + WS_H;
  double* scal = work + WS_S;
  double* Hk = H + (int64_t)k * ldh;
  const dim3 grid(nb), blk(KR_THREADS);
  // pass 0: h = V^T w, H[:, k] = h
  hipLaunchKernelGGL((k_mdot<Real, VEC>), grid, blk, 0, s, n, k1, (const Real*)V, ldv, (const Real*)w, part);
  hipLaunchKernelGGL(k_hreduce, dim3(k1), blk, 0, s, nb, (const double*)part, h, Hk, 0);
  // w -= V h, and the pass-1 dots V^T w of the updated w
  if (k1 <= 8)
    hipLaunchKernelGGL((k_mupdate_dot<Real, 8, VEC>), grid, blk, 0, s, n, k1, (const Real*)V, ldv, (const double*)h, w, part);
  else if (k1 <= 16)
    hipLaunchKernelGGL((k_mupdate_dot<Real, 16, VEC>), grid, blk, 0, s, n, k1, (const Real*)V, ldv, (const double*)h, w, part);
  else if (VEC == 1 && k1 <= KR_FUSED)   // 32 packs of 16 bytes would not leave two wavefronts per SIMD
    hipLaunchKernelGGL((k_mupdate_dot<Real, KR_FUSED, 1>), grid, blk, 0, s, n, k1, (const Real*)V, ldv, (const double*)h, w, part);
  else {
    hipLaunchKernelGGL((k_mupdate<Real, VEC>), grid, blk, 0, s, n, k1, (const Real*)V, ldv, (const double*)h, w, nullptr);
    hipLaunchKernelGGL((k_mdot<Real, VEC>), grid, blk, 0, s, n, k1, (const Real*)V, ldv, (const Real*)w, part);
  }
  // pass 1: H[:, k] += h, w -= V h with the norm partials of the result
  hipLaunchKernelGGL(k_hreduce, dim3(k1), blk, 0, s, nb, (const double*)part, h, Hk, 1);
  hipLaunchKernelGGL((k_mupdate<Real, VEC>), grid, blk, 0, s, n, k1, (const Real*)V, ldv, (const double*)h, w, npart);
  hipLaunchKernelGGL(k_finish, dim3(1), blk, 0, s, nb, k, ldh, (const double*)npart, H, cs, sn, g, res, scal);
  hipLaunchKernelGGL((k_scale_to<Real, VEC>), grid, blk, 0, s, n, (const Real*)w, (const double*)scal, V + (int64_t)(k + 1) * ldv);
}

template <typename Real>
void arnoldi(int64_t n, int k, int ldh, Real* V, int64_t ldv, Real* w, double* H, double* cs, double* sn, double* g, double* res,
             double* work, hipStream_t s) {
  if (kr_vec_ok<Real>(n, ldv, {V, w}))
    arnoldi_v<Real, 16 / sizeof(Real)>(n, k, ldh, V, ldv, w, H, cs, sn, g, res, work, s);
  else
    arnoldi_v<Real, 1>(n, k, ldh, V, ldv, w, H, cs, sn, g, res, work, s);
}

template <typename Real>
void xupdate(int64_t n, int k1, const Real* Z, int64_t ldz, const double* y, Real* x, hipStream_t s) {
  constexpr int VEC = 16 / sizeof(Real);
  if (kr_vec_ok<Real>(n, ldz, {Z, x}))
    hipLaunchKernelGGL((k_xupdate<Real, VEC>), dim3(kr_blocks(n / VEC)), dim3(KR_THREADS), 0, s, n, k1, Z, ldz, y, x);
  else
    hipLaunchKernelGGL((k_xupdate<Real, 1>), dim3(kr_blocks(n)), dim3(KR_THREADS), 0, s, n, k1, Z, ldz, y, x);
}

}  // namespace

extern "C" {

const char* bosss_hip_krylov_last_error(void) { return g_kr_err.c_str(); }

int bosss_hip_gmres_workspace(int32_t m, int64_t* doubles) {
  if (!doubles) { g_kr_err = "NULL argument"; return BOSSS_ERR_ARG; }
  if (m < 1 || m > KR_MAXM) { g_kr_err = "restart length must be 1.." + std::to_string(KR_MAXM); return BOSSS_ERR_ARG; }
  *doubles = WS_SIZE;
  return BOSSS_OK;
}

int bosss_hip_gmres_arnoldi_dev(bosss_ctx* ctx, int64_t n, int32_t k, int32_t m, int32_t precision, void* V, int64_t ldv, void* w,
                                double* H, double* cs, double* sn, double* g, double* res, double* work, void* stream) {
  if (int rc = kr_select(ctx)) return rc;
  if (!V || !w || !H || !cs || !sn || !g || !res || !work) { g_kr_err = "NULL argument"; return BOSSS_ERR_ARG; }
  if (m < 1 || m > KR_MAXM || k < 0 || k >= m) { g_kr_err = "need 0 <= k < m <= " + std::to_string(KR_MAXM); return BOSSS_ERR_ARG; }
  if (n < 1 || ldv < n) { g_kr_err = "need n >= 1 and ldv >= n"; return BOSSS_ERR_ARG; }
  hipStream_t s = (hipStream_t)stream;
  if (precision == BOSSS_FP64)
    arnoldi<double>(n, k, m + 1, (double*)V, ldv, (double*)w, H, cs, sn, g, res, work, s);
  else if (precision == BOSSS_FP32)
    arnoldi<float>(n, k, m + 1, (float*)V, ldv, (float*)w, H, cs, sn, g, res, work, s);
  else { g_kr_err = "precision must be BOSSS_FP64 or BOSSS_FP32"; return BOSSS_ERR_ARG; }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { g_kr_err = hipGetErrorString(e); return BOSSS_ERR_HIP; }
  return BOSSS_OK;
}

int bosss_hip_gmres_update_dev(bosss_ctx* ctx, int64_t n, int32_t k1, int32_t m, int32_t precision, const void* Z, int64_t ldz,
                               const double* H, const double* g, double* y, void* x, void* stream) {
  if (int rc = kr_select(ctx)) return rc;
  if (!Z || !H || !g || !y || !x) { g_kr_err = "NULL argument"; return BOSSS_ERR_ARG; }
  if (m < 1 || m > KR_MAXM || k1 < 0 || k1 > m) { g_kr_err = "need 0 <= k1 <= m <= " + std::to_string(KR_MAXM); return BOSSS_ERR_ARG; }
  if (n < 1 || ldz < n) { g_kr_err = "need n >= 1 and ldz >= n"; return BOSSS_ERR_ARG; }
  if (k1 == 0) return BOSSS_OK;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_trisolve, dim3(1), dim3(64), 0, s, k1, m + 1, H, g, y);
  if (precision == BOSSS_FP64)
    xupdate<double>(n, k1, (const double*)Z, ldz, (const double*)y, (double*)x, s);
  else if (precision == BOSSS_FP32)
    xupdate<float>(n, k1, (const float*)Z, ldz, (const double*)y, (float*)x, s);
  else { g_kr_err = "precision must be BOSSS_FP64 or BOSSS_FP32"; return BOSSS_ERR_ARG; }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { g_kr_err = hipGetErrorString(e); return BOSSS_ERR_HIP; }
  return BOSSS_OK;
}

int bosss_hip_block_gemv_uniform_dev(bosss_ctx* ctx, int nblk, int32_t n, int32_t precision, const void* AinvT, double alpha,
                                     const void* x, double beta, void* y, void* stream) {
  if (nblk <= 0) return BOSSS_OK;
  if (int rc = kr_select(ctx)) return rc;
  if (!AinvT || !x || !y) { g_kr_err = "NULL argument"; return BOSSS_ERR_ARG; }
  if (n < 1 || n > 128) { g_kr_err = "uniform block size must be 1..128"; return BOSSS_ERR_ARG; }
  const dim3 block(n <= 64 ? 64 : 128);
  hipStream_t s = (hipStream_t)stream;
  if (precision == BOSSS_FP64)
    hipLaunchKernelGGL(k_bdiag_gemv_t<double>, dim3(nblk), block, 0, s, nblk, n, (const double*)AinvT, alpha, (const double*)x, beta, (double*)y);
  else if (precision == BOSSS_FP32)
    hipLaunchKernelGGL(k_bdiag_gemv_t<float>, dim3(nblk), block, 0, s, nblk, n, (const float*)AinvT, alpha, (const float*)x, beta, (float*)y);
  else { g_kr_err = "precision must be BOSSS_FP64 or BOSSS_FP32"; return BOSSS_ERR_ARG; }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { g_kr_err = hipGetErrorString(e); return BOSSS_ERR_HIP; }
  return BOSSS_OK;
}

}  // extern "C"
