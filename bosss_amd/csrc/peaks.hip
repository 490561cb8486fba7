// Arithmetic ceilings measured on the device: dense FP64 and FP32 MFMA issue and FP64 FMA
// issue, the rates the roofline fractions of the DG kernels are quoted against (published:
// 78.6, 157.3 and 78.6 TFLOP/s), and the rate of FP64 MFMA and FMA interleaved, which shows
// whether the two run side by side or share one FP64 rate.  Every wavefront runs several
// independent accumulation chains, long enough that launch overhead is negligible; the sums
// are stored so that the chains are not removed.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/bosss_hip.h"

namespace bosss {
int ctx_device_stream(bosss_ctx* ctx, int* device, hipStream_t* stream);
}

namespace {

constexpr int PK_THREADS = 256;
constexpr int PK_BLOCKS = 4096;     // 16 384 wavefronts: 16 per SIMD on 256 CUs
constexpr int PK_ITERS = 1024;
constexpr int PK_CHAINS = 4;        // independent MFMA accumulators per wavefront
constexpr int PK_FMA_CHAINS = 8;    // independent FMA chains per lane

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(PK_THREADS) k_peak_mfma64(double* out, double seed) {
  f64x4 acc[PK_CHAINS];
#pragma unroll
  for (int c = 0; c < PK_CHAINS; c++) acc[c] = f64x4{0.0, 0.0, 0.0, 0.0};
  const double a = seed + threadIdx.x * 1e-9, b = seed - threadIdx.x * 1e-9;
  for (int i = 0; i < PK_ITERS; i++) {
#pragma unroll
    for (int c = 0; c < PK_CHAINS; c++) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < PK_CHAINS; c++) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
  out[(size_t)blockIdx.x * PK_THREADS + threadIdx.x] = s;
}

__global__ void __launch_bounds__(PK_THREADS) k_peak_mfma32(double* out, float seed) {
  f32x4 acc[PK_CHAINS];
#pragma unroll
  for (int c = 0; c < PK_CHAINS; c++) acc[c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const float a = seed + threadIdx.x * 1e-6f, b = seed - threadIdx.x * 1e-6f;
  for (int i = 0; i < PK_ITERS; i++) {
#pragma unroll
    for (int c = 0; c < PK_CHAINS; c++) acc[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[c], 0, 0, 0);
  }
  float s = 0.0f;
#pragma unroll
  for (int c = 0; c < PK_CHAINS; c++) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
  out[(size_t)blockIdx.x * PK_THREADS + threadIdx.x] = s;
}

__global__ void __launch_bounds__(PK_THREADS) k_peak_fma64(double* out, double seed) {
  double x[PK_FMA_CHAINS];
#pragma unroll
  for (int c = 0; c < PK_FMA_CHAINS; c++) x[c] = seed + c * 1e-3 + threadIdx.x * 1e-9;
  const double a = 0.999999, b = 1e-7;
  for (int i = 0; i < PK_ITERS * 4; i++) {
#pragma unroll
    for (int c = 0; c < PK_FMA_CHAINS; c++) x[c] = fma(x[c], a, b);
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < PK_FMA_CHAINS; c++) s += x[c];
  out[(size_t)blockIdx.x * PK_THREADS + threadIdx.x] = s;
}

// FP64 MFMA and FP64 FMA interleaved at equal flops: each v_mfma_f64_16x16x4f64 (2048 flops per
// wavefront) is followed by PK_MIX_FMA FMAs (16 x 64 lanes x 2 = 2048).  If the matrix core ran
// beside the FP64 VALU, this stream would approach the sum of the two single rates; if the two
// share one FP64 rate, it runs at about either single rate.
constexpr int PK_MIX_FMA = 16;

__global__ void __launch_bounds__(PK_THREADS) k_peak_mix64(double* out, double seed) {
  f64x4 acc[PK_CHAINS];
#pragma unroll
  for (int c = 0; c < PK_CHAINS; c++) acc[c] = f64x4{0.0, 0.0, 0.0, 0.0};
  double x[PK_FMA_CHAINS];
#pragma unroll
  for (int c = 0; c < PK_FMA_CHAINS; c++) x[c] = seed + c * 1e-3 + threadIdx.x * 1e-9;
  const double a = seed + threadIdx.x * 1e-9, b = seed - threadIdx.x * 1e-9;
  const double fa = 0.999999, fb = 1e-7;
  for (int i = 0; i < PK_ITERS; i++) {
#pragma unroll
    for (int c = 0; c < PK_CHAINS; c++) {
      acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < PK_MIX_FMA; j++) x[j % PK_FMA_CHAINS] = fma(x[j % PK_FMA_CHAINS], fa, fb);
      // issue order: one MFMA, then its 16 FMAs (not all MFMAs of the iteration first)
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x002, PK_MIX_FMA, 0);
    }
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < PK_CHAINS; c++) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
#pragma unroll
  for (int c = 0; c < PK_FMA_CHAINS; c++) s += x[c];
  out[(size_t)blockIdx.x * PK_THREADS + threadIdx.x] = s;
}

thread_local std::string g_peak_err;   // last error of this host thread

int pfail(int code, const char* what, hipError_t e) {
  g_peak_err = std::string(what) + ": " + hipGetErrorString(e);
  return code;
}

// Best of three launches (the first warms the clocks up) of `launch` on stream s, in ms.
template <class F>
int best_of_three(hipStream_t s, F launch, float* best) {
  hipEvent_t t0, t1;
  hipError_t e = hipEventCreate(&t0);
  if (e != hipSuccess) return pfail(BOSSS_ERR_HIP, "hipEventCreate", e);
  e = hipEventCreate(&t1);
  if (e != hipSuccess) {
    (void)hipEventDestroy(t0);
    return pfail(BOSSS_ERR_HIP, "hipEventCreate", e);
  }
  *best = 1e30f;
  for (int rep = 0; rep < 3 && e == hipSuccess; rep++) {
    (void)hipEventRecord(t0, s);
    launch();
    (void)hipEventRecord(t1, s);
    e = hipEventSynchronize(t1);
    float ms = 0.0f;
    if (e == hipSuccess) (void)hipEventElapsedTime(&ms, t0, t1);
    if (ms > 0.0f && ms < *best) *best = ms;
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  return e == hipSuccess ? BOSSS_OK : pfail(BOSSS_ERR_HIP, "peak kernel", e);
}

}  // namespace

extern "C" {

const char* bosss_hip_peaks_last_error(void) { return g_peak_err.c_str(); }

int bosss_hip_measure_peaks(bosss_ctx* ctx, double* tflops_fp64_mfma, double* tflops_fp32_mfma, double* tflops_fp64_fma) {
  if (!ctx || !tflops_fp64_mfma || !tflops_fp32_mfma || !tflops_fp64_fma) {
    g_peak_err = "NULL argument";
    return BOSSS_ERR_ARG;
  }
  int dev = 0;
  hipStream_t s = nullptr;
  int rc = bosss::ctx_device_stream(ctx, &dev, &s);
  if (rc != BOSSS_OK) return rc;
  hipError_t e = hipSetDevice(dev);
  if (e != hipSuccess) return pfail(BOSSS_ERR_HIP, "hipSetDevice", e);
  double* out = nullptr;
  e = hipMalloc(&out, sizeof(double) * PK_BLOCKS * PK_THREADS);
  if (e != hipSuccess) return pfail(BOSSS_ERR_HIP, "hipMalloc", e);
  const double waves = (double)PK_BLOCKS * PK_THREADS / 64.0;
  // FLOPs: one 16x16x4 MFMA = 2 * 16 * 16 * 4; one FMA = 2 per lane
  const double mfma_flops = waves * PK_ITERS * PK_CHAINS * 2048.0;
  const double fma_flops = (double)PK_BLOCKS * PK_THREADS * PK_ITERS * 4.0 * PK_FMA_CHAINS * 2.0;
  double* res[3] = {tflops_fp64_mfma, tflops_fp32_mfma, tflops_fp64_fma};
  const double flops[3] = {mfma_flops, mfma_flops, fma_flops};
  for (int k = 0; k < 3 && rc == BOSSS_OK; k++) {
    float best = 0.0f;
    rc = best_of_three(
        s,
        [&] {
          if (k == 0) hipLaunchKernelGGL(k_peak_mfma64, dim3(PK_BLOCKS), dim3(PK_THREADS), 0, s, out, 1.0);
          if (k == 1) hipLaunchKernelGGL(k_peak_mfma32, dim3(PK_BLOCKS), dim3(PK_THREADS), 0, s, out, 1.0f);
          if (k == 2) hipLaunchKernelGGL(k_peak_fma64, dim3(PK_BLOCKS), dim3(PK_THREADS), 0, s, out, 1.0);
        },
        &best);
    if (rc == BOSSS_OK) *res[k] = flops[k] / (best * 1e-3) / 1e12;
  }
  e = hipFree(out);
  if (rc == BOSSS_OK && e != hipSuccess) return pfail(BOSSS_ERR_HIP, "hipFree", e);
  return rc;
}

int bosss_hip_measure_fp64_mix(bosss_ctx* ctx, double* tflops_fp64_mix) {
  if (!ctx || !tflops_fp64_mix) {
    g_peak_err = "NULL argument";
    return BOSSS_ERR_ARG;
  }
  int dev = 0;
  hipStream_t s = nullptr;
  int rc = bosss::ctx_device_stream(ctx, &dev, &s);
  if (rc != BOSSS_OK) return rc;
  hipError_t e = hipSetDevice(dev);
  if (e != hipSuccess) return pfail(BOSSS_ERR_HIP, "hipSetDevice", e);
  double* out = nullptr;
  e = hipMalloc(&out, sizeof(double) * PK_BLOCKS * PK_THREADS);
  if (e != hipSuccess) return pfail(BOSSS_ERR_HIP, "hipMalloc", e);
  const double waves = (double)PK_BLOCKS * PK_THREADS / 64.0;
  const double flops = waves * PK_ITERS * PK_CHAINS * (2048.0 + PK_MIX_FMA * 128.0);
  float best = 0.0f;
  rc = best_of_three(
      s, [&] { hipLaunchKernelGGL(k_peak_mix64, dim3(PK_BLOCKS), dim3(PK_THREADS), 0, s, out, 1.0); }, &best);
  if (rc == BOSSS_OK) *tflops_fp64_mix = flops / (best * 1e-3) / 1e12;
  e = hipFree(out);
  if (rc == BOSSS_OK && e != hipSuccess) return pfail(BOSSS_ERR_HIP, "hipFree", e);
  return rc;
}

}  // extern "C"
