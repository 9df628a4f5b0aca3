// sg_probe.hip — VALU issue-rate probe of libsg (measurement only; no
// counterpart in the reference).
//
// The sweep kernels are bound by 32-bit VALU issue, so their roofline needs
// the device's sustained VALU rate for the instructions they are made of.
// The probe fills every CU (2048 work-groups of 256 threads: 8 waves per
// SIMD) with threads that each run 16 independent dependency chains of one
// instruction kind:
//   kind 0: v_bitop3_b32 (3-input boolean, the Metropolis adder tree),
//   kind 1: v_mad_u64_u32 (the Philox multiply).
// Each chain step is one instruction of that kind, so a launch issues
// threads x iters x 16 lane-operations of it; the host divides by the launch's
// HIP-event duration.  The chains' final values are stored, so none of the
// work is dead.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/sg.h"

namespace {

constexpr int kProbeThreads = 256;
constexpr int kProbeBlocks = 2048;
constexpr int kChains = 16;

__global__ __launch_bounds__(kProbeThreads) void k_probe_bitop3(uint32_t *out, uint32_t seed,
                                                                int iters) {
  const uint32_t t = blockIdx.x * kProbeThreads + threadIdx.x;
  uint32_t a[kChains], b[kChains];
#pragma unroll
  for (int i = 0; i < kChains; ++i) {
    a[i] = seed * (t + 1) + 0x9E3779B9u * (i + 1);
    b[i] = (seed ^ t) * 0x85EBCA6Bu + 0xC2B2AE35u * (i + 1);
  }
  const uint32_t k = seed ^ 0x27D4EB2Fu;
  for (int n = 0; n < iters; n += 2) {
    // each chain depends only on itself: 16 independent instructions per
    // step, a 3-input XOR (0x96) then a majority (0xE8), so that no two steps
    // fold into one
#pragma unroll
    for (int i = 0; i < kChains; ++i) a[i] = __builtin_amdgcn_bitop3_b32(a[i], b[i], k, 0x96);
#pragma unroll
    for (int i = 0; i < kChains; ++i) a[i] = __builtin_amdgcn_bitop3_b32(a[i], b[i], k, 0xE8);
  }
  uint32_t r = 0;
#pragma unroll
  for (int i = 0; i < kChains; ++i) r += a[i];
  out[t] = r;
}

__global__ __launch_bounds__(kProbeThreads) void k_probe_mad64(uint32_t *out, uint32_t seed,
                                                               int iters) {
  const uint32_t t = blockIdx.x * kProbeThreads + threadIdx.x;
  uint64_t a[kChains];
#pragma unroll
  for (int i = 0; i < kChains; ++i) a[i] = (uint64_t)(seed * (t + 1)) + 0x9E3779B9u * (i + 1);
  for (int n = 0; n < iters; ++n) {
    // a = lo32(a) * M + a: one v_mad_u64_u32 per chain and step
#pragma unroll
    for (int i = 0; i < kChains; ++i) a[i] = (uint64_t)(uint32_t)a[i] * 0xD2511F53u + a[i];
  }
  uint32_t r = 0;
#pragma unroll
  for (int i = 0; i < kChains; ++i) r += (uint32_t)a[i] ^ (uint32_t)(a[i] >> 32);
  out[t] = r;
}

thread_local std::string g_probe_error;

}  // namespace

extern "C" {

sg_status sg_probe_valu(int32_t device, int32_t kind, int32_t iters, double *lane_ops,
                        double *ms) {
  g_probe_error.clear();
  if (!lane_ops || !ms || kind < 0 || kind > 1 || iters < 2 || (iters & 1)) {
    g_probe_error = "bad probe arguments (kind 0 or 1, even iters >= 2)";
    return SG_ERR_INVALID;
  }
  auto fail = [](const char *what, hipError_t e) {
    g_probe_error = std::string(what) + ": " + hipGetErrorString(e);
    return SG_ERR_DEVICE;
  };
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return fail("hipSetDevice", e);
  uint32_t *out = nullptr;
  e = hipMalloc((void **)&out, sizeof(uint32_t) * kProbeThreads * kProbeBlocks);
  if (e != hipSuccess) {
    g_probe_error = std::string("hipMalloc: ") + hipGetErrorString(e);
    return SG_ERR_NOMEM;
  }
  hipEvent_t a = nullptr, b = nullptr;
  hipStream_t st = nullptr;
  sg_status rc = SG_OK;
  float t = 0;
  do {
    if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) {
      rc = fail("hipStreamCreate", e);
      break;
    }
    if ((e = hipEventCreate(&a)) != hipSuccess || (e = hipEventCreate(&b)) != hipSuccess) {
      rc = fail("hipEventCreate", e);
      break;
    }
    auto launch = [&](int n) {
      if (kind == 0)
        hipLaunchKernelGGL(k_probe_bitop3, dim3(kProbeBlocks), dim3(kProbeThreads), 0, st, out,
                           0x5EEDu, n);
      else
        hipLaunchKernelGGL(k_probe_mad64, dim3(kProbeBlocks), dim3(kProbeThreads), 0, st, out,
                           0x5EEDu, n);
      return hipGetLastError();
    };
    if ((e = launch(iters)) != hipSuccess) {   // warm-up: code load, clocks
      rc = fail("probe launch", e);
      break;
    }
    if ((e = hipEventRecord(a, st)) != hipSuccess || (e = launch(iters)) != hipSuccess ||
        (e = hipEventRecord(b, st)) != hipSuccess || (e = hipEventSynchronize(b)) != hipSuccess ||
        (e = hipEventElapsedTime(&t, a, b)) != hipSuccess) {
      rc = fail("probe timing", e);
      break;
    }
  } while (false);
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  if (st) (void)hipStreamDestroy(st);
  (void)hipFree(out);
  if (rc != SG_OK) return rc;
  *lane_ops = (double)kProbeThreads * kProbeBlocks * (double)iters * kChains;
  *ms = t;
  return SG_OK;
}

const char *sg_probe_last_error(void) { return g_probe_error.c_str(); }

}  // extern "C"
