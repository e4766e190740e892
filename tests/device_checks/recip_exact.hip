// Exhaustive checks of the integration kernels' division helpers against IEEE division:
//   vh::recip_rn(z) == 1.0f / z for every positive normal z below 2^126 (2,113,929,216 values);
//   vh::div_rn(a, b) == a / b for every finite a with |a| >= 2^-76 or a == +0 and every
//   b = 1..256 (the weights of 16-bit storage), and for 2^20 sampled a against every b of a
//   sampled set up to 2^24 (FP32 weights).
// Prints the mismatch counts; exit status 0 iff there are none.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "vh_internal.h"

__global__ void k_recip(uint32_t base, uint32_t end, unsigned long long* bad) {
  const uint32_t u = base + blockIdx.x * 256u + threadIdx.x;
  if (u >= end) return;
  const float z = __builtin_bit_cast(float, u);
  if (__builtin_bit_cast(uint32_t, 1.0f / z) != __builtin_bit_cast(uint32_t, vh::recip_rn(z))) atomicAdd(bad, 1ull);
}

// a: every bit pattern in [base, base + n) that is finite and in the fast domain; b = 1..bmax
__global__ void k_div(uint32_t base, uint32_t n, int bmax, unsigned long long* bad) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const uint32_t u = base + i;
  if ((u & 0x7f800000u) == 0x7f800000u) return;  // inf / nan
  const float a = __builtin_bit_cast(float, u);
  if (!(__builtin_fabsf(a) >= 0x1p-76f || u == 0u)) return;
  unsigned cnt = 0;
  for (int ib = 1; ib <= bmax; ++ib) {
    const float b = (float)ib;
    cnt += __builtin_bit_cast(uint32_t, a / b) != __builtin_bit_cast(uint32_t, vh::div_rn(a, b));
  }
  if (cnt) atomicAdd(bad, (unsigned long long)cnt);
}

// sampled a (hash of the index) against b = 1 + 65 k, k = 0..258111 (up to 2^24)
__global__ void k_div_wide(uint32_t n, unsigned long long* bad) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  uint32_t h = i * 0x9e3779b9u;
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  const uint32_t u = (h & 0x807fffffu) | ((96u + (h >> 23) % 64u) << 23);  // |a| in [2^-31, 2^33)
  const float a = __builtin_bit_cast(float, u);
  unsigned cnt = 0;
  for (uint32_t k = 0; k < 258112u; k += 1u + (i & 7u)) {
    const float b = (float)(1u + 65u * k);
    cnt += __builtin_bit_cast(uint32_t, a / b) != __builtin_bit_cast(uint32_t, vh::div_rn(a, b));
  }
  if (cnt) atomicAdd(bad, (unsigned long long)cnt);
}

int main() {
  unsigned long long* bad = nullptr;
  if (hipMalloc(&bad, 3 * 8) != hipSuccess || hipMemset(bad, 0, 3 * 8) != hipSuccess) return 2;
  const uint32_t lo = 0x00800000u, hi = 0x7e800000u;  // [2^-126, 2^126)
  const uint32_t chunk = 1u << 28;
  for (uint64_t b = lo; b < hi; b += chunk) {
    const uint32_t n = (uint32_t)((hi - b) < chunk ? (hi - b) : chunk);
    hipLaunchKernelGGL(k_recip, dim3((n + 255) / 256), dim3(256), 0, 0, (uint32_t)b, (uint32_t)(b + n), bad);
  }
  for (uint64_t b = 0; b < (1ull << 32); b += chunk)
    hipLaunchKernelGGL(k_div, dim3(chunk / 256), dim3(256), 0, 0, (uint32_t)b, chunk, 256, bad + 1);
  hipLaunchKernelGGL(k_div_wide, dim3((1u << 20) / 256), dim3(256), 0, 0, 1u << 20, bad + 2);
  if (hipDeviceSynchronize() != hipSuccess) return 2;
  unsigned long long h[3] = {0, 0, 0};
  if (hipMemcpy(h, bad, sizeof h, hipMemcpyDeviceToHost) != hipSuccess) return 2;
  std::printf("recip_rn vs 1.0f/z: %u values checked, %llu mismatches\n", hi - lo, h[0]);
  std::printf("div_rn vs a/b, all finite a in the fast domain, b = 1..256: %llu mismatches\n", h[1]);
  std::printf("div_rn vs a/b, 2^20 sampled a, b sampled up to 2^24: %llu mismatches\n", h[2]);
  return (h[0] || h[1] || h[2]) ? 1 : 0;
}
