// Exhaustive check of vh::recip_rn (the reciprocal of the integration kernels) against the
// IEEE quotient 1.0f / z for every positive normal z below 2^126 (2,113,929,216 values).
// Prints the number of mismatches; exit status 0 iff there are none.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "vh_internal.h"

__global__ void k_check(uint32_t base, uint32_t end, unsigned long long* bad, uint32_t* first) {
  const uint32_t u = base + blockIdx.x * 256u + threadIdx.x;
  if (u >= end) return;
  const float z = __builtin_bit_cast(float, u);
  const float want = 1.0f / z;
  const float got = vh::recip_rn(z);
  if (__builtin_bit_cast(uint32_t, want) != __builtin_bit_cast(uint32_t, got)) {
    atomicAdd(bad, 1ull);
    atomicMin(first, u);
  }
}

int main() {
  const uint32_t lo = 0x00800000u, hi = 0x7e800000u;  // [2^-126, 2^126)
  unsigned long long* bad = nullptr;
  uint32_t* first = nullptr;
  if (hipMalloc(&bad, 8) != hipSuccess || hipMalloc(&first, 4) != hipSuccess) return 2;
  if (hipMemset(bad, 0, 8) != hipSuccess || hipMemset(first, 0xff, 4) != hipSuccess) return 2;
  const uint32_t chunk = 1u << 28;
  for (uint64_t b = lo; b < hi; b += chunk) {
    const uint32_t n = (uint32_t)((hi - b) < chunk ? (hi - b) : chunk);
    hipLaunchKernelGGL(k_check, dim3((n + 255) / 256), dim3(256), 0, 0, (uint32_t)b, (uint32_t)(b + n), bad, first);
  }
  if (hipDeviceSynchronize() != hipSuccess) return 2;
  unsigned long long nbad = 0;
  uint32_t f = 0;
  if (hipMemcpy(&nbad, bad, 8, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(&f, first, 4, hipMemcpyDeviceToHost) != hipSuccess)
    return 2;
  std::printf("recip_rn vs 1.0f/z: %u values checked, %llu mismatches", hi - lo, nbad);
  if (nbad) std::printf(" (first z bits 0x%08x)", f);
  std::printf("\n");
  return nbad ? 1 : 0;
}
