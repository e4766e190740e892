// Host check of the F16R scale from an FP32 value (csrc/vh_math.h): for every finite binary16
// h, f16r_scale_of / f16r_inv_scale_of of its FP32 value equal the bit forms
// f16r_scale_from_bits / f16r_inv_scale_from_bits, and their product is 1. Prints
// "ok <values>" or the first counterexample.
#include <cinttypes>
#include <cstdio>
#include <cstring>

#include "vh_math.h"

using namespace vh;

static uint32_t bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

int main() {
  int64_t n = 0;
  for (uint32_t h = 0; h < 65536; ++h) {
    if (((h >> 10) & 31u) == 31u) continue;  // inf / NaN: never stored
    const float hf = f16_bits_to_f32((uint16_t)h);
    const float a = f16r_scale_of(hf), b = f16r_scale_from_bits((uint16_t)h);
    const float c = f16r_inv_scale_of(hf), d = f16r_inv_scale_from_bits((uint16_t)h);
    if (bits(a) != bits(b) || bits(c) != bits(d) || a * c != 1.0f) {
      std::printf("mismatch h=0x%04x scale %a vs %a inverse %a vs %a\n", h, a, b, c, d);
      return 1;
    }
    ++n;
  }
  std::printf("ok %" PRId64 "\n", n);
  return 0;
}
