// math_check.cpp -- checks the kernels' own atan2f (csrc/itsca_math.h) against
// the reference's libm (glibc atan2f, reached through std::atan2 on floats in
// MathUtils.h:39 and Tracklet.cxx:36).
//
// TEST INFRASTRUCTURE ONLY.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../tracking-itsu_amd/csrc/itsca_math.h"

extern "C" long itsca_check_atan2f(long n, uint64_t seed)
{
  uint64_t s = seed;
  long bad = 0;
  for (long i = 0; i < n; ++i) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const uint32_t a = (uint32_t)z, b = (uint32_t)(z >> 32);
    float x, y;
    switch (i & 3) {
      case 0:  // any bit pattern
        std::memcpy(&x, &a, 4);
        std::memcpy(&y, &b, 4);
        break;
      case 1:  // cluster coordinates (|x|, |y| < ~100 cm)
        x = (float)(int32_t)a * 5e-8f;
        y = (float)(int32_t)b * 5e-8f;
        break;
      case 2:  // cluster differences
        x = (float)(int32_t)a * 1e-11f;
        y = (float)(int32_t)b * 5e-11f;
        break;
      default:  // large ratios
        x = (float)(int32_t)a * 1e-12f;
        y = (float)(int32_t)b * 1e-5f;
        break;
    }
    const float g = ::atan2f(y, x), m = itsca::fdlibmAtan2f(y, x);
    if (std::memcmp(&g, &m, 4) != 0 && !(std::isnan(g) && std::isnan(m))) ++bad;
  }
  return bad;
}
