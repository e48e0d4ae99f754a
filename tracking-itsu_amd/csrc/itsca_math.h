// itsca_math.h -- cut constants, binning and the project's own atan2f.
//
// Shared bit-for-bit between the gfx950 kernels (hipcc, -ffp-contract=off,
// correctly rounded fp32 divide/sqrt, denormals kept) and host C++ (g++ on
// x86-64 without -march, i.e. SSE2 and no FMA).  Every expression below keeps
// the reference's evaluation order so that each float result is identical.
//
// Reference: include/ITSReconstruction/CA/Constants.h:29-97 (constants),
//            include/ITSReconstruction/CA/MathUtils.h:37-60 (phi, r, cross),
//            include/ITSReconstruction/CA/IndexTableUtils.h:47-62 (binning),
//            src/CA/TrackingUtils.cxx:30-48 (search window).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ITSCA_HD __host__ __device__ __forceinline__
#define ITSCA_CONST __constant__
#else
#define ITSCA_HD inline
#define ITSCA_CONST
#endif

namespace itsca {

constexpr int kLayers = 7;          // Constants.h:38
constexpr int kTrackletLayers = 6;  // Constants.h:39
constexpr int kCellLayers = 5;      // Constants.h:40
constexpr int kUnused = -1;         // Constants.h:41
constexpr int kZBins = 20;          // Constants.h:77
constexpr int kPhiBins = 20;        // Constants.h:78
constexpr int kBins = kZBins * kPhiBins;  // 400 real bins; slot 400 holds phi == 2pi
constexpr int kTableSize = kBins + 1;     // 401 entries per index table
constexpr int kCellsMinLevel = 5;         // Constants.h:73

constexpr float kPi = 3.14159265359f;     // Constants.h:31
constexpr float kTwoPi = 2.0f * kPi;      // Constants.h:32
constexpr float kFloatMin = 1e-20f;       // Constants.h:33
constexpr float kZCut = 0.5f;             // Constants.h:68
constexpr float kPhiCut = 0.3f;           // Constants.h:69
constexpr float kCellMaxDeltaTanLambda = 0.025f;  // Constants.h:58
constexpr float kCellMaxDeltaPhi = 0.14f;         // Constants.h:67
constexpr float kInversePhiBinSize = kPhiBins / kTwoPi;  // Constants.h:79 (int / float -> float)

// Per-layer tables.  Accessors take the index so that device code reads them
// from registers after unrolling rather than from memory.
ITSCA_HD float layerHalfZ(int l)  // Constants.h:43-46
{
  return l < 3 ? 16.333f : (l < 5 ? 42.140f : 73.745f);
}
ITSCA_HD float layerR(int l)  // Constants.h:47-50
{
  switch (l) {
    case 0: return 2.33959f;
    case 1: return 3.14076f;
    case 2: return 3.91924f;
    case 3: return 19.6213f;
    case 4: return 24.5597f;
    case 5: return 34.388f;
    default: return 39.3329f;
  }
}
// Constants.h:80-84: 0.5 * ZBins / halfZ evaluated in double, stored as float.
ITSCA_HD float inverseZBinSize(int l)
{
  return l < 3 ? (float)(0.5 * kZBins / 16.333f)
               : (l < 5 ? (float)(0.5 * kZBins / 42.140f) : (float)(0.5 * kZBins / 73.745f));
}
ITSCA_HD float trackletMaxDeltaZ(int l)  // Constants.h:54-57
{
  return l < 2 ? 0.1f : 0.3f;
}
ITSCA_HD float cellMaxDeltaZ(int l)  // Constants.h:59-62
{
  switch (l) {
    case 0: return 0.2f;
    case 1: return 0.4f;
    case 2: return 0.5f;
    case 3: return 0.6f;
    default: return 3.0f;
  }
}
ITSCA_HD float cellMaxDCA(int l)  // Constants.h:63-66
{
  switch (l) {
    case 0: return 0.05f;
    case 1: return 0.04f;
    case 2: return 0.05f;
    case 3: return 0.2f;
    default: return 0.4f;
  }
}
ITSCA_HD float neighbourMaxNormalDelta(int l)  // Constants.h:70-71
{
  switch (l) {
    case 0: return 0.002f;
    case 1: return 0.009f;
    case 2: return 0.002f;
    default: return 0.005f;
  }
}
ITSCA_HD float neighbourMaxCurvatureDelta(int l)  // Constants.h:72
{
  switch (l) {
    case 0: return 0.008f;
    case 1: return 0.0025f;
    case 2: return 0.003f;
    default: return 0.0035f;
  }
}

ITSCA_HD int32_t floatBits(float f) { return __builtin_bit_cast(int32_t, f); }
ITSCA_HD float bitsFloat(int32_t i) { return __builtin_bit_cast(float, i); }
ITSCA_HD float fabsBits(float f) { return bitsFloat(floatBits(f) & 0x7fffffff); }

// ---------------------------------------------------------------------------
// atan(x) and atan2(y, x) in single precision, following the published fdlibm
// 5.3 algorithm (Sun Microsystems, 1993; float version by Cygnus Support):
// reduce |x| into one of four intervals around 0, 0.5, 1, 1.5 and infinity
// using atan(a) + atan((x-a)/(1+a*x)), then evaluate an 11-term odd minimax
// polynomial split into even and odd halves.  The coefficients are the
// published fdlibm float ones.  This is the algorithm behind glibc 2.35's
// atan2f on x86-64, which the reference reaches through std::atan2(float,
// float) (MathUtils.h:39, Tracklet.cxx:36); evaluated with IEEE single
// operations and no contraction it reproduces that function bit for bit.
// ---------------------------------------------------------------------------
ITSCA_HD float fdlibmAtanf(float x)
{
  const float atanhi0 = bitsFloat(0x3eed6338), atanhi1 = bitsFloat(0x3f490fda),
              atanhi2 = bitsFloat(0x3f7b985e), atanhi3 = bitsFloat(0x3fc90fda);
  const float atanlo0 = bitsFloat(0x31ac3769), atanlo1 = bitsFloat(0x33222168),
              atanlo2 = bitsFloat(0x33140fb4), atanlo3 = bitsFloat(0x33a22168);
  const float a0 = bitsFloat(0x3eaaaaab), a1 = bitsFloat(0xbe4ccccd), a2 = bitsFloat(0x3e124925),
              a3 = bitsFloat(0xbde38e38), a4 = bitsFloat(0x3dba2e6e), a5 = bitsFloat(0xbd9d8795),
              a6 = bitsFloat(0x3d886b35), a7 = bitsFloat(0xbd6ef16b), a8 = bitsFloat(0x3d4bda59),
              a9 = bitsFloat(0xbd15a221), a10 = bitsFloat(0x3c8569d7);
  const int32_t hx = floatBits(x);
  const int32_t ix = hx & 0x7fffffff;
  int id;
  if (ix >= 0x4c000000) {  // |x| >= 2^25
    if (ix > 0x7f800000) return x + x;  // NaN
    return hx > 0 ? atanhi3 + atanlo3 : -atanhi3 - atanlo3;
  }
  if (ix < 0x3ee00000) {  // |x| < 0.4375
    if (ix < 0x31000000) return x;  // |x| < 2^-29: atan(x) rounds to x
    id = -1;
  } else {
    x = fabsBits(x);
    if (ix < 0x3f980000) {    // |x| < 1.1875
      if (ix < 0x3f300000) {  // 7/16 <= |x| < 11/16
        id = 0;
        x = (2.0f * x - 1.0f) / (2.0f + x);
      } else {  // 11/16 <= |x| < 19/16
        id = 1;
        x = (x - 1.0f) / (x + 1.0f);
      }
    } else {
      if (ix < 0x401c0000) {  // |x| < 2.4375
        id = 2;
        x = (x - 1.5f) / (1.0f + 1.5f * x);
      } else {  // 2.4375 <= |x| < 2^25
        id = 3;
        x = -1.0f / x;
      }
    }
  }
  const float z = x * x;
  const float w = z * z;
  const float s1 = z * (a0 + w * (a2 + w * (a4 + w * (a6 + w * (a8 + w * a10)))));
  const float s2 = w * (a1 + w * (a3 + w * (a5 + w * (a7 + w * a9))));
  if (id < 0) return x - x * (s1 + s2);
  const float hi = id == 0 ? atanhi0 : id == 1 ? atanhi1 : id == 2 ? atanhi2 : atanhi3;
  const float lo = id == 0 ? atanlo0 : id == 1 ? atanlo1 : id == 2 ? atanlo2 : atanlo3;
  const float r = hi - ((x * (s1 + s2) - lo) - x);
  return hx < 0 ? -r : r;
}

ITSCA_HD float fdlibmAtan2f(float y, float x)
{
  const float tiny = 1.0e-30f;
  const float pi_o_4 = bitsFloat(0x3f490fdb), pi_o_2 = bitsFloat(0x3fc90fdb),
              pi = bitsFloat(0x40490fdb), pi_lo = bitsFloat(0xb3bbbd2e);
  const int32_t hx = floatBits(x), hy = floatBits(y);
  const int32_t ix = hx & 0x7fffffff, iy = hy & 0x7fffffff;
  if (ix > 0x7f800000 || iy > 0x7f800000) return x + y;  // NaN
  if (hx == 0x3f800000) return fdlibmAtanf(y);          // x == 1
  const int m = ((hy >> 31) & 1) | ((hx >> 30) & 2);     // 2*sign(x) + sign(y)
  if (iy == 0) {
    if (m < 2) return y;
    return m == 2 ? pi + tiny : -pi - tiny;
  }
  if (ix == 0) return hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
  if (ix == 0x7f800000) {
    if (iy == 0x7f800000) {
      switch (m) {
        case 0: return pi_o_4 + tiny;
        case 1: return -pi_o_4 - tiny;
        case 2: return 3.0f * pi_o_4 + tiny;
        default: return -3.0f * pi_o_4 - tiny;
      }
    }
    switch (m) {
      case 0: return 0.0f;
      case 1: return -0.0f;
      case 2: return pi + tiny;
      default: return -pi - tiny;
    }
  }
  if (iy == 0x7f800000) return hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
  const int32_t k = (iy - ix) >> 23;
  float z;
  if (k > 60) {
    z = pi_o_2 + 0.5f * pi_lo;
  } else if (hx < 0 && k < -60) {
    z = 0.0f;
  } else {
    z = fdlibmAtanf(fabsBits(y / x));
  }
  switch (m) {
    case 0: return z;
    case 1: return bitsFloat(floatBits(z) ^ (int32_t)0x80000000);
    case 2: return pi - (z - pi_lo);
    default: return (z - pi_lo) - pi;
  }
}

// MathUtils.h:47-51
ITSCA_HD float normalizedPhi(float phi)
{
  return (phi < 0) ? phi + kTwoPi : (phi > kTwoPi) ? phi - kTwoPi : phi;
}
// MathUtils.h:37-40 composed with :47-51 (Cluster.cxx:37-39)
ITSCA_HD float clusterPhi(float dx, float dy)
{
  return normalizedPhi(fdlibmAtan2f(-dy, -dx) + kPi);
}
// The reference's float -> int conversions (IndexTableUtils.h:49-55) as its
// x86-64 build performs them (cvttss2si): truncation toward zero, and
// INT32_MIN for NaN and for values outside the int range.  The gfx950
// conversion saturates instead (and gives 0 for NaN), so out-of-range inputs
// are mapped explicitly; in range both are the same truncation.
ITSCA_HD int truncToIntX86(float v)
{
  return (v >= -2147483648.0f && v < 2147483648.0f) ? (int)v : (int)0x80000000;
}
// IndexTableUtils.h:47-51 (float -> int truncation toward zero).  A z below
// the layer gives a negative index (a negative bin in phi row 0).
ITSCA_HD int zBinIndex(int layer, float z)
{
  return truncToIntX86((z + layerHalfZ(layer)) * inverseZBinSize(layer));
}
// z sub-bin of the tracklet search index (no counterpart in the reference):
// z bin zb of layer l split into S parts.  For a fixed zb it is monotone in z
// (truncation of a monotone float expression, then a clamp into [0, S)), which
// is all the search relies on.
ITSCA_HD int zSubIndex(int layer, float z, int zb, int S)
{
  const int f = (int)((z + layerHalfZ(layer)) * (inverseZBinSize(layer) * (float)S));
  const int sub = f - zb * S;
  return sub < 0 ? 0 : (sub >= S ? S - 1 : sub);
}
// IndexTableUtils.h:53-56
ITSCA_HD int phiBinIndex(float phi) { return truncToIntX86(phi * kInversePhiBinSize); }
// phi sub-row of the tracklet search index (no counterpart in the reference):
// phi row `row` split into P parts.  For a fixed row it is monotone in phi
// (truncation of a monotone float expression, then a clamp into [0, P)); phi
// stays within (-8, 14) wherever it is called.
ITSCA_HD int phiSubIndex(float phi, int row, int P)
{
  const int f = (int)(phi * (kInversePhiBinSize * (float)P));
  const int sub = f - row * P;
  return sub < 0 ? 0 : (sub >= P ? P - 1 : sub);
}
// IndexTableUtils.h:58-62 (row stride PhiBins, clamped to the 401st slot; a
// negative z index in phi row 0 gives a negative bin, which the reference's
// sort puts before bin 0 and its index table counts inside bin 0's range).
// The sum wraps like the x86 integer add it stands for.
ITSCA_HD int binIndex(int zIndex, int phiIndex)
{
  const int b = (int)((uint32_t)phiIndex * (uint32_t)kPhiBins + (uint32_t)zIndex);
  return b < kBins ? b : kBins;
}

}  // namespace itsca
