/*
 * pac_variant.h -- ball element type of a compilation.
 *
 * The enumeration / filter / emptiness kernels and the pipeline driver are
 * compiled once per input type from the same source:
 *   vf32  FP32 balls (float4), as the reference reads them
 *         (AlphaComplexParallel.cu:63-67); exact stage in 512-bit integers
 *         (16 cooperating lanes), values bounded by WL = 16 limbs;
 *   vf64  FP64 balls (double4), continuous coordinates; exact stage in
 *         1024-bit integers (32 lanes), values bounded by WL = 24 limbs
 *         (a coordinate's lowest set bit can lie ~2^-70 below its units,
 *         and the degree-8 tet filter needs 8 times that many bits).
 * Each compilation lives in its own namespace pac::vf32 / pac::vf64, so the
 * two sets of kernels and helpers link side by side.  Build with
 * -DPAC_BALL_F64=1 for the FP64 set.
 */
#ifndef PAC_VARIANT_H
#define PAC_VARIANT_H

#include <hip/hip_runtime.h>

#ifndef PAC_BALL_F64
#define PAC_BALL_F64 0
#endif

#if PAC_BALL_F64
#define PAC_V vf64
#else
#define PAC_V vf32
#endif

namespace pac {
namespace PAC_V {
#if PAC_BALL_F64
using Ball = double4;
using Real = double;
constexpr int WL = 24;
#else
using Ball = float4;
using Real = float;
constexpr int WL = 16;
#endif
constexpr bool kBallF64 = PAC_BALL_F64 != 0;
}  // namespace PAC_V
}  // namespace pac

#endif
