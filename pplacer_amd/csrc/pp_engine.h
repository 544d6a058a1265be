// Scalar operations of the stage-2 engine in its two forms.  The FP64 form
// rounds every product and sum separately (the library is built with
// -ffp-contract=off), as the C stubs do on x86-64; where it sums the site
// logs in site order (logs) they are correctly rounded (pp_crmath.h), as the
// stubs' libm returns them.  The FP32 form uses fused multiply-adds and the
// hardware log.
#pragma once

#include "pp_crmath.h"
#include "pp_device.h"

namespace pp {

template <typename Real> struct RealOps;
template <> struct RealOps<double> {
  static __device__ __forceinline__ double madd(double acc, double a, double b) { return acc + a * b; }
  static __device__ __forceinline__ double logd(double u) { return log(u); }
  static __device__ __forceinline__ double logs(double u) { return cr::log(u); }
};
template <> struct RealOps<float> {
  static __device__ __forceinline__ float madd(float acc, float a, float b) { return fmaf(a, b, acc); }
  static __device__ __forceinline__ double logd(float u) { return (double)__logf(u); }
  static __device__ __forceinline__ double logs(float u) { return logd(u); }
};

}  // namespace pp
