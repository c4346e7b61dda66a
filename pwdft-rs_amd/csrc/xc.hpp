// PZ-LDA exchange-correlation of one grid point (src/potential/xc.rs:129-141, 223-288), shared by
// the pointwise grid kernels and the SCF's mixing step (which evaluates it on the mixed density).
#pragma once

#include <hip/hip_runtime.h>

namespace pwk {
namespace xc {

// src/consts.rs:15-49 (RHO_FLOOR src/functionals/mod.rs:5)
constexpr double kPiD = 3.141592653589793;
constexpr double kHaToEv = 27.211386245988;
constexpr double kBohr = 0.529177210903;
constexpr double kBohr3 = kBohr * kBohr * kBohr;
constexpr double kRhoFloor = 1e-30;

__device__ __forceinline__ void lda_point(double rho, double& exc, double& vxc) {
  if (rho < kRhoFloor) {
    exc = 0.0;
    vxc = 0.0;
    return;
  }
  double rb = rho * kBohr3;
  double ex_ha = -0.75 * cbrt(3.0 * rb / kPiD);
  double vx_ha = (4.0 / 3.0) * ex_ha;
  double rs = cbrt(3.0 / (4.0 * kPiD * rb));
  double ec, vc;
  if (rs >= 1.0) {
    const double gamma = -0.1423, b1 = 1.0529, b2 = 0.3334;
    double srs = sqrt(rs);
    double den = 1.0 + b1 * srs + b2 * rs;
    ec = gamma / den;
    double dec = -gamma * (b1 / (2.0 * srs) + b2) / (den * den);
    vc = fma(rs / 3.0, -dec, ec);
  } else {
    const double a = 0.0311, b = -0.048, c = 0.0020, d = -0.0116;
    double lr = log(rs);
    ec = fma(d, rs, fma(c * rs, lr, fma(a, lr, b)));
    double dec = fma(c, lr + 1.0, a / rs) + d;
    vc = fma(rs / 3.0, -dec, ec);
  }
  exc = ex_ha * kHaToEv + ec * kHaToEv;
  vxc = vx_ha * kHaToEv + vc * kHaToEv;
}

__device__ __forceinline__ void lda_point(float rho, float& exc, float& vxc) {
  if (rho < 1e-30f) {
    exc = 0.0f;
    vxc = 0.0f;
    return;
  }
  const float pi = 3.14159265358979f;
  float rb = rho * (float)kBohr3;
  float ex_ha = -0.75f * cbrtf(3.0f * rb / pi);
  float vx_ha = (4.0f / 3.0f) * ex_ha;
  float rs = cbrtf(3.0f / (4.0f * pi * rb));
  float ec, vc;
  if (rs >= 1.0f) {
    const float gamma = -0.1423f, b1 = 1.0529f, b2 = 0.3334f;
    float srs = sqrtf(rs);
    float den = 1.0f + b1 * srs + b2 * rs;
    ec = gamma / den;
    float dec = -gamma * (b1 / (2.0f * srs) + b2) / (den * den);
    vc = ec - rs / 3.0f * dec;
  } else {
    const float a = 0.0311f, b = -0.048f, c = 0.0020f, d = -0.0116f;
    float lr = logf(rs);
    ec = a * lr + b + c * rs * lr + d * rs;
    float dec = a / rs + c * (lr + 1.0f) + d;
    vc = ec - rs / 3.0f * dec;
  }
  exc = (ex_ha + ec) * (float)kHaToEv;
  vxc = (vx_ha + vc) * (float)kHaToEv;
}

}  // namespace xc
}  // namespace pwk
