// sin and cos of an FP64 argument of bounded size, for the Helmholtz kernel evaluations.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

// Largest |x| for sincos_red: n = rint(2x/pi) stays below 2^20, where the three-part
// reduction below is accurate to rounding.  The host checks |k| times the largest
// distance of a computation against it.
constexpr double SINCOS_RED_MAX = 1.6e6;

// sin x and cos x for |x| <= SINCOS_RED_MAX.
// x = n pi/2 + y with n = rint(2x/pi): y from three fused multiply-adds with pi/2 split
// into three doubles p1 + p2 + p3 (Cody and Waite's additive reduction; the split is
// computed from the decimal digits of pi and equals pi/2 to 5.6e-50).  sin y and cos y
// come from their Taylor series through y^15 and y^16 (coefficients +-1/k!, rounded to
// nearest): on |y| <= pi/4 the first omitted terms are below 5e-17 and 3e-18.  The
// quadrant n mod 4 then maps (sin y, cos y) to (sin x, cos x).
__host__ __device__ __forceinline__ void sincos_red(double x, double& s, double& c) {
  const double n = rint(x * 0x1.45f306dc9c883p-1);  // 2/pi
  double y = fma(-n, 0x1.921fb54442d18p+0, x);        // p1
  y = fma(-n, 0x1.1a62633145c07p-54, y);              // p2
  y = fma(-n, -0x1.f1976b7ed8fbcp-110, y);            // p3
  const double z = y * y;
  double ps = -0x1.ae7f3e733b81fp-41;  // -1/15!
  ps = fma(ps, z, 0x1.6124613a86d09p-33);   // 1/13!
  ps = fma(ps, z, -0x1.ae64567f544e4p-26);  // -1/11!
  ps = fma(ps, z, 0x1.71de3a556c734p-19);   // 1/9!
  ps = fma(ps, z, -0x1.a01a01a01a01ap-13);  // -1/7!
  ps = fma(ps, z, 0x1.1111111111111p-7);    // 1/5!
  ps = fma(ps, z, -0x1.5555555555555p-3);   // -1/3!
  const double sn = fma(ps * z, y, y);
  double pc = 0x1.ae7f3e733b81fp-45;        // 1/16!
  pc = fma(pc, z, -0x1.93974a8c07c9dp-37);  // -1/14!
  pc = fma(pc, z, 0x1.1eed8eff8d898p-29);   // 1/12!
  pc = fma(pc, z, -0x1.27e4fb7789f5cp-22);  // -1/10!
  pc = fma(pc, z, 0x1.a01a01a01a01ap-16);   // 1/8!
  pc = fma(pc, z, -0x1.6c16c16c16c17p-10);  // -1/6!
  pc = fma(pc, z, 0x1.5555555555555p-5);    // 1/4!
  pc = fma(pc, z, -0.5);                    // -1/2!
  const double cn = fma(pc, z, 1.0);
  const int q = (int)n & 3;  // n mod 4, also for negative n
  const double sq = (q & 1) ? cn : sn, cq = (q & 1) ? sn : cn;
  s = (q & 2) ? -sq : sq;
  c = ((q + 1) & 2) ? -cq : cq;
}
