// Device/host math shared by every cell model of the fibre solver.
//
// The helpers follow the code the reference's CellML generator emits for its
// FastMonodomainSolver kernels (core/src/cellml/source_code_generator/
// 03_generator_vc.cpp:304-617), written here as templates over the arithmetic
// type so the same model source serves the FP64 and the FP32-state paths:
//   - exponential: (1 + x/1024)^1024 by ten squarings (03_generator_vc.cpp:385-407)
//   - logarithmic: Taylor about 1, 3 or 9 + one Newton step (03_generator_vc.cpp:431-490)
//   - powN / powReciprocalN: recursive products (03_generator_vc.cpp:541-611)
// Libraries are compiled with -ffp-contract=off for the bit-exact variants, so
// each + - * / below rounds exactly as the reference's g++ -std=c++17 build does.
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FMS_HD __host__ __device__ __forceinline__
#else
// host-only compilation of the model headers (CPU checks of generated models)
#define FMS_HD inline
#endif

namespace fms {

// exp/log policy: the reference's default approximation (approximateExponentialFunction=True)
// IEEE double division without the operand scaling of the general sequence.
// The device's correctly rounded f64 division is v_div_scale (x2), v_rcp_f64,
// four FMAs refining the reciprocal, a multiply, the remainder FMA,
// v_div_fmas and v_div_fixup; v_div_scale leaves its operand unchanged and
// v_div_fmas is a plain FMA unless the operands are denormal, zero or
// non-finite, the numerator is below 2^-969 or the exponents differ by 768 or
// more, and v_div_fixup only handles those cases. For every other operand
// pair the sequence below performs the same operations on the same values
// and returns the same, correctly rounded, quotient with three instructions
// fewer; with a constant divisor the reciprocal refinement folds and the
// quotient costs a multiply and two FMAs. Model states and rates are far from
// those ranges (the model tests compare the rates and Heun steps with the
// host's IEEE division bit for bit).
FMS_HD double div_normal(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double nb = -b;
  const double r0 = __builtin_amdgcn_rcp(b);
  const double r1 = __builtin_fma(r0, __builtin_fma(nb, r0, 1.0), r0);
  const double r2 = __builtin_fma(r1, __builtin_fma(nb, r1, 1.0), r1);
  const double q = a * r2;
  return __builtin_fma(__builtin_fma(nb, q, a), r2, q);
#else
  return a / b;
#endif
}

// FP64 arithmetic type of the model right-hand sides: + - * as double,
// / through div_normal
struct FD {
  double v;
  FD() = default;
  FMS_HD constexpr FD(double x) : v(x) {}
  FMS_HD FD &operator+=(FD o) { v = v + o.v; return *this; }
  FMS_HD FD &operator-=(FD o) { v = v - o.v; return *this; }
  FMS_HD FD &operator*=(FD o) { v = v * o.v; return *this; }
};
FMS_HD FD operator+(FD a, FD b) { return FD(a.v + b.v); }
FMS_HD FD operator-(FD a, FD b) { return FD(a.v - b.v); }
FMS_HD FD operator*(FD a, FD b) { return FD(a.v * b.v); }
FMS_HD FD operator/(FD a, FD b) { return FD(div_normal(a.v, b.v)); }
FMS_HD FD operator-(FD a) { return FD(-a.v); }
FMS_HD bool operator<(FD a, FD b) { return a.v < b.v; }
FMS_HD bool operator>(FD a, FD b) { return a.v > b.v; }
FMS_HD bool operator<=(FD a, FD b) { return a.v <= b.v; }
FMS_HD bool operator>=(FD a, FD b) { return a.v >= b.v; }
FMS_HD bool operator==(FD a, FD b) { return a.v == b.v; }
FMS_HD bool operator!=(FD a, FD b) { return a.v != b.v; }

struct ExpApprox {
  template <class T> static FMS_HD T exp(T x) {
    x = T(1.0) + x * T(1. / 1024);   // x/1024 exactly (power of two)
#pragma unroll
    for (int i = 0; i < 10; i++) x *= x;
    return x;
  }
  template <class T> static FMS_HD T log(T x) {
    T t1 = x - T(1);
    T t1_2 = t1 * t1;
    T t1_4 = t1_2 * t1_2;
    T t1_6 = t1_4 * t1_2;
    T result1 = t1 - T(0.5) * t1_2 + T(1. / 3) * t1_2 * t1 - T(0.25) * t1_4 + T(0.2) * t1_4 * t1 - T(1. / 6) * t1_6;
    T t3 = x - T(3);
    T t3_2 = t3 * t3;
    T t3_4 = t3_2 * t3_2;
    T result3 = T(1.0986122886681098) + T(1. / 3) * t3 - T(0.05555555555555555) * t3_2 +
                T(0.012345679012345678) * t3_2 * t3 - T(0.0030864197530864196) * t3_4;
    T t9 = x - T(9);
    T t9_2 = t9 * t9;
    T t9_4 = t9_2 * t9_2;
    T result9 = T(2.1972245773362196) + T(1. / 9) * t9 - T(0.006172839506172839) * t9_2 +
                T(0.0004572473708276177) * t9_2 * t9 - T(3.8103947568968146e-05) * t9_4;
    T result = result1;
    if (x >= T(2) && x < T(6)) result = result3;
    if (x >= T(6)) result = result9;
    result -= (T(1) - x / exp(result));
    return result;
  }
};

// exp/log policy: the device math library (approximateExponentialFunction=False)
struct ExpLibm {
  static FMS_HD double exp(double x) { return ::exp(x); }
  static FMS_HD float exp(float x) { return ::expf(x); }
  static FMS_HD double log(double x) { return ::log(x); }
  static FMS_HD float log(float x) { return ::logf(x); }
  static FMS_HD FD exp(FD x) { return FD(::exp(x.v)); }
  static FMS_HD FD log(FD x) { return FD(::log(x.v)); }
};

// Natural logarithm from IEEE double operations only (no library call beyond
// the exact frexp), so that a C build with the same operations returns the
// same bits on the host. Models whose source calls an accurate log outside the
// generator's approximation policy use it: the hand-written TK2014 kernel calls
// Vc::log (core/src/specialized_solver/fast_monodomain_solver/
// tk2014.cpp.disabled:242,269).
//   x = m 2^k, m in [sqrt(1/2), sqrt(2)); s = (m - 1)/(m + 1) (m - 1 exact);
//   log m = 2 atanh s = 2s + s z P(z), z = s^2 <= 0.0295,
//   P(z) = sum_{n=1..10} 2 z^(n-1) / (2n + 1) (truncation below 2^-54 relative);
//   log x = k ln2_hi + (log m + k ln2_lo), ln2_hi with 32 significant bits.
FMS_HD double log_accurate(double x) {
  if (!(x > 0.0)) return x == 0.0 ? -HUGE_VAL : (x - x) / (x - x);   // -inf, or NaN for x < 0 / NaN
  if (x == HUGE_VAL) return x;
  int k;
  double m = frexp(x, &k);   // x = m 2^k, m in [0.5, 1)
  if (m < 0.7071067811865476) {
    m = 2.0 * m;
    k -= 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  double p = 0.09523809523809523;
  p = p * z + 0.10526315789473684;
  p = p * z + 0.11764705882352941;
  p = p * z + 0.13333333333333333;
  p = p * z + 0.15384615384615385;
  p = p * z + 0.18181818181818182;
  p = p * z + 0.2222222222222222;
  p = p * z + 0.2857142857142857;
  p = p * z + 0.4;
  p = p * z + 0.6666666666666666;
  const double log_m = 2.0 * s + s * z * p;
  const double dk = (double)k;
  return dk * 0.6931471803691238 + (log_m + dk * 1.9082149292705877e-10);
}
FMS_HD FD log_accurate(FD x) { return FD(log_accurate(x.v)); }
FMS_HD float log_accurate(float x) { return (float)log_accurate((double)x); }

template <class T> FMS_HD T pow2(T x) { return x * x; }
template <class T> FMS_HD T pow3(T x) { return x * (pow2(x)); }
template <class T> FMS_HD T pow4(T x) { return pow2(pow2(x)); }
template <class T> FMS_HD T pow5(T x) { return pow2(pow3(x)); }
template <class T> FMS_HD T pow6(T x) { return pow3(pow3(x)); }
template <class T> FMS_HD T pow7(T x) { return pow3(pow4(x)); }
template <class T> FMS_HD T pow8(T x) { return pow4(pow4(x)); }
template <class T> FMS_HD T powReciprocal1(T x) { return T(1.) / (x); }
template <class T> FMS_HD T powReciprocal2(T x) { return T(1.) / (pow2(x)); }
template <class T> FMS_HD T powReciprocal3(T x) { return T(1.) / (x * (pow2(x))); }
template <class T> FMS_HD T powReciprocal4(T x) { return T(1.) / (pow2(pow2(x))); }

// ternary of the CellML C subset: both branches are values, pick one
template <class T> FMS_HD T select(bool c, T a, T b) { return c ? a : b; }

}  // namespace fms
