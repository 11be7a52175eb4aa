// Hodgkin-Huxley 1952 membrane model, resting potential shifted to -75 mV.
//
// Textbook HH rate functions with E_R = -75 mV, E_Na = E_R + 115, E_K = E_R - 12,
// E_L = E_R + 10.613, C_m = 1 uF/cm^2, in the operation order of the model file
// the reference loads (testing/unit_testing/input/hodgkin_huxley_1952.c:53-66)
// after its generator rewrites exp -> exponential and pow(x,3|4) -> pow3|pow4
// (core/src/cellml/source_code_generator/03_generator_vc.cpp:13-302).
//
// Parameters (per fibre): p[0] = i_Stim (CONSTANTS[2]; parametersUsedAsConstant
// = [2] in examples/electrophysiology/monodomain/hodgkin_huxley_fast/
// settings_fast_fiber.py:162), p[1..3] = g_Na, g_K, g_L (CONSTANTS[3..5]).
#pragma once

#include "fms_math.h"

namespace fms {

struct ModelHH {
  static constexpr int NS = 4;   // V, m, h, n
  static constexpr int NA = 9;   // algebraics of the model file
  static constexpr int NP = 4;   // i_Stim, g_Na, g_K, g_L
  static constexpr int STIM_PARAM = 0;
  // algorithmic ops of one rhs evaluation, counted as the code generator counts
  // them for the model file: 23 add/sub, 22 mul, 10 div, 6 exp
  static constexpr double OPS_RHS = 61;
  static constexpr double DIVS_RHS = 10;
  static constexpr double EXPS_RHS = 6;
  static constexpr const char *name() { return "hodgkin_huxley_1952"; }
  static FMS_HD double y0(int k) {
    const double v[NS] = {-75, 0.05, 0.6, 0.325};
    return v[k];
  }
  static FMS_HD double p0(int k) {
    const double v[NP] = {0, 120, 36, 0.3};
    return v[k];
  }
  // every algebraic is computed (no parametersUsedAsAlgebraic)
  static FMS_HD constexpr bool alg_is_param(int k) { return (void)k, false; }

  // rates r[0..NS) and algebraics a[0..NA) at state y; T is the arithmetic type
  template <class T, class E>
  static FMS_HD void rhs(const T *y, const T *p, double t, T *r, T *a) {
    (void)t;
    const T constant0 = T(-75);
    const T constant1 = T(1);
    const T constant6 = constant0 + T(115.000);
    const T constant7 = constant0 - T(12.0000);
    const T constant8 = constant0 + T(10.6130);
    const T i_stim = p[0], g_na = p[1], g_k = p[2], g_l = p[3];
    a[1] = (T(-0.100000) * (y[0] + T(50.0000))) / (E::exp(-(y[0] + T(50.0000)) / T(10.0000)) - T(1.00000));
    a[5] = T(4.00000) * E::exp(-(y[0] + T(75.0000)) / T(18.0000));
    r[1] = a[1] * (T(1.00000) - y[1]) - a[5] * y[1];
    a[2] = T(0.0700000) * E::exp(-(y[0] + T(75.0000)) / T(20.0000));
    a[6] = T(1.00000) / (E::exp(-(y[0] + T(45.0000)) / T(10.0000)) + T(1.00000));
    r[2] = a[2] * (T(1.00000) - y[2]) - a[6] * y[2];
    a[3] = (T(-0.0100000) * (y[0] + T(65.0000))) / (E::exp(-(y[0] + T(65.0000)) / T(10.0000)) - T(1.00000));
    a[7] = T(0.125000) * E::exp((y[0] + T(75.0000)) / T(80.0000));
    r[3] = a[3] * (T(1.00000) - y[3]) - a[7] * y[3];
    a[0] = g_na * pow3(y[1]) * y[2] * (y[0] - constant6);
    a[4] = g_k * pow4(y[3]) * (y[0] - constant7);
    a[8] = g_l * (y[0] - constant8);
    r[0] = -(-i_stim + a[0] + a[4] + a[8]) / constant1;
  }
};

}  // namespace fms
