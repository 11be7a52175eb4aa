/*
 * Oracle cell model: Hodgkin-Huxley 1952 with the resting potential shifted to
 * E_R = -75 mV (test infrastructure only, see fms_oracle_core.h).
 *
 * Restates testing/unit_testing/input/hodgkin_huxley_1952.c:36-68 in the form the
 * reference's Vc generator gives it (GEN/03_generator_vc.cpp:13-302, 937-1016):
 * CONSTANTS become local constants, exp -> exponential, pow(x,3) -> pow3(x) =
 * x*(x*x), pow(x,4) -> pow4(x) = (x*x)*(x*x), and CONSTANTS[2] (i_Stim) is the
 * only parameter (parametersUsedAsConstant = [2], as in
 * examples/electrophysiology/monodomain/hodgkin_huxley_fast/settings_fast_fiber.py:162).
 * Extension: with NP = 4 the conductances g_Na, g_K, g_L (CONSTANTS[3..5]) are
 * parameters 1..3 as well (parametersUsedAsConstant = [2,3,4,5]).
 */
#define NS 4
#define NA 9
#define NP 4
#define MODEL_NAME "hodgkin_huxley_1952"

static const double model_y0[NS] = {-75, 0.05, 0.6, 0.325};

/* T: double, or fmso_vw (four nodes per operation, each lane the same IEEE
   operations as the scalar form, as the reference's Vc double_v evaluates W nodes) */
#define HH_RHS(NAME, T, EXPF, POW3, POW4)                                                         \
  static inline void NAME(const T *y, const T *p, double t, T *r, T *a)                         \
  {                                                                                               \
    (void)t;                                                                                      \
    const double constant0 = -75;                                                                 \
    const double constant1 = 1;                                                                   \
    const double constant6 = constant0 + 115.000;                                                 \
    const double constant7 = constant0 - 12.0000;                                                 \
    const double constant8 = constant0 + 10.6130;                                                 \
    const T i_stim = p[0], g_na = p[1], g_k = p[2], g_l = p[3];                                  \
    a[1] = (-0.100000 * (y[0] + 50.0000)) / (EXPF(-(y[0] + 50.0000) / 10.0000) - 1.00000);       \
    a[5] = 4.00000 * EXPF(-(y[0] + 75.0000) / 18.0000);                                          \
    r[1] = a[1] * (1.00000 - y[1]) - a[5] * y[1];                                                 \
    a[2] = 0.0700000 * EXPF(-(y[0] + 75.0000) / 20.0000);                                        \
    a[6] = 1.00000 / (EXPF(-(y[0] + 45.0000) / 10.0000) + 1.00000);                              \
    r[2] = a[2] * (1.00000 - y[2]) - a[6] * y[2];                                                 \
    a[3] = (-0.0100000 * (y[0] + 65.0000)) / (EXPF(-(y[0] + 65.0000) / 10.0000) - 1.00000);      \
    a[7] = 0.125000 * EXPF((y[0] + 75.0000) / 80.0000);                                          \
    r[3] = a[3] * (1.00000 - y[3]) - a[7] * y[3];                                                 \
    a[0] = g_na * POW3(y[1]) * y[2] * (y[0] - constant6);                                         \
    a[4] = g_k * POW4(y[3]) * (y[0] - constant7);                                                 \
    a[8] = g_l * (y[0] - constant8);                                                              \
    r[0] = -(-i_stim + a[0] + a[4] + a[8]) / constant1;                                           \
  }

static inline void model_rhs_a(const double *y, const double *p, double t, double *r, double *a);
static inline void model_rhs_e(const double *y, const double *p, double t, double *r, double *a);
#ifndef FMSO_SCALAR_ONLY
#define FMSO_MODEL_RHS_VW model_rhs_a_vw
#endif

#include "fms_oracle_core.h"

HH_RHS(model_rhs_a, double, exponential_approx, pow3_, pow4_)
HH_RHS(model_rhs_e, double, exp, pow3_, pow4_)
#ifdef FMSO_MODEL_RHS_VW
HH_RHS(model_rhs_a_vw, fmso_vw, exponential_approx_vw, pow3_vw, pow4_vw)
#endif

/* default parameter vector of the model file: i_Stim = 0, g_Na = 120, g_K = 36, g_L = 0.3 */
void fmso_default_params(double *p)
{
  p[0] = 0; p[1] = 120; p[2] = 36; p[3] = 0.3;
}
