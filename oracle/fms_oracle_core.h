/*
 * FP64 CPU oracle of the OpenDiHu FastMonodomainSolver split step.
 *
 * TEST INFRASTRUCTURE ONLY. This code is the checker the HIP path is compared
 * against and the CPU baseline timed beside it. Nothing in the product
 * (opendihu_amd/, include/) links, loads or calls it.
 *
 * It restates the reference's CPU "vc" path (paths relative to the reference
 * checkout; FMS/ = core/src/specialized_solver/fast_monodomain_solver/,
 * GEN/ = core/src/cellml/source_code_generator/):
 *   - Heun step of the generated kernel: GEN/03_generator_vc.cpp:1017-1130
 *       y* = y + dt*f(y); clamp; y += 0.5*dt*(f(y) + f(y*)); clamp
 *   - approximate exponential (1+x/1024)^1024: GEN/03_generator_vc.cpp:385-407
 *   - approximate logarithm (Taylor about 1/3/9 + one Newton step):
 *       GEN/03_generator_vc.cpp:431-490
 *   - integer powers powN / powReciprocalN: GEN/03_generator_vc.cpp:541-611
 *   - stimulation state machine: FMS/fast_monodomain_solver_compute.tpp:509-627
 *   - row assembly + Thomas solve (consistent P1 mass, IE/CN):
 *       FMS/fast_monodomain_solver_compute.tpp:256-507 (one solve per call,
 *       nTimeSteps1D is not looped over, as in the reference)
 *   - Strang order: FMS/fast_monodomain_solver_compute.tpp:110-126
 *   - step-count rule of TimeSteppingScheme:
 *       core/src/time_stepping_scheme/00_time_stepping_scheme.cpp:15-70
 *   - RepeatedCall outer loop: core/src/time_stepping_scheme/repeated_call.tpp:45-80
 *
 * Extensions the reference cannot run (uneven fibre lengths, per-fibre
 * prefactor and parameters, current-pulse stimulus, spike detector, lumped
 * mass for the nested-path known-answer tests) are marked "extension".
 *
 * The translation unit that includes this header defines the cell model:
 *   NS, NA, NP                       states, algebraics, parameters
 *   static void model_rhs_a(const double *y, const double *p, double t,
 *                           double *rates, double *alg)   approximate exp/log
 *   static void model_rhs_e(...)                          libm exp/log
 *   static const double model_y0[NS];                      initial states
 *   MODEL_NAME                                              string literal
 * and optionally FMSO_MODEL_RHS_V4, the name of a four-lane form of model_rhs_a
 * over fmso_v4 (declared below). The reaction then runs four nodes per vector
 * operation, as the reference's Vc code runs W nodes per double_v
 * (GEN/03_generator_vc.cpp:1017-1130), each lane with the scalar operations.
 * Compile with -ffp-contract=off so every + - * / rounds as written.
 */
#ifndef FMS_ORACLE_CORE_H
#define FMS_ORACLE_CORE_H

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- math helpers of the generated code (GEN/03_generator_vc.cpp) ---- */

static inline double exponential_approx(double x)
{
  /* GEN/03_generator_vc.cpp:391-396: exp(x) ~ (1 + x/1024)^1024 */
  x = 1.0 + x / 1024.;
  for (int i = 0; i < 10; i++)
    x *= x;
  return x;
}

static inline double logarithmic_approx(double x)
{
  /* GEN/03_generator_vc.cpp:436-489: Taylor about 1, 3 or 9, then one
     Newton step against the approximate exponential. */
  double t1 = x - 1;
  double t1_2 = t1 * t1;
  double t1_4 = t1_2 * t1_2;
  double t1_6 = t1_4 * t1_2;
  double result1 = t1 - 0.5 * t1_2 + 1. / 3 * t1_2 * t1 - 0.25 * t1_4 + 0.2 * t1_4 * t1 - 1. / 6 * t1_6;
  double t3 = x - 3;
  double t3_2 = t3 * t3;
  double t3_4 = t3_2 * t3_2;
  double result3 = 1.0986122886681098 + 1. / 3 * t3 - 0.05555555555555555 * t3_2 + 0.012345679012345678 * t3_2 * t3 - 0.0030864197530864196 * t3_4;
  double t9 = x - 9;
  double t9_2 = t9 * t9;
  double t9_4 = t9_2 * t9_2;
  double result9 = 2.1972245773362196 + 1. / 9 * t9 - 0.006172839506172839 * t9_2 + 0.0004572473708276177 * t9_2 * t9 - 3.8103947568968146e-05 * t9_4;
  double result = result1;
  if (x >= 2 && x < 6) result = result3;
  if (x >= 6) result = result9;
  result -= (1 - x / exponential_approx(result));
  return result;
}

/* accurate natural logarithm for the hand-written TK2014 kernel, which calls
   Vc::log (FMS/tk2014.cpp.disabled:242,269) outside the generator's exp/log
   approximation: x = m 2^k, m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
   log m = 2 atanh s = 2s + s z P(z) with z = s^2 and P the first ten terms
   2 z^(n-1)/(2n+1) of the atanh series, log x = k ln2_hi + (log m + k ln2_lo).
   Plain IEEE operations and the exact frexp: no libm rounding enters */
static inline double fmso_log_accurate(double x)
{
  if (!(x > 0.0)) return x == 0.0 ? -HUGE_VAL : (x - x) / (x - x);
  if (x == HUGE_VAL) return x;
  int k;
  double m = frexp(x, &k);
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

/* four lanes of doubles (GCC vector extension): element-wise IEEE operations,
   the W = 4 double_v of the reference's Vc build on AVX2 */
typedef double fmso_v4 __attribute__((vector_size(32)));

static inline fmso_v4 exponential_approx_v4(fmso_v4 x)
{
  x = 1.0 + x / 1024.;
  for (int i = 0; i < 10; i++)
    x *= x;
  return x;
}
static inline fmso_v4 pow3_v4(fmso_v4 x) { return x * (x * x); }
static inline fmso_v4 pow4_v4(fmso_v4 x) { fmso_v4 s = x * x; return s * s; }

/* lane masks of a comparison and the lane select m ? a : b */
typedef __typeof__((fmso_v4){0} > (fmso_v4){0}) fmso_v4m;
static inline fmso_v4 fmso_vsel(fmso_v4m m, fmso_v4 a, fmso_v4 b)
{
  return (fmso_v4)((m & (fmso_v4m)a) | (~m & (fmso_v4m)b));
}

/* scalar-or-lane helpers of the four-lane model forms: broadcast a scalar, and
   c ? a : b with c a scalar condition or a lane mask */
static inline fmso_v4 fmso_v4_id(fmso_v4 x) { return x; }
static inline fmso_v4 fmso_v4_bc(double x) { return (fmso_v4){x, x, x, x}; }
static inline fmso_v4 fmso_v4_sel_s(int c, fmso_v4 a, fmso_v4 b) { return c ? a : b; }
#define FMSO_BC(x) _Generic((x), fmso_v4: fmso_v4_id, default: fmso_v4_bc)(x)
#define FMSO_SEL(c, a, b) _Generic((c), int: fmso_v4_sel_s, default: fmso_vsel)((c), FMSO_BC(a), FMSO_BC(b))

static inline fmso_v4 fmso_log_accurate_v4(fmso_v4 x)
{
  fmso_v4 r;
  for (int l = 0; l < 4; l++) r[l] = fmso_log_accurate(x[l]);
  return r;
}

static inline fmso_v4 logarithmic_approx_v4(fmso_v4 x)
{
  /* logarithmic_approx per lane: the three Taylor branches selected by lane */
  fmso_v4 t1 = x - 1;
  fmso_v4 t1_2 = t1 * t1;
  fmso_v4 t1_4 = t1_2 * t1_2;
  fmso_v4 t1_6 = t1_4 * t1_2;
  fmso_v4 result1 = t1 - 0.5 * t1_2 + 1. / 3 * t1_2 * t1 - 0.25 * t1_4 + 0.2 * t1_4 * t1 - 1. / 6 * t1_6;
  fmso_v4 t3 = x - 3;
  fmso_v4 t3_2 = t3 * t3;
  fmso_v4 t3_4 = t3_2 * t3_2;
  fmso_v4 result3 = 1.0986122886681098 + 1. / 3 * t3 - 0.05555555555555555 * t3_2 + 0.012345679012345678 * t3_2 * t3 - 0.0030864197530864196 * t3_4;
  fmso_v4 t9 = x - 9;
  fmso_v4 t9_2 = t9 * t9;
  fmso_v4 t9_4 = t9_2 * t9_2;
  fmso_v4 result9 = 2.1972245773362196 + 1. / 9 * t9 - 0.006172839506172839 * t9_2 + 0.0004572473708276177 * t9_2 * t9 - 3.8103947568968146e-05 * t9_4;
  fmso_v4 result = fmso_vsel((x >= 2) & (x < 6), result3, result1);
  result = fmso_vsel(x >= 6, result9, result);
  result -= (1 - x / exponential_approx_v4(result));
  return result;
}
#ifdef FMSO_MODEL_RHS_V4
static inline void FMSO_MODEL_RHS_V4(const fmso_v4 *y, const fmso_v4 *p, double t, fmso_v4 *r,
                                     fmso_v4 *a);
#endif

/* integer powers, built as the generator builds them: powN = pow(N/2)(pow(N-N/2)) */
static inline double pow2_(double x) { return x * x; }
static inline double pow3_(double x) { return x * (pow2_(x)); }
static inline double pow4_(double x) { return pow2_(pow2_(x)); }
static inline double pow5_(double x) { return pow2_(pow3_(x)); }
static inline double pow6_(double x) { return pow3_(pow3_(x)); }
static inline double powR1_(double x) { return 1. / (x); }
static inline double powR2_(double x) { return 1. / (pow2_(x)); }
static inline double powR3_(double x) { return 1. / (x * (pow2_(x))); }
static inline double powR4_(double x) { return 1. / (pow2_(pow2_(x))); }

/* ---- problem description (extension: runtime sizes, uneven fibres) ---- */

enum { FMSO_IE = 0, FMSO_CN = 1 };
enum { FMSO_CONSISTENT_MASS = 0, FMSO_LUMPED_MASS = 1 };
enum { FMSO_EXP_APPROX = 0, FMSO_EXP_LIBM = 1 };
enum { FMSO_STIM_NONE = 0, FMSO_STIM_FIRING_TABLE = 1, FMSO_STIM_CURRENT = 2 };

typedef struct {
  int n_fibers;
  const int *n_nodes;            /* [n_fibers], >= 2 each */
  const double *elem_len;        /* [sum(n_nodes-1)] element lengths, fibre after fibre */
  const double *prefactor;       /* [n_fibers] sigma/(Am*Cm) */
  const double *params;          /* [n_fibers*NP] */
  int scheme;                    /* FMSO_IE / FMSO_CN */
  int mass;                      /* FMSO_CONSISTENT_MASS (FMS) / FMSO_LUMPED_MASS */
  int exp_mode;                  /* FMSO_EXP_APPROX / FMSO_EXP_LIBM */
  double value_for_stimulated_point;
  int stim_mode;
  const int *stim_begin;         /* [n_fibers] first stimulated node (fibre-local) */
  const int *stim_end;           /* [n_fibers] one past the last stimulated node */
  /* firing-table stimulation (reference semantics) */
  int fe_rows, fe_cols;
  const unsigned char *firing_events;      /* [fe_rows*fe_cols] */
  const int *motor_unit;                   /* [n_fibers] */
  const double *call_frequency;            /* [n_fibers] setSpecificStatesCallFrequency */
  const double *repeat_after_first_call;   /* [n_fibers] */
  const double *call_enable_begin;         /* [n_fibers] */
  const double *jitter;                    /* concatenated jitter factors */
  const int *jitter_offset, *jitter_count; /* [n_fibers] */
  /* current-pulse stimulation (extension) */
  const double *stim_onset, *stim_duration, *stim_amplitude;  /* [n_fibers] */
  int stim_param;                /* index of the parameter that carries the current */
  /* spike detector (extension): upward crossing of the threshold */
  double spike_threshold;
  /* onlyComputeIfHasBeenStimulated (compute.tpp:201-205), applied per fibre:
     the reaction of a fibre runs from the Heun step of its first stimulation on */
  int only_if_stimulated;
} fmso_problem;

typedef struct {
  double *states;                /* SoA [NS][n_total] */
  double *last_check_time;       /* [n_fibers] */
  double *current_jitter;        /* [n_fibers] */
  int *jitter_index;             /* [n_fibers] */
  unsigned char *currently_stimulating;   /* [n_fibers] */
  unsigned char *has_been_stimulated;     /* [n_fibers] */
  int *first_spike_step;         /* [n_total] or NULL */
  int *spike_count;              /* [n_total] or NULL */
  int64_t step_counter;          /* split steps done so far (spike-step index) */
  int *stim_log_count;           /* number of logged stimulation onsets, may be NULL */
  double *stim_log_time;         /* [stim_log_capacity] onset times */
  int *stim_log_fiber;           /* [stim_log_capacity] */
  int stim_log_capacity;
} fmso_state;

static long fmso_total_nodes(const fmso_problem *p)
{
  long n = 0;
  for (int f = 0; f < p->n_fibers; f++) n += p->n_nodes[f];
  return n;
}

/* ---- stimulation state machine, FMS/fast_monodomain_solver_compute.tpp:509-627 ---- */

static int fmso_stim_check(const fmso_problem *p, fmso_state *s, int f, double t)
{
  const double freq = p->call_frequency[f];
  const double repeat = p->repeat_after_first_call[f];
  const double begin = p->call_enable_begin[f];
  int check = 0;
  if (t >= s->last_check_time[f] + 1. / (freq + s->current_jitter[f]) && t >= begin - 1e-13) {
    check = 1;
    if (repeat != 0 && t - (s->last_check_time[f] + 1. / (freq + s->current_jitter[f])) > repeat) {
      s->last_check_time[f] += 1. / (freq + s->current_jitter[f]);
      double jf = 0.0;
      if (p->jitter_count[f] > 0)
        jf = p->jitter[p->jitter_offset[f] + s->jitter_index[f] % p->jitter_count[f]];
      s->current_jitter[f] = jf * freq;
      s->jitter_index[f]++;
      check = 0;
    }
  }
  int idx = (int)round(t * freq);
  unsigned long row = (unsigned long)(long)idx % (unsigned long)p->fe_rows;
  int col = p->motor_unit[f] % p->fe_cols;
  int stim = check && p->firing_events[row * p->fe_cols + col];
  if (stim) {
    s->has_been_stimulated[f] = 1;
    if (!s->currently_stimulating[f]) {
      s->currently_stimulating[f] = 1;
      if (s->stim_log_count) {
        int k = *s->stim_log_count;
        if (k < s->stim_log_capacity) { s->stim_log_time[k] = t; s->stim_log_fiber[k] = f; }
        *s->stim_log_count = k + 1;
      }
    }
  } else {
    s->currently_stimulating[f] = 0;
  }
  return stim;
}

/* current pulse (extension): on at Heun steps whose start time lies in [onset, onset+duration) */
static inline int fmso_current_on(const fmso_problem *p, int f, double t)
{
  return t >= p->stim_onset[f] - 1e-13 && t < p->stim_onset[f] + p->stim_duration[f] - 1e-13;
}

/* ---- one Heun step of one node, GEN/03_generator_vc.cpp:1017-1130 ---- */

static inline void fmso_heun_node(double *y, const double *prm, double t, double dt,
                                  int clamp, double vstim, int exp_mode)
{
  double r[NS], ys[NS], rs[NS], a[NA > 0 ? NA : 1], as[NA > 0 ? NA : 1];
  if (exp_mode == FMSO_EXP_APPROX) model_rhs_a(y, prm, t, r, a); else model_rhs_e(y, prm, t, r, a);
  for (int k = 0; k < NS; k++) ys[k] = y[k] + dt * r[k];
  if (clamp) ys[0] = vstim;
  if (exp_mode == FMSO_EXP_APPROX) model_rhs_a(ys, prm, t, rs, as); else model_rhs_e(ys, prm, t, rs, as);
  for (int k = 0; k < NS; k++) y[k] += 0.5 * dt * (r[k] + rs[k]);
  if (clamp) y[0] = vstim;
}

/* reaction half step of one fibre: n0D Heun steps starting at t0 */
static void fmso_compute0D_fiber(const fmso_problem *p, fmso_state *s, int f, long off, long ntot,
                                 double t0, double dt, int n0D, const unsigned char *flags)
{
  const int n = p->n_nodes[f];
  double prm[NP > 0 ? NP : 1], prm_on[NP > 0 ? NP : 1];
  for (int k = 0; k < NP; k++) prm[k] = prm_on[k] = p->params[(long)f * NP + k];
  for (int j = 0; j < n0D; j++) {
    const double t = t0 + j * dt;
    if (!(flags[j] & 2)) continue;
    const int on = flags[j] & 1;
    if (p->stim_mode == FMSO_STIM_CURRENT && NP > 0)
      prm_on[p->stim_param] = on ? p->stim_amplitude[f] : prm[p->stim_param];
    int i = 0;
#ifdef FMSO_MODEL_RHS_V4
    /* models with a four-lane right-hand side: groups of four nodes, each lane
       the scalar Heun step's operations (approximate exponential only) */
    if (p->exp_mode == FMSO_EXP_APPROX) {
      for (; i + 4 <= n; i += 4) {
        fmso_v4 y[NS], r[NS], ys[NS], rs[NS], a[NA > 0 ? NA : 1], as[NA > 0 ? NA : 1], pl[NP > 0 ? NP : 1];
        int clamp[4];
        const double *pp[4];
        for (int k = 0; k < NS; k++) memcpy(&y[k], s->states + (long)k * ntot + off + i, sizeof(fmso_v4));
        for (int l = 0; l < 4; l++) {
          const int in_set = i + l >= p->stim_begin[f] && i + l < p->stim_end[f];
          clamp[l] = p->stim_mode == FMSO_STIM_FIRING_TABLE && on && in_set;
          pp[l] = (p->stim_mode == FMSO_STIM_CURRENT && in_set) ? prm_on : prm;
        }
        for (int k = 0; k < NP; k++) pl[k] = (fmso_v4){pp[0][k], pp[1][k], pp[2][k], pp[3][k]};
        FMSO_MODEL_RHS_V4(y, pl, t, r, a);
        for (int k = 0; k < NS; k++) ys[k] = y[k] + dt * r[k];
        for (int l = 0; l < 4; l++)
          if (clamp[l]) ys[0][l] = p->value_for_stimulated_point;
        FMSO_MODEL_RHS_V4(ys, pl, t, rs, as);
        for (int k = 0; k < NS; k++) y[k] += 0.5 * dt * (r[k] + rs[k]);
        for (int l = 0; l < 4; l++)
          if (clamp[l]) y[0][l] = p->value_for_stimulated_point;
        for (int k = 0; k < NS; k++) memcpy(s->states + (long)k * ntot + off + i, &y[k], sizeof(fmso_v4));
      }
    }
#endif
    for (; i < n; i++) {
      double y[NS];
      for (int k = 0; k < NS; k++) y[k] = s->states[(long)k * ntot + off + i];
      const int in_set = i >= p->stim_begin[f] && i < p->stim_end[f];
      const int clamp = p->stim_mode == FMSO_STIM_FIRING_TABLE && on && in_set;
      const double *pp = (p->stim_mode == FMSO_STIM_CURRENT && in_set) ? prm_on : prm;
      fmso_heun_node(y, pp, t, dt, clamp, p->value_for_stimulated_point, p->exp_mode);
      for (int k = 0; k < NS; k++) s->states[(long)k * ntot + off + i] = y[k];
    }
  }
}

/* ---- row assembly + Thomas, FMS/fast_monodomain_solver_compute.tpp:336-486 ---- */

static void fmso_thomas_fiber(double *u, int n, const double *h, double prefactor, double dt,
                              int scheme, int mass, double *cA, double *dA)
{
  const int ie = scheme == FMSO_IE;
  const double wo = mass == FMSO_LUMPED_MASS ? 0.0 : 1. / 6;   /* off-diagonal mass weight */
  const double wd = mass == FMSO_LUMPED_MASS ? 0.5 : 1. / 3;   /* diagonal mass weight */
  for (int i = 0; i < n; i++) {
    double a = 0, b = 0, c = 0, d = 0;
    double u_center = u[i];
    if (i > 0) {
      double u_previous = u[i - 1];
      double h_left = h[i - 1];
      double k_left = 1. / h_left * (1) * prefactor;
      double m_left = mass == FMSO_LUMPED_MASS ? h_left * wo : h_left * 1. / 6;
      if (ie) a = (k_left - 1 / dt * m_left);
      else    a = (k_left / 2. - 1 / dt * m_left);
      double k_right = 1. / h_left * (-1) * prefactor;
      double m_right = mass == FMSO_LUMPED_MASS ? h_left * wd : h_left * 1. / 3;
      if (ie) {
        b += (k_right - 1 / dt * m_right);
        d += (-1 / dt * m_left) * u_previous + (-1 / dt * m_right) * u_center;
      } else {
        b += (k_right / 2. - 1 / dt * m_right);
        d += (-k_left / 2. - 1 / dt * m_left) * u_previous + (-k_right / 2. - 1 / dt * m_right) * u_center;
      }
    }
    if (i < n - 1) {
      double u_next = u[i + 1];
      double h_right = h[i];
      double k_right = 1. / h_right * (1) * prefactor;
      double m_right = mass == FMSO_LUMPED_MASS ? h_right * wo : h_right * 1. / 6;
      if (ie) c = (k_right - 1 / dt * m_right);
      else    c = (k_right / 2. - 1 / dt * m_right);
      double k_left = 1. / h_right * (-1) * prefactor;
      double m_left = mass == FMSO_LUMPED_MASS ? h_right * wd : h_right * 1. / 3;
      if (ie) {
        b += (k_left - 1 / dt * m_left);
        d += (-1 / dt * m_left) * u_center + (-1 / dt * m_right) * u_next;
      } else {
        b += (k_left / 2. - 1 / dt * m_left);
        d += (-k_left / 2. - 1 / dt * m_left) * u_center + (-k_right / 2. - 1 / dt * m_right) * u_next;
      }
    }
    if (i == 0) {
      cA[i] = c / b;
      dA[i] = d / b;
    } else {
      if (i != n - 1)
        cA[i] = c / (b - cA[i - 1] * a);
      dA[i] = (d - dA[i - 1] * a) / (b - cA[i - 1] * a);
    }
  }
  u[n - 1] = dA[n - 1];
  double previous = dA[n - 1];
  for (int i = n - 2; i >= 0; i--) {
    double r = dA[i] - cA[i] * previous;
    u[i] = r;
    previous = r;
  }
}

/* ---- computeMonodomain: FMS/fast_monodomain_solver_compute.tpp:110-126 ---- */

static void fmso_fiber_flags(const fmso_problem *p, fmso_state *s, int f, double t0, double dt,
                             int n0D, unsigned char *flags)
{
  for (int j = 0; j < n0D; j++) {
    const double t = t0 + j * dt;
    int stim = 0;
    if (p->stim_mode == FMSO_STIM_FIRING_TABLE) {
      stim = fmso_stim_check(p, s, f, t);
    } else if (p->stim_mode == FMSO_STIM_CURRENT) {
      stim = fmso_current_on(p, f, t);
      if (stim) s->has_been_stimulated[f] = 1;
    }
    const int compute = !p->only_if_stimulated || s->has_been_stimulated[f];
    flags[j] = (unsigned char)((stim ? 1 : 0) | (compute ? 2 : 0));
  }
}

static void fmso_spikes(const fmso_problem *p, fmso_state *s, long lo, long hi, long ntot,
                        const double *vm_before)
{
  for (long i = lo; i < hi; i++) {
    const double v = s->states[i];
    if (vm_before[i - lo] < p->spike_threshold && v >= p->spike_threshold) {
      if (s->spike_count) s->spike_count[i]++;
      if (s->first_spike_step && s->first_spike_step[i] < 0) s->first_spike_step[i] = (int)s->step_counter;
    }
  }
  (void)ntot;
}

/* Advance n_split Strang steps starting at t0. The stimulation state machine of a
   fibre runs in time order (it does not depend on the states), so it is
   evaluated first for each half step; fibres are independent, so they are
   distributed over threads (the reference distributes them over ranks,
   FMS/fast_monodomain_solver_initialization.tpp:260). */
int fmso_advance(const fmso_problem *p, fmso_state *s, double t0, double dts, int n_split,
                 double dt0D, int n0D, double dt1D, int n1D, int n_threads)
{
  (void)n1D;
  const long ntot = fmso_total_nodes(p);
  long *offs = (long *)malloc(sizeof(long) * (p->n_fibers + 1));
  long *eoffs = (long *)malloc(sizeof(long) * (p->n_fibers + 1));
  offs[0] = 0; eoffs[0] = 0;
  int nmax = 2;
  for (int f = 0; f < p->n_fibers; f++) {
    offs[f + 1] = offs[f] + p->n_nodes[f];
    eoffs[f + 1] = eoffs[f] + p->n_nodes[f] - 1;
    if (p->n_nodes[f] > nmax) nmax = p->n_nodes[f];
  }
  /* firing-table stimulation writes a global log in time order per fibre;
     the log is per run, so keep the run single-threaded when it is on */
  int nt = n_threads > 0 ? n_threads : 1;
  if (s->stim_log_count && p->stim_mode == FMSO_STIM_FIRING_TABLE) nt = 1;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
#endif
  for (int f = 0; f < p->n_fibers; f++) {
    unsigned char *flags = (unsigned char *)malloc(n0D > 0 ? n0D : 1);
    double *cA = (double *)malloc(sizeof(double) * nmax);
    double *dA = (double *)malloc(sizeof(double) * nmax);
    double *vb = (double *)malloc(sizeof(double) * nmax);
    const int n = p->n_nodes[f];
    for (int step = 0; step < n_split; step++) {
      const double t = t0 + step * dts;
      const double tmid = t + 0.5 * dts;
      if (p->spike_threshold == p->spike_threshold)
        memcpy(vb, s->states + offs[f], sizeof(double) * n);
      fmso_fiber_flags(p, s, f, t, dt0D, n0D, flags);
      fmso_compute0D_fiber(p, s, f, offs[f], ntot, t, dt0D, n0D, flags);
      fmso_thomas_fiber(s->states + offs[f], n, p->elem_len + eoffs[f], p->prefactor[f], dt1D,
                        p->scheme, p->mass, cA, dA);
      fmso_fiber_flags(p, s, f, tmid, dt0D, n0D, flags);
      fmso_compute0D_fiber(p, s, f, offs[f], ntot, tmid, dt0D, n0D, flags);
      if (p->spike_threshold == p->spike_threshold) {
        fmso_state sl = *s;
        sl.step_counter = s->step_counter + step + 1;
        fmso_spikes(p, &sl, offs[f], offs[f] + n, ntot, vb);
      }
    }
    free(flags); free(cA); free(dA); free(vb);
  }
  s->step_counter += n_split;
  free(offs); free(eoffs);
  return 0;
}

/* ---- exported single-purpose entry points used by the tests ---- */

int fmso_model_info(int *n_states, int *n_alg, int *n_params)
{
  *n_states = NS; *n_alg = NA; *n_params = NP;
  return 0;
}

void fmso_initial_states(double *y0)
{
  for (int k = 0; k < NS; k++) y0[k] = model_y0[k];
}

void fmso_rhs(const double *y, const double *prm, double t, double *rates, double *alg, int exp_mode)
{
  double a[NA > 0 ? NA : 1];
  if (exp_mode == FMSO_EXP_APPROX) model_rhs_a(y, prm, t, rates, a); else model_rhs_e(y, prm, t, rates, a);
  if (alg) for (int k = 0; k < NA; k++) alg[k] = a[k];
}

/* n Heun steps of one cell, no diffusion */
void fmso_heun_cell(double *y, const double *prm, double t0, double dt, int n, int exp_mode)
{
  for (int j = 0; j < n; j++) fmso_heun_node(y, prm, t0 + j * dt, dt, 0, 0.0, exp_mode);
}

/* n explicit Euler steps of one cell (used for the reference's CellML known-answer tests) */
void fmso_euler_cell(double *y, const double *prm, double t0, double dt, long n, int exp_mode)
{
  double r[NS], a[NA > 0 ? NA : 1];
  for (long j = 0; j < n; j++) {
    if (exp_mode == FMSO_EXP_APPROX) model_rhs_a(y, prm, t0 + j * dt, r, a); else model_rhs_e(y, prm, t0 + j * dt, r, a);
    for (int k = 0; k < NS; k++) y[k] += dt * r[k];
  }
}

void fmso_thomas(double *u, int n, const double *h, double prefactor, double dt, int scheme, int mass)
{
  double *cA = (double *)malloc(sizeof(double) * n);
  double *dA = (double *)malloc(sizeof(double) * n);
  fmso_thomas_fiber(u, n, h, prefactor, dt, scheme, mass, cA, dA);
  free(cA); free(dA);
}

/* ---- TimeSteppingScheme step-count rule, 00_time_stepping_scheme.cpp:15-70 ---- */

typedef struct {
  double start, end, target, width;
  int n;
  int significant;
} fmso_tss;

static void fmso_tss_set_n(fmso_tss *s, int n)
{
  s->n = n;
  s->width = (s->end - s->start) / n;
}

static void fmso_tss_set_width(fmso_tss *s, double w)
{
  s->target = w;
  long double n = ceill((long double)(s->end - s->start) / w - 1e-1);
  s->n = n < 1 ? 1 : (int)n;
  fmso_tss_set_n(s, s->n);
}

void fmso_tss_init(fmso_tss *s, double end_time, double width, int n_steps)
{
  s->start = 0.0; s->end = end_time;
  if (width > 0) {
    s->target = width;
    fmso_tss_set_width(s, width);
    s->significant = 1;
  } else {
    s->significant = 0; s->target = 0.0;
    fmso_tss_set_n(s, n_steps);
  }
}

void fmso_tss_set_time_span(fmso_tss *s, double start, double end)
{
  s->start = start; s->end = end;
  if (s->significant) {
    if (s->target > s->end - s->start) s->width = s->end - s->start;
    fmso_tss_set_width(s, s->width);
  } else {
    fmso_tss_set_n(s, s->n);
  }
}

/* RepeatedCall<FastMonodomainSolver>: per call k the values computeMonodomain
   receives (repeated_call.tpp:45-80, fast_monodomain_solver_compute.tpp:64-85).
   rc/strang/heun/ie widths are the settings' timeStepWidth values; end times
   likewise. Writes up to max_calls rows of
   (startTime, dtSplit, nSplit, dt0D, n0D, dt1D, n1D) and returns the number of calls. */
int fmso_schedule(double rc_end, double rc_width, double strang_end, double strang_width,
                  double heun_end, double heun_width, double ie_end, double ie_width,
                  int max_calls, double *out)
{
  fmso_tss rc, st, he, ie;
  fmso_tss_init(&rc, rc_end, rc_width, 0);
  fmso_tss_init(&st, strang_end, strang_width, 0);
  fmso_tss_init(&he, heun_end, heun_width, 0);
  fmso_tss_init(&ie, ie_end, ie_width, 0);
  const double span = rc.end - rc.start;
  double t = rc.start;
  int k = 0;
  for (int step = 0; step < rc.n; ) {
    fmso_tss_set_time_span(&st, t, t + rc.width);
    const double t0 = st.start, dts = st.width;
    fmso_tss_set_time_span(&he, t0, t0 + 0.5 * dts);
    fmso_tss_set_time_span(&ie, t0, t0 + dts);
    if (k < max_calls) {
      double *o = out + 7 * (long)k;
      o[0] = t0; o[1] = dts; o[2] = st.n; o[3] = he.width; o[4] = he.n; o[5] = ie.width; o[6] = ie.n;
    }
    k++;
    step++;
    t = rc.start + (double)step / rc.n * span;
  }
  return k;
}

const char *fmso_model_name(void) { return MODEL_NAME; }

/* the accurate logarithm over an array (checked against the host's log) */
void fmso_log_accurate_n(const double *x, double *y, long n)
{
  for (long i = 0; i < n; i++) y[i] = fmso_log_accurate(x[i]);
}

#endif /* FMS_ORACLE_CORE_H */
