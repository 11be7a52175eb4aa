// Discrete adjoint of the split step (Hodgkin-Huxley fibres, current-pulse
// stimulation): reverse Heun steps with the transposed reaction Jacobian and a
// reverse tridiagonal solve, parameter gradients for g_Na, g_K, g_L per fibre
// and for the diffusion prefactor sigma/(Am Cm) per fibre.
//
// Forward split step (fast_monodomain_solver_compute.tpp:110-126):
//   y^a = H_{n0D}(y, t)  ->  V^b = A^{-1} B V^a  ->  y' = H_{n0D}(y^b, t + dt/2)
// Heun step y' = y + dt/2 (f(y) + f(y*)), y* = y + dt f(y). With lambda' the
// adjoint of y', mu = J(y*)^T lambda':
//   lambda = lambda' + dt/2 mu + dt/2 J(y)^T (lambda' + dt mu)
//   theta_bar += dt/2 [ f_theta(y)^T (lambda' + dt mu) + f_theta(y*)^T lambda' ]
// Diffusion: A and B are symmetric (consistent P1 mass and stiffness), so with
// w = A^{-1} lambda_b:  lambda_a = B w,  pf_bar += w^T (dB/dpf V^a - dA/dpf V^b),
// dA/dpf = K1 (IE) or K1/2 (CN), dB/dpf = 0 (IE) or -K1/2 (CN), K1 the stiffness
// matrix for prefactor 1.
// The derivatives are those of the implemented forward (discrete adjoint): for
// the approximate exponential E(x) = (1 + x/1024)^1024, E'(x) = E(x)/(1 + x/1024).
#pragma once

#include "fms_fused.h"
#include "model_hh.h"

namespace fms {

// a / b for the derivative terms of the linearisation (never for a rate, whose
// bits must equal the forward's): the device reciprocal refined by one Newton
// step (which squares its relative error), then one multiply; four instructions
// instead of the eight of the correctly rounded sequence. The derivatives need
// no correct rounding: the gradient's agreement with the dense adjoint (1e-9)
// and with central differences is what the tests hold it to.
FMS_HD double deriv_div(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double r0 = __builtin_amdgcn_rcp(b);
  const double r1 = __builtin_fma(r0, __builtin_fma(-b, r0, 1.0), r0);
  return a * r1;
#else
  return a / b;
#endif
}

template <class E> struct ExpDeriv;
template <> struct ExpDeriv<ExpApprox> {
  static FMS_HD FD d(FD x, FD ex) { return FD(deriv_div(ex.v, (FD(1.0) + x * FD(1.0 / 1024)).v)); }
  // power-series coefficients of E(x) - 1 = sum_k c_k x^k, c_k = binom(1024, k) / 1024^k
  static FMS_HD constexpr double c(int k) {
    double v = 1.0;
    for (int j = 1; j < k; j++) v = v * (1024.0 - j) / ((j + 1) * 1024.0);
    return v;
  }
};
template <> struct ExpDeriv<ExpLibm> {
  static FMS_HD FD d(FD x, FD ex) { (void)x; return ex; }
  static FMS_HD constexpr double c(int k) {   // 1 / k!
    double v = 1.0;
    for (int j = 2; j <= k; j++) v /= j;
    return v;
  }
};

// u(x) = x / (E(x) - 1) has a removable singularity at x = 0: the HH rates
// alpha_m = u(-(V+50)/10) and alpha_n = u(-(V+65)/10) / 10. The quotient rule
// u' = (E - 1 - x E') / (E - 1)^2 loses its accuracy near 0: the numerator is
// about -x^2/2, while E - 1 and x E' each carry ~1e-13 absolute rounding error
// (the approximate exponential raises a rounded 1 + x/1024 to the 1024th
// power), so the relative error is ~2e-13 / x^2 - O(1) a few hundredths of a
// millivolt from the singular voltage, a point every action potential's
// upstroke passes. For |x| < 0.05, u' is taken from its own power series
// u'(x) = sum_j d_j x^j, degree 7 (truncation ~1e-18 relative at |x| = 0.05;
// the quotient rule's error there is 1e-10). The coefficients follow at
// compile time from the series of E: with s(x) = (E(x) - 1)/x = sum_k c_{k+1} x^k,
// u = 1/s term by term, d_j = (j+1) u_{j+1}.
template <class E>
constexpr double du_coeff(int j) {
  double s[9] = {}, u[9] = {};
  for (int k = 0; k < 9; k++) s[k] = ExpDeriv<E>::c(k + 1);
  u[0] = 1.0 / s[0];
  for (int m = 1; m < 9; m++) {
    double acc = 0.0;
    for (int k = 1; k <= m; k++) acc += s[k] * u[m - k];
    u[m] = -acc / s[0];
  }
  return (j + 1) * u[j + 1];
}

// against the exact rational coefficients (binomial series inverted in fractions)
static_assert(du_coeff<ExpApprox>(0) == -0.49951171875, "d0 = -binom(1024,2)/1024^2");
static_assert(du_coeff<ExpApprox>(1) - 0.16666650772094727 < 1e-17 &&
                  du_coeff<ExpApprox>(1) - 0.16666650772094727 > -1e-17, "d1");
static_assert(du_coeff<ExpApprox>(3) - -0.0055554495918386237 < 1e-18 &&
                  du_coeff<ExpApprox>(3) - -0.0055554495918386237 > -1e-18, "d3");
static_assert(du_coeff<ExpApprox>(7) - -6.6127054049909082e-06 < 1e-20 &&
                  du_coeff<ExpApprox>(7) - -6.6127054049909082e-06 > -1e-20, "d7");
static_assert(du_coeff<ExpLibm>(0) == -0.5, "Bernoulli: B1 = -1/2");
static_assert(du_coeff<ExpLibm>(2) < 1e-16 && du_coeff<ExpLibm>(2) > -1e-16, "B3 = 0 (to the rounding of the inversion)");
static_assert(du_coeff<ExpLibm>(5) - 1.0 / 5040 < 1e-17 && du_coeff<ExpLibm>(5) - 1.0 / 5040 > -1e-17, "B6 / 5! = 1/5040");

template <class E>
__device__ __forceinline__ double du_near_zero(double x) {
  constexpr double d0 = du_coeff<E>(0), d1 = du_coeff<E>(1), d2 = du_coeff<E>(2), d3 = du_coeff<E>(3);
  constexpr double d4 = du_coeff<E>(4), d5 = du_coeff<E>(5), d6 = du_coeff<E>(6), d7 = du_coeff<E>(7);
  return d0 + x * (d1 + x * (d2 + x * (d3 + x * (d4 + x * (d5 + x * (d6 + x * d7))))));
}

__device__ __forceinline__ bool near_removable(double x) { return x < 0.05 && x > -0.05; }

// Linearisation of the HH right-hand side at y: the rates r(y), computed with
// the operations of ModelHH::rhs in its order (so y* = y + dt r(y) has the
// forward's bits), and the partial derivatives that J(y)^T w and
// f_theta(y)^T w need, from the same exponentials.
struct HHLin {
  double r[4];
  double f0V, f0m, f0h, f0n;       // d f0 / d(V, m, h, n)
  double f1V, f1m, f2V, f2h, f3V, f3n;
  double gna, gk, gl;              // d f0 / d(g_Na, g_K, g_L)
  FMS_HD void jt(const double *w, double *out) const {
    out[0] = w[0] * f0V + w[1] * f1V + w[2] * f2V + w[3] * f3V;
    out[1] = w[0] * f0m + w[1] * f1m;
    out[2] = w[0] * f0h + w[2] * f2h;
    out[3] = w[0] * f0n + w[3] * f3n;
  }
  FMS_HD void gp(const double *w, double *out) const {
    out[0] = w[0] * gna;
    out[1] = w[0] * gk;
    out[2] = w[0] * gl;
  }
};

template <class E>
__device__ __forceinline__ HHLin hh_linearize(const double *y, const double *p) {
  // FD arithmetic (the forward's division), so r(y) has the forward's bits
  HHLin L;
  const FD V = y[0], m = y[1], h = y[2], n = y[3];
  const FD i_stim = p[0], g_na = p[1], g_k = p[2], g_l = p[3];
  const FD constant0 = FD(-75), constant1 = FD(1);
  const FD constant6 = constant0 + FD(115.000), constant7 = constant0 - FD(12.0000);
  const FD constant8 = constant0 + FD(10.6130);
  // rates, in ModelHH::rhs's operation order
  const FD em = E::exp(-(V + FD(50.0000)) / FD(10.0000));
  const FD a1 = (FD(-0.100000) * (V + FD(50.0000))) / (em - FD(1.00000));
  const FD ebm = E::exp(-(V + FD(75.0000)) / FD(18.0000));
  const FD a5 = FD(4.00000) * ebm;
  const FD r1 = a1 * (FD(1.00000) - m) - a5 * m;
  const FD eah = E::exp(-(V + FD(75.0000)) / FD(20.0000));
  const FD a2 = FD(0.0700000) * eah;
  const FD ebh = E::exp(-(V + FD(45.0000)) / FD(10.0000));
  const FD a6 = FD(1.00000) / (ebh + FD(1.00000));
  const FD r2 = a2 * (FD(1.00000) - h) - a6 * h;
  const FD en = E::exp(-(V + FD(65.0000)) / FD(10.0000));
  const FD a3 = (FD(-0.0100000) * (V + FD(65.0000))) / (en - FD(1.00000));
  const FD ebn = E::exp((V + FD(75.0000)) / FD(80.0000));
  const FD a7 = FD(0.125000) * ebn;
  const FD r3 = a3 * (FD(1.00000) - n) - a7 * n;
  const FD m3 = pow3(m), n4 = pow4(n);
  const FD a0 = g_na * m3 * h * (V - constant6);
  const FD a4 = g_k * n4 * (V - constant7);
  const FD a8 = g_l * (V - constant8);
  const FD r0 = -(-i_stim + a0 + a4 + a8) / constant1;
  L.r[0] = r0.v; L.r[1] = r1.v; L.r[2] = r2.v; L.r[3] = r3.v;
  // derivatives (E'(x) of the exponential actually used)
  const FD xm = -(V + FD(50.0000)) / FD(10.0000), xbm = -(V + FD(75.0000)) / FD(18.0000);
  const FD xah = -(V + FD(75.0000)) / FD(20.0000), xbh = -(V + FD(45.0000)) / FD(10.0000);
  const FD xn = -(V + FD(65.0000)) / FD(10.0000), xbn = (V + FD(75.0000)) / FD(80.0000);
  // d alpha_m / dV = -0.1 u'(xm), d alpha_n / dV = -0.01 u'(xn) (u above): the
  // quotient rule u' = (E - 1 - x E') / (E - 1)^2, and the series within
  // |x| < 0.05 of a singular voltage
  const FD den_m = em - FD(1.0);
  double dum = deriv_div((den_m - xm * ExpDeriv<E>::d(xm, em)).v, (den_m * den_m).v);
  const FD dbm = FD(4.0) * ExpDeriv<E>::d(xbm, ebm) * FD(-1.0 / 18.0);
  const FD dah = FD(0.07) * ExpDeriv<E>::d(xah, eah) * FD(-1.0 / 20.0);
  const FD dbh = FD(deriv_div((-(ExpDeriv<E>::d(xbh, ebh) * FD(-0.1))).v, ((ebh + FD(1.0)) * (ebh + FD(1.0))).v));
  const FD den_n = en - FD(1.0);
  double dun = deriv_div((den_n - xn * ExpDeriv<E>::d(xn, en)).v, (den_n * den_n).v);
  // a lane is near at most one of the two singular voltages (1.5 apart in x).
  // The polynomial is evaluated on every lane and selected: a branch here (even
  // a wave-uniform one) splits the unrolled per-node code into blocks that the
  // scheduler cannot interleave, which measured 1.9x slower at one wave per SIMD.
  const bool near_m = near_removable(xm.v), near_n = near_removable(xn.v);
  const double dus = du_near_zero<E>(near_m ? xm.v : xn.v);
  dum = near_m ? dus : dum;
  dun = near_n ? dus : dun;
  const FD dam = FD(-0.1 * dum), dan = FD(-0.01 * dun);
  const FD dbn = FD(0.125) * ExpDeriv<E>::d(xbn, ebn) * FD(1.0 / 80.0);
  const FD m2 = m * m, n3 = n * n * n;
  L.f0V = (-(g_na * m3 * h + g_k * n4 + g_l)).v;
  L.f0m = (-(g_na * FD(3.0) * m2 * h * (V - constant6))).v;
  L.f0h = (-(g_na * m3 * (V - constant6))).v;
  L.f0n = (-(g_k * FD(4.0) * n3 * (V - constant7))).v;
  L.f1V = (dam * (FD(1.0) - m) - dbm * m).v;
  L.f1m = (-a1 - a5).v;
  L.f2V = (dah * (FD(1.0) - h) - dbh * h).v;
  L.f2h = (-a2 - a6).v;
  L.f3V = (dan * (FD(1.0) - n) - dbn * n).v;
  L.f3n = (-a3 - a7).v;
  L.gna = (-m3 * h * (V - constant6)).v;
  L.gk = (-n4 * (V - constant7)).v;
  L.gl = (-(V - constant8)).v;
  return L;
}

// reverse of one Heun step from its input y (no clamp):
// lam[4] in: adjoint of the output, out: adjoint of y; gacc[3] += parameter gradient
template <class E>
__device__ __forceinline__ void hh_heun_step_reverse(const double *y, const double *p, double dt, double *lam, double *gacc) {
  const HHLin Ly = hh_linearize<E>(y, p);
  double ys[4];
  for (int k = 0; k < 4; k++) ys[k] = y[k] + dt * Ly.r[k];
  const HHLin Ls = hh_linearize<E>(ys, p);
  double mu[4], gs[3], nu[4], jy[4], g3[3];
  Ls.jt(lam, mu);
  Ls.gp(lam, gs);
  for (int k = 0; k < 4; k++) nu[k] = lam[k] + dt * mu[k];
  Ly.jt(nu, jy);
  Ly.gp(nu, g3);
  for (int k = 0; k < 4; k++) lam[k] = lam[k] + 0.5 * dt * mu[k] + 0.5 * dt * jy[k];
  for (int k = 0; k < 3; k++) gacc[k] += 0.5 * dt * (g3[k] + gs[k]);
}

// current-pulse stimulus parameter of one Heun step, as the forward sets it
// (the pulse is a constant of each Heun step; fms_fused.h fused_node_half).
// Selects per element, not a store at a run-time index, so pj stays in registers.
template <class E>
FMS_HD void stim_params(const DevStim &st, int f, bool inset, const double *p, double t, double *pj) {
  const bool pulse = st.mode == 2 && current_on(st, f, t) && inset;
#pragma unroll
  for (int k = 0; k < 4; k++) pj[k] = pulse && k == st.stim_param ? st.amplitude[f] : p[k];
}

// reverse of one reaction half of a thread's NPT nodes, each from its state
// y0[q] at t0: l[q] in: adjoint of the half's output, out: adjoint of y0[q];
// g[q] += dL/d(gNa, gK, gL). Heun step j is reversed from its recomputed input
// with the stimulus parameter of that step (a per-lane select), so all lanes
// take the same path. The node loop is the innermost one: the NPT nodes'
// independent linearisations then share one basic block, which the scheduler
// interleaves (a node loop around the run-time loop over the Heun steps gives
// one block per node and a single dependency chain at a time).
template <class E, int NPT>
__device__ __forceinline__ void reverse_half(const DevStim &st, int f, const bool *inset, const double (*y0)[4],
                                             const double *p, double t0, double dt, int n0D, double (*l)[4],
                                             double (*g)[3]) {
  for (int j = n0D - 1; j >= 0; j--) {
    double y[NPT][4];
#pragma unroll
    for (int q = 0; q < NPT; q++)
#pragma unroll
      for (int k = 0; k < 4; k++) y[q][k] = y0[q][k];
    for (int jj = 0; jj < j; jj++) {
#pragma unroll
      for (int q = 0; q < NPT; q++) {
        double pj[4];
        stim_params<E>(st, f, inset[q], p, t0 + jj * dt, pj);
        heun_node<ModelHH, double, E, false>(y[q][0], y[q] + 1, pj, t0 + jj * dt, dt, false, 0.0);
      }
    }
#pragma unroll
    for (int q = 0; q < NPT; q++) {
      double pj[4];
      stim_params<E>(st, f, inset[q], p, t0 + j * dt, pj);
      hh_heun_step_reverse<E>(y[q], pj, dt, l[q], g[q]);
    }
  }
}

// Reverse of kc split steps of one segment, one workgroup per fibre, steps
// s = kc-1 .. 0 (rows of `steps`, 4 doubles each). Per step: the sensors'
// dL/dV, reverse of half 2 from y^b = (V^b, gates after half 1), reverse
// diffusion w = A^{-1} lambda_V, lambda_V = B w (A, B symmetric),
// pf_bar += w^T (dB/dpf V^a - dA/dpf V^b), reverse of half 1 from y0. lambda and
// the per-node parameter gradients stay in registers for the whole segment.
// Tape: [kc][9][N] = y0 (4), y after half 1 (V^a and the gates, 4), V^b.
template <class E, int T, int NPT, int MMAX>
__global__ __launch_bounds__(T) void k_adj_fused(DevFibers fb, DevStim st, const double *__restrict__ tape,
                                                 const double *__restrict__ steps, int kc, int n0D, int ie,
                                                 const int *__restrict__ sensor_slot,
                                                 const double *__restrict__ dtrace, int n_sensors, double *lam,
                                                 double *gacc, double *pf_bar, int raise_priority) {
  extern __shared__ double sh[];
  // with raise_priority the reverse sweep's waves win issue over the next
  // segment's recompute that runs beside it on the same CUs
  if (raise_priority) __builtin_amdgcn_s_setprio(2);
  const int f = blockIdx.x;
  const int tid = threadIdx.x;
  const long long N = fb.n_total;
  const int off = fb.fiber_off[f];
  const int n = fb.fiber_off[f + 1] - off;
  const double pf = fb.prefactor[f];
  double *red = sh;                          // partitioned solve's block interfaces (pcr_lds_doubles)
  double *lw = sh + pcr_lds_doubles<T>();    // n: lambda_V, then w
  double *lv = lw + n;       // n: the diffusion's V (IE: V^b, CN: V^a + V^b), for the neighbours
  double *h = lv + n;        // n - 1 element lengths, staged once (read after the first step's barrier)
  for (int i = tid; i < n - 1; i += T) h[i] = fb.elem_len[off - f + i];
  double p[4];
#pragma unroll
  for (int k = 0; k < 4; k++) p[k] = fb.params[(long long)f * 4 + k];
  const int sb = fb.stim_begin[f], se = fb.stim_end[f];
  double l[NPT][4], g[NPT][3];
  int slot[NPT];
  bool inset[NPT];
  double pfacc = 0;
  // the fibre's running prefactor gradient, held by thread 0 for the segment:
  // the same additions in the same order as adding to pf_bar[f] every step
  double pftotal = tid == 0 ? pf_bar[f] : 0.0;
  // lanes past the fibre's end shadow its last node: they compute, never store
#pragma unroll
  for (int q = 0; q < NPT; q++) {
    const int i = tid + q * T < n ? tid + q * T : n - 1;
#pragma unroll
    for (int k = 0; k < 4; k++) l[q][k] = lam[k * N + off + i];
#pragma unroll
    for (int k = 0; k < 3; k++) g[q][k] = gacc[k * N + off + i];
    slot[q] = sensor_slot[off + i];
    inset[q] = i >= sb && i < se;
  }
  for (int s = kc - 1; s >= 0; s--) {
    const double *row = steps + 4 * s;
    const double t = row[0], dts = row[1], dt0 = row[2], dt1 = row[3];
    // all of the step's tape entries and sensor adjoints are loaded up front
    // (loading them a step ahead instead measured 1.9x slower: the extra live
    // registers push the linearisation into AGPRs)
    const double *tp = tape + (long long)s * 9 * N;
    double yb[NPT][4], y0[NPT][4], va[NPT], dl[NPT];
#pragma unroll
    for (int q = 0; q < NPT; q++) {
      const long long node = off + (tid + q * T < n ? tid + q * T : n - 1);
      yb[q][0] = tp[8 * N + node];
      yb[q][1] = tp[5 * N + node];
      yb[q][2] = tp[6 * N + node];
      yb[q][3] = tp[7 * N + node];
      y0[q][0] = tp[node];
      y0[q][1] = tp[N + node];
      y0[q][2] = tp[2 * N + node];
      y0[q][3] = tp[3 * N + node];
      va[q] = tp[4 * N + node];
      dl[q] = dtrace[((long long)s * fb.n_fibers + f) * n_sensors + (slot[q] >= 0 ? slot[q] : 0)];
    }
    // sensors, reverse of half 2 from y^b (V after the diffusion, gates after half 1)
#pragma unroll
    for (int q = 0; q < NPT; q++) {
      const bool valid = tid + q * T < n;
      const int i = valid ? tid + q * T : n - 1;
      l[q][0] += slot[q] >= 0 ? dl[q] : 0.0;
      if (valid) lv[i] = ie ? yb[q][0] : va[q] + yb[q][0];
    }
    reverse_half<E, NPT>(st, f, inset, yb, p, t + 0.5 * dts, dt0, n0D, l, g);
#pragma unroll
    for (int q = 0; q < NPT; q++)
      if (tid + q * T < n) lw[tid + q * T] = l[q][0];
    __syncthreads();
    // reverse diffusion
    lds_thomas_part<T, MMAX, true>(lw, n, h, pf, dt1, ie != 0, red, tid);
#pragma unroll
    for (int q = 0; q < NPT; q++) {
      const bool valid = tid + q * T < n;
      const int i = valid ? tid + q * T : n - 1;
      double a, b, c, d;
      assemble_row(i, n, lw, h, pf, dt1, ie != 0, a, b, c, d);   // d = (B w)_i
      const double vi = lv[i];
      double k1v = 0;
      if (i > 0) k1v += (lv[i - 1] - vi) / h[i - 1];
      if (i < n - 1) k1v += (lv[i + 1] - vi) / h[i];
      pfacc += valid ? lw[i] * (ie ? -k1v : -0.5 * k1v) : 0.0;
      l[q][0] = d;
    }
    // this step's prefactor gradient, summed in a fixed order (butterfly within
    // each wave, then the waves in order) and added to the fibre's total step by
    // step, so the result does not depend on the checkpoint segmentation
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) pfacc += __shfl_xor(pfacc, o, 64);
    if ((tid & 63) == 0) red[tid >> 6] = pfacc;
    pfacc = 0;
    __syncthreads();
    if (tid == 0) {
      double sum = red[0];
#pragma unroll
      for (int q = 1; q < T / 64; q++) sum += red[q];
      pftotal += sum;
    }
    // (every read of lw and lv above precedes that barrier, and the next step
    // writes red only after the barrier that follows its first loop, so thread
    // 0's read of red needs no barrier of its own)
    // reverse of half 1
    reverse_half<E, NPT>(st, f, inset, y0, p, t, dt0, n0D, l, g);
  }
  if (tid == 0) pf_bar[f] = pftotal;
#pragma unroll
  for (int q = 0; q < NPT; q++) {
    const int i = tid + q * T;
    if (i < n) {
#pragma unroll
      for (int k = 0; k < 4; k++) lam[k * N + off + i] = l[q][k];
#pragma unroll
      for (int k = 0; k < 3; k++) gacc[k * N + off + i] = g[q][k];
    }
  }
}

// per-fibre sums of the per-node parameter-gradient accumulators
__global__ void k_adj_reduce(DevFibers fb, const double *gacc, double *grad) {
  const int f = blockIdx.x;
  const int off = fb.fiber_off[f], n = fb.fiber_off[f + 1] - off;
  __shared__ double part[3][256];
  double s0 = 0, s1 = 0, s2 = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    s0 += gacc[off + i];
    s1 += gacc[fb.n_total + off + i];
    s2 += gacc[2 * fb.n_total + off + i];
  }
  part[0][threadIdx.x] = s0;
  part[1][threadIdx.x] = s1;
  part[2][threadIdx.x] = s2;
  __syncthreads();
  if (threadIdx.x < 3) {
    double s = 0;
    for (int q = 0; q < (int)blockDim.x; q++) s += part[threadIdx.x][q];
    grad[f * 3 + threadIdx.x] = s;
  }
}

}  // namespace fms
