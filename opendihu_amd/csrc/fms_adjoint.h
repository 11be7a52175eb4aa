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

template <class E> struct ExpDeriv;
template <> struct ExpDeriv<ExpApprox> {
  static FMS_HD double d(double x, double ex) { return ex / (1.0 + x / 1024.); }
};
template <> struct ExpDeriv<ExpLibm> {
  static FMS_HD double d(double x, double ex) { (void)x; return ex; }
};

// J(y)^T w (out[0..3]) and f_theta(y)^T w for theta = (g_Na, g_K, g_L) (gp[0..2])
template <class E>
FMS_HD void hh_vjp(const double *y, const double *p, const double *w, double *out, double *gp) {
  const double V = y[0], m = y[1], h = y[2], n = y[3];
  const double gna = p[1], gk = p[2], gl = p[3];
  const double ENa = -75 + 115.000, EK = -75 - 12.0000, EL = -75 + 10.6130, Cm = 1;
  // rate functions and their V-derivatives, as the model file writes them
  const double xm = -(V + 50.0000) / 10.0000, exm = E::exp(xm), dexm = ExpDeriv<E>::d(xm, exm) * (-0.1);
  const double num_m = -0.100000 * (V + 50.0000), den_m = exm - 1.00000;
  const double am = num_m / den_m, dam = (-0.1 * den_m - num_m * dexm) / (den_m * den_m);
  const double xbm = -(V + 75.0000) / 18.0000, exbm = E::exp(xbm);
  const double bm = 4.00000 * exbm, dbm = 4.00000 * ExpDeriv<E>::d(xbm, exbm) * (-1.0 / 18.0);
  const double xah = -(V + 75.0000) / 20.0000, exah = E::exp(xah);
  const double ah = 0.0700000 * exah, dah = 0.0700000 * ExpDeriv<E>::d(xah, exah) * (-1.0 / 20.0);
  const double xbh = -(V + 45.0000) / 10.0000, exbh = E::exp(xbh);
  const double bh = 1.00000 / (exbh + 1.00000);
  const double dbh = -(ExpDeriv<E>::d(xbh, exbh) * (-0.1)) / ((exbh + 1.0) * (exbh + 1.0));
  const double xn = -(V + 65.0000) / 10.0000, exn = E::exp(xn), dexn = ExpDeriv<E>::d(xn, exn) * (-0.1);
  const double num_n = -0.0100000 * (V + 65.0000), den_n = exn - 1.00000;
  const double an = num_n / den_n, dan = (-0.01 * den_n - num_n * dexn) / (den_n * den_n);
  const double xbn = (V + 75.0000) / 80.0000, exbn = E::exp(xbn);
  const double bn = 0.125000 * exbn, dbn = 0.125000 * ExpDeriv<E>::d(xbn, exbn) / 80.0;
  const double m2 = m * m, m3 = m2 * m, n3 = n * n * n, n4 = n3 * n;
  const double df0dV = -(gna * m3 * h + gk * n4 + gl) / Cm;
  const double df0dm = -(gna * 3.0 * m2 * h * (V - ENa)) / Cm;
  const double df0dh = -(gna * m3 * (V - ENa)) / Cm;
  const double df0dn = -(gk * 4.0 * n3 * (V - EK)) / Cm;
  out[0] = w[0] * df0dV + w[1] * (dam * (1.0 - m) - dbm * m) + w[2] * (dah * (1.0 - h) - dbh * h) +
           w[3] * (dan * (1.0 - n) - dbn * n);
  out[1] = w[0] * df0dm + w[1] * (-am - bm);
  out[2] = w[0] * df0dh + w[2] * (-ah - bh);
  out[3] = w[0] * df0dn + w[3] * (-an - bn);
  gp[0] = w[0] * (-m3 * h * (V - ENa) / Cm);
  gp[1] = w[0] * (-n4 * (V - EK) / Cm);
  gp[2] = w[0] * (-(V - EL) / Cm);
}

// reverse of n0D Heun steps of one node starting from state y0 at t0;
// lam[4] in: adjoint of the output, out: adjoint of y0; gacc[3] += parameter gradient
template <class E>
FMS_HD void hh_heun_reverse(const double *y0, const double *p, double t0, double dt, int n0D, double *lam,
                            double *gacc) {
  for (int j = n0D - 1; j >= 0; j--) {
    double y[4] = {y0[0], y0[1], y0[2], y0[3]};
    double g3[3];
    for (int q = 0; q < j; q++) heun_node<ModelHH, double, E, false>(y[0], y + 1, p, t0 + q * dt, dt, false, 0.0);
    double r[4], a[9], ys[4];
    ModelHH::rhs<double, E>(y, p, t0 + j * dt, r, a);
    for (int k = 0; k < 4; k++) ys[k] = y[k] + dt * r[k];
    double mu[4], gs[3];
    hh_vjp<E>(ys, p, lam, mu, gs);
    double nu[4];
    for (int k = 0; k < 4; k++) nu[k] = lam[k] + dt * mu[k];
    double jy[4];
    hh_vjp<E>(y, p, nu, jy, g3);
    for (int k = 0; k < 4; k++) lam[k] = lam[k] + 0.5 * dt * mu[k] + 0.5 * dt * jy[k];
    for (int k = 0; k < 3; k++) gacc[k] += 0.5 * dt * (g3[k] + gs[k]);
  }
}

// current-pulse stimulus parameter of one Heun step, as the forward sets it
// (the pulse is a constant of each Heun step; fms_fused.h fused_node_half)
template <class E>
FMS_HD void stim_params(const DevStim &st, int f, bool inset, const double *p, double t, double *pj) {
  for (int k = 0; k < 4; k++) pj[k] = p[k];
  if (st.mode == 2) pj[st.stim_param] = (current_on(st, f, t) && inset) ? st.amplitude[f] : p[st.stim_param];
}

// forward reaction half of one node (n0D Heun steps from t0), the forward's arithmetic
template <class E>
__device__ __forceinline__ void forward_half(const DevStim &st, int f, bool inset, double *y, const double *p,
                                             double t0, double dt, int n0D) {
  for (int j = 0; j < n0D; j++) {
    double pj[4];
    stim_params<E>(st, f, inset, p, t0 + j * dt, pj);
    heun_node<ModelHH, double, E, false>(y[0], y + 1, pj, t0 + j * dt, dt, false, 0.0);
  }
}

// reverse of one reaction half of one node starting from state y0 at t0:
// l[4] in: adjoint of the half's output, out: adjoint of y0; g[3] += dL/d(gNa, gK, gL).
// Heun step j is reversed from its recomputed input with the stimulus
// parameter of that step (a per-lane select), so all lanes take the same path.
template <class E>
__device__ __forceinline__ void reverse_half(const DevStim &st, int f, bool inset, const double *y0, const double *p,
                                             double t0, double dt, int n0D, double *l, double *g) {
  for (int j = n0D - 1; j >= 0; j--) {
    double y[4] = {y0[0], y0[1], y0[2], y0[3]};
    forward_half<E>(st, f, inset, y, p, t0, dt, j);
    double pj[4];
    stim_params<E>(st, f, inset, p, t0 + j * dt, pj);
    hh_heun_reverse<E>(y, pj, t0 + j * dt, dt, 1, l, g);
  }
}

// Reverse of kc split steps of one segment, one workgroup per fibre, steps
// s = kc-1 .. 0 (rows of `steps`, 4 doubles each). Per step: the sensors'
// dL/dV, reverse of half 2 from y^b = (V^b, gates after half 1; the gates are
// recomputed from the tape's y0), reverse diffusion w = A^{-1} lambda_V,
// lambda_V = B w (A, B symmetric), pf_bar += w^T (dB/dpf V^a - dA/dpf V^b),
// reverse of half 1 from y0. lambda and the per-node parameter gradients stay
// in registers for the whole segment. Tape: [kc][6][N] = y0 (4), V^a, V^b.
template <class E, int T, int NPT, int MMAX>
__global__ __launch_bounds__(T) void k_adj_fused(DevFibers fb, DevStim st, const double *__restrict__ tape,
                                                 const double *__restrict__ steps, int kc, int n0D, int ie,
                                                 const int *__restrict__ sensor_slot,
                                                 const double *__restrict__ dtrace, int n_sensors, double *lam,
                                                 double *gacc, double *pf_bar) {
  extern __shared__ double sh[];
  const int f = blockIdx.x;
  const int tid = threadIdx.x;
  const long long N = fb.n_total;
  const int off = fb.fiber_off[f];
  const int n = fb.fiber_off[f + 1] - off;
  const double *h = fb.elem_len + off - f;
  const double pf = fb.prefactor[f];
  double *red = sh;          // 16*T: partitioned solve
  double *lw = sh + 16 * T;  // n: lambda_V, then w
  double p[4];
  for (int k = 0; k < 4; k++) p[k] = fb.params[(long long)f * 4 + k];
  const int sb = fb.stim_begin[f], se = fb.stim_end[f];
  double l[NPT][4], g[NPT][3];
  double pfacc = 0;
  // lanes past the fibre's end shadow its last node: they compute, never store
#pragma unroll
  for (int q = 0; q < NPT; q++) {
    const int i = tid + q * T < n ? tid + q * T : n - 1;
#pragma unroll
    for (int k = 0; k < 4; k++) l[q][k] = lam[k * N + off + i];
#pragma unroll
    for (int k = 0; k < 3; k++) g[q][k] = gacc[k * N + off + i];
  }
  for (int s = kc - 1; s >= 0; s--) {
    const double *row = steps + 4 * s;
    const double t = row[0], dts = row[1], dt0 = row[2], dt1 = row[3];
    const double *tp = tape + (long long)s * 6 * N;
    // sensors, reverse of half 2
#pragma unroll
    for (int q = 0; q < NPT; q++) {
      const bool valid = tid + q * T < n;
      const int i = valid ? tid + q * T : n - 1;
      const long long node = off + i;
      const bool inset = i >= sb && i < se;
      const int slot = sensor_slot[node];
      const double dl = dtrace[((long long)s * fb.n_fibers + f) * n_sensors + (slot >= 0 ? slot : 0)];
      l[q][0] += slot >= 0 ? dl : 0.0;
      double y[4] = {tp[node], tp[N + node], tp[2 * N + node], tp[3 * N + node]};
      forward_half<E>(st, f, inset, y, p, t, dt0, n0D);
      y[0] = tp[5 * N + node];
      reverse_half<E>(st, f, inset, y, p, t + 0.5 * dts, dt0, n0D, l[q], g[q]);
      if (valid) lw[i] = l[q][0];
    }
    __syncthreads();
    // reverse diffusion
    lds_thomas_part<T, MMAX, true>(lw, n, h, pf, dt1, ie != 0, red, tid);
    const double *Va = tp + 4 * N + off, *Vb = tp + 5 * N + off;
#pragma unroll
    for (int q = 0; q < NPT; q++) {
      const bool valid = tid + q * T < n;
      const int i = valid ? tid + q * T : n - 1;
      double a, b, c, d;
      assemble_row(i, n, lw, h, pf, dt1, ie != 0, a, b, c, d);   // d = (B w)_i
      const double vi = ie ? Vb[i] : Va[i] + Vb[i];
      double k1v = 0;
      if (i > 0) k1v += ((ie ? Vb[i - 1] : Va[i - 1] + Vb[i - 1]) - vi) / h[i - 1];
      if (i < n - 1) k1v += ((ie ? Vb[i + 1] : Va[i + 1] + Vb[i + 1]) - vi) / h[i];
      pfacc += valid ? lw[i] * (ie ? -k1v : -0.5 * k1v) : 0.0;
      l[q][0] = d;
    }
    // this step's prefactor gradient, summed in a fixed order and added to the
    // fibre's total step by step (independent of the checkpoint segmentation)
    red[tid] = pfacc;
    pfacc = 0;
    __syncthreads();
    if (tid == 0) {
      double sum = 0;
      for (int q = 0; q < T; q++) sum += red[q];
      pf_bar[f] += sum;
    }
    __syncthreads();
    // reverse of half 1
#pragma unroll
    for (int q = 0; q < NPT; q++) {
      const int i = tid + q * T < n ? tid + q * T : n - 1;
      const long long node = off + i;
      const bool inset = i >= sb && i < se;
      const double y0[4] = {tp[node], tp[N + node], tp[2 * N + node], tp[3 * N + node]};
      reverse_half<E>(st, f, inset, y0, p, t, dt0, n0D, l[q], g[q]);
    }
  }
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
