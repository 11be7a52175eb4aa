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

struct DevTape {
  const double *tape;     // [K][9][N]: y at the start of half 1 (4), V before diffusion (1), y after it (4)
  long long N;
};

// reverse of one reaction half for all nodes: which = 0 (first half, from tape y0)
// or 1 (second half, from tape y^b)
template <class E>
__global__ __launch_bounds__(256) void k_adj_heun(DevFibers fb, DevStim st, DevTape tp, int s, int which, double t0,
                                                  double dt, int n0D, double *lam, double *gacc) {
  const long long N = fb.n_total;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int f = fb.node_fiber[i];
  const int li = (int)(i - fb.fiber_off[f]);
  const double *base = tp.tape + (long long)s * 9 * N + (which ? 5 * N : 0);
  double y0[4] = {base[i], base[N + i], base[2 * N + i], base[3 * N + i]};
  double p[4];
  for (int k = 0; k < 4; k++) p[k] = fb.params[(long long)f * 4 + k];
  // current pulses: the parameter during the pulse (a constant of each Heun step)
  const bool inset = li >= fb.stim_begin[f] && li < fb.stim_end[f];
  double l[4] = {lam[i], lam[N + i], lam[2 * N + i], lam[3 * N + i]};
  double g[3] = {0, 0, 0};
  if (st.mode == 2 && inset) {
    // pulses switch within a half step only at Heun-step boundaries: reverse step by step
    for (int j = n0D - 1; j >= 0; j--) {
      double y[4] = {y0[0], y0[1], y0[2], y0[3]};
      double pj[4] = {p[0], p[1], p[2], p[3]};
      for (int q = 0; q < j; q++) {
        double pq[4] = {p[0], p[1], p[2], p[3]};
        pq[st.stim_param] = current_on(st, f, t0 + q * dt) ? st.amplitude[f] : p[st.stim_param];
        heun_node<ModelHH, double, E, false>(y[0], y + 1, pq, t0 + q * dt, dt, false, 0.0);
      }
      pj[st.stim_param] = current_on(st, f, t0 + j * dt) ? st.amplitude[f] : p[st.stim_param];
      hh_heun_reverse<E>(y, pj, t0 + j * dt, dt, 1, l, g);
    }
  } else {
    hh_heun_reverse<E>(y0, p, t0, dt, n0D, l, g);
  }
  lam[i] = l[0]; lam[N + i] = l[1]; lam[2 * N + i] = l[2]; lam[3 * N + i] = l[3];
  gacc[i] += g[0]; gacc[N + i] += g[1]; gacc[2 * N + i] += g[2];
}

// lambda_V[sensor] += dL/dV(sensor) of split step s
__global__ void k_adj_sensors(DevFibers fb, const int *sensor_nodes, int n_sensors, const double *dtrace, int s,
                              double *lam) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= fb.n_fibers * n_sensors) return;
  const int f = idx / n_sensors;
  lam[fb.fiber_off[f] + sensor_nodes[idx]] += dtrace[(long long)s * fb.n_fibers * n_sensors + idx];
}

// reverse diffusion of one fibre per workgroup (Vm adjoint in LDS, partitioned solve)
template <int T, int MMAX>
__global__ __launch_bounds__(T) void k_adj_thomas(DevFibers fb, DevTape tp, int s, double dt, int ie, double *lam,
                                                  double *pf_bar) {
  extern __shared__ double sh[];
  const int f = blockIdx.x;
  const int tid = threadIdx.x;
  const int off = fb.fiber_off[f];
  const int n = fb.fiber_off[f + 1] - off;
  const long long N = fb.n_total;
  const double *h = fb.elem_len + off - f;
  const double pf = fb.prefactor[f];
  double *red = sh;
  double *lw = sh + 16 * T;        // n: rhs lambda_b, then w
  double *lr = lw + n;             // n: result B w
  double *part = lr + n;           // T partial sums
  for (int i = tid; i < n; i += T) lw[i] = lam[off + i];
  __syncthreads();
  lds_thomas_part<T, MMAX, true>(lw, n, h, pf, dt, ie != 0, red, tid);   // w = A^{-1} lambda_b
  const double *Va = tp.tape + (long long)s * 9 * N + 4 * N + off;
  const double *Vb = tp.tape + (long long)s * 9 * N + 5 * N + off;
  double acc = 0;
  for (int i = tid; i < n; i += T) {
    double a, b, c, d;
    assemble_row(i, n, lw, h, pf, dt, ie != 0, a, b, c, d);   // d = (B w)_i
    lr[i] = d;
    // (K1 v)_i for v = V^b (IE) or V^a + V^b (CN)
    double k1v = 0;
    const double vi = ie ? Vb[i] : Va[i] + Vb[i];
    if (i > 0) k1v += ((ie ? Vb[i - 1] : Va[i - 1] + Vb[i - 1]) - vi) / h[i - 1];
    if (i < n - 1) k1v += ((ie ? Vb[i + 1] : Va[i + 1] + Vb[i + 1]) - vi) / h[i];
    acc += lw[i] * (ie ? -k1v : -0.5 * k1v);
  }
  part[tid] = acc;
  __syncthreads();
  for (int i = tid; i < n; i += T) lam[off + i] = lr[i];
  if (tid == 0) {
    double sum = 0;
    for (int q = 0; q < T; q++) sum += part[q];
    pf_bar[f] += sum;
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
