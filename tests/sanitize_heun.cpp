// Host build of the device model code under AddressSanitizer and
// UndefinedBehaviorSanitizer: the model's right-hand side, the Heun step of the
// reaction kernels (heun_node, FP64 and FP32 gates, with and without the Vm
// clamp and the algebraics), the row assembly and the serial tridiagonal solve,
// over random states. Any out-of-bounds access, uninitialised-use pattern the
// sanitizers see, signed overflow or invalid shift aborts the program. Prints
// a checksum of the results so that nothing is optimised away.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "fms_kernels.h"
#include FMS_MODEL_HEADER

using M = FMS_MODEL_TYPE;

int main(int argc, char **argv) {
  const int n_states = argc > 1 ? atoi(argv[1]) : 2000;
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> u(0.5, 1.5), v(-90.0, 40.0);
  double y0[M::NS], p[M::NP > 0 ? M::NP : 1];
  for (int k = 0; k < M::NS; k++) y0[k] = M::y0(k);
  for (int k = 0; k < M::NP; k++) p[k] = M::p0(k);
  double sum = 0.0;
  long nonfinite = 0;
  for (int s = 0; s < n_states; s++) {
    double V = y0[0] == 0 ? v(rng) : y0[0] * u(rng);
    double g[M::NS - 1];
    float gf[M::NS - 1];
    for (int k = 1; k < M::NS; k++) {
      g[k - 1] = y0[k] == 0 ? u(rng) - 0.5 : y0[k] * u(rng);
      gf[k - 1] = (float)g[k - 1];
    }
    double Vf = V, alg[M::NA];
    const bool clamp = (s % 7) == 0;
    for (int j = 0; j < 3; j++) {
      fms::heun_node<M, double, fms::ExpApprox, true>(V, g, p, j * 1e-3, 1.25e-5, clamp, 20.0, alg);
      fms::heun_node<M, float, fms::ExpApprox, false>(Vf, gf, p, j * 1e-3, 1.25e-5, clamp, 20.0);
      fms::heun_node<M, double, fms::ExpLibm, false>(V, g, p, j * 1e-3, 1.25e-5, false, 20.0);
    }
    sum += V + Vf;
    for (int k = 0; k < M::NS - 1; k++) sum += g[k] + gf[k];
    for (int k = 0; k < M::NA; k++)
      if (!M::alg_is_param(k)) sum += std::isfinite(alg[k]) ? 0.0 : 1.0;
    nonfinite += !std::isfinite(V);
  }
  // row assembly and the serial solve on random fibres, both schemes
  for (int f = 0; f < 50; f++) {
    const int n = 2 + (int)(rng() % 3000);
    std::vector<double> vm(n), h(n - 1), cA(n), dA(n);
    for (int i = 0; i < n; i++) vm[i] = v(rng);
    for (int e = 0; e < n - 1; e++) h[e] = 0.01 * u(rng);
    fms::thomas_serial_fiber(vm.data(), n, h.data(), 0.007656, 2.5e-5, f % 2 == 0, cA.data(), dA.data());
    for (int i = 0; i < n; i++) sum += vm[i];
  }
  printf("checksum %.17g nonfinite %ld\n", sum, nonfinite);
  return 0;
}
