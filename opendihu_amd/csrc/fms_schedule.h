// RepeatedCall<FastMonodomainSolver> time-step bookkeeping on the host.
//
// The reference drives the fast solver from RepeatedCall::advanceTimeSpan
// (core/src/time_stepping_scheme/repeated_call.tpp:45-80): call k covers
// [t_k, t_k + dt_rc] with t_k = t_start + k/N * span. Each nested scheme
// recomputes its own step width for that span with the rule of
// TimeSteppingScheme::setTimeSpan / setTimeStepWidth
// (core/src/time_stepping_scheme/00_time_stepping_scheme.cpp:15-70):
// n = max(1, ceil(span/width - 0.1)), width = span/n, and the stored target
// width follows the last width. The fast solver reads Strang's width, then sets
// Heun to [t, t + dt_split/2] and the diffusion scheme to [t, t + dt_split]
// (fast_monodomain_solver_compute.tpp:75-85). Because t_k + dt - t_k is not dt
// in floating point, every call gets slightly different widths; this table
// reproduces them exactly.
#pragma once

#include <cmath>

namespace fms_host {

struct StepScheme {
  double start = 0.0, end = 1.0, target = 0.0, width = 0.0;
  int n = 1;
  bool significant = false;

  void set_number(int steps) {
    n = steps;
    width = (end - start) / n;
  }
  void set_width(double w) {
    target = w;
    const long double k = std::ceil((long double)(end - start) / w - 1e-1);
    n = k < 1 ? 1 : (int)k;
    set_number(n);
  }
  void init(double end_time, double w, int steps) {
    start = 0.0;
    end = end_time;
    if (w > 0) {
      target = w;
      set_width(w);
      significant = true;
    } else {
      significant = false;
      target = 0.0;
      set_number(steps);
    }
  }
  void set_span(double s, double e) {
    start = s;
    end = e;
    if (significant) {
      if (target > end - start) width = end - start;
      set_width(width);
    } else {
      set_number(n);
    }
  }
};

}  // namespace fms_host

extern "C" int fms_schedule(double rc_end, double rc_width, double strang_end, double strang_width,
                            double heun_end, double heun_width, double ie_end, double ie_width, int max_calls,
                            double *out) {
  FMS_API_BEGIN
  if (rc_width <= 0 || strang_width <= 0 || heun_width <= 0 || ie_width <= 0 || max_calls < 0)
    return fail(FMS_ERR_ARGUMENT, "time step widths must be positive");
  fms_host::StepScheme rc, st, he, ie;
  rc.init(rc_end, rc_width, 0);
  st.init(strang_end, strang_width, 0);
  he.init(heun_end, heun_width, 0);
  ie.init(ie_end, ie_width, 0);
  const double span = rc.end - rc.start;
  double t = rc.start;
  int calls = 0;
  for (int step = 0; step < rc.n;) {
    st.set_span(t, t + rc.width);
    const double t0 = st.start, dts = st.width;
    he.set_span(t0, t0 + 0.5 * dts);
    ie.set_span(t0, t0 + dts);
    if (calls < max_calls && out) {
      double *o = out + 7 * (long long)calls;
      o[0] = t0;
      o[1] = dts;
      o[2] = st.n;
      o[3] = he.width;
      o[4] = he.n;
      o[5] = ie.width;
      o[6] = ie.n;
    }
    calls++;
    step++;
    t = rc.start + (double)step / rc.n * span;
  }
  return calls;
  FMS_API_END
}
