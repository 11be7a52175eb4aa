// Host runtime and C ABI of the MI355X FastMonodomainSolver fibre library.
//
// One shared library per cell model: the model header is selected at compile
// time (FMS_MODEL_HEADER / FMS_MODEL_TYPE), the way the reference compiles one
// library per CellML file (fast_monodomain_solver_gpu.tpp:12-206) - here ahead
// of time with hipcc for gfx950 instead of g++ -fopenmp at run time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/opendihu_fms.h"
#include "fms_adjoint.h"
#include "fms_fused.h"
#include "fms_kernels.h"

#ifndef FMS_MODEL_HEADER
#define FMS_MODEL_HEADER "model_hh.h"
#define FMS_MODEL_TYPE fms::ModelHH
#endif
#include FMS_MODEL_HEADER

using Model = FMS_MODEL_TYPE;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      return fail(FMS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));                 \
  } while (0)

template <class T> int dalloc(T **p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  HIP_TRY(hipMalloc((void **)p, n * sizeof(T)));
  return 0;
}

template <class T> int h2d(T *d, const T *h, size_t n, hipStream_t s) {
  if (n == 0) return 0;
  HIP_TRY(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  return 0;
}

constexpr int FUSED_T = 256;
constexpr int PART_T = 256;

}  // namespace

struct FmsSolver {
  // configuration
  int F = 0;
  long long N = 0;
  int scheme = FMS_IMPLICIT_EULER, precision = FMS_FP64, exp_mode = FMS_EXP_APPROXIMATE;
  int thomas = FMS_THOMAS_PARTITIONED, kernels = FMS_KERNELS_SPLIT;
  double vstim = 20.0;
  int nmax = 0;
  std::vector<int> h_off;     // [F+1]
  std::vector<int> h_order;   // fibres longest first
  hipStream_t stream = nullptr;

  // device state
  int *d_off = nullptr, *d_node_fiber = nullptr, *d_sb = nullptr, *d_se = nullptr, *d_order = nullptr;
  double *d_h = nullptr, *d_pf = nullptr, *d_params = nullptr;
  double *d_vm = nullptr;
  void *d_gates = nullptr;    // double or float [(NS-1)*N]
  double *d_cA = nullptr, *d_dA = nullptr;  // serial Thomas scratch (split kernels)
  unsigned char *d_flags = nullptr;
  int flags_cap = 0;
  double *d_steps = nullptr;
  int steps_cap = 0;

  // stimulation
  int stim_mode = FMS_STIM_NONE;
  unsigned char *d_fe = nullptr;
  int *d_mu = nullptr, *d_joff = nullptr, *d_jcnt = nullptr, *d_jidx = nullptr;
  double *d_freq = nullptr, *d_rep = nullptr, *d_beg = nullptr, *d_jit = nullptr, *d_last = nullptr,
         *d_curjit = nullptr;
  unsigned char *d_curr = nullptr, *d_hasbeen = nullptr;
  int *d_log_count = nullptr, *d_log_fiber = nullptr;
  double *d_log_time = nullptr;
  int log_cap = 0;
  double *d_onset = nullptr, *d_dur = nullptr, *d_amp = nullptr;
  int stim_param = 0;
  int only_if_stimulated = 0;
  int fe_rows = 1, fe_cols = 1;

  // spikes
  int spikes = 0;
  double spike_th = 0.0;
  double *d_vprev = nullptr;
  int *d_sfirst = nullptr, *d_scount = nullptr;
  long long step_counter = 0;

  // schedule knobs (fms_set_tuning): steps per fused launch, gate residency
  int tune_chunk = 512;
  int tune_resident = -1;   // -1 auto, 0 stream gates from HBM, 1 keep gates in registers

  // timing: events on the solver's stream and launch counters
  hipEvent_t events[8] = {};
  long long n_launch_fused = 0, n_launch_reaction = 0, n_launch_thomas = 0;

  // sensors, traces and the adjoint (checkpoints every ckpt_interval steps, tape of one segment)
  int n_sensors = 0;
  int *d_sensor_nodes = nullptr;   // [F][S] fibre-local
  int *d_sensor_slot = nullptr;    // [N]
  double *d_trace = nullptr, *d_dtrace = nullptr;
  long long trace_cap = 0;
  double *d_ckpt = nullptr, *d_tape = nullptr, *d_lam = nullptr, *d_gacc = nullptr, *d_pfbar = nullptr,
         *d_grad = nullptr;
  long long ckpt_cap = 0, tape_cap = 0;
  std::vector<double> adj_steps;   // host copy of the last adjoint forward's step table
  int adj_n0D = 1, adj_interval = 0, adj_K = 0;

  // algebraics for transfer
  int *d_alg_idx = nullptr;
  int n_alg_idx = 0;
  double *d_alg = nullptr;

  fms::DevFibers fibers() const {
    fms::DevFibers f;
    f.n_fibers = F;
    f.n_total = N;
    f.fiber_off = d_off;
    f.node_fiber = d_node_fiber;
    f.elem_len = d_h;
    f.prefactor = d_pf;
    f.params = d_params;
    f.stim_begin = d_sb;
    f.stim_end = d_se;
    return f;
  }

  fms::DevStim stim() const {
    fms::DevStim s;
    memset(&s, 0, sizeof(s));
    s.mode = stim_mode;
    s.fe_rows = fe_rows;
    s.fe_cols = fe_cols;
    s.firing_events = d_fe;
    s.motor_unit = d_mu;
    s.freq = d_freq;
    s.repeat = d_rep;
    s.enable_begin = d_beg;
    s.jitter = d_jit;
    s.jitter_off = d_joff;
    s.jitter_cnt = d_jcnt;
    s.last_check = d_last;
    s.cur_jitter = d_curjit;
    s.jitter_idx = d_jidx;
    s.currently = d_curr;
    s.has_been = d_hasbeen;
    s.log_count = log_cap > 0 ? d_log_count : nullptr;
    s.log_time = d_log_time;
    s.log_fiber = d_log_fiber;
    s.log_cap = log_cap;
    s.onset = d_onset;
    s.duration = d_dur;
    s.amplitude = d_amp;
    s.stim_param = stim_param;
    s.only_if_stimulated = only_if_stimulated;
    return s;
  }
};

namespace {

void free_all(FmsSolver *s) {
  void *ptrs[] = {s->d_off, s->d_node_fiber, s->d_sb, s->d_se, s->d_order, s->d_h, s->d_pf, s->d_params,
                  s->d_vm, s->d_gates, s->d_cA, s->d_dA, s->d_flags, s->d_steps, s->d_fe, s->d_mu,
                  s->d_joff, s->d_jcnt, s->d_jidx, s->d_freq, s->d_rep, s->d_beg, s->d_jit, s->d_last,
                  s->d_curjit, s->d_curr, s->d_hasbeen, s->d_log_count, s->d_log_fiber, s->d_log_time,
                  s->d_onset, s->d_dur, s->d_amp, s->d_vprev, s->d_sfirst, s->d_scount, s->d_alg_idx,
                  s->d_alg, s->d_sensor_nodes, s->d_sensor_slot, s->d_trace, s->d_dtrace, s->d_ckpt, s->d_tape,
                  s->d_lam, s->d_gacc, s->d_pfbar, s->d_grad};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (hipEvent_t e : s->events)
    if (e) (void)hipEventDestroy(e);
  if (s->stream) (void)hipStreamDestroy(s->stream);
}

// ---------------------------------------------------------------------------
// launchers, templated over gate type and exp policy
// ---------------------------------------------------------------------------

template <class G, class E>
int launch_reaction(FmsSolver *s, double t0, double dt, int n0D, bool store_alg) {
  const int threads = 256;
  const long long blocks = (s->N + threads - 1) / threads;
  hipLaunchKernelGGL((fms::k_reaction<Model, G, E>), dim3((unsigned)blocks), dim3(threads), 0, s->stream,
                     s->fibers(), s->stim(), s->d_vm, (G *)s->d_gates, s->d_flags, t0, dt, n0D, s->vstim,
                     store_alg ? s->d_alg : nullptr, s->d_alg_idx, s->n_alg_idx);
  HIP_TRY(hipGetLastError());
  s->n_launch_reaction++;
  return 0;
}

int launch_flags(FmsSolver *s, double t0, double dt, int n0D) {
  const int threads = 64;
  hipLaunchKernelGGL(fms::k_stim_flags, dim3((s->F + threads - 1) / threads), dim3(threads), 0, s->stream,
                     s->stim(), s->F, s->d_flags, t0, dt, n0D);
  HIP_TRY(hipGetLastError());
  return 0;
}

int launch_thomas(FmsSolver *s, double dt1) {
  const int ie = s->scheme == FMS_IMPLICIT_EULER;
  if (s->thomas == FMS_THOMAS_SERIAL) {
    const int threads = 64;
    hipLaunchKernelGGL(fms::k_thomas_serial, dim3((s->F + threads - 1) / threads), dim3(threads), 0, s->stream,
                       s->fibers(), s->d_vm, s->d_cA, s->d_dA, dt1, ie);
  } else {
    const size_t lds = (size_t)(16 * PART_T + s->nmax) * sizeof(double);
    hipLaunchKernelGGL((fms::k_thomas_part<PART_T, 16>), dim3(s->F), dim3(PART_T), lds, s->stream, s->fibers(),
                       s->d_vm, dt1, ie);
  }
  HIP_TRY(hipGetLastError());
  s->n_launch_thomas++;
  return 0;
}

int launch_spikes(FmsSolver *s) {
  const int threads = 256;
  const long long blocks = (s->N + threads - 1) / threads;
  hipLaunchKernelGGL(fms::k_spikes, dim3((unsigned)blocks), dim3(threads), 0, s->stream, s->N, s->d_vm,
                     s->d_vprev, s->d_sfirst, s->d_scount, s->spike_th, (int)s->step_counter);
  HIP_TRY(hipGetLastError());
  return 0;
}

template <class G, class E>
int run_split(FmsSolver *s, const double *steps, int K, int n0D) {
  int rc;
  for (int k = 0; k < K; k++) {
    const double t = steps[4 * k], dts = steps[4 * k + 1], dt0 = steps[4 * k + 2], dt1 = steps[4 * k + 3];
    const double tmid = t + 0.5 * dts;
    if ((rc = launch_flags(s, t, dt0, n0D))) return rc;
    if ((rc = launch_reaction<G, E>(s, t, dt0, n0D, false))) return rc;
    if ((rc = launch_thomas(s, dt1))) return rc;
    if ((rc = launch_flags(s, tmid, dt0, n0D))) return rc;
    if ((rc = launch_reaction<G, E>(s, tmid, dt0, n0D, s->d_alg && k == K - 1))) return rc;
    s->step_counter++;
    if (s->spikes && (rc = launch_spikes(s))) return rc;
  }
  return 0;
}

template <class G, class E, int NPT, int MMAX>
int launch_fused_npt(FmsSolver *s, const double *d_steps, int K, int n0D, const fms::DevTrace &tr) {
  constexpr int T = FUSED_T;
  fms::DevSpikes spk;
  spk.enabled = s->spikes;
  spk.threshold = s->spike_th;
  spk.vm_prev = s->d_vprev;
  spk.first_step = s->d_sfirst;
  spk.count = s->d_scount;
  spk.step0 = (int)s->step_counter;
  const size_t lds = (size_t)(16 * T + s->nmax) * sizeof(double) + 2 * (size_t)n0D + 16;
  hipLaunchKernelGGL((fms::k_fused<Model, G, E, T, MMAX, NPT>), dim3(s->F), dim3(T), lds, s->stream, s->fibers(),
                     s->stim(), s->d_vm, (G *)s->d_gates, d_steps, K, n0D, s->vstim,
                     s->scheme == FMS_IMPLICIT_EULER ? 1 : 0, s->thomas == FMS_THOMAS_SERIAL ? 1 : 0, s->d_order,
                     spk, tr);
  HIP_TRY(hipGetLastError());
  s->step_counter += K;
  s->n_launch_fused++;
  return 0;
}

// gates resident in registers when the whole fibre fits the workgroup's
// registers (small models), streamed from HBM otherwise; the partitioned
// solve needs ceil(n / min(T, n/2)) <= MMAX rows per thread
template <class G, class E>
int launch_fused(FmsSolver *s, const double *d_steps, int K, int n0D, const fms::DevTrace &tr = fms::DevTrace{}) {
  constexpr int T = FUSED_T;
  const int npt = (s->nmax + T - 1) / T;
  if constexpr (Model::NS <= 8) {
    const bool resident = s->tune_resident != 0;
    if (resident && npt <= 4) return launch_fused_npt<G, E, 4, 8>(s, d_steps, K, n0D, tr);
    if (resident && npt <= 8) return launch_fused_npt<G, E, 8, 8>(s, d_steps, K, n0D, tr);
  }
  if (npt <= 8) return launch_fused_npt<G, E, 0, 8>(s, d_steps, K, n0D, tr);
  return launch_fused_npt<G, E, 0, 16>(s, d_steps, K, n0D, tr);
}

template <class G, class E>
int run(FmsSolver *s, const double *h_steps, int K, int n0D, bool fused_only);

template <class G, class E>
int run(FmsSolver *s, const double *h_steps, int K, int n0D) {
  if (s->kernels == FMS_KERNELS_SPLIT) return run_split<G, E>(s, h_steps, K, n0D);
  // the fused kernel does not return algebraics: with algebraicsForTransfer
  // requested, the last step of the call runs on the split kernels (same
  // arithmetic, bit-identical states)
  if (s->d_alg) {
    int rc = 0;
    if (K > 1 && (rc = run<G, E>(s, h_steps, K - 1, n0D, true))) return rc;
    if (s->thomas == FMS_THOMAS_SERIAL && !s->d_cA) {
      if (dalloc(&s->d_cA, (size_t)s->N) || dalloc(&s->d_dA, (size_t)s->N)) return FMS_ERR_HIP;
    }
    return run_split<G, E>(s, h_steps + 4 * (size_t)(K - 1), 1, n0D);
  }
  return run<G, E>(s, h_steps, K, n0D, true);
}

template <class G, class E>
int run(FmsSolver *s, const double *h_steps, int K, int n0D, bool fused_only) {
  (void)fused_only;
  // fused: upload the step table (4 doubles + one look-ahead record) and launch in chunks
  const int chunk = s->tune_chunk;
  if (s->steps_cap < 4 * (K + 1)) {
    if (s->d_steps) HIP_TRY(hipFree(s->d_steps));
    s->steps_cap = 4 * (K + 1);
    if (dalloc(&s->d_steps, (size_t)s->steps_cap)) return FMS_ERR_HIP;
  }
  std::vector<double> tab(4 * (size_t)(K + 1));
  memcpy(tab.data(), h_steps, sizeof(double) * 4 * K);
  for (int q = 0; q < 4; q++) tab[4 * (size_t)K + q] = h_steps[4 * (size_t)(K - 1) + q];
  HIP_TRY(hipMemcpyAsync(s->d_steps, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
  // the fused kernel reads step s+1's record for the first half of step s+1;
  // a launch covers steps [k0, k0+kc) and then finishes the last step itself
  int rc;
  for (int k0 = 0; k0 < K; k0 += chunk) {
    const int kc = std::min(chunk, K - k0);
    if ((rc = launch_fused<G, E>(s, s->d_steps + 4 * (size_t)k0, kc, n0D))) return rc;
  }
  return 0;
}

int dispatch(FmsSolver *s, const double *h_steps, int K, int n0D) {
  if (K <= 0) return 0;
  if (n0D < 1) return fail(FMS_ERR_ARGUMENT, "n0D must be >= 1");
  if (s->flags_cap < s->F * n0D) {
    if (s->d_flags) HIP_TRY(hipFree(s->d_flags));
    s->flags_cap = s->F * n0D;
    if (dalloc(&s->d_flags, (size_t)s->flags_cap)) return FMS_ERR_HIP;
  }
  if (s->precision == FMS_FP64) {
    if (s->exp_mode == FMS_EXP_APPROXIMATE) return run<double, fms::ExpApprox>(s, h_steps, K, n0D);
    return run<double, fms::ExpLibm>(s, h_steps, K, n0D);
  }
  if (s->exp_mode == FMS_EXP_APPROXIMATE) return run<float, fms::ExpApprox>(s, h_steps, K, n0D);
  return run<float, fms::ExpLibm>(s, h_steps, K, n0D);
}

size_t gate_bytes(const FmsSolver *s) { return s->precision == FMS_FP64 ? sizeof(double) : sizeof(float); }

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" {

int fms_last_error(char *buf, int len) {
  if (!buf || len <= 0) return FMS_ERR_ARGUMENT;
  snprintf(buf, (size_t)len, "%s", g_last_error.c_str());
  return 0;
}

int fms_model_info(int *n_states, int *n_algebraics, int *n_params, char *name, int name_len) {
  if (n_states) *n_states = Model::NS;
  if (n_algebraics) *n_algebraics = Model::NA;
  if (n_params) *n_params = Model::NP;
  if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s", Model::name());
  return 0;
}

int fms_model_ops(double *ops_rhs, double *divs_rhs, double *exps_rhs) {
  if (ops_rhs) *ops_rhs = Model::OPS_RHS;
  if (divs_rhs) *divs_rhs = Model::DIVS_RHS;
  if (exps_rhs) *exps_rhs = Model::EXPS_RHS;
  return 0;
}

int fms_model_defaults(double *states, double *params) {
  if (states)
    for (int k = 0; k < Model::NS; k++) states[k] = Model::y0(k);
  if (params)
    for (int k = 0; k < Model::NP; k++) params[k] = Model::p0(k);
  return 0;
}

int fms_create(const FmsConfig *c, FmsSolver **out) {
  if (!c || !out) return fail(FMS_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  if (c->n_fibers <= 0 || !c->n_nodes || !c->element_lengths || !c->prefactor)
    return fail(FMS_ERR_ARGUMENT, "n_fibers, n_nodes, element_lengths and prefactor are required");
  if (c->scheme != FMS_IMPLICIT_EULER && c->scheme != FMS_CRANK_NICOLSON)
    return fail(FMS_ERR_ARGUMENT, "scheme must be FMS_IMPLICIT_EULER or FMS_CRANK_NICOLSON");
  if (c->precision != FMS_FP64 && c->precision != FMS_MIXED) return fail(FMS_ERR_ARGUMENT, "bad precision");
  if (c->exp_mode != FMS_EXP_APPROXIMATE && c->exp_mode != FMS_EXP_LIBM) return fail(FMS_ERR_ARGUMENT, "bad exp_mode");
  if (c->thomas != FMS_THOMAS_SERIAL && c->thomas != FMS_THOMAS_PARTITIONED) return fail(FMS_ERR_ARGUMENT, "bad thomas");
  if (c->kernels != FMS_KERNELS_SPLIT && c->kernels != FMS_KERNELS_FUSED) return fail(FMS_ERR_ARGUMENT, "bad kernels");
  FmsSolver *s = new FmsSolver();
  s->F = c->n_fibers;
  s->scheme = c->scheme;
  s->precision = c->precision;
  s->exp_mode = c->exp_mode;
  s->thomas = c->thomas;
  s->kernels = c->kernels;
  s->vstim = c->value_for_stimulated_point;
  s->h_off.resize(s->F + 1);
  s->h_off[0] = 0;
  long long total = 0;
  for (int f = 0; f < s->F; f++) {
    const int n = c->n_nodes[f];
    if (n < 2) {
      delete s;
      return fail(FMS_ERR_ARGUMENT, "every fibre needs at least 2 nodes");
    }
    s->nmax = std::max(s->nmax, n);
    total += n;
    if (total > 0x7fffffffLL) {
      delete s;
      return fail(FMS_ERR_ARGUMENT, "more than 2^31-1 nodes in one solver");
    }
    s->h_off[f + 1] = (int)total;
  }
  s->N = total;
  if (s->nmax > 4096 && (s->thomas == FMS_THOMAS_PARTITIONED || s->kernels == FMS_KERNELS_FUSED)) {
    delete s;
    return fail(FMS_ERR_UNSUPPORTED, "fibres longer than 4096 nodes need thomas=SERIAL and kernels=SPLIT");
  }
  // fibres longest first for the fused kernel (LPT over the dispatcher)
  s->h_order.resize(s->F);
  for (int f = 0; f < s->F; f++) s->h_order[f] = f;
  std::stable_sort(s->h_order.begin(), s->h_order.end(),
                   [&](int a, int b) { return c->n_nodes[a] > c->n_nodes[b]; });

  int rc = 0;
  auto chk = [&](int r) {
    if (r && !rc) rc = r;
  };
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return fail(FMS_ERR_HIP, "hipStreamCreate failed");
  }
  const long long N = s->N;
  const int F = s->F;
  chk(dalloc(&s->d_off, F + 1));
  chk(dalloc(&s->d_node_fiber, (size_t)N));
  chk(dalloc(&s->d_sb, F));
  chk(dalloc(&s->d_se, F));
  chk(dalloc(&s->d_order, F));
  chk(dalloc(&s->d_h, (size_t)(N - F)));
  chk(dalloc(&s->d_pf, F));
  chk(dalloc(&s->d_params, (size_t)F * (Model::NP > 0 ? Model::NP : 1)));
  chk(dalloc(&s->d_vm, (size_t)N));
  chk(dalloc(&s->d_hasbeen, (size_t)F));
  chk(hipMalloc(&s->d_gates, (size_t)(Model::NS - 1) * N * gate_bytes(s) + 8) == hipSuccess ? 0 : FMS_ERR_HIP);
  if (s->thomas == FMS_THOMAS_SERIAL && s->kernels == FMS_KERNELS_SPLIT) {
    chk(dalloc(&s->d_cA, (size_t)N));
    chk(dalloc(&s->d_dA, (size_t)N));
  }
  if (rc) {
    free_all(s);
    delete s;
    return fail(FMS_ERR_HIP, "device allocation failed: " + g_last_error);
  }
  std::vector<int> node_fiber((size_t)N);
  for (int f = 0; f < F; f++)
    for (int i = s->h_off[f]; i < s->h_off[f + 1]; i++) node_fiber[i] = f;
  std::vector<int> zeros(F, 0);
  std::vector<double> params((size_t)F * (Model::NP > 0 ? Model::NP : 1));
  for (int f = 0; f < F; f++)
    for (int k = 0; k < Model::NP; k++)
      params[(size_t)f * Model::NP + k] = c->parameters ? c->parameters[(size_t)f * Model::NP + k] : Model::p0(k);
  chk(h2d(s->d_off, s->h_off.data(), F + 1, s->stream));
  chk(h2d(s->d_node_fiber, node_fiber.data(), (size_t)N, s->stream));
  chk(h2d(s->d_sb, zeros.data(), F, s->stream));
  chk(h2d(s->d_se, zeros.data(), F, s->stream));
  chk(h2d(s->d_order, s->h_order.data(), F, s->stream));
  if (!rc && hipMemsetAsync(s->d_hasbeen, 0, F, s->stream) != hipSuccess) rc = FMS_ERR_HIP;
  chk(h2d(s->d_h, c->element_lengths, (size_t)(N - F), s->stream));
  chk(h2d(s->d_pf, c->prefactor, F, s->stream));
  chk(h2d(s->d_params, params.data(), params.size(), s->stream));
  // initial states, state-major
  std::vector<double> y0(Model::NS);
  for (int k = 0; k < Model::NS; k++) y0[k] = c->initial_states ? c->initial_states[k] : Model::y0(k);
  {
    std::vector<double> v((size_t)N, y0[0]);
    chk(h2d(s->d_vm, v.data(), (size_t)N, s->stream));
    if (s->precision == FMS_FP64) {
      std::vector<double> g((size_t)(Model::NS - 1) * N);
      for (int k = 1; k < Model::NS; k++) std::fill(g.begin() + (size_t)(k - 1) * N, g.begin() + (size_t)k * N, y0[k]);
      chk(h2d((double *)s->d_gates, g.data(), g.size(), s->stream));
    } else {
      std::vector<float> g((size_t)(Model::NS - 1) * N);
      for (int k = 1; k < Model::NS; k++)
        std::fill(g.begin() + (size_t)(k - 1) * N, g.begin() + (size_t)k * N, (float)y0[k]);
      chk(h2d((float *)s->d_gates, g.data(), g.size(), s->stream));
    }
  }
  if (!rc && hipStreamSynchronize(s->stream) != hipSuccess) rc = fail(FMS_ERR_HIP, "upload failed");
  if (rc) {
    free_all(s);
    delete s;
    return rc;
  }
  *out = s;
  return 0;
}

int fms_destroy(FmsSolver *s) {
  if (!s) return 0;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  free_all(s);
  delete s;
  return 0;
}

int fms_set_stimulation(FmsSolver *s, const FmsStimulation *st) {
  if (!s || !st) return fail(FMS_ERR_ARGUMENT, "null argument");
  const int F = s->F;
  if (st->mode == FMS_STIM_NONE) {
    s->stim_mode = FMS_STIM_NONE;
    return 0;
  }
  if (!st->stim_begin || !st->stim_end) return fail(FMS_ERR_ARGUMENT, "stim_begin/stim_end required");
  int rc = 0;
  auto chk = [&](int r) {
    if (r && !rc) rc = r;
  };
  chk(h2d(s->d_sb, st->stim_begin, F, s->stream));
  chk(h2d(s->d_se, st->stim_end, F, s->stream));
  if (st->mode == FMS_STIM_FIRING_TABLE) {
    if (st->firing_rows <= 0 || st->firing_cols <= 0 || !st->firing_events || !st->motor_unit ||
        !st->call_frequency || !st->repeat_after_first_call || !st->call_enable_begin || !st->jitter_offset ||
        !st->jitter_count)
      return fail(FMS_ERR_ARGUMENT, "firing-table stimulation needs the table and per-fibre settings");
    int njit = 0;
    for (int f = 0; f < F; f++) njit = std::max(njit, st->jitter_offset[f] + st->jitter_count[f]);
    s->fe_rows = st->firing_rows;
    s->fe_cols = st->firing_cols;
    const size_t nfe = (size_t)st->firing_rows * st->firing_cols;
    if (!s->d_fe) {
      chk(dalloc(&s->d_mu, F)); chk(dalloc(&s->d_joff, F)); chk(dalloc(&s->d_jcnt, F)); chk(dalloc(&s->d_jidx, F));
      chk(dalloc(&s->d_freq, F)); chk(dalloc(&s->d_rep, F)); chk(dalloc(&s->d_beg, F)); chk(dalloc(&s->d_last, F));
      chk(dalloc(&s->d_curjit, F)); chk(dalloc(&s->d_curr, F));
      chk(dalloc(&s->d_log_count, 1));
    } else {
      HIP_TRY(hipFree(s->d_fe));
      if (s->d_jit) HIP_TRY(hipFree(s->d_jit));
      s->d_jit = nullptr;
    }
    chk(dalloc(&s->d_fe, nfe));
    chk(dalloc(&s->d_jit, (size_t)std::max(njit, 1)));
    if (rc) return rc;
    chk(h2d(s->d_fe, st->firing_events, nfe, s->stream));
    chk(h2d(s->d_mu, st->motor_unit, F, s->stream));
    chk(h2d(s->d_freq, st->call_frequency, F, s->stream));
    chk(h2d(s->d_rep, st->repeat_after_first_call, F, s->stream));
    chk(h2d(s->d_beg, st->call_enable_begin, F, s->stream));
    if (njit > 0 && st->jitter) chk(h2d(s->d_jit, st->jitter, (size_t)njit, s->stream));
    chk(h2d(s->d_joff, st->jitter_offset, F, s->stream));
    chk(h2d(s->d_jcnt, st->jitter_count, F, s->stream));
    // lastStimulationCheckTime = enableBegin - 1e-13 - 1/(f + 0)
    // (fast_monodomain_solver_initialization.tpp:342-344)
    std::vector<double> last(F);
    for (int f = 0; f < F; f++)
      last[f] = st->last_check_time ? st->last_check_time[f]
                                    : st->call_enable_begin[f] - 1e-13 - 1. / (st->call_frequency[f] + 0.0);
    chk(h2d(s->d_last, last.data(), F, s->stream));
    HIP_TRY(hipMemsetAsync(s->d_curjit, 0, F * sizeof(double), s->stream));
    HIP_TRY(hipMemsetAsync(s->d_jidx, 0, F * sizeof(int), s->stream));
    HIP_TRY(hipMemsetAsync(s->d_curr, 0, F, s->stream));
    HIP_TRY(hipMemsetAsync(s->d_hasbeen, 0, F, s->stream));
    HIP_TRY(hipMemsetAsync(s->d_log_count, 0, sizeof(int), s->stream));
    if (st->log_capacity > 0) {
      if (s->d_log_time) HIP_TRY(hipFree(s->d_log_time));
      if (s->d_log_fiber) HIP_TRY(hipFree(s->d_log_fiber));
      chk(dalloc(&s->d_log_time, (size_t)st->log_capacity));
      chk(dalloc(&s->d_log_fiber, (size_t)st->log_capacity));
    }
    s->log_cap = st->log_capacity;
  } else if (st->mode == FMS_STIM_CURRENT) {
    if (!st->onset || !st->duration || !st->amplitude) return fail(FMS_ERR_ARGUMENT, "current pulses need onset/duration/amplitude");
    if (st->stim_param < 0 || st->stim_param >= Model::NP) return fail(FMS_ERR_ARGUMENT, "stim_param out of range");
    if (!s->d_onset) {
      chk(dalloc(&s->d_onset, F)); chk(dalloc(&s->d_dur, F)); chk(dalloc(&s->d_amp, F));
    }
    if (rc) return rc;
    chk(h2d(s->d_onset, st->onset, F, s->stream));
    chk(h2d(s->d_dur, st->duration, F, s->stream));
    chk(h2d(s->d_amp, st->amplitude, F, s->stream));
    s->stim_param = st->stim_param;
  } else {
    return fail(FMS_ERR_ARGUMENT, "unknown stimulation mode");
  }
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(s->stream));
  s->stim_mode = st->mode;
  return 0;
}

int fms_advance(FmsSolver *s, double t0, double dts, int n_split, double dt0D, int n0D, double dt1D, int n1D) {
  (void)n1D;
  if (!s) return fail(FMS_ERR_ARGUMENT, "null solver");
  if (n_split <= 0) return 0;
  std::vector<double> tab(4 * (size_t)n_split);
  for (int k = 0; k < n_split; k++) {
    tab[4 * (size_t)k] = t0 + k * dts;   // currentTime = startTime + timeStepNo*timeStepWidthSplitting
    tab[4 * (size_t)k + 1] = dts;
    tab[4 * (size_t)k + 2] = dt0D;
    tab[4 * (size_t)k + 3] = dt1D;
  }
  return dispatch(s, tab.data(), n_split, n0D);
}

int fms_advance_schedule(FmsSolver *s, const double *steps, int n_steps, int n0D) {
  if (!s || (!steps && n_steps > 0)) return fail(FMS_ERR_ARGUMENT, "null argument");
  return dispatch(s, steps, n_steps, n0D);
}

int fms_synchronize(FmsSolver *s) {
  if (!s) return fail(FMS_ERR_ARGUMENT, "null solver");
  HIP_TRY(hipStreamSynchronize(s->stream));
  return 0;
}

int fms_total_nodes(FmsSolver *s, long long *n) {
  if (!s || !n) return fail(FMS_ERR_ARGUMENT, "null argument");
  *n = s->N;
  return 0;
}

int fms_get_states(FmsSolver *s, int k, double *out) {
  if (!s || !out) return fail(FMS_ERR_ARGUMENT, "null argument");
  if (k < 0 || k >= Model::NS) return fail(FMS_ERR_ARGUMENT, "state index out of range");
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (k == 0) {
    HIP_TRY(hipMemcpy(out, s->d_vm, (size_t)s->N * sizeof(double), hipMemcpyDeviceToHost));
  } else if (s->precision == FMS_FP64) {
    HIP_TRY(hipMemcpy(out, (double *)s->d_gates + (size_t)(k - 1) * s->N, (size_t)s->N * sizeof(double),
                      hipMemcpyDeviceToHost));
  } else {
    std::vector<float> tmp((size_t)s->N);
    HIP_TRY(hipMemcpy(tmp.data(), (float *)s->d_gates + (size_t)(k - 1) * s->N, (size_t)s->N * sizeof(float),
                      hipMemcpyDeviceToHost));
    for (long long i = 0; i < s->N; i++) out[i] = tmp[(size_t)i];
  }
  return 0;
}

int fms_set_states(FmsSolver *s, int k, const double *in) {
  if (!s || !in) return fail(FMS_ERR_ARGUMENT, "null argument");
  if (k < 0 || k >= Model::NS) return fail(FMS_ERR_ARGUMENT, "state index out of range");
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (k == 0) {
    HIP_TRY(hipMemcpy(s->d_vm, in, (size_t)s->N * sizeof(double), hipMemcpyHostToDevice));
  } else if (s->precision == FMS_FP64) {
    HIP_TRY(hipMemcpy((double *)s->d_gates + (size_t)(k - 1) * s->N, in, (size_t)s->N * sizeof(double),
                      hipMemcpyHostToDevice));
  } else {
    std::vector<float> tmp((size_t)s->N);
    for (long long i = 0; i < s->N; i++) tmp[(size_t)i] = (float)in[i];
    HIP_TRY(hipMemcpy((float *)s->d_gates + (size_t)(k - 1) * s->N, tmp.data(), (size_t)s->N * sizeof(float),
                      hipMemcpyHostToDevice));
  }
  return 0;
}

int fms_set_algebraics_for_transfer(FmsSolver *s, const int *idx, int count) {
  if (!s || count < 0 || (count > 0 && !idx)) return fail(FMS_ERR_ARGUMENT, "bad argument");
  for (int k = 0; k < count; k++)
    if (idx[k] < 0 || idx[k] >= Model::NA) return fail(FMS_ERR_ARGUMENT, "algebraic index out of range");
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (s->d_alg_idx) HIP_TRY(hipFree(s->d_alg_idx));
  if (s->d_alg) HIP_TRY(hipFree(s->d_alg));
  s->d_alg_idx = nullptr;
  s->d_alg = nullptr;
  s->n_alg_idx = count;
  if (count == 0) return 0;
  if (dalloc(&s->d_alg_idx, (size_t)count) || dalloc(&s->d_alg, (size_t)count * s->N)) return FMS_ERR_HIP;
  HIP_TRY(hipMemcpy(s->d_alg_idx, idx, count * sizeof(int), hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(s->d_alg, 0, (size_t)count * s->N * sizeof(double)));
  return 0;
}

int fms_get_algebraics(FmsSolver *s, double *out) {
  if (!s || !out) return fail(FMS_ERR_ARGUMENT, "null argument");
  if (!s->d_alg) return fail(FMS_ERR_STATE, "no algebraics for transfer configured");
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(out, s->d_alg, (size_t)s->n_alg_idx * s->N * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int fms_enable_spike_detection(FmsSolver *s, double th) {
  if (!s) return fail(FMS_ERR_ARGUMENT, "null solver");
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (!s->d_vprev) {
    if (dalloc(&s->d_vprev, (size_t)s->N) || dalloc(&s->d_sfirst, (size_t)s->N) || dalloc(&s->d_scount, (size_t)s->N))
      return FMS_ERR_HIP;
  }
  HIP_TRY(hipMemcpy(s->d_vprev, s->d_vm, (size_t)s->N * sizeof(double), hipMemcpyDeviceToDevice));
  std::vector<int> m1((size_t)s->N, -1);
  HIP_TRY(hipMemcpy(s->d_sfirst, m1.data(), (size_t)s->N * sizeof(int), hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(s->d_scount, 0, (size_t)s->N * sizeof(int)));
  s->spikes = 1;
  s->spike_th = th;
  return 0;
}

int fms_get_spikes(FmsSolver *s, int *first, int *count) {
  if (!s) return fail(FMS_ERR_ARGUMENT, "null solver");
  if (!s->spikes) return fail(FMS_ERR_STATE, "spike detection is not enabled");
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (first) HIP_TRY(hipMemcpy(first, s->d_sfirst, (size_t)s->N * sizeof(int), hipMemcpyDeviceToHost));
  if (count) HIP_TRY(hipMemcpy(count, s->d_scount, (size_t)s->N * sizeof(int), hipMemcpyDeviceToHost));
  return 0;
}

int fms_get_stimulation_log(FmsSolver *s, int *count, double *times, int *fibers, int cap) {
  if (!s || !count) return fail(FMS_ERR_ARGUMENT, "null argument");
  *count = 0;
  if (s->stim_mode != FMS_STIM_FIRING_TABLE || s->log_cap <= 0) return 0;
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(count, s->d_log_count, sizeof(int), hipMemcpyDeviceToHost));
  const int n = std::min(std::min(*count, s->log_cap), cap);
  if (n > 0 && times) HIP_TRY(hipMemcpy(times, s->d_log_time, n * sizeof(double), hipMemcpyDeviceToHost));
  if (n > 0 && fibers) HIP_TRY(hipMemcpy(fibers, s->d_log_fiber, n * sizeof(int), hipMemcpyDeviceToHost));
  return 0;
}

int fms_set_tuning(FmsSolver *s, int key, int value) {
  if (!s) return fail(FMS_ERR_ARGUMENT, "null solver");
  switch (key) {
    case FMS_TUNE_STEPS_PER_LAUNCH:
      if (value < 1) return fail(FMS_ERR_ARGUMENT, "steps per launch must be >= 1");
      s->tune_chunk = value;
      return 0;
    case FMS_TUNE_ONLY_IF_STIMULATED:
      if (value < 0 || value > 1) return fail(FMS_ERR_ARGUMENT, "onlyComputeIfHasBeenStimulated is 0 or 1");
      s->only_if_stimulated = value;
      return 0;
    case FMS_TUNE_RESIDENT_GATES:
      if (value < -1 || value > 1) return fail(FMS_ERR_ARGUMENT, "resident must be -1, 0 or 1");
      if (value == 1 && ((s->nmax + FUSED_T - 1) / FUSED_T > 8 || Model::NS > 8))
        return fail(FMS_ERR_UNSUPPORTED, "resident gates need fibres of <= 2048 nodes and <= 8 states");
      s->tune_resident = value;
      return 0;
    default:
      return fail(FMS_ERR_ARGUMENT, "unknown tuning key");
  }
}

int fms_record_event(FmsSolver *s, int slot) {
  if (!s || slot < 0 || slot >= 8) return fail(FMS_ERR_ARGUMENT, "event slot must be 0..7");
  if (!s->events[slot]) HIP_TRY(hipEventCreate(&s->events[slot]));
  HIP_TRY(hipEventRecord(s->events[slot], s->stream));
  return 0;
}

int fms_event_elapsed_ms(FmsSolver *s, int from, int to, double *ms) {
  if (!s || !ms || from < 0 || from >= 8 || to < 0 || to >= 8 || !s->events[from] || !s->events[to])
    return fail(FMS_ERR_ARGUMENT, "events not recorded");
  HIP_TRY(hipEventSynchronize(s->events[to]));
  float f = 0.f;
  HIP_TRY(hipEventElapsedTime(&f, s->events[from], s->events[to]));
  *ms = f;
  return 0;
}

int fms_launch_counts(FmsSolver *s, long long *fused, long long *reaction, long long *thomas) {
  if (!s) return fail(FMS_ERR_ARGUMENT, "null solver");
  if (fused) *fused = s->n_launch_fused;
  if (reaction) *reaction = s->n_launch_reaction;
  if (thomas) *thomas = s->n_launch_thomas;
  return 0;
}

}  // extern "C"

#include "fms_schedule.h"
#include "fms_reference_abi.h"

// ===========================================================================
// parameters, sensors and the discrete adjoint
// ===========================================================================
namespace {

template <class T> int ensure(T **p, long long &cap, long long n) {
  if (cap >= n && *p) return 0;
  if (*p) HIP_TRY(hipFree(*p));
  *p = nullptr;
  cap = 0;
  if (dalloc(p, (size_t)n)) return FMS_ERR_HIP;
  cap = n;
  return 0;
}

int upload_steps(FmsSolver *s, const double *steps, int K) {
  if (s->steps_cap < 4 * (K + 1)) {
    if (s->d_steps) HIP_TRY(hipFree(s->d_steps));
    s->d_steps = nullptr;
    s->steps_cap = 4 * (K + 1);
    if (dalloc(&s->d_steps, (size_t)s->steps_cap)) return FMS_ERR_HIP;
  }
  std::vector<double> tab(4 * (size_t)(K + 1));
  memcpy(tab.data(), steps, sizeof(double) * 4 * K);
  for (int q = 0; q < 4; q++) tab[4 * (size_t)K + q] = steps[4 * (size_t)(K - 1) + q];
  HIP_TRY(hipMemcpyAsync(s->d_steps, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
  return 0;
}

int copy_state(FmsSolver *s, double *dst, bool to_ckpt) {
  const size_t n = (size_t)s->N;
  const size_t gb = (size_t)(Model::NS - 1) * n * sizeof(double);
  if (to_ckpt) {
    HIP_TRY(hipMemcpyAsync(dst, s->d_vm, n * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(dst + n, s->d_gates, gb, hipMemcpyDeviceToDevice, s->stream));
  } else {
    HIP_TRY(hipMemcpyAsync(s->d_vm, dst, n * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_gates, dst + n, gb, hipMemcpyDeviceToDevice, s->stream));
  }
  return 0;
}

template <class E> int adjoint_backward(FmsSolver *s, const double *dtraces, double *grad_params, double *pf_bar) {
  if constexpr (!std::is_same<Model, fms::ModelHH>::value) {
    (void)s; (void)dtraces; (void)grad_params; (void)pf_bar;
    return fail(FMS_ERR_UNSUPPORTED, "the adjoint is implemented for the Hodgkin-Huxley model");
  } else {
    const long long N = s->N;
    const int F = s->F, S = s->n_sensors, K = s->adj_K, Kc = s->adj_interval, n0D = s->adj_n0D;
    const int ie = s->scheme == FMS_IMPLICIT_EULER;
    long long cap = 0;
    int rc;
    if ((rc = ensure(&s->d_dtrace, cap, (long long)K * F * S))) return rc;
    HIP_TRY(hipMemcpyAsync(s->d_dtrace, dtraces, (size_t)K * F * S * sizeof(double), hipMemcpyHostToDevice,
                           s->stream));
    long long c1 = 0, c2 = 0, c3 = 0, c4 = 0;
    if ((rc = ensure(&s->d_lam, c1, 4 * N)) || (rc = ensure(&s->d_gacc, c2, 3 * N)) ||
        (rc = ensure(&s->d_pfbar, c3, F)) || (rc = ensure(&s->d_grad, c4, 3 * (long long)F)))
      return rc;
    HIP_TRY(hipMemsetAsync(s->d_lam, 0, 4 * N * sizeof(double), s->stream));
    HIP_TRY(hipMemsetAsync(s->d_gacc, 0, 3 * N * sizeof(double), s->stream));
    HIP_TRY(hipMemsetAsync(s->d_pfbar, 0, F * sizeof(double), s->stream));
    if ((rc = ensure(&s->d_tape, s->tape_cap, (long long)Kc * 9 * N))) return rc;
    if ((rc = upload_steps(s, s->adj_steps.data(), K))) return rc;
    const int saved_spikes = s->spikes;
    s->spikes = 0;
    const int nseg = (K + Kc - 1) / Kc;
    fms::DevTape tp{s->d_tape, N};
    const int threads = 256;
    const unsigned blocks = (unsigned)((N + threads - 1) / threads);
    const size_t lds = (size_t)(16 * PART_T + 2 * s->nmax + PART_T) * sizeof(double);
    for (int seg = nseg - 1; seg >= 0; seg--) {
      const int k0 = seg * Kc, kc = std::min(Kc, K - k0);
      if ((rc = copy_state(s, s->d_ckpt + (size_t)seg * Model::NS * N, false))) break;
      fms::DevTrace tr{};
      tr.tape = s->d_tape;
      const long long save = s->step_counter;
      if ((rc = launch_fused<double, E>(s, s->d_steps + 4 * (size_t)k0, kc, n0D, tr))) break;
      s->step_counter = save;
      for (int q = kc - 1; q >= 0; q--) {
        const int k = k0 + q;
        const double *st = s->adj_steps.data() + 4 * (size_t)k;
        const double t = st[0], dts = st[1], dt0 = st[2], dt1 = st[3];
        hipLaunchKernelGGL(fms::k_adj_sensors, dim3((F * S + 255) / 256), dim3(256), 0, s->stream, s->fibers(),
                           s->d_sensor_nodes, S, s->d_dtrace, k, s->d_lam);
        hipLaunchKernelGGL((fms::k_adj_heun<E>), dim3(blocks), dim3(threads), 0, s->stream, s->fibers(), s->stim(),
                           tp, q, 1, t + 0.5 * dts, dt0, n0D, s->d_lam, s->d_gacc);
        hipLaunchKernelGGL((fms::k_adj_thomas<PART_T, 16>), dim3(F), dim3(PART_T), lds, s->stream, s->fibers(), tp, q,
                           dt1, ie, s->d_lam, s->d_pfbar);
        hipLaunchKernelGGL((fms::k_adj_heun<E>), dim3(blocks), dim3(threads), 0, s->stream, s->fibers(), s->stim(),
                           tp, q, 0, t, dt0, n0D, s->d_lam, s->d_gacc);
        if (hipGetLastError() != hipSuccess) {
          rc = fail(FMS_ERR_HIP, "adjoint kernel launch failed");
          break;
        }
      }
      if (rc) break;
    }
    s->spikes = saved_spikes;
    if (rc) return rc;
    hipLaunchKernelGGL(fms::k_adj_reduce, dim3(F), dim3(256), 0, s->stream, s->fibers(), s->d_gacc, s->d_grad);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (grad_params) HIP_TRY(hipMemcpy(grad_params, s->d_grad, 3 * (size_t)F * sizeof(double), hipMemcpyDeviceToHost));
    if (pf_bar) HIP_TRY(hipMemcpy(pf_bar, s->d_pfbar, (size_t)F * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
  }
}

}  // namespace

extern "C" {

int fms_set_parameters(FmsSolver *s, const double *params, const double *prefactor) {
  if (!s) return fail(FMS_ERR_ARGUMENT, "null solver");
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (params && Model::NP > 0)
    HIP_TRY(hipMemcpy(s->d_params, params, (size_t)s->F * Model::NP * sizeof(double), hipMemcpyHostToDevice));
  if (prefactor) HIP_TRY(hipMemcpy(s->d_pf, prefactor, (size_t)s->F * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

int fms_set_sensors(FmsSolver *s, const int *sensor_nodes, int n_sensors) {
  if (!s || n_sensors < 0 || (n_sensors > 0 && !sensor_nodes)) return fail(FMS_ERR_ARGUMENT, "bad sensors");
  const int F = s->F;
  std::vector<int> slot((size_t)s->N, -1);
  for (int f = 0; f < F; f++) {
    const int n = s->h_off[f + 1] - s->h_off[f];
    for (int j = 0; j < n_sensors; j++) {
      const int i = sensor_nodes[(size_t)f * n_sensors + j];
      if (i < 0 || i >= n) return fail(FMS_ERR_ARGUMENT, "sensor node outside its fibre");
      if (slot[(size_t)s->h_off[f] + i] >= 0) return fail(FMS_ERR_ARGUMENT, "duplicate sensor node");
      slot[(size_t)s->h_off[f] + i] = j;
    }
  }
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (s->d_sensor_nodes) HIP_TRY(hipFree(s->d_sensor_nodes));
  if (s->d_sensor_slot) HIP_TRY(hipFree(s->d_sensor_slot));
  s->d_sensor_nodes = nullptr;
  s->d_sensor_slot = nullptr;
  s->n_sensors = n_sensors;
  if (n_sensors == 0) return 0;
  if (dalloc(&s->d_sensor_nodes, (size_t)F * n_sensors) || dalloc(&s->d_sensor_slot, (size_t)s->N)) return FMS_ERR_HIP;
  HIP_TRY(hipMemcpy(s->d_sensor_nodes, sensor_nodes, (size_t)F * n_sensors * sizeof(int), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s->d_sensor_slot, slot.data(), (size_t)s->N * sizeof(int), hipMemcpyHostToDevice));
  return 0;
}

int fms_adjoint_forward(FmsSolver *s, const double *steps, int n_steps, int n0D, int checkpoint_interval,
                        double *traces) {
  if (!s || !steps || n_steps <= 0 || n0D < 1 || checkpoint_interval < 1)
    return fail(FMS_ERR_ARGUMENT, "bad adjoint forward arguments");
  if (s->n_sensors <= 0) return fail(FMS_ERR_STATE, "call fms_set_sensors first");
  if (s->precision != FMS_FP64) return fail(FMS_ERR_UNSUPPORTED, "the adjoint runs in FP64");
  if (s->stim_mode == FMS_STIM_FIRING_TABLE)
    return fail(FMS_ERR_UNSUPPORTED, "the Vm clamp of firing-table stimulation is not differentiable; use current pulses");
  if (s->only_if_stimulated) return fail(FMS_ERR_UNSUPPORTED, "onlyComputeIfHasBeenStimulated is not differentiable");
  const long long N = s->N;
  const int F = s->F, S = s->n_sensors, K = n_steps, Kc = checkpoint_interval;
  const int nseg = (K + Kc - 1) / Kc;
  int rc;
  if ((rc = ensure(&s->d_trace, s->trace_cap, (long long)K * F * S))) return rc;
  if ((rc = ensure(&s->d_ckpt, s->ckpt_cap, (long long)nseg * Model::NS * N))) return rc;
  if (s->flags_cap < s->F * n0D) {
    if (s->d_flags) HIP_TRY(hipFree(s->d_flags));
    s->flags_cap = s->F * n0D;
    if (dalloc(&s->d_flags, (size_t)s->flags_cap)) return FMS_ERR_HIP;
  }
  if ((rc = upload_steps(s, steps, K))) return rc;
  s->adj_steps.assign(steps, steps + 4 * (size_t)K);
  s->adj_n0D = n0D;
  s->adj_interval = Kc;
  s->adj_K = K;
  for (int seg = 0; seg < nseg; seg++) {
    const int k0 = seg * Kc, kc = std::min(Kc, K - k0);
    if ((rc = copy_state(s, s->d_ckpt + (size_t)seg * Model::NS * N, true))) return rc;
    fms::DevTrace tr{};
    tr.sensor_slot = s->d_sensor_slot;
    tr.trace = s->d_trace + (size_t)k0 * F * S;
    tr.n_sensors = S;
    if (s->exp_mode == FMS_EXP_APPROXIMATE) rc = launch_fused<double, fms::ExpApprox>(s, s->d_steps + 4 * (size_t)k0, kc, n0D, tr);
    else rc = launch_fused<double, fms::ExpLibm>(s, s->d_steps + 4 * (size_t)k0, kc, n0D, tr);
    if (rc) return rc;
  }
  if (traces) {
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(traces, s->d_trace, (size_t)K * F * S * sizeof(double), hipMemcpyDeviceToHost));
  }
  return 0;
}

int fms_adjoint_backward(FmsSolver *s, const double *dtraces, double *grad_params, double *grad_prefactor) {
  if (!s || !dtraces) return fail(FMS_ERR_ARGUMENT, "null argument");
  if (s->adj_K <= 0) return fail(FMS_ERR_STATE, "call fms_adjoint_forward first");
  if (s->exp_mode == FMS_EXP_APPROXIMATE) return adjoint_backward<fms::ExpApprox>(s, dtraces, grad_params, grad_prefactor);
  return adjoint_backward<fms::ExpLibm>(s, dtraces, grad_params, grad_prefactor);
}

}  // extern "C"
