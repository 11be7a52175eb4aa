// The two symbols OpenDiHu's FastMonodomainSolver binds when it loads a
// prebuilt "gpu" library (fast_monodomain_solver_gpu.tpp:184-205), on top of
// the native fms_* API. The reference's library is a process-wide singleton
// with sizes baked into the generated source (fast_monodomain_solver_gpu.tpp:
// 293-312); here the sizes come from fms_reference_set_dims() or the
// OPENDIHU_FMS_DIMS environment variable ("key=value,..." with the keys
// nFibersToCompute, nInstancesPerFiber, nParametersPerInstance,
// firingEventsNRows, firingEventsNColumns, frequencyJitterNColumns,
// nAlgebraicsForTransfer, nStatesForTransfer, scheme=IE|CN,
// precision=fp64|mixed, exp=approx|libm).
//
// Semantics follow the reference's gpu back-end where it is self-consistent:
//   * Vm and the gates live in the library between calls; initializeArrays
//     sets every instance to statesOneInstance (gpu.tpp:407-467).
//   * The stimulated nodes are NMJ-1..NMJ+1 (gpu.tpp:540).
//   * parameters[p*nInstances + i] are per instance; the fibre solver holds one
//     set per fibre (the reference host fills every instance of a fibre with
//     the same value, initialization.tpp:459-465). Parameters and prefactor are
//     taken on every call (gpu.tpp:481,490) and uploaded when they change; a
//     call whose instances of one fibre differ fails (FMS_ERR_UNSUPPORTED).
//   * A void entry clears fms_last_error first: an empty message after the
//     call means success.
//   * After the call, statesForTransfer[i*nInstances + inst] holds state
//     statesForTransferIndices[i] (index 0 is Vm) and algebraicsForTransfer the
//     requested algebraics of the last Heun step (gpu.tpp:401-461).
// It does not reproduce the generated code's defects (see the reference's
// generator: Heun adding the Euler increment twice, gpu.tpp:794 writing fibre
// 0's last node for every fibre); it computes what the vc path computes.
#pragma once

#include <string>
#include <vector>

namespace {

struct RefAbi {
  bool dims_set = false;
  FmsReferenceDims dims{};
  bool initialized = false;
  FmsSolver *solver = nullptr;
  std::vector<double> states_one;
  std::vector<int> alg_idx, state_idx;
  std::vector<unsigned char> firing;
  std::vector<double> jitter;
  std::vector<int> jit_off, jit_cnt, mu, nmj;
  std::vector<double> last_check, freq, repeat, begin;
  std::vector<float> last_lengths;
  std::vector<double> last_params;
  double last_prefactor = 0.0;
};

RefAbi g_ref;

bool ref_dims_from_env() {
  const char *env = getenv("OPENDIHU_FMS_DIMS");
  if (!env) return false;
  FmsReferenceDims d{};
  d.scheme = FMS_IMPLICIT_EULER;
  d.precision = FMS_FP64;
  d.exp_mode = FMS_EXP_APPROXIMATE;
  std::string s(env);
  size_t pos = 0;
  while (pos < s.size()) {
    size_t end = s.find(',', pos);
    if (end == std::string::npos) end = s.size();
    const std::string kv = s.substr(pos, end - pos);
    pos = end + 1;
    const size_t eq = kv.find('=');
    if (eq == std::string::npos) continue;
    const std::string k = kv.substr(0, eq), v = kv.substr(eq + 1);
    const int iv = atoi(v.c_str());
    if (k == "nFibersToCompute") d.n_fibers_to_compute = iv;
    else if (k == "nInstancesPerFiber") d.n_instances_per_fiber = iv;
    else if (k == "nParametersPerInstance") d.n_parameters_per_instance = iv;
    else if (k == "firingEventsNRows") d.firing_events_rows = iv;
    else if (k == "firingEventsNColumns") d.firing_events_cols = iv;
    else if (k == "frequencyJitterNColumns") d.frequency_jitter_cols = iv;
    else if (k == "nAlgebraicsForTransfer") d.n_algebraics_for_transfer = iv;
    else if (k == "nStatesForTransfer") d.n_states_for_transfer = iv;
    else if (k == "scheme") d.scheme = (v == "CN" || v == "CrankNicolson") ? FMS_CRANK_NICOLSON : FMS_IMPLICIT_EULER;
    else if (k == "precision") d.precision = (v == "mixed" || v == "single") ? FMS_MIXED : FMS_FP64;
    else if (k == "exp") d.exp_mode = v == "libm" ? FMS_EXP_LIBM : FMS_EXP_APPROXIMATE;
  }
  g_ref.dims = d;
  g_ref.dims_set = true;
  return true;
}

// parameters[p*N + inst] (float, N = F*n instances) as one set per fibre
// [F][NP]; parameters past n_params_given keep the model's defaults. The fibre
// solver holds one parameter set per fibre, so a call whose instances of one
// fibre differ is refused (FMS_ERR_UNSUPPORTED) rather than computed with the
// first instance's values.
bool ref_fibre_parameters(const float *parameters, int F, int n, int n_params_given, std::vector<double> &prm) {
  const long long N = (long long)F * n;
  prm.assign((size_t)F * (Model::NP > 0 ? Model::NP : 1), 0.0);
  for (int f = 0; f < F; f++)
    for (int k = 0; k < Model::NP; k++) {
      if (k >= n_params_given || !parameters) {
        prm[(size_t)f * Model::NP + k] = Model::p0(k);
        continue;
      }
      const float *row = parameters + (size_t)k * N + (size_t)f * n;
      for (int i = 1; i < n; i++)
        if (!(row[i] == row[0])) {
          fail(FMS_ERR_UNSUPPORTED, "computeMonodomain: parameter " + std::to_string(k) + " differs between nodes of fibre " +
                                        std::to_string(f) + " (node " + std::to_string(i) +
                                        "); the fibre solver takes one parameter set per fibre");
          return false;
        }
      prm[(size_t)f * Model::NP + k] = (double)row[0];
    }
  return true;
}

}  // namespace

extern "C" int fms_reference_set_dims(const FmsReferenceDims *d) {
  FMS_API_BEGIN
  if (!d) return fail(FMS_ERR_ARGUMENT, "null dims");
  if (d->n_fibers_to_compute <= 0 || d->n_instances_per_fiber < 2 || d->firing_events_rows <= 0 ||
      d->firing_events_cols <= 0 || d->n_states_for_transfer < 1 || d->n_parameters_per_instance < 0 ||
      d->frequency_jitter_cols < 0 || d->n_algebraics_for_transfer < 0)
    return fail(FMS_ERR_ARGUMENT, "invalid reference dimensions");
  if (g_ref.solver) {
    fms_destroy(g_ref.solver);
    g_ref.solver = nullptr;
  }
  g_ref = RefAbi();
  g_ref.dims = *d;
  g_ref.dims_set = true;
  return 0;
  FMS_API_END
}

extern "C" void initializeArrays(const double *statesOneInstance, const int *algebraicsForTransferIndices,
                                 const int *statesForTransferIndices, const char *firingEvents,
                                 const double *setSpecificStatesFrequencyJitter, const int *motorUnitNo,
                                 const int *fiberStimulationPointIndex, const double *lastStimulationCheckTime,
                                 const double *setSpecificStatesCallFrequency,
                                 const double *setSpecificStatesRepeatAfterFirstCall,
                                 const double *setSpecificStatesCallEnableBegin) {
  FMS_API_BEGIN
  g_last_error.clear();
  if (!g_ref.dims_set && !ref_dims_from_env()) {
    fail(FMS_ERR_STATE, "initializeArrays: sizes unknown (call fms_reference_set_dims or set OPENDIHU_FMS_DIMS)");
    return;
  }
  const FmsReferenceDims &d = g_ref.dims;
  const int F = d.n_fibers_to_compute;
  if (g_ref.solver) {
    fms_destroy(g_ref.solver);
    g_ref.solver = nullptr;
  }
  g_ref.states_one.assign(statesOneInstance, statesOneInstance + Model::NS);
  g_ref.alg_idx.assign(algebraicsForTransferIndices, algebraicsForTransferIndices + d.n_algebraics_for_transfer);
  g_ref.state_idx.assign(statesForTransferIndices, statesForTransferIndices + d.n_states_for_transfer);
  const size_t nfe = (size_t)d.firing_events_rows * d.firing_events_cols;
  g_ref.firing.resize(nfe);
  for (size_t i = 0; i < nfe; i++) g_ref.firing[i] = firingEvents[i] ? 1 : 0;
  g_ref.jitter.assign(setSpecificStatesFrequencyJitter,
                      setSpecificStatesFrequencyJitter + (size_t)F * d.frequency_jitter_cols);
  g_ref.jit_off.resize(F);
  g_ref.jit_cnt.assign(F, d.frequency_jitter_cols);
  for (int f = 0; f < F; f++) g_ref.jit_off[f] = f * d.frequency_jitter_cols;
  g_ref.mu.assign(motorUnitNo, motorUnitNo + F);
  g_ref.nmj.assign(fiberStimulationPointIndex, fiberStimulationPointIndex + F);
  g_ref.last_check.assign(lastStimulationCheckTime, lastStimulationCheckTime + F);
  g_ref.freq.assign(setSpecificStatesCallFrequency, setSpecificStatesCallFrequency + F);
  g_ref.repeat.assign(setSpecificStatesRepeatAfterFirstCall, setSpecificStatesRepeatAfterFirstCall + F);
  g_ref.begin.assign(setSpecificStatesCallEnableBegin, setSpecificStatesCallEnableBegin + F);
  g_ref.initialized = true;
  FMS_API_END_VOID
}

extern "C" void computeMonodomain(const float *parameters, double *algebraicsForTransfer, double *statesForTransfer,
                                  const float *elementLengths, double startTime, double timeStepWidthSplitting,
                                  int nTimeStepsSplitting, double dt0D, int nTimeSteps0D, double dt1D,
                                  int nTimeSteps1D, double prefactor, double valueForStimulatedPoint) {
  FMS_API_BEGIN
  g_last_error.clear();
  if (!g_ref.initialized) {
    fail(FMS_ERR_STATE, "computeMonodomain called before initializeArrays");
    return;
  }
  const FmsReferenceDims &d = g_ref.dims;
  const int F = d.n_fibers_to_compute, n = d.n_instances_per_fiber;
  const long long N = (long long)F * n;
  const size_t nlen = (size_t)F * (n - 1);
  // parameters and prefactor of this call (the reference passes both on every
  // call, gpu.tpp:481,490): one parameter set per fibre
  std::vector<double> prm;
  if (!ref_fibre_parameters(parameters, F, n, d.n_parameters_per_instance, prm)) return;
  if (!g_ref.solver) {
    std::vector<int> nn(F, n);
    std::vector<double> h(nlen), pf(F, prefactor);
    for (size_t i = 0; i < nlen; i++) h[i] = elementLengths[i];
    FmsConfig c{};
    c.n_fibers = F;
    c.n_nodes = nn.data();
    c.element_lengths = h.data();
    c.prefactor = pf.data();
    c.parameters = prm.data();
    c.initial_states = g_ref.states_one.data();
    c.scheme = d.scheme;
    c.precision = d.precision;
    c.exp_mode = d.exp_mode;
    c.thomas = FMS_THOMAS_PARTITIONED;
    c.kernels = FMS_KERNELS_FUSED;
    c.value_for_stimulated_point = valueForStimulatedPoint;
    if (fms_create(&c, &g_ref.solver)) return;
    std::vector<int> sb(F), se(F);
    for (int f = 0; f < F; f++) {
      sb[f] = std::max(0, g_ref.nmj[f] - 1);
      se[f] = std::min(n, g_ref.nmj[f] + 2);
    }
    FmsStimulation st{};
    st.mode = FMS_STIM_FIRING_TABLE;
    st.stim_begin = sb.data();
    st.stim_end = se.data();
    st.firing_rows = d.firing_events_rows;
    st.firing_cols = d.firing_events_cols;
    st.firing_events = g_ref.firing.data();
    st.motor_unit = g_ref.mu.data();
    st.call_frequency = g_ref.freq.data();
    st.repeat_after_first_call = g_ref.repeat.data();
    st.call_enable_begin = g_ref.begin.data();
    st.jitter = g_ref.jitter.data();
    st.jitter_offset = g_ref.jit_off.data();
    st.jitter_count = g_ref.jit_cnt.data();
    st.last_check_time = g_ref.last_check.data();
    if (fms_set_stimulation(g_ref.solver, &st)) return;
    if (!g_ref.alg_idx.empty() &&
        fms_set_algebraics_for_transfer(g_ref.solver, g_ref.alg_idx.data(), (int)g_ref.alg_idx.size()))
      return;
    g_ref.last_lengths.assign(elementLengths, elementLengths + nlen);
    g_ref.last_params = prm;
    g_ref.last_prefactor = prefactor;
  } else {
    if (memcmp(g_ref.last_lengths.data(), elementLengths, nlen * sizeof(float)) != 0) {
      // geometry moved (e.g. a coupled contraction solver): upload the new lengths
      std::vector<double> h(nlen);
      for (size_t i = 0; i < nlen; i++) h[i] = elementLengths[i];
      if (hipStreamSynchronize(g_ref.solver->stream) != hipSuccess ||
          hipMemcpy(g_ref.solver->d_h, h.data(), nlen * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        fail(FMS_ERR_HIP, "element length upload failed");
        return;
      }
      g_ref.last_lengths.assign(elementLengths, elementLengths + nlen);
    }
    // parameters (e.g. parametersUsedAsAlgebraic fed back by a coupled solver)
    // and the prefactor: uploaded when they differ from the previous call's
    const bool new_params = prm != g_ref.last_params;
    const bool new_pf = !(prefactor == g_ref.last_prefactor);
    if (new_params || new_pf) {
      std::vector<double> pf(F, prefactor);
      if (fms_set_parameters(g_ref.solver, new_params ? prm.data() : nullptr, new_pf ? pf.data() : nullptr)) return;
      g_ref.last_params = prm;
      g_ref.last_prefactor = prefactor;
    }
  }
  g_ref.solver->vstim = valueForStimulatedPoint;
  if (fms_advance(g_ref.solver, startTime, timeStepWidthSplitting, nTimeStepsSplitting, dt0D, nTimeSteps0D, dt1D,
                  nTimeSteps1D))
    return;
  for (size_t i = 0; i < g_ref.state_idx.size(); i++) {
    const int k = g_ref.state_idx[i];
    if (k < 0 || k >= Model::NS) {
      fail(FMS_ERR_ARGUMENT, "statesForTransfer index out of range");
      return;
    }
    if (fms_get_states(g_ref.solver, k, statesForTransfer + i * (size_t)N)) return;
  }
  if (!g_ref.alg_idx.empty() && algebraicsForTransfer) fms_get_algebraics(g_ref.solver, algebraicsForTransfer);
  FMS_API_END_VOID
}
