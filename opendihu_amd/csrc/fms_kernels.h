// HIP kernels of the FastMonodomainSolver split step for CDNA4 (gfx950).
//
// One Strang step (core/src/specialized_solver/fast_monodomain_solver/
// fast_monodomain_solver_compute.tpp:110-126) is
//     reaction half (Heun, t)  ->  diffusion (one tridiagonal solve, dt1D)
//     ->  reaction half (Heun, t + dt/2)
// Storage is state-major (SoA): Vm is one FP64 array over all nodes of all
// fibres (fibres concatenated), the remaining NS-1 states are NS-1 arrays of
// the gate type G (double, or float for the mixed-precision form), gate k at
// gates + k * gate_stride (the node count rounded up to 64, so that every
// array starts on a 512-byte boundary).
//
// Kernels here:
//   k_stim_flags      one thread per fibre: stimulation state machine per Heun step
//                     (k_stim_flags_pair: for two consecutive reaction halves)
//   k_reaction        one node per lane, coalesced SoA streaming, n0D Heun steps
//                     (PAIR: the second half of step q and the first of step q+1,
//                     which follow each other without a solve between them, in
//                     one pass over the states)
//   k_reaction_fibre  the same with one workgroup per fibre (large models)
//   k_thomas_serial   one lane per fibre, the reference's Thomas recurrence (bit-exact)
//   k_thomas_part     one workgroup per fibre, partitioned Thomas + PCR in LDS
//   k_fused           one workgroup per fibre, many split steps; Vm lives in LDS,
//                     gates in registers (RESIDENT) or streamed from HBM
#pragma once

#include "fms_math.h"
#include <stdint.h>

namespace fms {

// ----------------------------------------------------------------------------
// device views
// ----------------------------------------------------------------------------

struct DevFibers {
  int n_fibers;
  long long n_total;
  long long gate_stride;     // distance between two gate arrays: n_total rounded up to 64 nodes
  const int *fiber_off;      // [F+1] first node of each fibre
  const int *node_fiber;     // [N] fibre of each node
  const double *elem_len;    // [N-F] element lengths, fibre f starts at fiber_off[f]-f
  const double *prefactor;   // [F] sigma/(Am*Cm)
  const double *params;      // [F*NP]
  const int *stim_begin;     // [F] first stimulated node (fibre-local)
  const int *stim_end;       // [F] one past the last stimulated node
};

struct DevStim {
  int mode;                  // 0 none, 1 firing table (clamp Vm), 2 current pulse
  // firing table, reference semantics (fast_monodomain_solver_compute.tpp:509-627)
  int fe_rows, fe_cols;
  const unsigned char *firing_events;  // [rows*cols]
  const int *motor_unit;
  const double *freq, *repeat, *enable_begin;
  const double *jitter;
  const int *jitter_off, *jitter_cnt;
  double *last_check, *cur_jitter;
  int *jitter_idx;
  unsigned char *currently, *has_been;
  int *log_count;            // stimulation log (StimulationLogging::logStimulationBegin)
  double *log_time;
  int *log_fiber;
  int log_cap;
  // current pulse (extension)
  const double *onset, *duration, *amplitude;
  int stim_param;
  // onlyComputeIfHasBeenStimulated: a fibre's reaction runs from the Heun step
  // of its first stimulation on (fast_monodomain_solver_compute.tpp:201-205)
  int only_if_stimulated;
  // the reference's point-buffer order (vc_width > 0): nodes are swept in
  // buffers of vc_width consecutive nodes of the concatenated fibres; buffer b
  // belongs to fibre floor(b * vc_factor) (vc_factor = width / nodes of fibre 0,
  // compute.tpp:158-162), a fibre's stimulation is checked in its centre buffer
  // vc_centre[f] only (-1: none, compute.tpp:163-166), and the has-been flag a
  // buffer sees is the one set by the buffers swept before it (compute.tpp:172-205)
  int vc_width;
  double vc_factor;
  const int *vc_centre;
};

// ----------------------------------------------------------------------------
// Heun step of one node, reference form (03_generator_vc.cpp:1017-1130):
//   y* = y + dt f(y); [clamp y*_0]; y += 0.5 dt (f(y) + f(y*)); [clamp y_0]
// The arithmetic is FP64 in both precisions; G is the storage type of the
// gates (double, or float for the mixed form: gates are rounded to FP32 once
// per Heun step, Vm and the rates stay FP64, so the rate functions'
// removable singularities such as alpha_m at V = -50 mV are no more likely to
// be hit than in the FP64 reference).
// ----------------------------------------------------------------------------
// ALG: also return the algebraics of the second evaluation (the reference's
// algebraicsForTransfer); a compile-time switch so that the 71 algebraics of a
// large model are not kept live in the hot variant.
template <class M, class G, class E, bool ALG = false>
FMS_HD void heun_node(double &V, G *g, const double *p, double t, double dt, bool clamp, double vstim,
                      double *alg_out = nullptr) {
  // the rates are evaluated in FD (FP64, divisions via div_normal)
  FD y[M::NS], r[M::NS], ys[M::NS], rs[M::NS], a[M::NA], as[M::NA], pp[M::NP > 0 ? M::NP : 1];
#pragma unroll
  for (int k = 0; k < M::NP; k++) pp[k] = FD(p[k]);
  y[0] = FD(V);
#pragma unroll
  for (int k = 1; k < M::NS; k++) y[k] = FD((double)g[k - 1]);
  M::template rhs<FD, E>(y, pp, t, r, a);
#pragma unroll
  for (int k = 0; k < M::NS; k++) ys[k] = FD(y[k].v + dt * r[k].v);
  if (clamp) ys[0] = FD(vstim);
  M::template rhs<FD, E>(ys, pp, t, rs, as);
#pragma unroll
  for (int k = 0; k < M::NS; k++) y[k] = FD(y[k].v + 0.5 * dt * (r[k].v + rs[k].v));
  if (clamp) y[0] = FD(vstim);
  V = y[0].v;
#pragma unroll
  for (int k = 1; k < M::NS; k++) g[k - 1] = (G)y[k].v;
  if constexpr (ALG) {
#pragma unroll
    for (int k = 0; k < M::NA; k++) alg_out[k] = as[k].v;
  }
}

// ----------------------------------------------------------------------------
// stimulation state machine of one fibre at time t
// (fast_monodomain_solver_compute.tpp:529-623)
// ----------------------------------------------------------------------------
__device__ __forceinline__ int stim_check(const DevStim &s, int f, double t) {
  const double freq = s.freq[f];
  const double repeat = s.repeat[f];
  const double begin = s.enable_begin[f];
  int check = 0;
  if (t >= s.last_check[f] + 1. / (freq + s.cur_jitter[f]) && t >= begin - 1e-13) {
    check = 1;
    if (repeat != 0 && t - (s.last_check[f] + 1. / (freq + s.cur_jitter[f])) > repeat) {
      s.last_check[f] += 1. / (freq + s.cur_jitter[f]);
      double jf = 0.0;
      if (s.jitter_cnt[f] > 0) jf = s.jitter[s.jitter_off[f] + s.jitter_idx[f] % s.jitter_cnt[f]];
      s.cur_jitter[f] = jf * freq;
      s.jitter_idx[f]++;
      check = 0;
    }
  }
  const int idx = (int)round(t * freq);
  const unsigned long row = (unsigned long)(long)idx % (unsigned long)s.fe_rows;
  const int col = s.motor_unit[f] % s.fe_cols;
  const int stim = check && s.firing_events[row * s.fe_cols + col];
  if (stim) {
    s.has_been[f] = 1;
    if (!s.currently[f]) {
      s.currently[f] = 1;
      if (s.log_count) {
        const int k = atomicAdd(s.log_count, 1);
        if (k < s.log_cap) {
          s.log_time[k] = t;
          s.log_fiber[k] = f;
        }
      }
    }
  } else {
    s.currently[f] = 0;
  }
  return stim;
}

__device__ __forceinline__ int current_on(const DevStim &s, int f, double t) {
  return t >= s.onset[f] - 1e-13 && t < s.onset[f] + s.duration[f] - 1e-13;
}

// per Heun step flag byte: bit 0 = stimulate, bit 1 = compute the reaction
// (the fibre has been stimulated by this step), bit 2 = the fibre had been
// stimulated before the reaction half began (point-buffer order)
constexpr unsigned char FLAG_STIM = 1, FLAG_COMPUTE = 2, FLAG_HB0 = 4;

__device__ __forceinline__ unsigned char stim_flag(const DevStim &s, int f, double t) {
  int stim = 0;
  if (s.vc_width > 0 && s.vc_centre[f] < 0) {
    // no centre buffer: the reference never checks this fibre's stimulation
  } else if (s.mode == 1) {
    stim = stim_check(s, f, t);
  } else if (s.mode == 2) {
    stim = current_on(s, f, t);
    if (stim) s.has_been[f] = 1;
  }
  const bool compute = !s.only_if_stimulated || s.has_been[f];
  return (unsigned char)((stim ? FLAG_STIM : 0) | (compute ? FLAG_COMPUTE : 0));
}

// ----------------------------------------------------------------------------
// The split-step kernels read their times from the device step table
// (4 doubles per split step: startTime, dtSplit, dt0D, dt1D) through a cursor
// in device memory: cursor[0] = row of the first step of the current launch
// sequence, cursor[1] = the spike detector's step counter before row 0,
// cursor[2] = the table's address. A captured graph of G steps therefore
// replays over successive rows (k_cursor_advance) of whatever table the
// current call uploaded.
// ----------------------------------------------------------------------------
// spike detector state (optional): upward crossings of the threshold by Vm
// after each split step
struct DevSpikes {
  int enabled;
  double threshold;
  double *vm_prev;      // [N]
  int *first_step;      // [N]
  int *count;           // [N]
  int step0;            // fused kernel: global index of the step before the launch's first step
};

__device__ __forceinline__ void spike_update(const DevSpikes &spk, long long i, double v, int step) {
  const double prev = spk.vm_prev[i];
  if (prev < spk.threshold && v >= spk.threshold) {
    spk.count[i]++;
    if (spk.first_step[i] < 0) spk.first_step[i] = step;
  }
  spk.vm_prev[i] = v;
}

struct StepRef {
  const long long *cursor;
  int q;   // step within the launch sequence
  __device__ __forceinline__ const double *row() const {
    return reinterpret_cast<const double *>(cursor[2]) + 4 * (cursor[0] + q);
  }
};

// start time of a reaction half: the first half at startTime, the second at
// startTime + dtSplit/2 (fast_monodomain_solver_compute.tpp:110-126)
__device__ __forceinline__ double half_start(const double *row, int half) {
  return half ? row[0] + 0.5 * row[1] : row[0];
}

__global__ void k_cursor_set(long long *cursor, long long row, long long step_counter, const double *table) {
  if (threadIdx.x == 0) {
    cursor[0] = row;
    cursor[1] = step_counter;
    cursor[2] = reinterpret_cast<long long>(table);
  }
}

__global__ void k_cursor_advance(long long *cursor, int g) {
  if (threadIdx.x == 0) cursor[0] += g;
}

// one thread per fibre: flags[f*n0D + j] for the n0D Heun steps of a half step
__global__ void k_stim_flags(DevStim st, int n_fibers, unsigned char *flags, StepRef sr, int half, int n0D) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n_fibers) return;
  const double *row = sr.row();
  const double t0 = half_start(row, half), dt = row[2];
  const unsigned char hb0 = st.has_been[f] ? FLAG_HB0 : 0;
  for (int j = 0; j < n0D; j++) flags[(long long)f * n0D + j] = stim_flag(st, f, t0 + j * dt) | hb0;
}

// flags of two consecutive reaction halves, the second half of step q and the
// first half of step q+1, in the state machine's time order: flags[f*n0D + j]
// and flags[(n_fibers + f)*n0D + j]
__global__ void k_stim_flags_pair(DevStim st, int n_fibers, unsigned char *flags, StepRef sr, int n0D) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n_fibers) return;
  const double *row = sr.row();
  const double ta = half_start(row, 1), dta = row[2];
  const unsigned char hba = st.has_been[f] ? FLAG_HB0 : 0;
  for (int j = 0; j < n0D; j++) flags[(long long)f * n0D + j] = stim_flag(st, f, ta + j * dta) | hba;
  const double tb = half_start(row + 4, 0), dtb = row[4 + 2];
  const unsigned char hbb = st.has_been[f] ? FLAG_HB0 : 0;
  for (int j = 0; j < n0D; j++) flags[((long long)n_fibers + f) * n0D + j] = stim_flag(st, f, tb + j * dtb) | hbb;
}

// Point-buffer order (DevStim::vc_width > 0) with onlyComputeIfHasBeenStimulated:
// whether node i computes Heun step j of a reaction half. Its buffer b = i / W
// belongs to fibre fb = floor(b * W / n_0); buffers before fb's centre buffer
// see the has-been flag of the half's start, the centre buffer the flag after
// its own check of step j, buffers after it the flag after the whole half.
// fl points at the half's flags [F][n0D].
__device__ __forceinline__ bool vc_compute(const DevStim &st, int n_fibers, const unsigned char *fl, int n0D,
                                           long long i, int j) {
  const long long b = i / st.vc_width;
  int fb = (int)((double)b * st.vc_factor);
  fb = fb < 0 ? 0 : (fb >= n_fibers ? n_fibers - 1 : fb);
  const unsigned char *ff = fl + (long long)fb * n0D;
  if (ff[j] & FLAG_HB0) return true;
  const long long cb = st.vc_centre[fb];
  if (cb < 0 || b < cb) return false;
  return (b == cb ? ff[j] : ff[n0D - 1]) & FLAG_COMPUTE;
}

// ----------------------------------------------------------------------------
// reaction: one node per lane, n0D Heun steps, SoA coalesced streaming
//
// The heavy reaction code runs with every lane of the wave active: lanes past
// the last node recompute the last node and do not store, and a Heun step that
// a fibre skips (onlyComputeIfHasBeenStimulated) runs with dt = 0 on its lanes,
// which leaves every state unchanged (y + 0*r = y), instead of branching around
// it per lane. Waves in which no lane computes skip the step with a wave-uniform
// branch. Per-lane branches around the large models' reaction code (thousands
// of VALU instructions under register pressure) gave wrong and run-to-run
// different results on gfx950; this form gives the oracle's bits.
// ALG: also return the algebraics of the last Heun step (algebraicsForTransfer).
// PAIR: the two reaction halves between two diffusion solves in one pass, the
// second half of step q (flags[0 .. F*n0D)) and the first half of step q+1
// (flags[F*n0D .. 2*F*n0D)), with the spike detector of step q between them:
// the states are read and written once for both, the arithmetic is unchanged.
// ----------------------------------------------------------------------------
template <class M, class G, class E, bool ALG, bool PAIR = false>
__global__ __launch_bounds__(256) void k_reaction(DevFibers fb, DevStim st, double *__restrict__ vm,
                                                  G *__restrict__ gates, const unsigned char *__restrict__ flags,
                                                  StepRef sr, int half, int n0D, double vstim,
                                                  double *__restrict__ alg_transfer, const int *__restrict__ alg_idx,
                                                  int n_alg_transfer, DevSpikes spk) {
  static_assert(!(ALG && PAIR), "algebraics are returned by an unpaired last half");
  const long long N = fb.n_total;
  const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i0 < N;
  const long long i = valid ? i0 : N - 1;
  const int f = fb.node_fiber[i];
  const int li = (int)(i - fb.fiber_off[f]);
  const double *row = sr.row();
  const double t0 = half_start(row, PAIR ? 1 : half), dt = row[2];
  const double t1 = PAIR ? half_start(row + 4, 0) : 0.0, dt1 = PAIR ? row[4 + 2] : 0.0;
  const int nj = PAIR ? 2 * n0D : n0D;
  const int spike_step = PAIR ? (int)(sr.cursor[1] + sr.cursor[0] + sr.q + 1) : 0;
  // whether this lane's Heun step j computes (the per-fibre flag, or the
  // reference's point-buffer rule)
  auto lane_computes = [&](int j) {
    const bool second = PAIR && j >= n0D;
    const int jj = second ? j - n0D : j;
    const unsigned char *hf = flags + (second ? (long long)fb.n_fibers * n0D : 0LL);
    return st.vc_width > 0 && st.only_if_stimulated ? vc_compute(st, fb.n_fibers, hf, n0D, i, jj)
                                                    : (hf[(long long)f * n0D + jj] & FLAG_COMPUTE) != 0;
  };
  // a wave none of whose lanes computes a Heun step of this launch (fibres
  // before their first stimulation under onlyComputeIfHasBeenStimulated) leaves
  // every state as it is and stores no algebraics: only the spike detector
  // reads Vm
  bool any = false;
  for (int j = 0; j < nj && !any; j++) any = __builtin_amdgcn_ballot_w64(lane_computes(j)) != 0;
  if (!any) {
    if constexpr (PAIR) {
      if (spk.enabled && valid) spike_update(spk, i, vm[i], spike_step);
    }
    return;
  }
  double p[M::NP > 0 ? M::NP : 1];
#pragma unroll
  for (int k = 0; k < M::NP; k++) p[k] = fb.params[(long long)f * M::NP + k];
  const bool inset = li >= fb.stim_begin[f] && li < fb.stim_end[f];
  double V = vm[i];
  G g[M::NS - 1];
#pragma unroll
  for (int k = 0; k < M::NS - 1; k++) g[k] = gates[(long long)k * fb.gate_stride + i];
  const double pbase = M::NP > 0 ? p[st.stim_param] : 0.0;
  double alg[ALG ? M::NA : 1] = {};
  bool last_computes = false;   // the last Heun step computes: its algebraics are stored
  for (int j = 0; j < nj; j++) {
    const bool second = PAIR && j >= n0D;
    const int jj = second ? j - n0D : j;
    const unsigned char *hf = flags + (second ? (long long)fb.n_fibers * n0D : 0LL);
    const unsigned char fl = hf[(long long)f * n0D + jj];
    const bool compute = lane_computes(j);
    last_computes = compute;
    if (__builtin_amdgcn_ballot_w64(compute) != 0) {   // wave-uniform
      const int on = fl & FLAG_STIM;
      const bool clamp = compute && st.mode == 1 && on && inset;
      const double dtj = compute ? (second ? dt1 : dt) : 0.0;
      const double tj = second ? t1 + jj * dt1 : t0 + jj * dt;
      if (st.mode == 2 && M::NP > 0) p[st.stim_param] = (on && inset) ? st.amplitude[f] : pbase;
      if constexpr (ALG) {
        heun_node<M, G, E, true>(V, g, p, tj, dtj, clamp, vstim, alg);
      } else {
        heun_node<M, G, E, false>(V, g, p, tj, dtj, clamp, vstim);
      }
    }
    if constexpr (PAIR) {
      if (j == n0D - 1 && spk.enabled && valid) spike_update(spk, i, V, spike_step);
    }
  }
  if (valid) {
    vm[i] = V;
#pragma unroll
    for (int k = 0; k < M::NS - 1; k++) gates[(long long)k * fb.gate_stride + i] = g[k];
    if constexpr (ALG) {
      // a node whose last Heun step did not compute keeps the algebraics it had
      // (the reference calls no generated code for a skipped point buffer)
      if (last_computes)
        for (int k = 0; k < n_alg_transfer; k++) {
          const int a = alg_idx[k];
          double v = 0;
#pragma unroll
          for (int q = 0; q < M::NA; q++)
            if (q == a) v = alg[q];
          if (!M::alg_is_param(a)) alg_transfer[(long long)k * N + i] = v;
        }
    }
  }
}

// ----------------------------------------------------------------------------
// reaction with one workgroup per fibre (large models): the fibre's nodes in
// passes of T lanes aligned to 32-node segments of the arrays (lanes outside
// the fibre shadow its first or last node and store nothing), the fibre's
// flags and parameters uniform over the workgroup.
// PAIR as for k_reaction.
// ----------------------------------------------------------------------------
template <class M, class G, class E, int T, bool PAIR = false>
__global__ __launch_bounds__(T) void k_reaction_fibre(DevFibers fb, DevStim st, double *__restrict__ vm,
                                                      G *__restrict__ gates, const unsigned char *__restrict__ flags,
                                                      StepRef sr, int half, int n0D, double vstim,
                                                      const int *__restrict__ order, DevSpikes spk) {
  const int f = order ? order[blockIdx.x] : (int)blockIdx.x;
  const int off = fb.fiber_off[f];
  const int n = fb.fiber_off[f + 1] - off;
  const double *row = sr.row();
  const double t0 = half_start(row, PAIR ? 1 : half), dt = row[2];
  const double t1 = PAIR ? half_start(row + 4, 0) : 0.0, dt1 = PAIR ? row[4 + 2] : 0.0;
  const int sb = fb.stim_begin[f], se = fb.stim_end[f];
  const double amp = st.mode == 2 ? st.amplitude[f] : 0.0;
  const int nj = PAIR ? 2 * n0D : n0D;
  const int spike_step = PAIR ? (int)(sr.cursor[1] + sr.cursor[0] + sr.q + 1) : 0;
  const int lead = off & 31;
  // a fibre none of whose Heun steps in this launch computes (before its first
  // stimulation under onlyComputeIfHasBeenStimulated) keeps its states: only the
  // spike detector reads Vm
  bool any = false;
  for (int j = 0; j < nj; j++) {
    const bool second = PAIR && j >= n0D;
    const int jj = second ? j - n0D : j;
    any = any || (flags[(second ? (long long)fb.n_fibers * n0D : 0LL) + (long long)f * n0D + jj] & FLAG_COMPUTE);
  }
  if (!__builtin_amdgcn_readfirstlane((int)any)) {
    if constexpr (PAIR) {
      if (spk.enabled)
        for (int base = -lead; base < n; base += T) {
          const int lr = base + (int)threadIdx.x;
          if (lr >= 0 && lr < n) spike_update(spk, off + lr, vm[off + lr], spike_step);
        }
    }
    return;
  }
  // passes start on a 32-node boundary of the concatenated arrays (a 256-byte
  // FP64 or 128-byte FP32 segment): lanes before the fibre's first node shadow
  // it, lanes past its end shadow its last node, and shadows store nothing. A
  // fibre's node count plus its offset within the segment rarely needs one more
  // pass (1481 nodes: 24 passes either way).
  for (int base = -lead; base < n; base += T) {
    const int lr = base + (int)threadIdx.x;
    const bool valid = lr >= 0 && lr < n;
    const int li = lr < 0 ? 0 : (lr < n ? lr : n - 1);
    const long long i = off + li;
    const bool inset = li >= sb && li < se;
    double p[M::NP > 0 ? M::NP : 1];
#pragma unroll
    for (int k = 0; k < M::NP; k++) p[k] = fb.params[(long long)f * M::NP + k];
    const double pbase = M::NP > 0 ? p[st.stim_param] : 0.0;
    double V = vm[i];
    G g[M::NS - 1];
#pragma unroll
    for (int k = 0; k < M::NS - 1; k++) g[k] = gates[(long long)k * fb.gate_stride + i];
    for (int j = 0; j < nj; j++) {
      const bool second = PAIR && j >= n0D;
      const int jj = second ? j - n0D : j;
      const int fl = __builtin_amdgcn_readfirstlane(
          (int)flags[(second ? (long long)fb.n_fibers * n0D : 0LL) + (long long)f * n0D + jj]);
      if (fl & FLAG_COMPUTE) {   // wave-uniform: one fibre per workgroup
        const int on = fl & FLAG_STIM;
        const bool clamp = st.mode == 1 && on && inset;
        if (st.mode == 2 && M::NP > 0) p[st.stim_param] = (on && inset) ? amp : pbase;
        heun_node<M, G, E, false>(V, g, p, second ? t1 + jj * dt1 : t0 + jj * dt, second ? dt1 : dt, clamp, vstim);
      }
      if constexpr (PAIR) {
        if (j == n0D - 1 && spk.enabled && valid) spike_update(spk, i, V, spike_step);
      }
    }
    if (valid) {
      vm[i] = V;
#pragma unroll
      for (int k = 0; k < M::NS - 1; k++) gates[(long long)k * fb.gate_stride + i] = g[k];
    }
  }
}

// ----------------------------------------------------------------------------
// Row i of the split-step diffusion system for one fibre, exactly as
// fast_monodomain_solver_compute.tpp:336-429 assembles it (consistent P1 mass):
//   IE: (K - M/dt) x = -(M/dt) u        CN: (K/2 - M/dt) x = (-K/2 - M/dt) u
// u points at the fibre's Vm (global or LDS), h at its element lengths.
// ----------------------------------------------------------------------------
FMS_HD void assemble_row(int i, int n, const double *u, const double *h, double prefactor, double dt, bool ie,
                         double &a, double &b, double &c, double &d) {
  a = 0; b = 0; c = 0; d = 0;
  const double u_center = u[i];
  if (i > 0) {
    const double u_previous = u[i - 1];
    const double h_left = h[i - 1];
    const double k_left = div_normal(1., h_left) * (1) * prefactor;
    const double m_left = div_normal(h_left * 1., 6);
    const double k_right = div_normal(1., h_left) * (-1) * prefactor;
    const double m_right = div_normal(h_left * 1., 3);
    if (ie) {
      a = (k_left - div_normal(1, dt) * m_left);
      b += (k_right - div_normal(1, dt) * m_right);
      d += (-div_normal(1, dt) * m_left) * u_previous + (-div_normal(1, dt) * m_right) * u_center;
    } else {
      a = (k_left / 2. - div_normal(1, dt) * m_left);
      b += (k_right / 2. - div_normal(1, dt) * m_right);
      d += (-k_left / 2. - div_normal(1, dt) * m_left) * u_previous + (-k_right / 2. - div_normal(1, dt) * m_right) * u_center;
    }
  }
  if (i < n - 1) {
    const double u_next = u[i + 1];
    const double h_right = h[i];
    const double k_right = div_normal(1., h_right) * (1) * prefactor;
    const double m_right = div_normal(h_right * 1., 6);
    const double k_left = div_normal(1., h_right) * (-1) * prefactor;
    const double m_left = div_normal(h_right * 1., 3);
    if (ie) {
      c = (k_right - div_normal(1, dt) * m_right);
      b += (k_left - div_normal(1, dt) * m_left);
      d += (-div_normal(1, dt) * m_left) * u_center + (-div_normal(1, dt) * m_right) * u_next;
    } else {
      c = (k_right / 2. - div_normal(1, dt) * m_right);
      b += (k_left / 2. - div_normal(1, dt) * m_left);
      d += (-k_left / 2. - div_normal(1, dt) * m_left) * u_center + (-k_right / 2. - div_normal(1, dt) * m_right) * u_next;
    }
  }
}

// serial Thomas on one fibre, the reference recurrence (compute.tpp:431-486);
// cA/dA are n-long scratch arrays
FMS_HD void thomas_serial_fiber(double *u, int n, const double *h, double prefactor, double dt, bool ie,
                                double *cA, double *dA) {
  for (int i = 0; i < n; i++) {
    double a, b, c, d;
    assemble_row(i, n, u, h, prefactor, dt, ie, a, b, c, d);
    if (i == 0) {
      cA[i] = div_normal(c, b);
      dA[i] = div_normal(d, b);
    } else {
      if (i != n - 1) cA[i] = div_normal(c, b - cA[i - 1] * a);
      dA[i] = div_normal(d - dA[i - 1] * a, b - cA[i - 1] * a);
    }
  }
  u[n - 1] = dA[n - 1];
  double previous = dA[n - 1];
  for (int i = n - 2; i >= 0; i--) {
    const double r = dA[i] - cA[i] * previous;
    u[i] = r;
    previous = r;
  }
}

// one lane per fibre: bit-exact reference recurrence
__global__ __launch_bounds__(64) void k_thomas_serial(DevFibers fb, double *vm, double *cA, double *dA, StepRef sr,
                                                      int ie) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= fb.n_fibers) return;
  const double dt = sr.row()[3];
  const int off = fb.fiber_off[f];
  const int n = fb.fiber_off[f + 1] - off;
  thomas_serial_fiber(vm + off, n, fb.elem_len + off - f, fb.prefactor[f], dt, ie != 0, cA + off, dA + off);
}

// ----------------------------------------------------------------------------
// Partitioned tridiagonal solve of one fibre by a group of T threads (whole
// waves), Vm in LDS. Three levels:
//  1. Each of P <= T threads owns a contiguous chunk of >= 2 rows. A forward and
//     a backward elimination within the chunk (rows scaled to a unit diagonal)
//     express every row through the chunk's first and last unknowns; the
//     chunks' first and last rows form a tridiagonal system of 2T rows
//     (identity rows for idle threads), rows 2t and 2t+1 held by thread t.
//  2. Each wave's block of 128 rows is solved in registers for its right-hand
//     side and for its two couplings to the neighbouring blocks' unknowns xL
//     (the previous block's last) and xR (the next block's first): one odd-even
//     elimination step, then parallel cyclic reduction over the 64 even rows
//     with cross-lane shuffles (no LDS, no barrier). Every unknown of the block
//     becomes d - v xL - u xR.
//  3. The blocks' first and last unknowns form a chain of 2 W unknowns (W waves)
//     that every thread solves itself after one barrier; each thread then
//     recovers its two rows and its chunk's interior.
// Rows are assembled exactly as the reference does; only the elimination order
// differs from its serial recurrence. The matrix is strictly diagonally
// dominant for every dt (mass term), so no pivoting is needed.
//
// lv: n doubles of Vm in LDS (overwritten with the solution)
// red: pcr_lds_doubles<T>() doubles of LDS scratch (6 per wave)
// Requires n >= 2 and ceil(n / min(T, n/2)) <= MMAX. Ends with a barrier.
// GIVEN_RHS: lv holds the right-hand side itself (adjoint solves) instead of the
// Vm from which the right-hand side B*Vm is assembled.
// ----------------------------------------------------------------------------
// 1/b by the hardware reciprocal and two Newton steps (within about one ulp,
// not correctly rounded): the partitioned solve is held to a tolerance against
// the serial recurrence, not to its bits, and multiplies by reciprocals
__device__ __forceinline__ double recip_nr(double b) {
  const double r0 = __builtin_amdgcn_rcp(b);
  const double r1 = __builtin_fma(r0, __builtin_fma(-b, r0, 1.0), r0);
  return __builtin_fma(r1, __builtin_fma(-b, r1, 1.0), r1);
}

template <int T, int MMAX, bool GIVEN_RHS = false>
__device__ __forceinline__ void lds_thomas_part(double *lv, int n, const double *h, double prefactor, double dt,
                                                bool ie, double *red, int tid) {
  static_assert(T % 64 == 0 && T >= 64 && T <= 1024, "the solve works on whole waves");
  constexpr int W = T / 64;
  const int P = n / 2 < T ? n / 2 : T;
  double aa[MMAX], cc[MMAX], dd[MMAX];
  int lo = 0, cnt = 0;
  if (tid < P) {
    lo = (int)(((long long)tid * n) / P);
    cnt = (int)(((long long)(tid + 1) * n) / P) - lo;
  }
  double fa = 0, fc = 0, fd = 0, la = 0, lc = 0, ld = 0;
  if (cnt > 0) {
#pragma unroll
    for (int k = 0; k < MMAX; k++) {
      if (k < cnt) {
        double a, b, c, d;
        assemble_row(lo + k, n, lv, h, prefactor, dt, ie, a, b, c, d);
        if (GIVEN_RHS) d = lv[lo + k];
        if (k < 2) {
          const double rb = recip_nr(b);   // the first two rows share one reciprocal
          aa[k] = a * rb;
          cc[k] = c * rb;
          dd[k] = d * rb;
        } else {
          const double r = recip_nr(b - a * cc[k - 1]);
          dd[k] = r * (d - a * dd[k - 1]);
          aa[k] = -r * a * aa[k - 1];
          cc[k] = r * c;
        }
      }
    }
#pragma unroll
    for (int k = MMAX - 3; k >= 1; k--) {
      if (k <= cnt - 3) {
        dd[k] -= cc[k] * dd[k + 1];
        aa[k] -= cc[k] * aa[k + 1];
        cc[k] = -cc[k] * cc[k + 1];
      }
    }
    if (cnt >= 3) {
      const double r = recip_nr(1.0 - cc[0] * aa[1]);
      dd[0] = r * (dd[0] - cc[0] * dd[1]);
      aa[0] = r * aa[0];
      cc[0] = -r * cc[0] * cc[1];
    }
    fa = aa[0]; fc = cc[0]; fd = dd[0];
#pragma unroll
    for (int k = 0; k < MMAX; k++)
      if (k == cnt - 1) { la = aa[k]; lc = cc[k]; ld = dd[k]; }
  }

  // level 2. Row convention: a x_prev + x + c x_next + v xL + u xR = d.
  // Even row (chunk first) E, odd row (chunk last) O of this lane; the block's
  // couplings to the outside move to the v and u columns.
  const int lane = tid & 63, wave = tid >> 6;
  double aE = fa, cE = fc, dE = fd, vE = 0, uE = 0;
  double aO = la, cO = lc, dO = ld, vO = 0, uO = 0;
  if (lane == 0) { vE = aE; aE = 0; }
  if (lane == 63) { uO = cO; cO = 0; }
  // odd-even step: the even row eliminates x_{2l-1} (odd row of lane l-1) and
  // x_{2l+1} (its own odd row); the even rows then couple even unknowns only
  double A, C, D, V, U;
  {
    const int src = (lane + 63) & 63;
    double pa = __shfl(aO, src), pc = __shfl(cO, src), pd = __shfl(dO, src), pv = __shfl(vO, src),
           pu = __shfl(uO, src);
    if (lane == 0) pa = pc = pd = pv = pu = 0;
    const double al = -aE, ga = -cE;
    A = al * pa;
    C = ga * cO;
    const double r = recip_nr(1.0 + al * pc + ga * aO);
    D = r * (dE + al * pd + ga * dO);
    V = r * (vE + al * pv + ga * vO);
    U = r * (uE + al * pu + ga * uO);
    A *= r;
    C *= r;
  }
  // parallel cyclic reduction over the 64 even rows of the wave
#pragma unroll
  for (int s = 1; s < 64; s <<= 1) {
    const bool hm = lane >= s, hp = lane + s < 64;
    const int im = (lane - s) & 63, ip = (lane + s) & 63;
    const double Am = __shfl(A, im), Cm = __shfl(C, im), Dm = __shfl(D, im), Vm = __shfl(V, im), Um = __shfl(U, im);
    const double Ap = __shfl(A, ip), Cp = __shfl(C, ip), Dp = __shfl(D, ip), Vp = __shfl(V, ip), Up = __shfl(U, ip);
    // a row without a neighbour at distance s has al (ga) = 0, which zeroes the
    // wrapped-around lane's (finite) coefficients: no select per operand
    const double al = hm ? -A : 0.0, ga = hp ? -C : 0.0;
    const double r = recip_nr(1.0 + al * Cm + ga * Ap);
    const double nA = al * Am, nC = ga * Cp;
    const double nD = D + al * Dm + ga * Dp;
    const double nV = V + al * Vm + ga * Vp;
    const double nU = U + al * Um + ga * Up;
    A = nA * r; C = nC * r; D = nD * r; V = nV * r; U = nU * r;
  }
  // x_{2l} = D - V xL - U xR. Block interface: first unknown (lane 0's even
  // row) and last unknown (lane 63's odd row, x_127 = dO - aO x_126 - vO xL - uO xR)
  if (lane == 0) {
    red[6 * wave + 0] = D; red[6 * wave + 1] = V; red[6 * wave + 2] = U;
  }
  if (lane == 63) {
    red[6 * wave + 3] = dO - aO * D; red[6 * wave + 4] = vO - aO * V; red[6 * wave + 5] = uO - aO * U;
  }
  __syncthreads();
  // level 3: the chain F_w = df - vf L_{w-1} - uf F_{w+1}, L_w = dl - vl L_{w-1} - ul F_{w+1}
  // (L_{-1} = F_W = 0), solved by every thread: L_{w-1} = al_w' + be_w' F_w from the left,
  // F_w = ga_w + de_w F_{w+1}, L_w = al_w + be_w F_{w+1}
  double xL = 0.0, xR = 0.0;
  if (W > 1) {
    double gam[W], del[W], alp[W], bet[W];
    double al_prev = 0.0, be_prev = 0.0;
#pragma unroll
    for (int w = 0; w < W; w++) {
      const double df = red[6 * w], vf = red[6 * w + 1], uf = red[6 * w + 2];
      const double dl = red[6 * w + 3], vl = red[6 * w + 4], ul = red[6 * w + 5];
      const double r = recip_nr(1.0 + vf * be_prev);
      gam[w] = r * (df - vf * al_prev);
      del[w] = -r * uf;
      alp[w] = dl - vl * (al_prev + be_prev * gam[w]);
      bet[w] = -vl * be_prev * del[w] - ul;
      al_prev = alp[w];
      be_prev = bet[w];
    }
    double Fn = 0.0;   // F_{w+1}
#pragma unroll
    for (int w = W - 1; w >= 0; w--) {
      const double Lw = alp[w] + bet[w] * Fn;
      const double Fw = gam[w] + del[w] * Fn;
      if (w == wave + 1) xR = Fw;
      if (w == wave - 1) xL = Lw;
      Fn = Fw;
    }
  }
  const double xe = D - V * xL - U * xR;
  const double xn = __shfl(xe, (lane + 1) & 63);
  const double xo = dO - vO * xL - uO * xR - aO * xe - cO * xn;
  if (cnt > 0) {
    lv[lo] = xe;
    lv[lo + cnt - 1] = xo;
#pragma unroll
    for (int k = 1; k < MMAX - 1; k++)
      if (k < cnt - 1) lv[lo + k] = dd[k] - aa[k] * xe - cc[k] * xo;
  }
  __syncthreads();
}

// LDS doubles of the partitioned solve's block interfaces (6 per wave)
template <int T> constexpr int pcr_lds_doubles() { return 6 * (T / 64); }

// one workgroup per fibre: load Vm to LDS, partitioned solve, store. The
// element lengths are read from global memory inside the elimination (staging
// them in LDS with Vm measured slower here: 0.0243 against 0.0228 ms per launch
// on 1024 x 1481 nodes)
template <int T, int MMAX>
__global__ __launch_bounds__(T) void k_thomas_part(DevFibers fb, double *vm, StepRef sr, int ie) {
  extern __shared__ double sh[];
  const int f = blockIdx.x;
  const double dt = sr.row()[3];
  const int tid = threadIdx.x;
  const int off = fb.fiber_off[f];
  const int n = fb.fiber_off[f + 1] - off;
  double *red = sh;
  double *lv = sh + pcr_lds_doubles<T>();
  for (int i = tid; i < n; i += T) lv[i] = vm[off + i];
  __syncthreads();
  lds_thomas_part<T, MMAX>(lv, n, fb.elem_len + off - f, fb.prefactor[f], dt, ie != 0, red, tid);
  for (int i = tid; i < n; i += T) vm[off + i] = lv[i];
}

// model right-hand side at n independent points (SoA in and out)
template <class M, class E>
__global__ __launch_bounds__(256) void k_eval_rhs(const double *__restrict__ ys, const double *__restrict__ ps,
                                                  double t, int n, double *__restrict__ rs, double *__restrict__ as) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  FD y[M::NS], r[M::NS], a[M::NA], p[M::NP > 0 ? M::NP : 1];
#pragma unroll
  for (int k = 0; k < M::NS; k++) y[k] = FD(ys[(long long)k * n + i]);
#pragma unroll
  for (int k = 0; k < M::NP; k++) p[k] = FD(ps[(long long)k * n + i]);
  M::template rhs<FD, E>(y, p, t, r, a);
#pragma unroll
  for (int k = 0; k < M::NS; k++) rs[(long long)k * n + i] = r[k].v;
  if (as)
#pragma unroll
    for (int k = 0; k < M::NA; k++) as[(long long)k * n + i] = a[k].v;
}

// n_steps Heun steps at n independent points (SoA in place)
template <class M, class E>
__global__ __launch_bounds__(256) void k_eval_heun(double *__restrict__ ys, const double *__restrict__ ps,
                                                   double t0, double dt, int n_steps, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double V = ys[i];
  double g[M::NS - 1], p[M::NP > 0 ? M::NP : 1];
#pragma unroll
  for (int k = 1; k < M::NS; k++) g[k - 1] = ys[(long long)k * n + i];
#pragma unroll
  for (int k = 0; k < M::NP; k++) p[k] = ps[(long long)k * n + i];
  for (int j = 0; j < n_steps; j++) heun_node<M, double, E, false>(V, g, p, t0 + j * dt, dt, false, 0.0);
  ys[i] = V;
#pragma unroll
  for (int k = 1; k < M::NS; k++) ys[(long long)k * n + i] = g[k - 1];
}

// spike detector: upward crossings of the threshold after each split step
__global__ void k_spikes(long long N, const double *vm, double *vm_prev, int *first_step, int *count, double th,
                         StepRef sr) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int step = (int)(sr.cursor[1] + sr.cursor[0] + sr.q + 1);
  const double v = vm[i];
  const double p = vm_prev[i];
  if (p < th && v >= th) {
    count[i]++;
    if (first_step[i] < 0) first_step[i] = step;
  }
  vm_prev[i] = v;
}

}  // namespace fms
