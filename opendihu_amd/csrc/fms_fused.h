// Fused multi-step split-step kernel: one workgroup per fibre runs many
// Strang steps without leaving the chip.
//
//  * Vm of the fibre stays in LDS for the whole launch (<= 4096 nodes, 32 KiB).
//  * Gates either stay in registers (RESIDENT: thread t owns nodes t, t+T, ...,
//    NPT of them) or stream from HBM once per split step (large models).
//  * The second reaction half of step s and the first half of step s+1 run in
//    one pass over the nodes, so each split step reads and writes every gate
//    once and Vm never touches HBM between launches.
//  * Diffusion is the partitioned LDS solve (lds_thomas_part) or, for bit-exact
//    runs, the reference's serial recurrence executed by thread 0 in LDS.
//  * The stimulation state machine of the fibre runs on thread 0 in time order
//    and hands the per-Heun-step flags to the workgroup through LDS.
// Uneven fibre lengths: the host orders fibres longest first (LPT) so the
// dispatcher fills CUs evenly; short fibres simply leave threads idle.
#pragma once

#include "fms_kernels.h"

namespace fms {

// sensor traces and the adjoint tape (both optional)
struct DevTrace {
  const int *sensor_slot;   // [N] sensor number of each node, -1 elsewhere; nullptr: no traces
  double *trace;            // [K][F][n_sensors], this launch's first step first
  int n_sensors;
  double *tape;             // [K][2*NS+1][N]: y before half 1 (NS), y after half 1 (NS: Vm
                            // before the diffusion and the gates), Vm after it (1); nullptr: none
  int raise_priority;       // the waves win issue over a kernel running beside them
};

template <int NS, class G>
__device__ __forceinline__ void tape_store(const DevTrace &tr, long long N, int s, int slot0, long long i, double V,
                                           const G *g, int nstore) {
  double *t = tr.tape + ((long long)s * (2 * NS + 1) + slot0) * N + i;
  t[0] = V;
  for (int k = 1; k < nstore; k++) t[(long long)k * N] = (double)g[k - 1];
}

// tape entry with the gates read from their HBM copy (streamed path); g == nullptr
// stores Vm only
template <int NS, class G>
__device__ __forceinline__ void tape_store_streamed(const DevTrace &tr, long long N, long long gstride, int s,
                                                    int slot0, long long i, double V, const G *gates) {
  double *t = tr.tape + ((long long)s * (2 * NS + 1) + slot0) * N + i;
  t[0] = V;
  if (gates)
    for (int k = 1; k < NS; k++) t[(long long)k * N] = (double)gates[(long long)(k - 1) * gstride + i];
}

// One reaction half of the NPT nodes a thread owns (n0D Heun steps each, gates
// resident in registers). The node loop is the innermost one, so the nodes'
// independent Heun steps share one basic block that the scheduler interleaves
// (a node loop around the run-time loop over the Heun steps would give one
// block, and one dependency chain, per node). Each node takes its stimulus
// parameter by a select from the fibre's base parameters.
template <class M, class G, class E, int NPT>
__device__ __forceinline__ void fused_nodes_half(double *V, G (*g)[M::NS - 1], const double *p, const DevStim &st,
                                                 const bool *inset, const unsigned char *flags, double t0, double dt,
                                                 int n0D, double vstim) {
  for (int j = 0; j < n0D; j++) {
    // one fibre per workgroup: the flags are the same on every lane
    const int fl = __builtin_amdgcn_readfirstlane((int)flags[j]);
    if (!(fl & FLAG_COMPUTE)) continue;
    const int on = fl & FLAG_STIM;
#pragma unroll
    for (int q = 0; q < NPT; q++) {
      const bool clamp = st.mode == 1 && on && inset[q];
      const bool pulse = st.mode == 2 && on && inset[q];
      double pq[M::NP > 0 ? M::NP : 1];
#pragma unroll
      for (int k = 0; k < M::NP; k++) pq[k] = pulse && k == st.stim_param ? st.amplitude[0] : p[k];
      heun_node<M, G, E, false>(V[q], g[q], pq, t0 + j * dt, dt, clamp, vstim);
    }
  }
}

// One reaction half of one node with its gates streamed from HBM: load the
// gates, n0D Heun steps, store the gates (if `store`), return Vm. The caller
// runs it on every lane of the wave (lanes past the fibre's end pass
// store = false).
template <class M, class G, class E>
__device__ __forceinline__ double streamed_half_body(G *__restrict__ gates, long long gstride, long long node, bool store,
                                                     double V, const double *__restrict__ fparams, int stim_mode,
                                                     int stim_param, double amplitude, bool inset,
                                                     const unsigned char *flags, double t0, double dt, int n0D,
                                                     double vstim) {
  G g[M::NS - 1];
#pragma unroll
  for (int k = 0; k < M::NS - 1; k++) g[k] = gates[(long long)k * gstride + node];
  double p[M::NP > 0 ? M::NP : 1];
#pragma unroll
  for (int k = 0; k < M::NP; k++) p[k] = fparams[k];
  const double pbase = M::NP > 0 ? p[stim_param] : 0.0;
  for (int j = 0; j < n0D; j++) {
    const int fl = __builtin_amdgcn_readfirstlane((int)flags[j]);
    if (!(fl & FLAG_COMPUTE)) continue;
    const int on = fl & FLAG_STIM;
    const bool clamp = stim_mode == 1 && on && inset;
    if (stim_mode == 2 && M::NP > 0) p[stim_param] = (on && inset) ? amplitude : pbase;
    heun_node<M, G, E, false>(V, g, p, t0 + j * dt, dt, clamp, vstim);
  }
  if (store) {
#pragma unroll
    for (int k = 0; k < M::NS - 1; k++) gates[(long long)k * gstride + node] = g[k];
  }
  return V;
}

// Large models run the half as a called function: inlined into the fused
// kernel (beside its LDS solve and barriers) the 56-state model's reaction
// computed wrong, in some builds run-to-run different, states on gfx950 with
// hipcc of ROCm 7.2 under both the default and the ILP scheduler, while the
// same code as a called function, or in a kernel of its own (k_reaction,
// k_reaction_fibre), gives the oracle's bits. The call costs scratch spills
// (the callee has 256 VGPRs and no AGPRs), so for large models the split
// kernels are the faster exact schedule and the autotuner picks them. Small
// models (at most 8 states) inline it: the same bits as the call, 3.6 % faster
// (Hodgkin-Huxley, 4096 fibres of 400-2000 nodes).
template <class M, class G, class E>
__device__ __noinline__ double streamed_half_call(G *__restrict__ gates, long long gstride, long long node, bool store,
                                                  double V, const double *__restrict__ fparams, int stim_mode,
                                                  int stim_param, double amplitude, bool inset,
                                                  const unsigned char *flags, double t0, double dt, int n0D,
                                                  double vstim) {
  return streamed_half_body<M, G, E>(gates, gstride, node, store, V, fparams, stim_mode, stim_param, amplitude, inset,
                                     flags, t0, dt, n0D, vstim);
}

template <class M, class G, class E>
__device__ __forceinline__ double streamed_half(G *__restrict__ gates, long long gstride, long long node, bool store,
                                                double V, const double *__restrict__ fparams, int stim_mode,
                                                int stim_param, double amplitude, bool inset,
                                                const unsigned char *flags, double t0, double dt, int n0D,
                                                double vstim) {
  if constexpr (M::NS <= 8)
    return streamed_half_body<M, G, E>(gates, gstride, node, store, V, fparams, stim_mode, stim_param, amplitude,
                                       inset, flags, t0, dt, n0D, vstim);
  else
    return streamed_half_call<M, G, E>(gates, gstride, node, store, V, fparams, stim_mode, stim_param, amplitude,
                                       inset, flags, t0, dt, n0D, vstim);
}

// NPT > 0: gates resident in registers, NPT nodes per thread; NPT == 0: streamed
template <class M, class G, class E, int T, int MMAX, int NPT>
__global__ __launch_bounds__(T) void k_fused(DevFibers fb, DevStim st, double *__restrict__ vm,
                                             G *__restrict__ gates, const double *__restrict__ steps, int K,
                                             int n0D, double vstim, int ie, int serial, const int *__restrict__ order,
                                             DevSpikes spk, DevTrace tr) {
  extern __shared__ double sh[];
  if (tr.raise_priority) __builtin_amdgcn_s_setprio(2);
  const int f = order ? order[blockIdx.x] : (int)blockIdx.x;
  const int tid = threadIdx.x;
  const long long N = fb.n_total, GS = fb.gate_stride;
  const int off = fb.fiber_off[f];
  const int n = fb.fiber_off[f + 1] - off;
  const double pf = fb.prefactor[f];
  double *red = sh;   // block interfaces of the partitioned solve (pcr_lds_doubles), or the serial solve's cA/dA (16*T doubles)
  double *lv = sh + (serial ? 16 * T : pcr_lds_doubles<T>());   // n doubles: Vm
  double *h = lv + n;   // n - 1 element lengths, staged once for all the launch's solves
  unsigned char *sflags = (unsigned char *)(h + n);   // 2*n0D flags

  for (int i = tid; i < n; i += T) {
    lv[i] = vm[off + i];
    if (i < n - 1) h[i] = fb.elem_len[off - f + i];
  }

  double p[M::NP > 0 ? M::NP : 1];
#pragma unroll
  for (int k = 0; k < M::NP; k++) p[k] = fb.params[(long long)f * M::NP + k];
  const int sb = fb.stim_begin[f], se = fb.stim_end[f];
  // the current-pulse amplitude is read through a per-fibre view
  DevStim stf = st;
  if (st.mode == 2) stf.amplitude = st.amplitude + f;
  const double *fparams = fb.params + (long long)f * (M::NP > 0 ? M::NP : 1);
  const double amp = st.mode == 2 ? st.amplitude[f] : 0.0;

  constexpr int NQ = NPT > 0 ? NPT : 1;
  G greg[NQ][M::NS - 1];
  double vprev[NQ];
  int sfirst[NQ], scount[NQ];
  bool inset[NQ];
  if constexpr (NPT > 0) {
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      // lanes past the fibre's end shadow its last node (computed, never stored)
      const int i = tid + q * T < n ? tid + q * T : n - 1;
#pragma unroll
      for (int k = 0; k < M::NS - 1; k++) greg[q][k] = gates[(long long)k * GS + off + i];
      inset[q] = i >= sb && i < se;
      vprev[q] = 0;
      sfirst[q] = 0;
      scount[q] = 0;
      if (spk.enabled) {
        vprev[q] = spk.vm_prev[off + i];
        sfirst[q] = spk.first_step[off + i];
        scount[q] = spk.count[off + i];
      }
    }
  }
  // streamed gates with spike detection: bit b of `below` says whether the
  // thread's node of pass b (node b*T + tid) was below the threshold at the
  // last check; the count and first step are read and written only at a
  // crossing, the previous Vm once at the end of the launch
  unsigned below = 0;
  if constexpr (NPT == 0) {
    if (spk.enabled)
      for (int base = 0, b = 0; base + tid < n; base += T, b++)
        below |= (spk.vm_prev[off + base + tid] < spk.threshold ? 1u : 0u) << b;
  }

  // flags of the first half step of step 0
  if (tid == 0)
    for (int j = 0; j < n0D; j++) sflags[j] = stim_flag(st, f, steps[0] + j * steps[2]);
  __syncthreads();

  // pass 0: first reaction half of step 0
  {
    const double t0 = steps[0], dt0 = steps[2];
    // every lane runs the reaction: lanes past the fibre's end shadow its last
    // node and store nothing (a shadow may read that node's Vm before or after
    // its owner updates it; the shadow's result is discarded)
    if constexpr (NPT > 0) {
      double V[NQ];
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        const bool valid = tid + q * T < n;
        const int i = valid ? tid + q * T : n - 1;
        V[q] = lv[i];
        if (tr.tape && valid) tape_store<M::NS>(tr, N, 0, 0, off + i, V[q], greg[q], M::NS);
      }
      fused_nodes_half<M, G, E, NQ>(V, greg, p, stf, inset, sflags, t0, dt0, n0D, vstim);
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        const bool valid = tid + q * T < n;
        if (tr.tape && valid) tape_store<M::NS>(tr, N, 0, M::NS, off + tid + q * T, V[q], greg[q], M::NS);
        if (valid) lv[tid + q * T] = V[q];
      }
    } else {
      for (int base = 0; base < n; base += T) {
        const bool valid = base + tid < n;
        const int i = valid ? base + tid : n - 1;
        double V = lv[i];
        if (tr.tape && valid) tape_store_streamed<M::NS>(tr, N, GS, 0, 0, off + i, V, gates);
        V = streamed_half<M, G, E>(gates, GS, off + i, valid, V, fparams, st.mode, st.stim_param, amp, i >= sb && i < se,
                                   sflags, t0, dt0, n0D, vstim);
        if (tr.tape && valid) tape_store_streamed<M::NS>(tr, N, GS, 0, M::NS, off + i, V, gates);
        if (valid) lv[i] = V;
      }
    }
  }
  __syncthreads();

  for (int s = 0; s < K; s++) {
    const double *st_s = steps + 4 * s;
    const double t_mid = st_s[0] + 0.5 * st_s[1];
    const double dt0 = st_s[2], dt1 = st_s[3];
    const bool more = s + 1 < K;

    // diffusion (one solve per split step, as the reference)
    if (serial) {
      if (tid == 0) thomas_serial_fiber(lv, n, h, pf, dt1, ie != 0, red, red + n);
      __syncthreads();
    } else {
      lds_thomas_part<T, MMAX>(lv, n, h, pf, dt1, ie != 0, red, tid);
    }

    // flags: second half of step s, then first half of step s+1
    if (tid == 0) {
      for (int j = 0; j < n0D; j++) sflags[j] = stim_flag(st, f, t_mid + j * dt0);
      if (more) {
        const double t1 = st_s[4], dt0n = st_s[6];
        for (int j = 0; j < n0D; j++) sflags[n0D + j] = stim_flag(st, f, t1 + j * dt0n);
      }
    }
    __syncthreads();

    // the thread's nodes: second half of step s, traces and spikes, first half
    // of step s+1 (a shadow lane computes and stores nothing)
    if constexpr (NPT > 0) {
      double V[NQ];
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        const bool valid = tid + q * T < n;
        const int i = valid ? tid + q * T : n - 1;
        V[q] = lv[i];
        if (tr.tape && valid) tape_store<M::NS>(tr, N, s, 2 * M::NS, off + i, V[q], greg[q], 1);
      }
      fused_nodes_half<M, G, E, NQ>(V, greg, p, stf, inset, sflags, t_mid, dt0, n0D, vstim);
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        const bool valid = tid + q * T < n;
        const int i = valid ? tid + q * T : n - 1;
        if (tr.sensor_slot && valid) {
          const int slot = tr.sensor_slot[off + i];
          if (slot >= 0) tr.trace[((long long)s * fb.n_fibers + f) * tr.n_sensors + slot] = V[q];
        }
        if (spk.enabled) {
          const bool cross = vprev[q] < spk.threshold && V[q] >= spk.threshold;
          scount[q] += cross ? 1 : 0;
          sfirst[q] = (cross && sfirst[q] < 0) ? spk.step0 + s + 1 : sfirst[q];
          vprev[q] = V[q];
        }
        if (more && tr.tape && valid) tape_store<M::NS>(tr, N, s + 1, 0, off + i, V[q], greg[q], M::NS);
      }
      if (more) fused_nodes_half<M, G, E, NQ>(V, greg, p, stf, inset, sflags + n0D, st_s[4], st_s[6], n0D, vstim);
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        const bool valid = tid + q * T < n;
        if (more && tr.tape && valid) tape_store<M::NS>(tr, N, s + 1, M::NS, off + tid + q * T, V[q], greg[q], M::NS);
        if (valid) lv[tid + q * T] = V[q];
      }
    } else {
      // streamed gates (large models): one reaction half per pass over the
      // nodes, so that one pass holds the registers of one half only
      for (int base = 0; base < n; base += T) {
        const bool valid = base + tid < n;
        const int i = valid ? base + tid : n - 1;
        double V = lv[i];
        if (tr.tape && valid) tape_store_streamed<M::NS>(tr, N, GS, s, 2 * M::NS, off + i, V, (const G *)nullptr);
        V = streamed_half<M, G, E>(gates, GS, off + i, valid, V, fparams, st.mode, st.stim_param, amp, i >= sb && i < se,
                                   sflags, t_mid, dt0, n0D, vstim);
        if (tr.sensor_slot && valid) {
          const int slot = tr.sensor_slot[off + i];
          if (slot >= 0) tr.trace[((long long)s * fb.n_fibers + f) * tr.n_sensors + slot] = V;
        }
        if (spk.enabled && valid) {
          const unsigned bit = 1u << (base / T);
          if ((below & bit) && V >= spk.threshold) {
            spk.count[off + i]++;
            if (spk.first_step[off + i] < 0) spk.first_step[off + i] = spk.step0 + s + 1;
          }
          below = V < spk.threshold ? (below | bit) : (below & ~bit);
        }
        if (valid) lv[i] = V;
      }
      if (more) {
        for (int base = 0; base < n; base += T) {
          const bool valid = base + tid < n;
          const int i = valid ? base + tid : n - 1;
          double V = lv[i];
          if (tr.tape && valid) tape_store_streamed<M::NS>(tr, N, GS, s + 1, 0, off + i, V, gates);
          V = streamed_half<M, G, E>(gates, GS, off + i, valid, V, fparams, st.mode, st.stim_param, amp,
                                     i >= sb && i < se, sflags + n0D, st_s[4], st_s[6], n0D, vstim);
          if (tr.tape && valid) tape_store_streamed<M::NS>(tr, N, GS, s + 1, M::NS, off + i, V, gates);
          if (valid) lv[i] = V;
        }
      }
    }
    __syncthreads();
  }

  for (int i = tid; i < n; i += T) {
    vm[off + i] = lv[i];
    // streamed gates: the last check's Vm is the launch's final Vm
    if (NPT == 0 && spk.enabled) spk.vm_prev[off + i] = lv[i];
  }
  if constexpr (NPT > 0) {
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const int i = tid + q * T;
      if (i < n) {
#pragma unroll
        for (int k = 0; k < M::NS - 1; k++) gates[(long long)k * GS + off + i] = greg[q][k];
        if (spk.enabled) {
          spk.vm_prev[off + i] = vprev[q];
          spk.first_step[off + i] = sfirst[q];
          spk.count[off + i] = scount[q];
        }
      }
    }
  }
}

}  // namespace fms
