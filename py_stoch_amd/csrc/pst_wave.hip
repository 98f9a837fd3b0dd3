// pst_wave.hip -- the large-network kernel: one wavefront per simulation
// (engine in pst_engine_wave.hpp), mixed precision.
//
// Every wavefront takes simulations one at a time from the job's work queue
// (one atomic per simulation) and runs each to its end; jobs are visited in
// order.  LDS holds the state of the wavefront's current simulation.
#include <cstdio>
#include <cstring>

#include "pst_engine_wave.hpp"

namespace pst {

// Four wavefronts per SIMD: 128 VGPRs and ~400 bytes of scratch spill per lane
// instead of 227 VGPRs at two, which hides more latency than the spills cost
// (first 48 random networks: 1,508 -> 1,264 ms; three: 1,339 ms; five: 1,622 ms).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4)))
sim_wave_kernel(const WaveJob* __restrict__ jobs, int n_jobs, unsigned int* __restrict__ queue,
                unsigned long long* __restrict__ stats) {
  __shared__ WaveLds lds[4];
  const int lane = (int)(threadIdx.x & 63u);
  WaveLds& L = lds[threadIdx.x >> 6];
  Counters cnt = {};
  EngineWave eng;
  for (int job = 0; job < n_jobs; ++job) {
    const WaveJob& J = jobs[job];
    for (;;) {
      unsigned int s = 0;
      if (lane == 0) s = atomicAdd(&queue[job], 1u);
      s = __shfl(s, 0, 64);
      if ((int64_t)s >= (int64_t)J.p.n_sims) break;
      const unsigned long long ev_start = cnt.events;
      eng.start(J, L, lane, (int64_t)s);
      while (!eng.step(J, L, lane, cnt)) {
      }
      const int fl = __shfl(eng.flag, 0, 64);
      cnt.fevents += (fl != 0) ? cnt.events - ev_start : 0ull;  // summed over lanes below
      if (lane == 0) {
        if (J.p.flags) J.p.flags[s] = eng.flag;
        cnt.flagged += (eng.flag != 0);
      }
    }
  }
  const unsigned long long vals[ST_EXACT + 1] = {cnt.events, cnt.leaps, cnt.rejected, cnt.ssa, cnt.flagged,
                                                 cnt.w.uniforms, cnt.w.ptrs, cnt.w.inv, cnt.w.decisions,
                                                 cnt.w.samples, cnt.fevents, cnt.w.asym, cnt.w.exact};
#pragma unroll
  for (int i = 0; i <= ST_EXACT; ++i) {
    const unsigned long long sum = wave_sum(vals[i]);
    if (lane == 0) atomicAdd(&stats[i], sum);
  }
}

int launch_wave(const void* jobs, int n_jobs, unsigned int* queue, unsigned long long* stats, int waves,
                hipStream_t stream) {
  const int grid = (waves + 3) / 4;
  if (grid <= 0) return 0;
  hipLaunchKernelGGL(sim_wave_kernel, dim3(grid), dim3(256), 0, stream, (const WaveJob*)jobs, n_jobs, queue, stats);
  return (int)hipGetLastError();
}

int blocks_per_cu_wave() {
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, sim_wave_kernel, 256, 0) != hipSuccess) return 1;
  return blocks > 0 ? blocks : 1;
}

// Sparse table of a network for the wave kernel; false (with a message) if
// the network exceeds its limits.
bool make_wave_tab(const HostNet& h, WaveTab* t, char* why, size_t why_len) {
  std::memset(t, 0, sizeof(*t));
  const int N = h.N, M = h.M;
  if (N > WAVE_MAX_N || M > WAVE_MAX_M) {
    snprintf(why, why_len, "N=%d, M=%d exceeds %d species / %d reactions", N, M, WAVE_MAX_N, WAVE_MAX_M);
    return false;
  }
  t->N = N;
  t->M = M;
  t->kappa = h.kappa;
  for (int j = 0; j < M; ++j) {
    int nr = 0, nv = 0;
    for (int i = 0; i < N; ++i) {
      const int p = h.ord[i * M + j];
      if (p > 0) {
        if (nr == WAVE_MAX_REACT) {
          snprintf(why, why_len, "reaction %d has more than %d species in its propensity", j, WAVE_MAX_REACT);
          return false;
        }
        t->rs[j][nr] = (uint8_t)i;
        t->ro[j][nr] = (uint8_t)p;
        t->rst[j][nr] = (uint8_t)h.stp[i * M + j];
        ++nr;
      }
      const int v = h.v[i * M + j];
      if (v != 0) {
        if (nv == WAVE_MAX_V) {
          snprintf(why, why_len, "reaction %d changes more than %d species", j, WAVE_MAX_V);
          return false;
        }
        t->vs[j][nv] = (uint8_t)i;
        t->vv[j][nv] = (int8_t)v;
        ++nv;
      }
    }
    t->nr[j] = (uint8_t)nr;
    t->nv[j] = (uint8_t)nv;
  }
  int e = 0;
  for (int i = 0; i < N; ++i) {
    t->tstart[i] = (uint16_t)e;
    for (int j = 0; j < M; ++j) {
      const int v = h.v[i * M + j];
      if (v != 0) {
        if (e == WAVE_MAX_T) {
          snprintf(why, why_len, "more than %d non-zero stoichiometry entries", WAVE_MAX_T);
          return false;
        }
        t->tj[e] = (uint16_t)j;
        t->tv[e] = (int8_t)v;
        ++e;
      }
    }
    t->H[i] = h.H[i];
    t->Ht[i] = h.Ht[i];
    t->irs[i] = h.irs[i];
  }
  t->tstart[N] = (uint16_t)e;
  t->nnz = e;
  for (int q = 0; q < h.kappa; ++q) t->E[q] = h.E[q];
  return true;
}

}  // namespace pst
