// pst_driver.hpp -- the persistent, wavefront-uniform simulation driver shared
// by the FP64 and mixed-precision engines.
//
// A launch processes a list of jobs (network + run settings + parameter rows
// + output array).  Each wavefront works on one job at a time, so the network
// table is uniform across its lanes; its first `lanes` lanes (JobParams; 64
// unless the launch has fewer simulations than 64 per SIMD) take simulations
// of that job from a device work queue (one atomic per wavefront per refill)
// and take the next one as soon as theirs ends, so uneven step counts do not
// leave lanes idle until the job's queue is empty.  When the job is exhausted
// and every lane of the wavefront has finished, the wavefront moves to the
// next job.
//
// Per lane the driver keeps a mode -- DECIDE (next step is a leap decision,
// P1_P2.py:96-294), SSA (inside a P3 burst, P3.py:71-151, or pure SSA) or
// DONE.  Every pass of its loop runs one step for every active lane (the
// engine's step() serves both modes with shared code), then, while at least
// SSA_LANES_MIN lanes are in an SSA burst, up to `ssa_slice` further SSA-only
// steps for them: a wavefront lasts as long as its slowest simulation, and a
// burst taken one step per pass would add 100 passes to that simulation's
// path (8 steps per pass: Lotka-Volterra 81 -> 70 ms, Schlogl 110 -> 76 ms).
// A pass in which no lane decides runs up to 1 + ssa_slice SSA steps (2 +
// 2 ssa_slice for pure SSA: 68 -> 66 ms on Schlogl, 65,536 simulations).
// Every simulation executes exactly the reference's sequence
// (decision, burst of ssa_burst SSA steps, decision, ...); only the
// interleaving of different simulations within a wavefront changes.
#pragma once
#include "pst_device.hpp"
#include "pst_args.hpp"

namespace pst {

enum Mode { MODE_DECIDE = 0, MODE_SSA = 1, MODE_DONE = 2 };

struct Counters {
  unsigned long long events;
  unsigned long long fevents;  // events of simulations that ended flagged
  unsigned int leaps, rejected, ssa, flagged;
  Work w;
};

__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0ull; }

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// x[idx] for a run-time idx in [0, NP), as an OR of masked elements: a select
// chain is folded back into an indexed load by the compiler, which then moves
// the per-lane array out of registers (into LDS or scratch).
template <int NP>
__device__ __forceinline__ int32_t pick_i(const int32_t (&x)[NP], int idx) {
  int32_t r = 0;
#pragma unroll
  for (int i = 0; i < NP; ++i) r |= x[i] & -(int32_t)(idx == i);
  return r;
}

// ---------------------------------------------------------------------------
// Trajectory output.  JobParams::layout selects
//   LAYOUT_ROWS    O[q][f][o_stride], the reference's [kappa][ita][U]
//                  (cuTauLeaping.py:126): one 4-byte store per sample.  Lanes
//                  cross sample times at different moments, so every store
//                  dirties a 32-byte sector of its own.
//   LAYOUT_BLOCKS  O[q][f / 8][o_stride][f % 8] (ceil(ita / 8) blocks per
//                  species): the eight consecutive samples of one
//                  simulation fill one 32-byte sector.  A lane gathers its
//                  simulation's samples of the first Ref::kOutLds recorded
//                  species in LDS and writes each block with two 16-byte
//                  stores once its last sample is known, so no sector is
//                  written twice (further species are stored sample by
//                  sample).  The LDS slot is per simulation lane of the
//                  workgroup: the decide and burst lanes of a pair
//                  (run_driver_paired) share it, so a simulation's
//                  partial block travels with it.
// The last block is padded with zeros.
// ---------------------------------------------------------------------------
constexpr int LAYOUT_ROWS = 0, LAYOUT_BLOCKS = 1;
constexpr int OUT_BLOCK = 8;
typedef int32_t I4 __attribute__((ext_vector_type(4)));  // one 16-byte store

template <int KB>
struct SampleOut {
  static constexpr int KS = KB > 0 ? KB : 1;
  static __device__ __forceinline__ int32_t* lds() {
    __shared__ int32_t buf[KS * OUT_BLOCK * 256];
    return buf;
  }
  static __device__ __forceinline__ int slot() { return (int)(((threadIdx.x >> 6) & 3u) * 64u + (threadIdx.x & 63u)); }
  template <class P>
  static __device__ __forceinline__ int64_t block_row(const P& p, int q, int f) {
    return ((int64_t)q * ((p.ita + OUT_BLOCK - 1) / OUT_BLOCK) + (f / OUT_BLOCK)) * p.o_stride * OUT_BLOCK;
  }
  // the simulation's base pointer: O + sim (rows) or O + 8 sim (blocks)
  template <class P>
  static __device__ __forceinline__ int32_t* base(const P& p, int64_t sim) {
    return p.O + (p.layout == LAYOUT_BLOCKS ? sim * OUT_BLOCK : sim);
  }
  // sample f of recorded species q
  template <class P>
  static __device__ __forceinline__ void put(const P& p, int32_t* ob, int q, int f, int32_t v) {
    if (p.layout != LAYOUT_BLOCKS) {
      glb(ob)[((int64_t)q * p.ita + f) * p.o_stride] = v;
    } else if (q < KB) {
      lds()[(q * OUT_BLOCK + (f & (OUT_BLOCK - 1))) * 256 + slot()] = v;
    } else {
      glb(ob)[block_row(p, q, f) + (f & (OUT_BLOCK - 1))] = v;
    }
  }
  // after sample f of every recorded species: write the gathered blocks
  // that sample f completes (the block's last sample, or the run's)
  template <class P>
  static __device__ __forceinline__ void row_done(const P& p, int32_t* ob, int kappa, int f) {
    if (p.layout != LAYOUT_BLOCKS) return;
    const int last = f & (OUT_BLOCK - 1);
    if (last != OUT_BLOCK - 1 && f != p.ita - 1) return;
#pragma unroll
    for (int q = 0; q < KB; ++q) {
      if (q < kappa) {
        int32_t w[OUT_BLOCK];
#pragma unroll
        for (int j = 0; j < OUT_BLOCK; ++j) {
          const int32_t v = lds()[(q * OUT_BLOCK + j) * 256 + slot()];
          w[j] = (j <= last) ? v : 0;
        }
        gptr<I4> dst = (gptr<I4>)glb(ob + block_row(p, q, f));
        dst[0] = I4{w[0], w[1], w[2], w[3]};
        dst[1] = I4{w[4], w[5], w[6], w[7]};
      }
    }
    // species stored sample by sample: zero the padding of the last block
    if (f == p.ita - 1)
      for (int q = KB; q < kappa; ++q)
        for (int j = last + 1; j < OUT_BLOCK; ++j) glb(ob)[block_row(p, q, f) + j] = 0;
  }
};

// ---------------------------------------------------------------------------
// Sample ranges (pst_plan_execute_range, graph-captured chunks).  A launch
// with f_end < ita stops each simulation once it has recorded sample
// f_end - 1 and saves its state to JobParams::state, word-major
// ([word][n_sims]): the engine registers, the mode, the simulation's events
// so far and, for the blocked layout, its gathered samples of the block in
// progress (the step that records sample f_end - 1 may record later samples
// of the same block too); a launch with resume = 1 takes every simulation up
// from there.  A simulation that ends saves MODE_DONE, so later ranges skip it.
// Stopping between two steps and going on from the saved registers changes
// nothing in a simulation's step sequence, so a run split into ranges gives
// the same bits as one launch.  The trajectory base pointer is recomputed on
// resume from the launch's own descriptor.  f_end = INT_MAX,
// resume = 0: one launch runs every simulation to its end.
// ---------------------------------------------------------------------------
// A value the optimizer may not reason about: the word offsets i * n of the
// saved-state records stay inside save/load instead of being hoisted out of
// the simulation loop as ~40 live 64-bit offsets (+68 VGPRs in the pure-SSA
// Schlogl kernel).
__device__ __forceinline__ int32_t opaque_i32(int32_t v) {
  asm volatile("" : "+v"(v));
  return v;
}

template <class Engine>
struct StateIO {
  using Out = typename Engine::Out;
  static constexpr int EW = (int)(sizeof(Engine) / 4);
  static constexpr int BW = Out::KS * OUT_BLOCK;  // the lane's LDS sample block words
  static constexpr int WORDS = EW + 3 + BW;
  static_assert(sizeof(Engine) % 4 == 0, "engine state is not whole words");
  template <class P>
  static __device__ __forceinline__ bool fits(const P& p) {
    return p.state != nullptr && p.state_words >= WORDS;
  }
  template <class P>
  static __device__ __forceinline__ void save(const P& p, int64_t sim, const Engine& eng, int mode,
                                              unsigned long long ev) {
    gptr<uint32_t> st = glb(p.state);
    const int64_t n = opaque_i32(p.n_sims);
    uint32_t w[EW];
    __builtin_memcpy(w, &eng, sizeof(Engine));
#pragma unroll
    for (int i = 0; i < EW; ++i) st[(int64_t)i * n + sim] = w[i];
    st[(int64_t)EW * n + sim] = (uint32_t)mode;
    st[(int64_t)(EW + 1) * n + sim] = (uint32_t)ev;
    st[(int64_t)(EW + 2) * n + sim] = (uint32_t)(ev >> 32);
    if (p.layout == LAYOUT_BLOCKS && mode != MODE_DONE) {
#pragma unroll
      for (int i = 0; i < BW; ++i) st[(int64_t)(EW + 3 + i) * n + sim] = (uint32_t)Out::lds()[i * 256 + Out::slot()];
    }
  }
  template <class P>
  static __device__ __forceinline__ int load(const P& p, int64_t sim, Engine& eng, unsigned long long& ev) {
    gptr<uint32_t> st = glb(p.state);
    const int64_t n = opaque_i32(p.n_sims);
    uint32_t w[EW];
#pragma unroll
    for (int i = 0; i < EW; ++i) w[i] = st[(int64_t)i * n + sim];
    __builtin_memcpy(&eng, w, sizeof(Engine));
    ev = ((unsigned long long)st[(int64_t)(EW + 2) * n + sim] << 32) | st[(int64_t)(EW + 1) * n + sim];
    const int mode = (int)st[(int64_t)EW * n + sim];
    if (p.layout == LAYOUT_BLOCKS && mode != MODE_DONE) {
#pragma unroll
      for (int i = 0; i < BW; ++i) Out::lds()[i * 256 + Out::slot()] = (int32_t)st[(int64_t)(EW + 3 + i) * n + sim];
    }
    return mode;
  }
};

// Engine interface (see EngineF64 / EngineF32):
//   int  start(job, sim)            -> initial mode, state loaded
//   int  step<SSA_ONLY>(job, cnt, m) -> next mode (DONE may set flag)
//   int  flag
constexpr int SSA_LANES_MIN = 1;

template <class Engine, class Ref>
__device__ __forceinline__ void finish_sim(const Ref& jr, const Engine& eng, int64_t sim, Counters& cnt,
                                           unsigned long long ev_start) {
  int32_t* fl = jr.p().flags;
  if (fl) glb(fl)[sim] = eng.flag;
  cnt.flagged += (eng.flag != 0);
  cnt.fevents += (eng.flag != 0) ? cnt.events - ev_start : 0ull;
}

// After a step: record an ended simulation, or stop one that has recorded
// the range's last sample.  Returns the lane's mode.
template <class Engine, class Ref>
__device__ __forceinline__ int settle(const Ref& jr, const Engine& eng, int64_t sim, Counters& cnt,
                                      unsigned long long ev_start, int mode, int f_stop) {
  using SIO = StateIO<Engine>;
  if (mode == MODE_DONE) {
    finish_sim(jr, eng, sim, cnt, ev_start);
    if (SIO::fits(jr.p())) SIO::save(jr.p(), sim, eng, MODE_DONE, cnt.events - ev_start);
    return MODE_DONE;
  }
  if (eng.F >= f_stop) {
    if (SIO::fits(jr.p())) {
      SIO::save(jr.p(), sim, eng, mode, cnt.events - ev_start);
      return MODE_DONE;
    }
  }
  return mode;
}

// A simulation taken from the queue: started, or resumed from its saved
// state.  Returns its mode (MODE_DONE: nothing left to do in this range).
template <class Engine, class Ref>
__device__ __forceinline__ int begin_sim(const Ref& jr, Engine& eng, int64_t sim, Counters& cnt,
                                         unsigned long long& ev_start, bool resume, int f_stop) {
  using SIO = StateIO<Engine>;
  if (resume && SIO::fits(jr.p())) {
    unsigned long long ev;
    const int mode = SIO::load(jr.p(), sim, eng, ev);
    ev_start = cnt.events - ev;
    eng.Ob = Engine::Out::base(jr.p(), sim);
    if (mode == MODE_DONE || eng.F >= f_stop) return MODE_DONE;
    // Without tau-leaping a simulation is only ever in an SSA step (the
    // engines' start and step never return DECIDE then).  Saying so here
    // keeps the mode a compile-time fact in network kernels, whose method is
    // a constant, so their leap code is dropped (the pure-SSA Schlogl
    // kernel shrinks to about a quarter of its machine code).
    return jr.p().method != 0 ? (int)MODE_SSA : mode;
  }
  ev_start = cnt.events;
  return eng.start(jr, sim);
}

template <class Ref>
__device__ __forceinline__ int range_stop(const Ref& jr, int f_end) {
  return f_end < jr.p().ita ? f_end : jr.p().ita;
}

template <class Engine, class Ref>
__device__ __forceinline__ void run_driver(const typename Ref::Dev* __restrict__ jobs, int n_jobs,
                                           unsigned int* __restrict__ queue, unsigned long long* __restrict__ stats,
                                           int f_end, int resume) {
  const int lane = (int)(threadIdx.x & 63u);
  Counters cnt = {};
  Engine eng;
  int mode = MODE_DONE;
  int64_t sim = -1;
  unsigned long long ev_start = 0;  // cnt.events when this lane's simulation began
  int job = 0;
  bool exhausted = false;

  Ref jr;
  jr.load(jobs);
  int f_stop = range_stop(jr, f_end);

  for (;;) {
    // ---- refill lanes whose simulation ended from the job's queue ----
    const bool want = (mode == MODE_DONE) && !exhausted && lane < jr.p().lanes;
    const unsigned long long wmask = __builtin_amdgcn_ballot_w64(want);
    if (wmask != 0ull) {
      const int n = __builtin_popcountll(wmask);
      const int leader = __builtin_ctzll(wmask);
      unsigned int base = 0;
      if (lane == leader) base = atomicAdd(&queue[job], (unsigned int)n);
      base = __shfl(base, leader, 64);
      const int n_sims = jr.p().n_sims;
      if ((int64_t)base + n >= (int64_t)n_sims) exhausted = true;
      if (want) {
        const int rank = __builtin_popcountll(wmask & ((1ull << lane) - 1ull));
        const int64_t s = (int64_t)base + rank;
        if (s < n_sims) {
          sim = s;
          mode = begin_sim(jr, eng, s, cnt, ev_start, resume != 0, f_stop);
        }
      }
    }
    if (!wave_any(mode != MODE_DONE)) {
      // this wavefront is idle and its job has no work left: next job
      job += 1;
      if (job >= n_jobs) break;
      jr.load(jobs + job);
      f_stop = range_stop(jr, f_end);
      exhausted = false;
      continue;
    }

    jr.refresh();
    if (wave_any(mode == MODE_DECIDE)) {
      if (mode != MODE_DONE) {
        mode = eng.template step<false>(jr, cnt, mode);
        mode = settle(jr, eng, sim, cnt, ev_start, mode, f_stop);
      }
      // SSA-only steps while lanes are inside a burst
      const int slice = jr.p().ssa_slice;
      for (int s = 0; s < slice; ++s) {
        if (__builtin_popcountll(__builtin_amdgcn_ballot_w64(mode == MODE_SSA)) < SSA_LANES_MIN) break;
        jr.refresh();
        if (mode == MODE_SSA) {
          mode = eng.template step<true>(jr, cnt, mode);
          mode = settle(jr, eng, sim, cnt, ev_start, mode, f_stop);
        }
      }
    } else {
      // no lane decides (pure SSA, or every live lane inside a burst): up to
      // 1 + ssa_slice SSA steps before the next refill, 2 + 2 ssa_slice for
      // pure SSA, where no lane waits to decide
      const int slice = jr.p().method == 2 ? 2 * jr.p().ssa_slice + 1 : jr.p().ssa_slice;
      for (int s = 0; s <= slice; ++s) {
        if (!wave_any(mode == MODE_SSA)) break;
        if (s > 0) jr.refresh();
        if (mode == MODE_SSA) {
          mode = eng.template step<true>(jr, cnt, mode);
          mode = settle(jr, eng, sim, cnt, ev_start, mode, f_stop);
        }
      }
    }
  }

  const unsigned long long vals[ST_EXACT + 1] = {cnt.events, cnt.leaps, cnt.rejected, cnt.ssa, cnt.flagged,
                                                 cnt.w.uniforms, cnt.w.ptrs, cnt.w.inv, cnt.w.decisions,
                                                 cnt.w.samples, cnt.fevents, cnt.w.asym, cnt.w.exact};
  // (32-bit per-lane counters: a lane's share of one launch stays far below 2^32)
#pragma unroll
  for (int i = 0; i <= ST_EXACT; ++i) {
    const unsigned long long sum = wave_sum(vals[i]);
    if (lane == 0) atomicAdd(&stats[i], sum);
  }
}

// ---------------------------------------------------------------------------
// Paired driver: a workgroup of 512 lanes is four decide wavefronts (0-3) and
// four burst wavefronts (4-7), two wavefronts per SIMD.  A decide wavefront
// runs its simulations like run_driver but takes no SSA-only steps: when a
// leap decision switches a simulation into an SSA fallback burst
// (P1_P2.py:157-173, then P3.py:71-151), the lane writes the simulation's
// state to an LDS slot and lane l of burst wavefront w + 4 takes it, runs the
// burst's remaining SSA steps back to back and hands the state back.  So the
// leap passes of a decide wavefront no longer wait for other lanes' SSA steps,
// a burst no longer waits for leap passes, and each SIMD overlaps the two
// instruction streams (one wavefront per SIMD leaves half its issue slots to
// latency).  Every simulation still executes exactly the reference's
// sequence; only where its steps run changes.
//
// Slot protocol (one slot per decide lane): FREE -> SENT (decide lane, after
// writing the state) -> BACK (burst lane, after writing the state back) ->
// FREE (decide lane, after reading it).  A decide lane with a SENT slot is
// AWAY and skips passes; a decide wavefront moves to its next job only when
// none of its lanes is away, so every simulation a burst wavefront holds
// belongs to its partner's current job.  A burst wavefront never waits for
// its partner except while it has nothing to run, and it leaves once the
// partner has finished every job (DONE flag), when no slot can be SENT.
// ---------------------------------------------------------------------------
enum { MODE_AWAY = 3 };
enum { SLOT_FREE = 0u, SLOT_SENT = 1u, SLOT_BACK = 2u };
constexpr int PAIR_SLOTS = 256;  // decide lanes per workgroup

__device__ __forceinline__ uint32_t lds_load_acquire(uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_store_release(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// A simulation in transit: the engine state, its id, its events so far
// (for the flagged-events counter), the job index and the next mode.
template <class Engine>
struct SlotIO {
  static constexpr int EW = (int)(sizeof(Engine) / 4);
  static constexpr int WORDS = EW + 6;
  static_assert(sizeof(Engine) % 4 == 0, "engine state is not whole words");
  // word-major [word][slot]: consecutive lanes hit consecutive banks
  static __device__ __forceinline__ void put(uint32_t* data, int slot, const Engine& eng, int64_t sim,
                                             unsigned long long ev, int job, int mode) {
    uint32_t w[EW];
    __builtin_memcpy(w, &eng, sizeof(Engine));
#pragma unroll
    for (int i = 0; i < EW; ++i) data[i * PAIR_SLOTS + slot] = w[i];
    data[(EW + 0) * PAIR_SLOTS + slot] = (uint32_t)sim;
    data[(EW + 1) * PAIR_SLOTS + slot] = (uint32_t)((uint64_t)sim >> 32);
    data[(EW + 2) * PAIR_SLOTS + slot] = (uint32_t)ev;
    data[(EW + 3) * PAIR_SLOTS + slot] = (uint32_t)(ev >> 32);
    data[(EW + 4) * PAIR_SLOTS + slot] = (uint32_t)job;
    data[(EW + 5) * PAIR_SLOTS + slot] = (uint32_t)mode;
  }
  static __device__ __forceinline__ void get(const uint32_t* data, int slot, Engine& eng, int64_t& sim,
                                             unsigned long long& ev, int& mode) {
    uint32_t w[EW];
#pragma unroll
    for (int i = 0; i < EW; ++i) w[i] = data[i * PAIR_SLOTS + slot];
    __builtin_memcpy(&eng, w, sizeof(Engine));
    sim = (int64_t)(((uint64_t)data[(EW + 1) * PAIR_SLOTS + slot] << 32) | data[(EW + 0) * PAIR_SLOTS + slot]);
    ev = ((unsigned long long)data[(EW + 3) * PAIR_SLOTS + slot] << 32) | data[(EW + 2) * PAIR_SLOTS + slot];
    mode = (int)data[(EW + 5) * PAIR_SLOTS + slot];
  }
  static __device__ __forceinline__ int job(const uint32_t* data, int slot) {
    return (int)data[(EW + 4) * PAIR_SLOTS + slot];
  }
};

template <class Engine, class Ref>
__device__ __forceinline__ void run_driver_paired(const typename Ref::Dev* __restrict__ jobs, int n_jobs,
                                                  unsigned int* __restrict__ queue,
                                                  unsigned long long* __restrict__ stats, int f_end, int resume) {
  using IO = SlotIO<Engine>;
  __shared__ uint32_t s_data[IO::WORDS * PAIR_SLOTS];
  __shared__ uint32_t s_state[PAIR_SLOTS];
  __shared__ uint32_t s_done[4];
  const int tid = (int)threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;  // wave-uniform
  const int pair = wave & 3;
  const int slot = pair * 64 + lane;
  if (tid < PAIR_SLOTS) s_state[tid] = SLOT_FREE;
  if (tid < 4) s_done[tid] = 0u;
  __syncthreads();  // the only barrier: every lane of the workgroup reaches it

  Counters cnt = {};
  Engine eng;
  int64_t sim = -1;
  unsigned long long ev_start = 0;  // cnt.events when this lane's simulation began (here)
  Ref jr;

  if (wave < 4) {
    // ---- decide wavefront ----
    int mode = MODE_DONE;
    int job = 0;
    bool exhausted = false;
    jr.load(jobs);
    int f_stop = range_stop(jr, f_end);
    for (;;) {
      // a burst handed back (ended simulations were recorded by the burst lane)
      if (mode == MODE_AWAY && lds_load_acquire(&s_state[slot]) == SLOT_BACK) {
        unsigned long long ev;
        IO::get(s_data, slot, eng, sim, ev, mode);
        ev_start = cnt.events - ev;
        s_state[slot] = SLOT_FREE;
        if (mode != MODE_DONE) mode = settle(jr, eng, sim, cnt, ev_start, mode, f_stop);
      }
      const bool want = (mode == MODE_DONE) && !exhausted && lane < jr.p().lanes;
      const unsigned long long wmask = __builtin_amdgcn_ballot_w64(want);
      if (wmask != 0ull) {
        const int n = __builtin_popcountll(wmask);
        const int leader = __builtin_ctzll(wmask);
        unsigned int base = 0;
        if (lane == leader) base = atomicAdd(&queue[job], (unsigned int)n);
        base = __shfl(base, leader, 64);
        const int n_sims = jr.p().n_sims;
        if ((int64_t)base + n >= (int64_t)n_sims) exhausted = true;
        if (want) {
          const int rank = __builtin_popcountll(wmask & ((1ull << lane) - 1ull));
          const int64_t s = (int64_t)base + rank;
          if (s < n_sims) {
            sim = s;
            mode = begin_sim(jr, eng, s, cnt, ev_start, resume != 0, f_stop);
          }
        }
      }
      if (!wave_any(mode != MODE_DONE)) {
        job += 1;
        if (job >= n_jobs) break;
        jr.load(jobs + job);
        f_stop = range_stop(jr, f_end);
        exhausted = false;
        continue;
      }
      if (wave_any(mode == MODE_DECIDE)) {
        jr.refresh();
        if (mode == MODE_DECIDE) {
          mode = eng.template step<false>(jr, cnt, mode);
          mode = settle(jr, eng, sim, cnt, ev_start, mode, f_stop);
        }
      } else if (!wave_any(mode == MODE_SSA)) {
        __builtin_amdgcn_s_sleep(2);  // every live lane is away: wait for a burst to come back
      }
      // a burst starts (or a forced-SSA simulation starts): hand it over
      if (mode == MODE_SSA) {
        IO::put(s_data, slot, eng, sim, cnt.events - ev_start, job, MODE_SSA);
        lds_store_release(&s_state[slot], SLOT_SENT);
        mode = MODE_AWAY;
      }
    }
    lds_store_release(&s_done[pair], 1u);
  } else {
    // ---- burst wavefront ----
    // (Both roles keep the default issue priority: raising the burst
    // wavefront's slowed the Lotka-Volterra workload from 39.9 to 42.3 ms.)
    bool busy = false;
    int bjob = -1;
    int f_stop = 0x7fffffff;
    for (;;) {
      const bool pick = !busy && lds_load_acquire(&s_state[slot]) == SLOT_SENT;
      const unsigned long long pmask = __builtin_amdgcn_ballot_w64(pick);
      if (pmask != 0ull) {
        // all simulations the partner sends belong to its current job
        const int pj = __builtin_amdgcn_readlane(IO::job(s_data, slot), __builtin_ctzll(pmask));
        if (pj != bjob) {
          jr.load(jobs + pj);
          bjob = pj;
          f_stop = range_stop(jr, f_end);
        }
        if (pick) {
          unsigned long long ev;
          int m;
          IO::get(s_data, slot, eng, sim, ev, m);
          ev_start = cnt.events - ev;
          busy = true;
        }
      }
      if (!wave_any(busy)) {
        if (lds_load_acquire(&s_done[pair]) != 0u) break;
        __builtin_amdgcn_s_sleep(4);
        continue;
      }
      jr.refresh();
      if (busy) {
        int m = eng.template step<true>(jr, cnt, MODE_SSA);
        if (m == MODE_DONE) m = settle(jr, eng, sim, cnt, ev_start, m, f_stop);
        // back to the decide lane at the burst's end, or at the range's
        // last sample (the decide lane saves it)
        if (m != MODE_SSA || eng.F >= f_stop) {
          IO::put(s_data, slot, eng, sim, cnt.events - ev_start, bjob, m);
          lds_store_release(&s_state[slot], SLOT_BACK);
          busy = false;
        }
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

}  // namespace pst
