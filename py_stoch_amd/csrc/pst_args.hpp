// pst_args.hpp -- kernel argument blocks shared by the launch code and the kernels.
//
// A network is passed to the kernels as a dense, zero-padded species x reaction
// table (NetTab) inside a job descriptor in device memory: every entry is the same for
// all lanes of a launch, so the compiler reads it with scalar loads into SGPRs
// and every branch on it is a scalar branch.  Per-simulation data (state, time,
// sample cursor, Philox counter) lives in VGPRs for the whole run.
#pragma once
#ifndef __HIPCC_RTC__
#include <stdint.h>
#include <hip/hip_runtime.h>
#endif

namespace pst {

// Dense network table with capacity NP species x MP reactions.  Rows beyond N
// and columns beyond M are zero, which makes them inert in every loop.
template <int NP, int MP>
struct NetTab {
  int8_t v[NP * MP];    // net change V = Post - Pre, species-major (d_V, P1_P2.py:8)
  int8_t ord[NP * MP];  // number of propensity factors of species i in reaction j
  int8_t stp[NP * MP];  // factor step: 1 -> x(x-1)(x-2)..., 0 -> x^p
  int8_t H[NP];         // highest order of reaction (d_H, P1_P2.py:9)
  int8_t Ht[NP];        // d_H_type (P1_P2.py:10)
  int8_t irs[NP];       // 1 if the species has any Pre entry (I_rs, P1_P2.py:338-346)
  int8_t E[NP];         // recorded species (d_E, P1_P2.py:13)
};

enum StatIdx {
  ST_EVENTS = 0, ST_LEAPS, ST_REJECTED, ST_SSA, ST_FLAGGED,
  ST_UNIFORMS, ST_PTRS, ST_INV, ST_DECISIONS, ST_SAMPLES, ST_FEVENTS, ST_ASYM, ST_EXACT,
  ST_COUNT = 16
};

// Settings of one job of a launch (uniform across the wavefronts working on it)
struct JobParams {
  int32_t N, M;
  int32_t kappa, ita;
  int32_t n_c;          // critical threshold (tl_parser.py:43)
  int32_t ssa_burst;    // SSA steps per fallback burst (P3.py:71)
  int32_t method;       // 0 tau-leaping, 1 forced SSA, 2 pure SSA
  int32_t n_sims;
  int32_t sims_per_set; // simulations sharing one rates/x0 row
  int32_t ssa_slice;    // SSA steps per pass while other lanes wait to decide (launch tuning)
  int32_t n_rates;      // length of a rates row: M, or the parameter count of compiled hazards
  uint32_t max_steps;   // 0: none; else leap decisions + SSA steps per simulation
  int32_t lanes;        // simulations per wavefront, 1..64 (launch tuning)
  int32_t layout;       // trajectory layout of O: LAYOUT_ROWS or LAYOUT_BLOCKS (pst_driver.hpp)
  double eta;           // tl_parser.py:44
  double t_max;
  uint64_t seed;
  uint64_t sim_offset;
  int64_t o_stride;     // fastest-axis length of O (>= n_sims)
  const double* rates;  // [n_sets][M]
  const int32_t* x0;    // [n_sets][N]
  const float* I;       // [ita] sample times
  int32_t* O;           // rows: [kappa][ita][o_stride]; blocks: [kappa][ceil(ita/8)][o_stride][8]
  int32_t* flags;       // [n_sims] or null
  uint32_t* state;      // [state_words][n_sims] saved simulations between sample ranges, or null
  int32_t state_words;  // words per simulation in state (at least the engine's record)
  int32_t reserved2;
};

// The sample range of a launch (pst_driver.hpp): stop simulations at sample
// f_end (saving their state), and/or resume them from their saved state.
struct Range {
  int32_t f_end;
  int32_t resume;
};
constexpr Range FULL_RUN = {0x7fffffff, 0};

// Copy a job's settings into VGPRs (an empty asm on each 32-bit word hides
// their uniformity).  They are read only by per-lane arithmetic (addresses,
// sample cursor, queue bounds), and held in SGPRs they would exceed the
// scalar register file and be spilled and reloaded lane by lane.
__device__ __forceinline__ void load_params_vgpr(JobParams* dst, const JobParams* src) {
  constexpr int NWORDS = (int)(sizeof(JobParams) / 4);
  uint32_t w[NWORDS];
  __builtin_memcpy(w, src, sizeof(JobParams));
#pragma unroll
  for (int i = 0; i < NWORDS; ++i) asm volatile("" : "+v"(w[i]));
  __builtin_memcpy(dst, w, sizeof(JobParams));
}

template <int NP, int MP>
struct JobDev {
  NetTab<NP, MP> tab;
  JobParams p;
};

// Kernel-side view of the current job.  The table bytes are copied four to a
// 32-bit word into SGPRs and hidden from the optimizer at the top of every
// pass of the simulation loop (refresh(): an empty asm on a VGPR copy, then
// readfirstlane), so the per-use byte extracts stay cheap scalar bit-field
// extracts inside the loop instead of ~64 hoisted, long-lived SGPRs that
// spill.  Scalar loads inside the loop are avoided: with one wavefront per
// SIMD their latency is not hidden.
template <int NP, int MP>
struct JobRef {
  static constexpr int kNP = NP, kMP = MP;
  static constexpr int kKP = MP;          // rates per row: one per reaction
  static constexpr bool kCustom = false;  // propensities from the mass-action table
  static constexpr int kOutLds = NP < 4 ? NP : 4;  // recorded species gathered in LDS (blocked layout)
  using Dev = JobDev<NP, MP>;
  static constexpr int NW = (NP * MP + 3) / 4;
  static constexpr int NS = (NP + 3) / 4;
  uint32_t v[NW], ord[NW], stp[NW];
  uint32_t H[NS], Ht[NS], irs[NS], E[NS];
  JobParams P;

  __device__ __forceinline__ void load(const JobDev<NP, MP>* jd) {
    __builtin_memcpy(v, jd->tab.v, sizeof(jd->tab.v));
    __builtin_memcpy(ord, jd->tab.ord, sizeof(jd->tab.ord));
    __builtin_memcpy(stp, jd->tab.stp, sizeof(jd->tab.stp));
    __builtin_memcpy(H, jd->tab.H, sizeof(jd->tab.H));
    __builtin_memcpy(Ht, jd->tab.Ht, sizeof(jd->tab.Ht));
    __builtin_memcpy(irs, jd->tab.irs, sizeof(jd->tab.irs));
    __builtin_memcpy(E, jd->tab.E, sizeof(jd->tab.E));
    load_params_vgpr(&P, &jd->p);
  }
  static __device__ __forceinline__ uint32_t opaque(uint32_t w) {
    asm volatile("" : "+v"(w));
    return __builtin_amdgcn_readfirstlane(w);
  }
  __device__ __forceinline__ void refresh() {
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      v[w] = opaque(v[w]);
      ord[w] = opaque(ord[w]);
      stp[w] = opaque(stp[w]);
    }
#pragma unroll
    for (int w = 0; w < NS; ++w) {
      H[w] = opaque(H[w]);
      Ht[w] = opaque(Ht[w]);
      irs[w] = opaque(irs[w]);
      E[w] = opaque(E[w]);
    }
  }
  static __device__ __forceinline__ int sbyte(const uint32_t* w, int idx) {
    return (int)(int8_t)(w[idx >> 2] >> (8 * (idx & 3)));
  }
  __device__ __forceinline__ const JobParams& p() const { return P; }
  __device__ __forceinline__ int V(int i, int j) const { return sbyte(v, i * MP + j); }
  __device__ __forceinline__ int Ord(int i, int j) const { return sbyte(ord, i * MP + j); }
  __device__ __forceinline__ int Stp(int i, int j) const { return sbyte(stp, i * MP + j); }
  __device__ __forceinline__ int Hv(int i) const { return sbyte(H, i); }
  __device__ __forceinline__ int Htv(int i) const { return sbyte(Ht, i); }
  __device__ __forceinline__ int Irs(int i) const { return sbyte(irs, i); }
  __device__ __forceinline__ int Ev(int q) const { return sbyte(E, q); }
};

// Kernel-side view of a job whose network is compiled into the kernel (the
// library generates `Net` and compiles it at run time, as the reference
// compiles each model's propensity code, cuTauLeaping.py:34-35): every table
// entry and every network-level setting is a constant, so the loops over
// species and reactions become straight-line code.  Net provides
//   static constexpr int N, M, kappa, n_c, method, ssa_burst;
//   static constexpr double eta;
//   static constexpr int8_t v[N*M], ord[N*M], stp[N*M], H[N], Ht[N], irs[N], E[N];
//   static constexpr bool custom; static constexpr int n_rates;
//   template <class R> static void props(x, c, a)   (used when custom)
// The job descriptor is the JobParams block alone.
template <class Net>
struct FixedParams : JobParams {
  static constexpr int N = Net::N, M = Net::M;
  static constexpr int kappa = Net::kappa, n_c = Net::n_c, method = Net::method, ssa_burst = Net::ssa_burst;
  static constexpr double eta = Net::eta;
};

template <class Net>
struct FixedRef {
  static constexpr int kNP = Net::N, kMP = Net::M;
  static constexpr int kKP = Net::n_rates;   // rates per row
  static constexpr bool kCustom = Net::custom;  // compiled hazard expressions
  static constexpr int kOutLds = Net::kappa < 4 ? Net::kappa : 4;  // recorded species gathered in LDS
  template <class R>
  __device__ __forceinline__ static void props(const int32_t (&x)[kNP], const R (&c)[kKP], R (&a)[kMP]) {
    Net::template props<R>(x, c, a);
  }
  using Dev = JobParams;
  FixedParams<Net> P;
  __device__ __forceinline__ void load(const JobParams* jp) { load_params_vgpr(&P, jp); }
  __device__ __forceinline__ void refresh() {}
  __device__ __forceinline__ const FixedParams<Net>& p() const { return P; }
  __device__ __forceinline__ static constexpr int V(int i, int j) { return Net::v[i * kMP + j]; }
  __device__ __forceinline__ static constexpr int Ord(int i, int j) { return Net::ord[i * kMP + j]; }
  __device__ __forceinline__ static constexpr int Stp(int i, int j) { return Net::stp[i * kMP + j]; }
  __device__ __forceinline__ static constexpr int Hv(int i) { return Net::H[i]; }
  __device__ __forceinline__ static constexpr int Htv(int i) { return Net::Ht[i]; }
  __device__ __forceinline__ static constexpr int Irs(int i) { return Net::irs[i]; }
  __device__ __forceinline__ static constexpr int Ev(int q) { return Net::E[q]; }
};

#ifndef __HIPCC_RTC__
// host-side copy of a network in dense form, built by the C-ABI layer
struct HostNet {
  int N, M, kappa;
  int8_t v[64 * 256];
  int8_t ord[64 * 256];
  int8_t stp[64 * 256];
  int8_t H[64], Ht[64], irs[64];
  int8_t E[64];
};

// size classes compiled into the library
constexpr int NUM_CLASSES = 3;
constexpr int CLASS_NP[NUM_CLASSES] = {4, 8, 16};
constexpr int CLASS_MP[NUM_CLASSES] = {4, 8, 16};

template <int NP, int MP>
inline NetTab<NP, MP> make_tab(const HostNet& h) {
  NetTab<NP, MP> t;
  for (int i = 0; i < NP * MP; ++i) { t.v[i] = 0; t.ord[i] = 0; t.stp[i] = 0; }
  for (int i = 0; i < NP; ++i) { t.H[i] = 0; t.Ht[i] = 0; t.irs[i] = 0; t.E[i] = 0; }
  for (int i = 0; i < h.N; ++i) {
    for (int j = 0; j < h.M; ++j) {
      t.v[i * MP + j] = h.v[i * h.M + j];
      t.ord[i * MP + j] = h.ord[i * h.M + j];
      t.stp[i * MP + j] = h.stp[i * h.M + j];
    }
    t.H[i] = h.H[i];
    t.Ht[i] = h.Ht[i];
    t.irs[i] = h.irs[i];
  }
  for (int q = 0; q < h.kappa; ++q) t.E[q] = h.E[q];
  return t;
}

// Launchers, one per translation unit (precision).  `jobs` is a device array
// of JobDev<NP,MP> for the size class `cls`, `queue` n_jobs zeroed counters,
// `waves` the number of wavefronts to launch.
int launch_fp64(int cls, const void* jobs, int n_jobs, unsigned int* queue, unsigned long long* stats,
                int waves, Range range, hipStream_t stream);
int launch_mixed(int cls, const void* jobs, int n_jobs, unsigned int* queue, unsigned long long* stats,
                 int waves, Range range, hipStream_t stream);
// the mixed-precision Poisson sampler on given means and words (validation)
int launch_poisson_probe(const float* lam, const uint32_t* words, int64_t n, int32_t* k, int32_t* path,
                         hipStream_t stream);
// resident 256-lane blocks per CU for a class (occupancy)
int blocks_per_cu_fp64(int cls);
int blocks_per_cu_mixed(int cls);

// the large-network kernel (one wavefront per simulation, mixed precision):
// `jobs` is a device array of WaveJob
struct WaveTab;
int launch_wave(const void* jobs, int n_jobs, unsigned int* queue, unsigned long long* stats, int waves,
                hipStream_t stream);
int blocks_per_cu_wave();
bool make_wave_tab(const HostNet& h, WaveTab* t, char* why, size_t why_len);

// smallest compiled size class holding (N, M, kappa), or -1
inline int size_class(int N, int M, int kappa) {
  for (int c = 0; c < NUM_CLASSES; ++c)
    if (N <= CLASS_NP[c] && M <= CLASS_MP[c] && kappa <= CLASS_NP[c]) return c;
  return -1;
}

#endif  // !__HIPCC_RTC__

}  // namespace pst
