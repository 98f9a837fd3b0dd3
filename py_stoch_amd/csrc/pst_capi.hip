// pst_capi.hip -- the extern "C" entry points declared in include/pst.h.
//
// Validation, table densification, device allocation and launch.  The
// reference's host driver (cuTauLeaping.py:18-84) JIT-compiles two kernels per
// run, uploads constants, launches P1_P2 and P3 once per iteration with a
// gpuarray.sum and a scalar copy back each time; here one persistent launch
// per size class runs every simulation of every job to the end and only the
// trajectories (and 16 counters) come back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include <mutex>
#include <string>

#include "../../include/pst.h"
#include "pst_args.hpp"
#include "pst_engine_wave.hpp"
#include "pst_rtc.hpp"

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(e == hipErrorOutOfMemory ? PST_ENOMEM : PST_EHIP, "%s: %s", what, hipGetErrorString(e));
}

// largest network of any kernel (the large-network kernel, mixed precision)
constexpr int kMaxSpecies = pst::WAVE_MAX_N;
constexpr int kMaxReactions = pst::WAVE_MAX_M;
constexpr int kSsaSlice = 7;
constexpr int kMaxParams = 256;

// A compiled hazard may use X(i) (0 <= i < N), C(k) (0 <= k < n_params),
// R(v), numbers, + - * / ( ) , and these functions; anything else
// (statements, preprocessor lines, comments, other names, an index that is
// not an integer literal in range) is rejected before the expression reaches
// the compiler.
bool hazard_ok(const char* h, int n_species, int n_params, std::string* why) {
  static const char* const kNames[] = {"X", "C", "R", "exp", "log", "sqrt", "pow", "fabs", "fmin", "fmax",
                                       "tanh", "exp2", "log2", "log10", "cbrt", "floor", "ceil"};
  int depth = 0;
  for (const char* p = h; *p;) {
    const char ch = *p;
    if (std::isalpha((unsigned char)ch) || ch == '_') {
      const char* q = p;
      while (std::isalnum((unsigned char)*q) || *q == '_') ++q;
      const std::string name(p, q);
      bool known = false;
      for (const char* k : kNames) known = known || name == k;
      if (!known) {
        *why = "unknown name '" + name + "'";
        return false;
      }
      p = q;
      if (name == "X" || name == "C") {
        // the argument is a decimal integer literal inside the table's range
        const int limit = name == "X" ? n_species : n_params;
        const char* r = p;
        if (*r != '(') {
          *why = name + " takes an index in parentheses";
          return false;
        }
        ++r;
        long idx = 0;
        int digits = 0;
        while (std::isdigit((unsigned char)*r) && digits < 9) idx = idx * 10 + (*r++ - '0'), ++digits;
        if (digits == 0 || *r != ')') {
          *why = name + "(...) needs a non-negative integer literal";
          return false;
        }
        if (idx >= limit) {
          *why = name + "(" + std::to_string(idx) + ") is out of range (" + (name == "X" ? "N" : "n_params") +
                 " = " + std::to_string(limit) + ")";
          return false;
        }
        p = r + 1;
      }
    } else if (std::isdigit((unsigned char)ch) || ch == '.') {
      // number, with an optional exponent
      while (std::isdigit((unsigned char)*p) || *p == '.') ++p;
      if (*p == 'e' || *p == 'E') {
        ++p;
        if (*p == '+' || *p == '-') ++p;
        while (std::isdigit((unsigned char)*p)) ++p;
      }
    } else if (ch == '(' || ch == ')') {
      depth += ch == '(' ? 1 : -1;
      if (depth < 0) {
        *why = "unbalanced parentheses";
        return false;
      }
      ++p;
    } else if (ch == '/' && (p[1] == '/' || p[1] == '*')) {
      *why = "comments are not allowed";
      return false;
    } else if (std::strchr("+-*/, ", ch)) {
      ++p;
    } else {
      *why = std::string("character '") + ch + "' is not allowed";
      return false;
    }
  }
  if (depth != 0) {
    *why = "unbalanced parentheses";
    return false;
  }
  return true;
}

__global__ void zero_u32(unsigned int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__global__ void zero_u64(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// blocks[kappa][ita8][bstride][8] -> rows[kappa][ita][rstride]: workgroup b
// takes block row r = b / chunks (species q = r / ita8, samples 8 (r % ita8)
// ...) and 256 simulations; each lane reads its simulation's 32 bytes with
// two 16-byte loads and writes up to eight rows, consecutive lanes to
// consecutive addresses.
__global__ void __launch_bounds__(256) blocks_to_rows(const int32_t* __restrict__ blocks, int64_t bstride,
                                                      int32_t* __restrict__ rows, int64_t rstride, int ita,
                                                      int ita8, int64_t n_sims, int64_t chunks) {
  const int64_t b = blockIdx.x;
  const int64_t r = b / chunks;
  const int64_t s = (b % chunks) * 256 + threadIdx.x;
  if (s >= n_sims) return;
  const int q = (int)(r / ita8);
  const int f0 = (int)(r % ita8) * 8;
  const int4* src = reinterpret_cast<const int4*>(blocks + (r * bstride + s) * 8);
  const int4 lo = src[0], hi = src[1];
  const int32_t w[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
  for (int j = 0; j < 8; ++j)
    if (f0 + j < ita) rows[((int64_t)q * ita + f0 + j) * rstride + s] = w[j];
}

// Check a pst_network / pst_config pair and build the dense host table.
// host_rows = false: rates, x0 and cfg->I are device arrays, and their
// values are not checked on the host.
int prepare(const pst_network* net, const pst_config* cfg, const double* rates, const int32_t* x0,
            int32_t n_sets, pst::HostNet* h, bool host_rows = true) {
  if (!net || !cfg || !rates || !x0) return fail(PST_EINVAL, "null network, config, rates or x0");
  const int N = net->n_species, M = net->n_reactions;
  if (N < 1 || M < 1) return fail(PST_EINVAL, "network needs at least one species and one reaction (N=%d, M=%d)", N, M);
  if (N > kMaxSpecies || M > kMaxReactions)
    return fail(PST_ETOOLARGE, "network has N=%d species and M=%d reactions; this build handles N<=%d, M<=%d",
                N, M, kMaxSpecies, kMaxReactions);
  if (cfg->method < 0 || cfg->method > 2) return fail(PST_EINVAL, "unknown method %d", cfg->method);
  if (cfg->precision < 0 || cfg->precision > 1) return fail(PST_EINVAL, "unknown precision %d", cfg->precision);
  if (cfg->ita < 1 || !cfg->I) return fail(PST_EINVAL, "need at least one sample time (ita=%d)", cfg->ita);
  if (cfg->kappa < 1 || !cfg->E) return fail(PST_EINVAL, "need at least one recorded species (kappa=%d)", cfg->kappa);
  if (cfg->kappa > N) return fail(PST_EINVAL, "kappa=%d exceeds the number of species %d", cfg->kappa, N);
  if (cfg->n_sims < 0) return fail(PST_EINVAL, "negative simulation count");
  if (cfg->sims_per_set < 1) return fail(PST_EINVAL, "sims_per_set must be >= 1");
  const int64_t need_sets = cfg->n_sims == 0 ? 0 : ((int64_t)cfg->n_sims + cfg->sims_per_set - 1) / cfg->sims_per_set;
  if (n_sets < need_sets) return fail(PST_EINVAL, "%d parameter rows given, %lld needed", n_sets, (long long)need_sets);
  if (cfg->n_c < 0) return fail(PST_EINVAL, "n_c must be >= 0");
  if (cfg->max_steps < 0 || cfg->max_steps > 0xffffffffll)
    return fail(PST_EINVAL, "max_steps must be in [0, 2^32)");
  if (cfg->ssa_burst < 1) return fail(PST_EINVAL, "ssa_burst must be >= 1");
  if (!(cfg->eta > 0.0) || !std::isfinite(cfg->eta)) return fail(PST_EINVAL, "eta must be a positive finite number");
  if (cfg->method == PST_METHOD_PURE_SSA && !std::isfinite(cfg->t_max))
    return fail(PST_EINVAL, "t_max must be finite for pure SSA");
  for (int f = 0; host_rows && f < cfg->ita; ++f)
    if (!std::isfinite(cfg->I[f])) return fail(PST_EINVAL, "sample time %d is not finite", f);
  if (net->a_size < 0 || (net->a_size > 0 && !net->A)) return fail(PST_EINVAL, "bad A table");
  if (net->v_size < 0 || (net->v_size > 0 && !net->V)) return fail(PST_EINVAL, "bad V table");
  if (!net->H || !net->H_type || !net->prop_order || !net->prop_step)
    return fail(PST_EINVAL, "null H, H_type, prop_order or prop_step");
  if (net->reserved != 0) return fail(PST_EINVAL, "reserved network field must be 0");
  const bool custom = net->hazards != nullptr;
  if (custom) {
    if (cfg->specialize != 1)
      return fail(PST_EINVAL, "hazard expressions are compiled into a network kernel: set specialize = 1");
    if (net->n_params < 0 || net->n_params > kMaxParams)
      return fail(PST_EINVAL, "n_params=%d out of range [0, %d]", net->n_params, kMaxParams);
    for (int j = 0; j < M; ++j) {
      std::string why;
      if (!net->hazards[j] || !net->hazards[j][0]) return fail(PST_EINVAL, "hazard %d is empty", j);
      if (!hazard_ok(net->hazards[j], N, net->n_params, &why)) return fail(PST_EINVAL, "hazard %d: %s", j, why.c_str());
    }
  }
  const int n_rates = custom ? net->n_params : M;

  std::memset(h, 0, sizeof(*h));
  h->N = N;
  h->M = M;
  h->kappa = cfg->kappa;
  for (int e = 0; e < net->a_size; ++e) {
    const int i = net->A[4 * e], j = net->A[4 * e + 1];
    if (i < 0 || i >= N || j < 0 || j >= M) return fail(PST_EINVAL, "A entry %d out of range", e);
    h->irs[i] = 1;
  }
  for (int e = 0; e < net->v_size; ++e) {
    const int i = net->V[4 * e], j = net->V[4 * e + 1], v = net->V[4 * e + 2];
    if (i < 0 || i >= N || j < 0 || j >= M) return fail(PST_EINVAL, "V entry %d out of range", e);
    if (v == 0 || v < -127 || v > 127) return fail(PST_EINVAL, "V entry %d has value %d (must be in [-127,127], non-zero)", e, v);
    if (h->v[i * M + j] != 0) return fail(PST_EINVAL, "V has two entries for species %d, reaction %d", i, j);
    if (e > 0) {
      const int pi = net->V[4 * (e - 1)], pj = net->V[4 * (e - 1) + 1];
      if (pi > i || (pi == i && pj >= j)) return fail(PST_EINVAL, "V entries must be species-major and ascending");
    }
    h->v[i * M + j] = (int8_t)v;
  }
  for (int i = 0; i < N; ++i) {
    const int hv = net->H[i], ht = net->H_type[i];
    if (hv < 0 || hv > 127 || ht < 0 || ht > 127) return fail(PST_EINVAL, "H/H_type of species %d out of range", i);
    h->H[i] = (int8_t)hv;
    h->Ht[i] = (int8_t)ht;
    for (int j = 0; j < M; ++j) {
      const int p = net->prop_order[i * M + j], s = net->prop_step[i * M + j];
      if (p < 0 || p > 127) return fail(PST_EINVAL, "prop_order[%d][%d]=%d out of range", i, j, p);
      if (s != 0 && s != 1) return fail(PST_EINVAL, "prop_step[%d][%d]=%d must be 0 or 1", i, j, s);
      h->ord[i * M + j] = (int8_t)p;
      h->stp[i * M + j] = (int8_t)s;
    }
  }
  for (int q = 0; q < cfg->kappa; ++q) {
    const int s = cfg->E[q];
    if (s < 0 || s >= N) return fail(PST_EINVAL, "recorded species %d out of range", s);
    h->E[q] = (int8_t)s;
  }
  for (int64_t r = 0; host_rows && r < need_sets; ++r) {
    for (int j = 0; j < n_rates; ++j) {
      const double k = rates[r * n_rates + j];
      if (!std::isfinite(k)) return fail(PST_EINVAL, "rate %d of row %lld is not finite", j, (long long)r);
      if (!custom && !(k >= 0.0)) return fail(PST_EINVAL, "rate %d of row %lld must be >= 0", j, (long long)r);
    }
    for (int i = 0; i < N; ++i)
      if (x0[r * N + i] < 0) return fail(PST_EINVAL, "initial count %d of row %lld is negative", i, (long long)r);
  }
  return PST_OK;
}

int check_device(int* cu_count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(PST_ENODEV, "no HIP device visible");
  int dev = 0;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return hip_fail(e, "hipGetDevice");
  hipDeviceProp_t p;
  if ((e = hipGetDeviceProperties(&p, dev)) != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
  if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
    return fail(PST_ENODEV, "device %d is %s; this library is built for gfx950 only", dev, p.gcnArchName);
  *cu_count = p.multiProcessorCount;
  return PST_OK;
}

// one job of a plan: validated tables on the host, parameters on the device
struct PlanJob {
  pst::HostNet net;
  pst::JobParams p;  // O / o_stride / flags filled at execute time
  int cls = -1;      // size class of the table-driven kernel, or -1
  int fixed = -1;    // group of the specialised kernel, or -1
  bool wave = false; // large-network kernel (one wavefront per simulation)
  std::vector<unsigned char> wave_tab;  // its sparse table (a WaveTab)
  std::vector<std::string> hazards;  // compiled hazard expressions (empty: mass action)
  double* d_rates = nullptr;
  int32_t* d_x0 = nullptr;
  float* d_I = nullptr;
  uint32_t* d_state = nullptr;  // saved simulations between sample ranges (allocated on first use)
  int32_t state_words = 0;
};

template <int NP, int MP>
void pack_jobs(const std::vector<PlanJob>& jobs, int cls, std::vector<unsigned char>& out, int* count) {
  std::vector<pst::JobDev<NP, MP>> v;
  for (const PlanJob& j : jobs) {
    if (j.cls != cls) continue;
    pst::JobDev<NP, MP> d;
    std::memset(&d, 0, sizeof(d));
    d.tab = pst::make_tab<NP, MP>(j.net);
    d.p = j.p;
    v.push_back(d);
  }
  out.resize(v.size() * sizeof(pst::JobDev<NP, MP>));
  if (!v.empty()) std::memcpy(out.data(), v.data(), out.size());
  *count = (int)v.size();
}

// Simulations per wavefront for one launch (jobs `ids`): 64 unless the launch
// holds fewer than 64 simulations per SIMD, then the smallest power of two
// that still needs no more wavefronts than the chip has SIMDs.  A wavefront
// issues the same instructions whatever its lane count, so spreading a small
// launch over every SIMD costs nothing per wavefront and leaves each one with
// fewer divergent paths and a shorter slowest-lane tail.
int lanes_for(const std::vector<PlanJob>& jobs, const std::vector<int>& ids, int simds) {
  auto need = [&](int L) {
    int64_t n = 0;
    for (int q : ids) n += ((int64_t)jobs[q].p.n_sims + L - 1) / L;
    return n;
  };
  int L = 64;
  while (L > 1 && need(L / 2) <= (int64_t)simds) L /= 2;
  return L;
}

}  // namespace

enum { KIND_TABLE = 0, KIND_FIXED = 1, KIND_WAVE = 2 };

// Which kernel runs a job: a network-specific kernel (specialize = 1, or
// compiled hazards), a table-driven size class, or -- networks beyond both,
// mixed precision -- the large-network kernel.  Negative: error status.
int kernel_kind(const pst_job& J, const pst::HostNet& h) {
  const bool fits_rtc = h.N <= pst::RTC_MAX_SPECIES && h.M <= pst::RTC_MAX_REACTIONS;
  if (J.net.hazards) {
    if (!fits_rtc)
      return fail(PST_ETOOLARGE, "compiled hazard expressions: network kernels take N<=%d, M<=%d (N=%d, M=%d)",
                  pst::RTC_MAX_SPECIES, pst::RTC_MAX_REACTIONS, h.N, h.M);
    return KIND_FIXED;
  }
  if (J.cfg.specialize && fits_rtc) return KIND_FIXED;
  if (pst::size_class(h.N, h.M, h.kappa) >= 0) return KIND_TABLE;
  if (J.cfg.precision != PST_PRECISION_MIXED)
    return fail(PST_ETOOLARGE,
                "N=%d, M=%d: FP64 kernels take N<=16 and M<=16 (M<=%d with specialize=1); larger networks run in "
                "mixed precision", h.N, h.M, pst::RTC_MAX_REACTIONS);
  std::vector<unsigned char> buf(sizeof(pst::WaveTab));
  char why[160];
  if (!pst::make_wave_tab(h, (pst::WaveTab*)buf.data(), why, sizeof(why)))
    return fail(PST_ETOOLARGE, "large-network kernel: %s", why);
  return KIND_WAVE;
}

// Network kernels are built without an occupancy request, and once more, on
// first need, asking for register use that fits kOccupancyWaves wavefronts
// per SIMD.  A launch that needs more wavefronts than the plain build holds
// resident uses the second build if it holds more: Schloegl tau-leaping,
// 262,144 simulations (238 VGPRs, two wavefronts per SIMD, against 168 VGPRs
// with a few spills, three): mixed 63.0 -> 61.4 ms, FP64 164.3 -> 153.0 ms.
// Launches the plain build holds at once keep it: Lotka-Volterra with 65,536
// simulations (one wavefront per SIMD) takes 57.6 ms plain and 66.1 ms with
// the request.  Asking for four wavefronts per SIMD was slower than three on
// every case.  Both builds give identical results.  (Fewer resident
// wavefronts than the limit is the separate waves_trim of the launch shape.)
constexpr int kOccupancyWaves = 3;

// A plan with several kernel launches (size classes, network groups, the
// large-network kernel) spreads them over the caller's stream and up to
// kSideStreams more, forked from and joined back to the caller's stream with
// events.  Every launch is a persistent grid sized for the whole chip, so on
// one stream a launch waits for the slowest simulation of the one before it;
// on separate streams its workgroups take the compute units the earlier
// launch releases.  The kernels share no state but the atomically updated
// counters, and no result depends on the order.
constexpr int kSideStreams = 3;

// jobs sharing one specialised kernel (same network and constant settings)
struct FixedGroup {
  pst::RtcKernel k;
  std::string occ_source;   // source of the occupancy build
  pst::RtcKernel k_occ;     // compiled on first need
  bool occ_tried = false;
  bool use_occ = false;     // this execute launches k_occ
  std::string pair_source;  // source of the paired build (decide + burst wavefronts)
  pst::RtcKernel k_pair;    // compiled on first need
  bool pair_tried = false;
  bool use_pair = false;    // this execute launches k_pair
  std::vector<int> jobs;
  void* d_jobs = nullptr;
  size_t d_bytes = 0;
  std::vector<unsigned char> uploaded;
};

struct pst_plan {
  int device = -1;
  int precision = 0;
  int cu_count = 0;
  std::vector<PlanJob> jobs;
  void* d_jobs[pst::NUM_CLASSES] = {nullptr, nullptr, nullptr};
  size_t d_jobs_bytes[pst::NUM_CLASSES] = {0, 0, 0};
  std::vector<unsigned char> uploaded[pst::NUM_CLASSES];  // last descriptors copied to d_jobs
  std::vector<FixedGroup> fixed;
  void* d_wave = nullptr;  // WaveJob array of the large-network jobs
  size_t d_wave_bytes = 0;
  std::vector<unsigned char> uploaded_wave;
  unsigned int* d_queue = nullptr;
  unsigned long long* d_stats = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool executed = false;      // the last execute enqueued every launch
  bool enqueued = false;      // ev1 marks the end of all work of the last execute
  int lanes_shift = 0;        // launch shape (pst_plan_autotune)
  int ssa_slice = kSsaSlice;
  int waves_trim = 0;
  int burst_waves = 0;        // 1: tau-leaping network kernels run bursts on paired wavefronts
  int range_pos = 0;          // samples every simulation has recorded in the pending ranges (0: none)
  uint64_t generation = 0;    // changes whenever the descriptors or the launch shape change
  hipGraph_t graph = nullptr; // sample-range chunks captured by pst_plan_graph_create
  hipGraphExec_t graph_exec = nullptr;
  uint64_t graph_generation = 0;
  // extra streams for plans with several kernel launches (created on first need)
  hipStream_t side[kSideStreams] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kSideStreams] = {};
};

// resident 256-lane blocks per CU for a kernel whose occupancy limit is bpc,
// after the plan's waves_trim (one block = one wavefront per SIMD)
static int trimmed(int bpc, int trim) { return std::max(1, bpc - trim); }

// Wait for every launch of the plan that may still read its descriptors: the
// caller's stream, the end of the previous execute (recorded on the stream
// that execute was given, which may be another one, after its side streams
// were joined back into it, also when it failed part way), and the side
// streams themselves.
static hipError_t sync_plan(pst_plan* p, hipStream_t s) {
  hipError_t e = hipStreamSynchronize(s);
  if (e == hipSuccess && p->enqueued) e = hipEventSynchronize(p->ev1);
  for (int i = 0; i < kSideStreams && e == hipSuccess; ++i)
    if (p->side[i]) e = hipStreamSynchronize(p->side[i]);
  return e;
}

// The kernel launches of one execute: size classes, network groups, then the
// large-network kernel, each on the stream next_stream() gives.
template <class NextStream>
static int enqueue_launches(pst_plan* p, const int* count, int n_wave, int64_t wave_sims, pst::Range range,
                            NextStream& next_stream) {
  hipError_t e;
  int qoff = 0;
  for (int c = 0; c < pst::NUM_CLASSES; ++c) {
    if (count[c] == 0) continue;
    const hipStream_t ls = next_stream();
    int64_t need = 0;
    for (const PlanJob& pj : p->jobs)
      if (pj.cls == c) need += ((int64_t)pj.p.n_sims + pj.p.lanes - 1) / pj.p.lanes;
    const int bpc = (p->precision == PST_PRECISION_MIXED) ? pst::blocks_per_cu_mixed(c) : pst::blocks_per_cu_fp64(c);
    const int64_t resident = (int64_t)trimmed(bpc, p->waves_trim) * p->cu_count * 4;
    const int waves = (int)(need < resident ? need : resident);
    // the queue counters of this class's jobs are queue[qoff .. qoff + count)
    const int rc = (p->precision == PST_PRECISION_MIXED)
                       ? pst::launch_mixed(c, p->d_jobs[c], count[c], p->d_queue + qoff, p->d_stats, waves, range, ls)
                       : pst::launch_fp64(c, p->d_jobs[c], count[c], p->d_queue + qoff, p->d_stats, waves, range, ls);
    if (rc < 0) return fail(PST_ETOOLARGE, "no compiled kernel for size class %d", c);
    if (rc != 0) return hip_fail((hipError_t)rc, "simulation kernel launch");
    qoff += count[c];
  }
  for (FixedGroup& g : p->fixed) {
    int64_t need = 0;
    for (int q : g.jobs) need += ((int64_t)p->jobs[q].p.n_sims + p->jobs[q].p.lanes - 1) / p->jobs[q].p.lanes;
    const pst::RtcKernel* k = g.use_pair ? &g.k_pair : g.use_occ ? &g.k_occ : &g.k;
    const int64_t resident = (int64_t)trimmed(k->blocks_per_cu, p->waves_trim) * p->cu_count * 4;
    const int waves = (int)(need < resident ? need : resident);
    const int cnt = (int)g.jobs.size();
    const hipStream_t ls = next_stream();
    if ((e = pst::rtc_launch(*k, (const pst::JobParams*)g.d_jobs, cnt, p->d_queue + qoff, p->d_stats, waves, range,
                             ls)) !=
        hipSuccess)
      return hip_fail(e, "specialised kernel launch");
    qoff += cnt;
  }
  if (n_wave > 0) {
    // one simulation per wavefront at a time
    const int64_t resident = (int64_t)trimmed(pst::blocks_per_cu_wave(), p->waves_trim) * p->cu_count * 4;
    const int waves = (int)(wave_sims < resident ? wave_sims : resident);
    const int rc = pst::launch_wave(p->d_wave, n_wave, p->d_queue + qoff, p->d_stats, waves, next_stream());
    if (rc != 0) return hip_fail((hipError_t)rc, "large-network kernel launch");
    qoff += n_wave;
  }
  return PST_OK;
}

// ---------------------------------------------------------------------------
// pst_run_noalloc: the library's only device memory for it is this array (one
// copy per device, held until the process ends): [0, 128) the counters,
// [128, 132) the work-queue counter, [kNoAllocJob, kNoAllocBytes) the job
// descriptor of one table-driven size class.  Every call on a device shares
// its array, so calls that overlap on one device race.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocBytes = PST_NOALLOC_SCRATCH_BYTES;
constexpr size_t kNoAllocQueue = 128;
constexpr size_t kNoAllocJob = 256;
static_assert(pst::ST_COUNT * sizeof(unsigned long long) <= kNoAllocQueue, "counters overlap the queue");
static_assert(sizeof(pst::JobDev<16, 16>) <= kNoAllocBytes - kNoAllocJob, "descriptor does not fit");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per
// device; the array's name is not a device address in host code).
static int scratch_on_device(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(PST_ENODEV, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = (unsigned char*)ptr;
  }
  *out = g_scratch_ptr[dev];
  return PST_OK;
}

// Writes the job descriptor (a kernel argument, so nothing is staged through
// host memory) behind the counters and clears the counters and the queue.
template <int NP, int MP>
__global__ void __launch_bounds__(256) noalloc_setup(pst::JobDev<NP, MP> jd, unsigned char* base) {
  constexpr int NW = (int)(sizeof(pst::JobDev<NP, MP>) / 4);
  static_assert(sizeof(pst::JobDev<NP, MP>) % 4 == 0, "descriptor is not whole words");
  const uint32_t* src = reinterpret_cast<const uint32_t*>(&jd);
  uint32_t* dst = reinterpret_cast<uint32_t*>(base + kNoAllocJob);
  for (int i = (int)threadIdx.x; i < NW; i += (int)blockDim.x) dst[i] = src[i];
  if (threadIdx.x < kNoAllocJob / 4) reinterpret_cast<uint32_t*>(base)[threadIdx.x] = 0u;
}

template <int NP, int MP>
static hipError_t noalloc_enqueue(const pst::HostNet& h, const pst::JobParams& P, unsigned char* base,
                                  hipStream_t s) {
  pst::JobDev<NP, MP> jd;
  std::memset(&jd, 0, sizeof(jd));
  jd.tab = pst::make_tab<NP, MP>(h);
  jd.p = P;
  hipLaunchKernelGGL((noalloc_setup<NP, MP>), dim3(1), dim3(256), 0, s, jd, base);
  return hipGetLastError();
}

extern "C" {

const char* pst_last_error(void) { return g_err; }

int pst_abi_version(void) { return PST_ABI_VERSION; }

int pst_blocks_to_rows(const int32_t* blocks, int64_t blocks_stride, int32_t* rows, int64_t rows_stride,
                       int32_t kappa, int32_t ita, int64_t n_sims, void* stream) {
  g_err[0] = 0;
  if (kappa < 1 || ita < 1 || n_sims < 0) return fail(PST_EINVAL, "need kappa >= 1, ita >= 1, n_sims >= 0");
  if (blocks_stride < n_sims || rows_stride < n_sims)
    return fail(PST_EINVAL, "strides (%lld, %lld) below n_sims %lld", (long long)blocks_stride, (long long)rows_stride,
                (long long)n_sims);
  if (n_sims == 0) return PST_OK;
  if (!blocks || !rows) return fail(PST_EINVAL, "null trajectory array");
  if (((uintptr_t)blocks & 15u) != 0) return fail(PST_EINVAL, "blocks array is not 16-byte aligned");
  const int ita8 = (ita + 7) / 8;
  const int64_t chunks = (n_sims + 255) / 256;
  const int64_t grid = (int64_t)kappa * ita8 * chunks;
  if (grid > 0x7fffffffLL) return fail(PST_ETOOLARGE, "%lld workgroups exceed one launch", (long long)grid);
  hipLaunchKernelGGL(blocks_to_rows, dim3((unsigned int)grid), dim3(256), 0, (hipStream_t)stream, blocks,
                     blocks_stride, rows, rows_stride, ita, ita8, n_sims, chunks);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? PST_OK : hip_fail(e, "blocks_to_rows launch");
}

void pst_max_network(int32_t* max_species, int32_t* max_reactions) {
  if (max_species) *max_species = kMaxSpecies;
  if (max_reactions) *max_reactions = kMaxReactions;
}

void pst_plan_destroy(pst_plan* p) {
  if (!p) return;
  for (PlanJob& j : p->jobs) {
    if (j.d_rates) (void)hipFree(j.d_rates);
    if (j.d_x0) (void)hipFree(j.d_x0);
    if (j.d_I) (void)hipFree(j.d_I);
  }
  for (int c = 0; c < pst::NUM_CLASSES; ++c)
    if (p->d_jobs[c]) (void)hipFree(p->d_jobs[c]);
  for (FixedGroup& g : p->fixed)
    if (g.d_jobs) (void)hipFree(g.d_jobs);
  if (p->d_wave) (void)hipFree(p->d_wave);
  if (p->d_queue) (void)hipFree(p->d_queue);
  if (p->d_stats) (void)hipFree(p->d_stats);
  if (p->ev0) (void)hipEventDestroy(p->ev0);
  if (p->ev1) (void)hipEventDestroy(p->ev1);
  for (int i = 0; i < kSideStreams; ++i) {
    if (p->side[i]) (void)hipStreamDestroy(p->side[i]);
    if (p->join[i]) (void)hipEventDestroy(p->join[i]);
  }
  if (p->fork) (void)hipEventDestroy(p->fork);
  for (PlanJob& j : p->jobs)
    if (j.d_state) (void)hipFree(j.d_state);
  if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  delete p;
}

int pst_batch_plan_create(const pst_job* jobs, int32_t n_jobs, pst_plan** out) {
  g_err[0] = 0;
  if (!out) return fail(PST_EINVAL, "null plan pointer");
  *out = nullptr;
  if (!jobs || n_jobs < 1) return fail(PST_EINVAL, "need at least one job");
  // validate every job on the host before touching the device
  std::vector<pst::HostNet> nets(n_jobs);
  std::vector<int> kinds(n_jobs, 0);
  for (int q = 0; q < n_jobs; ++q) {
    const pst_job& J = jobs[q];
    int rc = prepare(&J.net, &J.cfg, J.rates, J.x0, J.n_sets, &nets[q]);
    if (rc != PST_OK) {
      if (n_jobs > 1) {
        char msg[512];
        std::snprintf(msg, sizeof(msg), "job %d: %s", q, g_err);
        std::snprintf(g_err, sizeof(g_err), "%s", msg);
      }
      return rc;
    }
    if (J.cfg.precision != jobs[0].cfg.precision)
      return fail(PST_EINVAL, "job %d has precision %d, job 0 has %d: one precision per batch", q,
                  J.cfg.precision, jobs[0].cfg.precision);
    if (J.cfg.specialize != 0 && J.cfg.specialize != 1) return fail(PST_EINVAL, "job %d: specialize must be 0 or 1", q);
    if (J.cfg.layout != PST_LAYOUT_ROWS && J.cfg.layout != PST_LAYOUT_BLOCKS)
      return fail(PST_EINVAL, "job %d: unknown trajectory layout %d", q, J.cfg.layout);
    const int kd = kernel_kind(J, nets[q]);
    if (kd < 0) return kd;
    kinds[q] = kd;
  }
  int cu = 0;
  int rc = check_device(&cu);
  if (rc != PST_OK) return rc;
  pst_plan* p = new (std::nothrow) pst_plan();
  if (!p) return fail(PST_ENOMEM, "host allocation failed");
  (void)hipGetDevice(&p->device);
  p->cu_count = cu;
  p->precision = jobs[0].cfg.precision;
  p->jobs.resize(n_jobs);
  hipError_t e;
  for (int q = 0; q < n_jobs; ++q) {
    const pst_job& J = jobs[q];
    PlanJob& pj = p->jobs[q];
    pj.net = nets[q];
    pj.cls = kinds[q] == KIND_TABLE ? pst::size_class(pj.net.N, pj.net.M, pj.net.kappa) : -1;
    if (kinds[q] == KIND_WAVE) {
      pj.wave = true;
      pj.wave_tab.resize(sizeof(pst::WaveTab));
      char why[160];
      (void)pst::make_wave_tab(pj.net, (pst::WaveTab*)pj.wave_tab.data(), why, sizeof(why));
    }
    const int N = J.net.n_species, M = J.net.n_reactions;
    if (J.net.hazards)
      for (int j = 0; j < M; ++j) pj.hazards.emplace_back(J.net.hazards[j]);
    const int n_rates = J.net.hazards ? J.net.n_params : M;
    const int64_t sets = J.cfg.n_sims == 0 ? 1 : ((int64_t)J.cfg.n_sims + J.cfg.sims_per_set - 1) / J.cfg.sims_per_set;
    const int64_t rate_words = sets * (n_rates > 0 ? n_rates : 1);
    if ((e = hipMalloc(&pj.d_rates, sizeof(double) * rate_words)) != hipSuccess ||
        (e = hipMalloc(&pj.d_x0, sizeof(int32_t) * sets * N)) != hipSuccess ||
        (e = hipMalloc(&pj.d_I, sizeof(float) * J.cfg.ita)) != hipSuccess) {
      pst_plan_destroy(p);
      return hip_fail(e, "hipMalloc(job tables)");
    }
    if ((e = hipMemset(pj.d_rates, 0, sizeof(double) * rate_words)) != hipSuccess ||
        (n_rates > 0 &&
         (e = hipMemcpy(pj.d_rates, J.rates, sizeof(double) * sets * n_rates, hipMemcpyHostToDevice)) != hipSuccess) ||
        (e = hipMemcpy(pj.d_x0, J.x0, sizeof(int32_t) * sets * N, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(pj.d_I, J.cfg.I, sizeof(float) * J.cfg.ita, hipMemcpyHostToDevice)) != hipSuccess) {
      pst_plan_destroy(p);
      return hip_fail(e, "hipMemcpy(job tables)");
    }
    pst::JobParams& P = pj.p;
    std::memset(&P, 0, sizeof(P));
    P.N = N;
    P.M = M;
    P.kappa = J.cfg.kappa;
    P.ita = J.cfg.ita;
    P.n_c = J.cfg.n_c;
    P.ssa_burst = J.cfg.ssa_burst;
    P.method = J.cfg.method;
    P.n_sims = J.cfg.n_sims;
    P.sims_per_set = J.cfg.sims_per_set;
    P.ssa_slice = kSsaSlice;
    P.lanes = 64;
    P.n_rates = n_rates;
    P.max_steps = (uint32_t)J.cfg.max_steps;
    P.layout = J.cfg.layout;
    P.eta = J.cfg.eta;
    P.t_max = J.cfg.t_max;
    P.seed = J.cfg.seed;
    P.sim_offset = J.cfg.sim_offset;
    P.rates = pj.d_rates;
    P.x0 = pj.d_x0;
    P.I = pj.d_I;
    if (kinds[q] == KIND_FIXED) {
      pst::RtcKernel k;
      std::string err;
      const int rc2 = pst::rtc_kernel(pst::rtc_source(pj.net, P, p->precision, pj.hazards), p->precision, &k, &err);
      if (rc2 != PST_OK) {
        pst_plan_destroy(p);
        return fail(rc2, "job %d: %s", q, err.c_str());
      }
      int g = 0;
      while (g < (int)p->fixed.size() && p->fixed[g].k.fn != k.fn) ++g;
      if (g == (int)p->fixed.size()) {
        p->fixed.emplace_back();
        p->fixed[g].k = k;
        p->fixed[g].occ_source = pst::rtc_source(pj.net, P, p->precision, pj.hazards, kOccupancyWaves);
        p->fixed[g].pair_source = pst::rtc_source(pj.net, P, p->precision, pj.hazards, 0, true);
      }
      p->fixed[g].jobs.push_back(q);
      pj.fixed = g;
    }
  }
  if ((e = hipMalloc(&p->d_queue, sizeof(unsigned int) * n_jobs)) != hipSuccess ||
      (e = hipMalloc(&p->d_stats, sizeof(unsigned long long) * pst::ST_COUNT)) != hipSuccess) {
    pst_plan_destroy(p);
    return hip_fail(e, "hipMalloc(queue, stats)");
  }
  if ((e = hipEventCreate(&p->ev0)) != hipSuccess || (e = hipEventCreate(&p->ev1)) != hipSuccess) {
    pst_plan_destroy(p);
    return hip_fail(e, "hipEventCreate");
  }
  *out = p;
  return PST_OK;
}

int pst_plan_create(const pst_network* net, const pst_config* cfg, const double* rates, const int32_t* x0,
                    int32_t n_sets, pst_plan** out) {
  g_err[0] = 0;
  if (!net || !cfg) return fail(PST_EINVAL, "null network or config");
  pst_job j;
  j.net = *net;
  j.cfg = *cfg;
  j.rates = rates;
  j.x0 = x0;
  j.n_sets = n_sets;
  return pst_batch_plan_create(&j, 1, out);
}

// Everything an execute needs before its launches: buffer pointers, launch
// shape and kernel builds per group, and the job descriptors in device
// memory (uploaded, after the plan's earlier launches, only when they
// changed).  with_state: the descriptors point at the jobs' saved-state
// buffers (sample ranges).
static int prepare_plan(pst_plan* p, int32_t* const* O_dev, const int64_t* o_strides, int32_t* const* flags_dev,
                        hipStream_t s, bool with_state, int* count, int* n_wave_out, int64_t* wave_sims_out) {
  if (!O_dev || !o_strides) return fail(PST_EINVAL, "null trajectory buffer list");
  int dev = -1;
  (void)hipGetDevice(&dev);
  if (dev != p->device) return fail(PST_EINVAL, "plan was created on device %d, current device is %d", p->device, dev);
  const int n_jobs = (int)p->jobs.size();
  for (int q = 0; q < n_jobs; ++q) {
    PlanJob& pj = p->jobs[q];
    if (pj.p.n_sims > 0 && !O_dev[q]) return fail(PST_EINVAL, "job %d: null trajectory buffer", q);
    if (o_strides[q] < pj.p.n_sims)
      return fail(PST_EINVAL, "job %d: o_stride %lld < n_sims %d", q, (long long)o_strides[q], pj.p.n_sims);
    pj.p.O = O_dev[q];
    pj.p.o_stride = o_strides[q];
    pj.p.flags = flags_dev ? flags_dev[q] : nullptr;
    pj.p.state = with_state ? pj.d_state : nullptr;
    pj.p.state_words = with_state ? pj.state_words : 0;
  }
  hipError_t e;
  // simulations per wavefront, per launch, and SSA-only steps per pass
  const int simds = p->cu_count * 4;
  for (int c = 0; c < pst::NUM_CLASSES; ++c) {
    std::vector<int> ids;
    for (int q = 0; q < n_jobs; ++q)
      if (p->jobs[q].cls == c) ids.push_back(q);
    const int L = std::max(1, lanes_for(p->jobs, ids, simds) >> p->lanes_shift);
    for (int q : ids) p->jobs[q].p.lanes = L;
  }
  for (FixedGroup& g : p->fixed) {
    const int L = std::max(1, lanes_for(p->jobs, g.jobs, simds) >> p->lanes_shift);
    for (int q : g.jobs) p->jobs[q].p.lanes = L;
  }
  for (PlanJob& pj : p->jobs) pj.p.ssa_slice = p->ssa_slice;
  // network kernels: the occupancy build for launches the plain build cannot
  // hold resident at once
  for (FixedGroup& g : p->fixed) {
    int64_t need = 0;
    for (int q : g.jobs) need += ((int64_t)p->jobs[q].p.n_sims + p->jobs[q].p.lanes - 1) / p->jobs[q].p.lanes;
    g.use_occ = false;
    if (need <= (int64_t)g.k.blocks_per_cu * simds) continue;
    if (!g.occ_tried) {
      g.occ_tried = true;
      std::string err;
      const int rc = pst::rtc_kernel(g.occ_source, p->precision, &g.k_occ, &err);
      if (rc != PST_OK) return fail(rc, "network kernel (occupancy build): %s", err.c_str());
    }
    g.use_occ = g.k_occ.fn != nullptr && g.k_occ.blocks_per_cu > g.k.blocks_per_cu;
  }
  // network kernels of tau-leaping jobs: SSA bursts on paired wavefronts
  for (FixedGroup& g : p->fixed) {
    g.use_pair = false;
    if (!p->burst_waves || p->jobs[g.jobs[0]].p.method != PST_METHOD_TAU_LEAPING) continue;
    if (!g.pair_tried) {
      g.pair_tried = true;
      std::string err;
      const int rc = pst::rtc_kernel(g.pair_source, p->precision, &g.k_pair, &err, true);
      if (rc != PST_OK) return fail(rc, "network kernel (paired build): %s", err.c_str());
    }
    if (!g.k_pair.fn) return fail(PST_EHIP, "network kernel (paired build) is not available");
    g.use_pair = true;
  }
  // job descriptors, one device array per size class
  for (int c = 0; c < pst::NUM_CLASSES; ++c) count[c] = 0;
  for (int c = 0; c < pst::NUM_CLASSES; ++c) {
    std::vector<unsigned char> bytes;
    if (c == 0) pack_jobs<4, 4>(p->jobs, c, bytes, &count[c]);
    if (c == 1) pack_jobs<8, 8>(p->jobs, c, bytes, &count[c]);
    if (c == 2) pack_jobs<16, 16>(p->jobs, c, bytes, &count[c]);
    if (count[c] == 0) continue;
    if (bytes == p->uploaded[c]) continue;  // unchanged since the last execute
    if (p->d_jobs_bytes[c] < bytes.size()) {
      if (p->d_jobs[c]) (void)hipFree(p->d_jobs[c]);
      p->d_jobs[c] = nullptr;
      p->d_jobs_bytes[c] = 0;
      if ((e = hipMalloc(&p->d_jobs[c], bytes.size())) != hipSuccess) return hip_fail(e, "hipMalloc(jobs)");
      p->d_jobs_bytes[c] = bytes.size();
    }
    // the descriptor array may still be read by an earlier launch of this plan
    if ((e = sync_plan(p, s)) != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    if ((e = hipMemcpy(p->d_jobs[c], bytes.data(), bytes.size(), hipMemcpyHostToDevice)) != hipSuccess)
      return hip_fail(e, "hipMemcpy(jobs)");
    p->uploaded[c] = bytes;
    p->generation += 1;
  }
  // parameter blocks of the specialised kernels
  for (FixedGroup& g : p->fixed) {
    const int cnt = (int)g.jobs.size();
    std::vector<unsigned char> bytes(sizeof(pst::JobParams) * cnt);
    for (int i = 0; i < cnt; ++i)
      std::memcpy(bytes.data() + i * sizeof(pst::JobParams), &p->jobs[g.jobs[i]].p, sizeof(pst::JobParams));
    if (bytes == g.uploaded) continue;
    if (g.d_bytes < bytes.size()) {
      if (g.d_jobs) (void)hipFree(g.d_jobs);
      g.d_jobs = nullptr;
      g.d_bytes = 0;
      if ((e = hipMalloc(&g.d_jobs, bytes.size())) != hipSuccess) return hip_fail(e, "hipMalloc(jobs)");
      g.d_bytes = bytes.size();
    }
    if ((e = sync_plan(p, s)) != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    if ((e = hipMemcpy(g.d_jobs, bytes.data(), bytes.size(), hipMemcpyHostToDevice)) != hipSuccess)
      return hip_fail(e, "hipMemcpy(jobs)");
    g.uploaded = bytes;
    p->generation += 1;
  }
  // large-network jobs: parameters and sparse table per job
  int n_wave = 0;
  int64_t wave_sims = 0;
  {
    std::vector<unsigned char> bytes;
    for (const PlanJob& pj : p->jobs) {
      if (!pj.wave) continue;
      const size_t off = bytes.size();
      bytes.resize(off + sizeof(pst::WaveJob));
      pst::WaveJob* wj = (pst::WaveJob*)(bytes.data() + off);
      std::memcpy(&wj->p, &pj.p, sizeof(pst::JobParams));
      std::memcpy(&wj->t, pj.wave_tab.data(), sizeof(pst::WaveTab));
      ++n_wave;
      wave_sims += pj.p.n_sims;
    }
    if (n_wave > 0 && bytes != p->uploaded_wave) {
      if (p->d_wave_bytes < bytes.size()) {
        if (p->d_wave) (void)hipFree(p->d_wave);
        p->d_wave = nullptr;
        p->d_wave_bytes = 0;
        if ((e = hipMalloc(&p->d_wave, bytes.size())) != hipSuccess) return hip_fail(e, "hipMalloc(jobs)");
        p->d_wave_bytes = bytes.size();
      }
      if ((e = sync_plan(p, s)) != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
      if ((e = hipMemcpy(p->d_wave, bytes.data(), bytes.size(), hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "hipMemcpy(jobs)");
      p->uploaded_wave = bytes;
      p->generation += 1;
    }
  }
  *n_wave_out = n_wave;
  *wave_sims_out = wave_sims;
  return PST_OK;
}

// The launches of one execute (or one sample range) on stream s: the queue
// counters (and, unless accumulate, the event counters) cleared by fill
// kernels, the kernel launches spread over s and the side streams, ev0 / ev1
// around them.
static int enqueue_plan(pst_plan* p, hipStream_t s, const int* count, int n_wave, int64_t wave_sims,
                        pst::Range range, bool accumulate) {
  hipError_t e;
  const int n_jobs = (int)p->jobs.size();
  // The queue and counters are shared by every execute of the plan: the
  // previous one, on whichever stream it was given, must be done with them
  // before they are cleared.
  p->executed = false;
  if (p->enqueued && (e = hipStreamWaitEvent(s, p->ev1, 0)) != hipSuccess) return hip_fail(e, "hipStreamWaitEvent");
  if (!accumulate) hipLaunchKernelGGL(zero_u64, dim3(1), dim3(64), 0, s, p->d_stats, (int)pst::ST_COUNT);
  hipLaunchKernelGGL(zero_u32, dim3((n_jobs + 255) / 256), dim3(256), 0, s, p->d_queue, n_jobs);
  if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "zero kernels");
  if ((e = hipEventRecord(p->ev0, s)) != hipSuccess) return hip_fail(e, "hipEventRecord");
  // streams of the launches: launch i goes to the caller's stream (i % 4 ==
  // 0) or to side stream i % 4 - 1, which first waits for the zeroed counters
  int n_launch = (int)p->fixed.size() + (n_wave > 0 ? 1 : 0);
  for (int c = 0; c < pst::NUM_CLASSES; ++c) n_launch += count[c] > 0 ? 1 : 0;
  const int n_side = std::min(kSideStreams, n_launch - 1);
  if (n_side > 0) {
    if (!p->fork && (e = hipEventCreateWithFlags(&p->fork, hipEventDisableTiming)) != hipSuccess)
      return hip_fail(e, "hipEventCreate");
    for (int i = 0; i < n_side; ++i) {
      if (!p->side[i] && (e = hipStreamCreateWithFlags(&p->side[i], hipStreamNonBlocking)) != hipSuccess)
        return hip_fail(e, "hipStreamCreate");
      if (!p->join[i] && (e = hipEventCreateWithFlags(&p->join[i], hipEventDisableTiming)) != hipSuccess)
        return hip_fail(e, "hipEventCreate");
    }
    if ((e = hipEventRecord(p->fork, s)) != hipSuccess) return hip_fail(e, "hipEventRecord");
    for (int i = 0; i < n_side; ++i)
      if ((e = hipStreamWaitEvent(p->side[i], p->fork, 0)) != hipSuccess) return hip_fail(e, "hipStreamWaitEvent");
  }
  int li = 0;
  auto next_stream = [&]() -> hipStream_t {
    const int k = (li++) % (n_side + 1);
    return k == 0 ? s : p->side[k - 1];
  };
  const int rc_launch = enqueue_launches(p, count, n_wave, wave_sims, range, next_stream);
  // join the side streams back into the caller's stream and mark the end of
  // this execute, after a failed launch too: the next execute and
  // pst_plan_read_stats wait for ev1
  hipError_t ej = hipSuccess;
  for (int i = 0; i < n_side; ++i) {
    if ((e = hipEventRecord(p->join[i], p->side[i])) != hipSuccess && ej == hipSuccess) ej = e;
    if ((e = hipStreamWaitEvent(s, p->join[i], 0)) != hipSuccess && ej == hipSuccess) ej = e;
  }
  if ((e = hipEventRecord(p->ev1, s)) != hipSuccess && ej == hipSuccess) ej = e;
  if (e == hipSuccess) p->enqueued = true;
  if (rc_launch != PST_OK) return rc_launch;
  if (ej != hipSuccess) return hip_fail(ej, "hipEventRecord / hipStreamWaitEvent");
  p->executed = true;
  return PST_OK;
}

int pst_batch_plan_execute(pst_plan* p, int32_t* const* O_dev, const int64_t* o_strides, int32_t* const* flags_dev,
                           void* stream) {
  g_err[0] = 0;
  if (!p) return fail(PST_EINVAL, "null plan");
  hipStream_t s = (hipStream_t)stream;
  int count[pst::NUM_CLASSES];
  int n_wave = 0;
  int64_t wave_sims = 0;
  int rc = prepare_plan(p, O_dev, o_strides, flags_dev, s, false, count, &n_wave, &wave_sims);
  if (rc != PST_OK) return rc;
  p->range_pos = 0;  // a whole run abandons any pending sample ranges
  return enqueue_plan(p, s, count, n_wave, wave_sims, pst::FULL_RUN, false);
}

// Saved-state buffers for sample ranges: the words one simulation's record
// can take (the engine registers for the job's kernel -- rates, counts,
// stream key, block counter, clock, sample cursor, burst and step counters,
// trajectory pointer -- plus mode, events and the blocked layout's gathered
// samples), a bound over both precisions.
static int alloc_state(pst_plan* p) {
  for (PlanJob& pj : p->jobs) {
    if (pj.wave)
      return fail(PST_EINVAL, "sample ranges run the lane kernels only; a job of %d species and %d reactions "
                              "takes the large-network kernel", pj.net.N, pj.net.M);
    if (pj.d_state) continue;
    const int np = pj.cls >= 0 ? pst::CLASS_NP[pj.cls] : pj.net.N;
    const int kp = pj.cls >= 0 ? pst::CLASS_MP[pj.cls] : std::max(1, pj.p.n_rates);
    pj.state_words = 2 * kp + np + 32 + 8 * 4;  // + the gathered sample blocks of up to four species
    const size_t bytes = sizeof(uint32_t) * (size_t)pj.state_words * (size_t)std::max(1, pj.p.n_sims);
    hipError_t e = hipMalloc(&pj.d_state, bytes);
    if (e != hipSuccess) {
      pj.d_state = nullptr;
      return hip_fail(e, "hipMalloc(saved simulation states)");
    }
  }
  return PST_OK;
}

static int range_args_ok(pst_plan* p, int32_t f_end, int from) {
  if (f_end <= from) return fail(PST_EINVAL, "range end %d must exceed %d, the samples already recorded", f_end, from);
  for (size_t q = 0; q < p->jobs.size(); ++q) {
    const pst::JobParams& P = p->jobs[q].p;
    if (P.layout == PST_LAYOUT_BLOCKS && f_end < P.ita && f_end % 8 != 0)
      return fail(PST_EINVAL, "job %zu writes the blocked layout: a range must end on a multiple of 8 samples "
                              "or at the last sample (%d)", q, P.ita);
  }
  return PST_OK;
}

static int32_t max_ita(const pst_plan* p) {
  int32_t m = 0;
  for (const PlanJob& pj : p->jobs) m = std::max(m, pj.p.ita);
  return m;
}

int pst_plan_execute_range(pst_plan* p, int32_t* const* O_dev, const int64_t* o_strides, int32_t* const* flags_dev,
                           int32_t f_end, void* stream) {
  g_err[0] = 0;
  if (!p) return fail(PST_EINVAL, "null plan");
  int rc = range_args_ok(p, f_end, p->range_pos);
  if (rc == PST_OK) rc = alloc_state(p);
  if (rc != PST_OK) return rc;
  hipStream_t s = (hipStream_t)stream;
  int count[pst::NUM_CLASSES];
  int n_wave = 0;
  int64_t wave_sims = 0;
  rc = prepare_plan(p, O_dev, o_strides, flags_dev, s, true, count, &n_wave, &wave_sims);
  if (rc != PST_OK) return rc;
  const pst::Range r = {f_end, p->range_pos > 0 ? 1 : 0};
  rc = enqueue_plan(p, s, count, n_wave, wave_sims, r, false);
  if (rc != PST_OK) return rc;
  p->range_pos = f_end >= max_ita(p) ? 0 : f_end;
  return PST_OK;
}

int pst_plan_range_position(pst_plan* p, int32_t* samples_done) {
  g_err[0] = 0;
  if (!p || !samples_done) return fail(PST_EINVAL, "null plan or result");
  *samples_done = p->range_pos;
  return PST_OK;
}

int pst_plan_graph_create(pst_plan* p, int32_t* const* O_dev, const int64_t* o_strides, int32_t* const* flags_dev,
                          const int32_t* f_ends, int32_t n_ranges) {
  g_err[0] = 0;
  if (!p) return fail(PST_EINVAL, "null plan");
  if (!f_ends || n_ranges < 1) return fail(PST_EINVAL, "need at least one range end");
  if (f_ends[n_ranges - 1] < max_ita(p))
    return fail(PST_EINVAL, "the last range end %d must reach the last sample (%d)", f_ends[n_ranges - 1], max_ita(p));
  for (int r = 0; r < n_ranges; ++r) {
    const int rc = range_args_ok(p, f_ends[r], r > 0 ? f_ends[r - 1] : 0);
    if (rc != PST_OK) return rc;
  }
  const bool ranges = n_ranges > 1;
  if (ranges) {
    const int rc = alloc_state(p);
    if (rc != PST_OK) return rc;
  }
  hipError_t e;
  // a capture cannot upload descriptors: prepare them now, on a stream of our own
  hipStream_t cs = nullptr;
  if ((e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking)) != hipSuccess) return hip_fail(e, "hipStreamCreate");
  int count[pst::NUM_CLASSES];
  int n_wave = 0;
  int64_t wave_sims = 0;
  int rc = prepare_plan(p, O_dev, o_strides, flags_dev, cs, ranges, count, &n_wave, &wave_sims);
  if (rc == PST_OK && ((e = hipStreamSynchronize(cs)) != hipSuccess || (e = sync_plan(p, cs)) != hipSuccess))
    rc = hip_fail(e, "hipStreamSynchronize");
  if (rc != PST_OK) {
    (void)hipStreamDestroy(cs);
    return rc;
  }
  if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->graph_exec = nullptr;
  p->graph = nullptr;
  // every range's fill kernels and launches, in order, on the one stream:
  // the graph holds kernel nodes only (the fills are kernels, not memsets)
  if ((e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal)) != hipSuccess) {
    (void)hipStreamDestroy(cs);
    return hip_fail(e, "hipStreamBeginCapture");
  }
  for (int r = 0; r < n_ranges && rc == PST_OK; ++r) {
    const int n_jobs = (int)p->jobs.size();
    if (r == 0) hipLaunchKernelGGL(zero_u64, dim3(1), dim3(64), 0, cs, p->d_stats, (int)pst::ST_COUNT);
    hipLaunchKernelGGL(zero_u32, dim3((n_jobs + 255) / 256), dim3(256), 0, cs, p->d_queue, n_jobs);
    if ((e = hipGetLastError()) != hipSuccess) {
      rc = hip_fail(e, "zero kernels");
      break;
    }
    const pst::Range range = {ranges ? f_ends[r] : pst::FULL_RUN.f_end, r > 0 ? 1 : 0};
    auto one_stream = [&]() -> hipStream_t { return cs; };
    rc = enqueue_launches(p, count, n_wave, wave_sims, range, one_stream);
  }
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(cs, &g);
  if (rc == PST_OK && e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture");
  if (rc == PST_OK && (e = hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0)) != hipSuccess) {
    p->graph_exec = nullptr;
    rc = hip_fail(e, "hipGraphInstantiate");
  }
  (void)hipStreamDestroy(cs);
  if (rc != PST_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  p->graph = g;
  p->graph_generation = p->generation;
  p->range_pos = 0;
  return PST_OK;
}

int pst_plan_graph_launch(pst_plan* p, void* stream) {
  g_err[0] = 0;
  if (!p) return fail(PST_EINVAL, "null plan");
  if (!p->graph_exec) return fail(PST_EINVAL, "no graph: call pst_plan_graph_create first");
  if (p->graph_generation != p->generation)
    return fail(PST_EINVAL, "the plan's descriptors or launch shape changed since the graph was captured: "
                            "capture it again");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  p->executed = false;
  if (p->enqueued && (e = hipStreamWaitEvent(s, p->ev1, 0)) != hipSuccess) return hip_fail(e, "hipStreamWaitEvent");
  if ((e = hipEventRecord(p->ev0, s)) != hipSuccess) return hip_fail(e, "hipEventRecord");
  if ((e = hipGraphLaunch(p->graph_exec, s)) != hipSuccess) return hip_fail(e, "hipGraphLaunch");
  if ((e = hipEventRecord(p->ev1, s)) != hipSuccess) return hip_fail(e, "hipEventRecord");
  p->enqueued = true;
  p->executed = true;
  p->range_pos = 0;
  return PST_OK;
}

int pst_plan_execute(pst_plan* p, int32_t* O_dev, int64_t o_stride, int32_t* flags_dev, void* stream) {
  g_err[0] = 0;
  if (!p) return fail(PST_EINVAL, "null plan");
  if (p->jobs.size() != 1) return fail(PST_EINVAL, "plan holds %d jobs; use pst_batch_plan_execute", (int)p->jobs.size());
  if (!O_dev) return fail(PST_EINVAL, "null trajectory buffer");
  int32_t* O[1] = {O_dev};
  int64_t st[1] = {o_stride};
  int32_t* F[1] = {flags_dev};
  return pst_batch_plan_execute(p, O, st, F, stream);
}

int pst_plan_read_stats(pst_plan* p, pst_stats* st) {
  g_err[0] = 0;
  if (!p || !st) return fail(PST_EINVAL, "null plan or stats");
  if (!p->executed) return fail(PST_EINVAL, "plan has not been executed");
  hipError_t e = hipEventSynchronize(p->ev1);
  if (e != hipSuccess) return hip_fail(e, "simulation kernel");
  unsigned long long h[pst::ST_COUNT];
  if ((e = hipMemcpy(h, p->d_stats, sizeof(h), hipMemcpyDeviceToHost)) != hipSuccess) return hip_fail(e, "hipMemcpy(stats)");
  float ms = 0.0f;
  if ((e = hipEventElapsedTime(&ms, p->ev0, p->ev1)) != hipSuccess) return hip_fail(e, "hipEventElapsedTime");
  st->events = (int64_t)h[pst::ST_EVENTS];
  st->leaps = (int64_t)h[pst::ST_LEAPS];
  st->rejected_leaps = (int64_t)h[pst::ST_REJECTED];
  st->ssa_steps = (int64_t)h[pst::ST_SSA];
  st->flagged = (int64_t)h[pst::ST_FLAGGED];
  st->uniforms = (int64_t)h[pst::ST_UNIFORMS];
  st->ptrs_iterations = (int64_t)h[pst::ST_PTRS];
  st->inversion_steps = (int64_t)h[pst::ST_INV];
  st->leap_decisions = (int64_t)h[pst::ST_DECISIONS];
  st->sample_rows = (int64_t)h[pst::ST_SAMPLES];
  st->flagged_events = (int64_t)h[pst::ST_FEVENTS];
  st->asym_draws = (int64_t)h[pst::ST_ASYM];
  st->exact_draws = (int64_t)h[pst::ST_EXACT];
  st->kernel_ms = (double)ms;
  return PST_OK;
}

int pst_plan_autotune(pst_plan* p, int32_t* const* O_dev, const int64_t* o_strides, int32_t* const* flags_dev,
                      void* stream, double fraction, pst_launch* chosen) {
  g_err[0] = 0;
  if (!p || !chosen) return fail(PST_EINVAL, "null plan or result");
  if (!(fraction > 0.0 && fraction <= 1.0)) return fail(PST_EINVAL, "fraction %g outside (0, 1]", fraction);
  // the tuning horizon: the first ceil(fraction * ita) sample times of every job
  std::vector<int32_t> ita(p->jobs.size());
  for (size_t q = 0; q < p->jobs.size(); ++q) {
    ita[q] = p->jobs[q].p.ita;
    p->jobs[q].p.ita = std::max(1, std::min(ita[q], (int32_t)std::ceil(fraction * (double)ita[q])));
  }
  const int old_shift = p->lanes_shift, old_slice = p->ssa_slice, old_trim = p->waves_trim;
  const int old_bw = p->burst_waves;
  auto restore = [&]() {
    for (size_t q = 0; q < p->jobs.size(); ++q) p->jobs[q].p.ita = ita[q];
  };
  auto timed = [&](int shift, int slice, int trim, int bw, double* ms) -> int {
    p->lanes_shift = shift;
    p->ssa_slice = slice;
    p->waves_trim = trim;
    p->burst_waves = bw;
    const int rc = pst_batch_plan_execute(p, O_dev, o_strides, flags_dev, stream);
    if (rc != PST_OK) return rc;
    hipError_t e = hipEventSynchronize(p->ev1);
    if (e != hipSuccess) return hip_fail(e, "simulation kernel");
    float f = 0.0f;
    if ((e = hipEventElapsedTime(&f, p->ev0, p->ev1)) != hipSuccess) return hip_fail(e, "hipEventElapsedTime");
    *ms = (double)f;
    return PST_OK;
  };
  double ms = 0.0, def_ms = 0.0, best_ms = 0.0;
  int best_shift = 0, best_slice = kSsaSlice, best_trim = 0, best_bw = 0;
  int rc = timed(0, kSsaSlice, 0, 0, &ms);  // warm-up (first launch of each kernel)
  if (rc == PST_OK) rc = timed(0, kSsaSlice, 0, 0, &def_ms);
  best_ms = def_ms;
  // paired wavefronts apply to network kernels of tau-leaping jobs; the SSA
  // slice then matters only to the plan's other launches
  bool pairable = false, others = false;
  for (const FixedGroup& g : p->fixed) pairable = pairable || p->jobs[g.jobs[0]].p.method == PST_METHOD_TAU_LEAPING;
  for (const PlanJob& pj : p->jobs) others = others || pj.fixed < 0;
  const int slices[3] = {3, 7, 15};
  for (int bw = 0; bw <= (pairable ? 1 : 0) && rc == PST_OK; ++bw) {
    for (int trim = 0; trim <= 1 && rc == PST_OK; ++trim) {
      for (int shift = 0; shift <= 1 && rc == PST_OK; ++shift) {
        for (int k = 0; k < 3 && rc == PST_OK; ++k) {
          if (bw == 0 && trim == 0 && shift == 0 && slices[k] == kSsaSlice) continue;  // the default, timed above
          if (bw == 1 && !others && slices[k] != kSsaSlice) continue;
          rc = timed(shift, slices[k], trim, bw, &ms);
          if (rc != PST_OK && bw == 1) {  // no paired build for this network (e.g. its state does not fit LDS)
            rc = PST_OK;
            g_err[0] = 0;
            break;
          }
          // another shape must beat the default by 2 % (run-to-run noise)
          if (rc == PST_OK && ms < best_ms && ms < 0.98 * def_ms) {
            best_ms = ms;
            best_shift = shift;
            best_slice = slices[k];
            best_trim = trim;
            best_bw = bw;
          }
        }
      }
    }
  }
  restore();
  if (rc != PST_OK) {
    p->lanes_shift = old_shift;
    p->ssa_slice = old_slice;
    p->waves_trim = old_trim;
    p->burst_waves = old_bw;
    return rc;
  }
  p->lanes_shift = best_shift;
  p->ssa_slice = best_slice;
  p->waves_trim = best_trim;
  p->burst_waves = best_bw;
  chosen->lanes_shift = best_shift;
  chosen->ssa_slice = best_slice;
  chosen->waves_trim = best_trim;
  chosen->burst_waves = best_bw;
  chosen->kernel_ms = best_ms;
  chosen->default_ms = def_ms;
  return PST_OK;
}

int pst_plan_set_launch(pst_plan* p, int32_t lanes_shift, int32_t ssa_slice, int32_t waves_trim,
                        int32_t burst_waves) {
  g_err[0] = 0;
  if (!p) return fail(PST_EINVAL, "null plan");
  if (lanes_shift < 0 || lanes_shift > 6) return fail(PST_EINVAL, "lanes_shift %d outside [0, 6]", lanes_shift);
  if (ssa_slice < 0 || ssa_slice > 63) return fail(PST_EINVAL, "ssa_slice %d outside [0, 63]", ssa_slice);
  if (waves_trim < 0 || waves_trim > 7) return fail(PST_EINVAL, "waves_trim %d outside [0, 7]", waves_trim);
  if (burst_waves < 0 || burst_waves > 1) return fail(PST_EINVAL, "burst_waves %d outside [0, 1]", burst_waves);
  p->lanes_shift = lanes_shift;
  p->ssa_slice = ssa_slice;
  p->waves_trim = waves_trim;
  p->burst_waves = burst_waves;
  p->generation += 1;
  return PST_OK;
}

int pst_plan_set_stream(pst_plan* p, uint64_t seed, uint64_t sim_offset) {
  g_err[0] = 0;
  if (!p) return fail(PST_EINVAL, "null plan");
  for (PlanJob& j : p->jobs) {
    j.p.seed = seed;
    j.p.sim_offset = sim_offset;
  }
  p->generation += 1;
  return PST_OK;
}

int pst_batch_run(const pst_job* jobs, int32_t n_jobs, int32_t* const* O, int32_t* const* flags, pst_stats* stats) {
  g_err[0] = 0;
  if (!O) return fail(PST_EINVAL, "null trajectory array list");
  for (int q = 0; jobs && q < n_jobs; ++q)
    if (jobs[q].cfg.layout != PST_LAYOUT_ROWS)
      return fail(PST_EINVAL, "job %d: host arrays are returned as rows (layout must be PST_LAYOUT_ROWS)", q);
  pst_plan* p = nullptr;
  int rc = pst_batch_plan_create(jobs, n_jobs, &p);
  if (rc != PST_OK) return rc;
  std::vector<int32_t*> dO(n_jobs, nullptr), dF(n_jobs, nullptr);
  std::vector<int64_t> strides(n_jobs, 0);
  hipError_t e = hipSuccess;
  for (int q = 0; q < n_jobs && rc == PST_OK; ++q) {
    const pst_config& c = jobs[q].cfg;
    const size_t n_out = (size_t)c.kappa * c.ita * (size_t)c.n_sims;
    strides[q] = c.n_sims;
    if (c.n_sims > 0 && !O[q]) rc = fail(PST_EINVAL, "job %d: null trajectory array", q);
    else if (n_out > 0 && (e = hipMalloc(&dO[q], n_out * sizeof(int32_t))) != hipSuccess) rc = hip_fail(e, "hipMalloc(O)");
    else if (flags && flags[q] && c.n_sims > 0 &&
             (e = hipMalloc(&dF[q], (size_t)c.n_sims * sizeof(int32_t))) != hipSuccess)
      rc = hip_fail(e, "hipMalloc(flags)");
  }
  pst_stats st = {};
  if (rc == PST_OK) rc = pst_batch_plan_execute(p, dO.data(), strides.data(), dF.data(), nullptr);
  if (rc == PST_OK) rc = pst_plan_read_stats(p, &st);
  for (int q = 0; q < n_jobs && rc == PST_OK; ++q) {
    const pst_config& c = jobs[q].cfg;
    const size_t n_out = (size_t)c.kappa * c.ita * (size_t)c.n_sims;
    if (n_out > 0 && (e = hipMemcpy(O[q], dO[q], n_out * sizeof(int32_t), hipMemcpyDeviceToHost)) != hipSuccess)
      rc = hip_fail(e, "hipMemcpy(O)");
    else if (dF[q] && (e = hipMemcpy(flags[q], dF[q], (size_t)c.n_sims * sizeof(int32_t), hipMemcpyDeviceToHost)) !=
                          hipSuccess)
      rc = hip_fail(e, "hipMemcpy(flags)");
  }
  for (int q = 0; q < n_jobs; ++q) {
    if (dO[q]) (void)hipFree(dO[q]);
    if (dF[q]) (void)hipFree(dF[q]);
  }
  pst_plan_destroy(p);
  if (rc == PST_OK && stats) *stats = st;
  return rc;
}

static int run_host(const pst_network* net, const pst_config* cfg, const double* rates, const int32_t* x0,
                    int32_t n_sets, int32_t* O, int32_t* flags, pst_stats* stats) {
  if (!net || !cfg) return fail(PST_EINVAL, "null network or config");
  if (!O) return fail(PST_EINVAL, "null trajectory array");
  pst_job j;
  j.net = *net;
  j.cfg = *cfg;
  j.rates = rates;
  j.x0 = x0;
  j.n_sets = n_sets;
  int32_t* Os[1] = {O};
  int32_t* Fs[1] = {flags};
  return pst_batch_run(&j, 1, Os, Fs, stats);
}

int pst_tau_leaping_run(const pst_network* net, const pst_config* cfg, const double* rates, const int32_t* x0,
                        int32_t n_sets, int32_t* O, int32_t* flags, pst_stats* stats) {
  g_err[0] = 0;
  if (!cfg) return fail(PST_EINVAL, "null config");
  if (cfg->method == PST_METHOD_PURE_SSA)
    return fail(PST_EINVAL, "pst_tau_leaping_run takes TAU_LEAPING or FORCED_SSA; use pst_gillespie_run");
  return run_host(net, cfg, rates, x0, n_sets, O, flags, stats);
}

int pst_gillespie_run(const pst_network* net, const pst_config* cfg, const double* rates, const int32_t* x0,
                      int32_t n_sets, int32_t* O, int32_t* flags, pst_stats* stats) {
  g_err[0] = 0;
  if (!cfg) return fail(PST_EINVAL, "null config");
  pst_config c = *cfg;
  c.method = PST_METHOD_PURE_SSA;
  return run_host(net, &c, rates, x0, n_sets, O, flags, stats);
}

int pst_run_noalloc(const pst_network* net, const pst_config* cfg, const double* rates_dev, const int32_t* x0_dev,
                    int32_t n_sets, const float* I_dev, int32_t* O_dev, int64_t o_stride, int32_t* flags_dev,
                    void* stream, pst_stats* stats) {
  g_err[0] = 0;
  if (!net || !cfg) return fail(PST_EINVAL, "null network or config");
  if (!I_dev) return fail(PST_EINVAL, "null sample-time array");
  if (cfg->n_sims > 0 && !O_dev) return fail(PST_EINVAL, "null trajectory buffer");
  if (o_stride < cfg->n_sims) return fail(PST_EINVAL, "o_stride %lld < n_sims %d", (long long)o_stride, cfg->n_sims);
  if (net->hazards)
    return fail(PST_EINVAL, "pst_run_noalloc runs the library's table-driven kernels: compiled hazard expressions "
                            "need pst_plan_create");
  if (cfg->specialize != 0)
    return fail(PST_EINVAL, "pst_run_noalloc runs the library's table-driven kernels: specialize must be 0");
  if (cfg->layout != PST_LAYOUT_ROWS && cfg->layout != PST_LAYOUT_BLOCKS)
    return fail(PST_EINVAL, "unknown trajectory layout %d", cfg->layout);
  pst_config c = *cfg;
  c.I = I_dev;
  static thread_local pst::HostNet h;  // ~50 KB: kept off the stack, and no heap allocation
  int rc = prepare(net, &c, rates_dev, x0_dev, n_sets, &h, /*host_rows=*/false);
  if (rc != PST_OK) return rc;
  const int cls = pst::size_class(h.N, h.M, h.kappa);
  if (cls < 0)
    return fail(PST_ETOOLARGE, "pst_run_noalloc: N=%d, M=%d, kappa=%d is beyond the largest table-driven size class "
                               "(16 species, 16 reactions) that its %zu-byte device array holds",
                h.N, h.M, h.kappa, kNoAllocBytes);
  int cu = 0;
  if ((rc = check_device(&cu)) != PST_OK) return rc;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  unsigned char* base = nullptr;
  if ((rc = scratch_on_device(dev, &base)) != PST_OK) return rc;

  const int simds = cu * 4;
  int L = 64;  // simulations per wavefront, as for a plan (lanes_for)
  while (L > 1 && ((int64_t)c.n_sims + L / 2 - 1) / (L / 2) <= (int64_t)simds) L /= 2;
  pst::JobParams P;
  std::memset(&P, 0, sizeof(P));
  P.N = h.N;
  P.M = h.M;
  P.kappa = c.kappa;
  P.ita = c.ita;
  P.n_c = c.n_c;
  P.ssa_burst = c.ssa_burst;
  P.method = c.method;
  P.n_sims = c.n_sims;
  P.sims_per_set = c.sims_per_set;
  P.ssa_slice = kSsaSlice;
  P.lanes = L;
  P.n_rates = h.M;
  P.max_steps = (uint32_t)c.max_steps;
  P.eta = c.eta;
  P.t_max = c.t_max;
  P.seed = c.seed;
  P.sim_offset = c.sim_offset;
  P.o_stride = o_stride;
  P.layout = c.layout;
  P.rates = rates_dev;
  P.x0 = x0_dev;
  P.I = I_dev;
  P.O = O_dev;
  P.flags = flags_dev;

  hipStream_t s = (hipStream_t)stream;
  if (cls == 0) e = noalloc_enqueue<4, 4>(h, P, base, s);
  else if (cls == 1) e = noalloc_enqueue<8, 8>(h, P, base, s);
  else e = noalloc_enqueue<16, 16>(h, P, base, s);
  if (e != hipSuccess) return hip_fail(e, "descriptor kernel launch");
  const bool mixed = c.precision == PST_PRECISION_MIXED;
  const int bpc = mixed ? pst::blocks_per_cu_mixed(cls) : pst::blocks_per_cu_fp64(cls);
  const int64_t need = ((int64_t)c.n_sims + L - 1) / L;
  const int64_t resident = (int64_t)bpc * simds;
  const int waves = (int)(need < resident ? need : resident);
  const void* jobs = base + kNoAllocJob;
  unsigned int* queue = (unsigned int*)(base + kNoAllocQueue);
  unsigned long long* counters = (unsigned long long*)base;
  rc = mixed ? pst::launch_mixed(cls, jobs, 1, queue, counters, waves, pst::FULL_RUN, s)
             : pst::launch_fp64(cls, jobs, 1, queue, counters, waves, pst::FULL_RUN, s);
  if (rc != 0) return hip_fail((hipError_t)rc, "simulation kernel launch");
  if (stats) {
    unsigned long long hs[pst::ST_COUNT];
    if ((e = hipMemcpyAsync(hs, counters, sizeof(hs), hipMemcpyDeviceToHost, s)) != hipSuccess)
      return hip_fail(e, "hipMemcpyAsync(counters)");
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "simulation kernel");
    std::memset(stats, 0, sizeof(*stats));
    stats->events = (int64_t)hs[pst::ST_EVENTS];
    stats->leaps = (int64_t)hs[pst::ST_LEAPS];
    stats->rejected_leaps = (int64_t)hs[pst::ST_REJECTED];
    stats->ssa_steps = (int64_t)hs[pst::ST_SSA];
    stats->flagged = (int64_t)hs[pst::ST_FLAGGED];
    stats->uniforms = (int64_t)hs[pst::ST_UNIFORMS];
    stats->ptrs_iterations = (int64_t)hs[pst::ST_PTRS];
    stats->inversion_steps = (int64_t)hs[pst::ST_INV];
    stats->leap_decisions = (int64_t)hs[pst::ST_DECISIONS];
    stats->sample_rows = (int64_t)hs[pst::ST_SAMPLES];
    stats->flagged_events = (int64_t)hs[pst::ST_FEVENTS];
    stats->asym_draws = (int64_t)hs[pst::ST_ASYM];
    stats->exact_draws = (int64_t)hs[pst::ST_EXACT];
    stats->kernel_ms = 0.0;
  }
  return PST_OK;
}

int pst_poisson_probe_mixed(const float* lam, const uint32_t* words, int64_t n, int32_t* k, int32_t* path) {
  g_err[0] = 0;
  if (n < 0 || (n > 0 && (!lam || !words || !k || !path))) return fail(PST_EINVAL, "null array or negative count");
  if (n > ((int64_t)1 << 28)) return fail(PST_EINVAL, "at most 2^28 draws per call");
  if (n == 0) return PST_OK;
  int cu = 0;
  int rc = check_device(&cu);
  if (rc != PST_OK) return rc;
  float* d_lam = nullptr;
  uint32_t* d_w = nullptr;
  int32_t *d_k = nullptr, *d_p = nullptr;
  hipError_t e;
  if ((e = hipMalloc(&d_lam, n * sizeof(float))) != hipSuccess || (e = hipMalloc(&d_w, n * sizeof(uint32_t))) != hipSuccess ||
      (e = hipMalloc(&d_k, n * sizeof(int32_t))) != hipSuccess || (e = hipMalloc(&d_p, n * sizeof(int32_t))) != hipSuccess) {
    rc = hip_fail(e, "hipMalloc(probe)");
  } else if ((e = hipMemcpy(d_lam, lam, n * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess ||
             (e = hipMemcpy(d_w, words, n * sizeof(uint32_t), hipMemcpyHostToDevice)) != hipSuccess) {
    rc = hip_fail(e, "hipMemcpy(probe)");
  } else if ((rc = pst::launch_poisson_probe(d_lam, d_w, n, d_k, d_p, nullptr)) != 0) {
    rc = hip_fail((hipError_t)rc, "probe kernel launch");
  } else if ((e = hipMemcpy(k, d_k, n * sizeof(int32_t), hipMemcpyDeviceToHost)) != hipSuccess ||
             (e = hipMemcpy(path, d_p, n * sizeof(int32_t), hipMemcpyDeviceToHost)) != hipSuccess) {
    rc = hip_fail(e, "probe kernel");
  }
  if (d_lam) (void)hipFree(d_lam);
  if (d_w) (void)hipFree(d_w);
  if (d_k) (void)hipFree(d_k);
  if (d_p) (void)hipFree(d_p);
  return rc;
}

}  // extern "C"
