// pst_capi.hip -- the extern "C" entry points declared in include/pst.h.
//
// Validation, table densification, device allocation and launch.  The
// reference's host driver (cuTauLeaping.py:18-84) JIT-compiles two kernels per
// run, uploads constants, launches P1_P2 and P3 once per iteration with a
// gpuarray.sum and a scalar copy back each time; here one launch runs every
// simulation to the end and only the trajectories (and 8 counters) come back.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdarg>
#include <cstring>
#include <cmath>
#include <new>

#include "../../include/pst.h"
#include "pst_args.hpp"

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

constexpr int kMaxSpecies = 16;
constexpr int kMaxReactions = 16;

__global__ void zero_u64(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// Check a pst_network / pst_config pair and build the dense host table.
int prepare(const pst_network* net, const pst_config* cfg, const double* rates, const int32_t* x0,
            int32_t n_sets, pst::HostNet* h) {
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
  if (cfg->ssa_burst < 1) return fail(PST_EINVAL, "ssa_burst must be >= 1");
  if (!(cfg->eta > 0.0) || !std::isfinite(cfg->eta)) return fail(PST_EINVAL, "eta must be a positive finite number");
  if (cfg->method == PST_METHOD_PURE_SSA && !std::isfinite(cfg->t_max))
    return fail(PST_EINVAL, "t_max must be finite for pure SSA");
  for (int f = 0; f < cfg->ita; ++f)
    if (!std::isfinite(cfg->I[f])) return fail(PST_EINVAL, "sample time %d is not finite", f);
  if (net->a_size < 0 || (net->a_size > 0 && !net->A)) return fail(PST_EINVAL, "bad A table");
  if (net->v_size < 0 || (net->v_size > 0 && !net->V)) return fail(PST_EINVAL, "bad V table");
  if (!net->H || !net->H_type || !net->prop_order || !net->prop_step)
    return fail(PST_EINVAL, "null H, H_type, prop_order or prop_step");

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
  for (int64_t r = 0; r < need_sets; ++r) {
    for (int j = 0; j < M; ++j) {
      const double k = rates[r * M + j];
      if (!(k >= 0.0) || !std::isfinite(k)) return fail(PST_EINVAL, "rate %d of row %lld must be finite and >= 0", j, (long long)r);
    }
    for (int i = 0; i < N; ++i)
      if (x0[r * N + i] < 0) return fail(PST_EINVAL, "initial count %d of row %lld is negative", i, (long long)r);
  }
  return PST_OK;
}

}  // namespace

struct pst_plan {
  pst::HostNet net;
  pst::RunArgs ra;
  int device;
  int n_sets;
  double* d_rates = nullptr;
  int32_t* d_x0 = nullptr;
  float* d_I = nullptr;
  unsigned long long* d_stats = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool executed = false;
};

static int check_device() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(PST_ENODEV, "no HIP device visible");
  int dev = 0;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return hip_fail(e, "hipGetDevice");
  hipDeviceProp_t p;
  if ((e = hipGetDeviceProperties(&p, dev)) != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
  if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
    return fail(PST_ENODEV, "device %d is %s; this library is built for gfx950 only", dev, p.gcnArchName);
  return PST_OK;
}

extern "C" {

const char* pst_last_error(void) { return g_err; }

int pst_abi_version(void) { return PST_ABI_VERSION; }

void pst_max_network(int32_t* max_species, int32_t* max_reactions) {
  if (max_species) *max_species = kMaxSpecies;
  if (max_reactions) *max_reactions = kMaxReactions;
}

void pst_plan_destroy(pst_plan* p) {
  if (!p) return;
  if (p->d_rates) (void)hipFree(p->d_rates);
  if (p->d_x0) (void)hipFree(p->d_x0);
  if (p->d_I) (void)hipFree(p->d_I);
  if (p->d_stats) (void)hipFree(p->d_stats);
  if (p->ev0) (void)hipEventDestroy(p->ev0);
  if (p->ev1) (void)hipEventDestroy(p->ev1);
  delete p;
}

int pst_plan_create(const pst_network* net, const pst_config* cfg, const double* rates, const int32_t* x0,
                    int32_t n_sets, pst_plan** out) {
  g_err[0] = 0;
  if (!out) return fail(PST_EINVAL, "null plan pointer");
  *out = nullptr;
  pst_plan* p = new (std::nothrow) pst_plan();
  if (!p) return fail(PST_ENOMEM, "host allocation failed");
  int rc = prepare(net, cfg, rates, x0, n_sets, &p->net);
  if (rc != PST_OK) { delete p; return rc; }
  if ((rc = check_device()) != PST_OK) { delete p; return rc; }
  (void)hipGetDevice(&p->device);
  const int N = net->n_species, M = net->n_reactions;
  const int64_t sets = cfg->n_sims == 0 ? 1 : ((int64_t)cfg->n_sims + cfg->sims_per_set - 1) / cfg->sims_per_set;
  p->n_sets = (int)sets;
  hipError_t e;
  if ((e = hipMalloc(&p->d_rates, sizeof(double) * sets * M)) != hipSuccess) { pst_plan_destroy(p); return hip_fail(e, "hipMalloc(rates)"); }
  if ((e = hipMalloc(&p->d_x0, sizeof(int32_t) * sets * N)) != hipSuccess) { pst_plan_destroy(p); return hip_fail(e, "hipMalloc(x0)"); }
  if ((e = hipMalloc(&p->d_I, sizeof(float) * cfg->ita)) != hipSuccess) { pst_plan_destroy(p); return hip_fail(e, "hipMalloc(I)"); }
  if ((e = hipMalloc(&p->d_stats, sizeof(unsigned long long) * pst::ST_COUNT)) != hipSuccess) { pst_plan_destroy(p); return hip_fail(e, "hipMalloc(stats)"); }
  if ((e = hipMemcpy(p->d_rates, rates, sizeof(double) * sets * M, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(p->d_x0, x0, sizeof(int32_t) * sets * N, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(p->d_I, cfg->I, sizeof(float) * cfg->ita, hipMemcpyHostToDevice)) != hipSuccess) {
    pst_plan_destroy(p);
    return hip_fail(e, "hipMemcpy(tables)");
  }
  if ((e = hipEventCreate(&p->ev0)) != hipSuccess || (e = hipEventCreate(&p->ev1)) != hipSuccess) {
    pst_plan_destroy(p);
    return hip_fail(e, "hipEventCreate");
  }
  pst::RunArgs& ra = p->ra;
  std::memset(&ra, 0, sizeof(ra));
  ra.N = N;
  ra.M = M;
  ra.kappa = cfg->kappa;
  ra.ita = cfg->ita;
  ra.n_c = cfg->n_c;
  ra.ssa_burst = cfg->ssa_burst;
  ra.method = cfg->method;
  ra.n_sims = cfg->n_sims;
  ra.sims_per_set = cfg->sims_per_set;
  ra.eta = cfg->eta;
  ra.t_max = cfg->t_max;
  ra.seed = cfg->seed;
  ra.sim_offset = cfg->sim_offset;
  ra.rates = p->d_rates;
  ra.x0 = p->d_x0;
  ra.I = p->d_I;
  ra.stats = p->d_stats;
  ra.precision = cfg->precision;
  ra.ssa_slice = 4;
  *out = p;
  return PST_OK;
}

int pst_plan_execute(pst_plan* p, int32_t* O_dev, int64_t o_stride, int32_t* flags_dev, void* stream) {
  g_err[0] = 0;
  if (!p) return fail(PST_EINVAL, "null plan");
  if (!O_dev) return fail(PST_EINVAL, "null trajectory buffer");
  if (o_stride < p->ra.n_sims) return fail(PST_EINVAL, "o_stride %lld < n_sims %d", (long long)o_stride, p->ra.n_sims);
  int dev = -1;
  (void)hipGetDevice(&dev);
  if (dev != p->device) return fail(PST_EINVAL, "plan was created on device %d, current device is %d", p->device, dev);
  hipStream_t s = (hipStream_t)stream;
  pst::RunArgs ra = p->ra;
  ra.O = O_dev;
  ra.o_stride = o_stride;
  ra.flags = flags_dev;
  hipLaunchKernelGGL(zero_u64, dim3(1), dim3(64), 0, s, p->d_stats, (int)pst::ST_COUNT);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "zero_u64 launch");
  if ((e = hipEventRecord(p->ev0, s)) != hipSuccess) return hip_fail(e, "hipEventRecord");
  const int rc = (p->ra.precision == PST_PRECISION_MIXED) ? pst::launch_mixed(p->net, ra, s) : pst::launch_fp64(p->net, ra, s);
  if (rc < 0) return fail(PST_ETOOLARGE, "no compiled kernel for N=%d, M=%d, kappa=%d", p->net.N, p->net.M, p->net.kappa);
  if (rc != 0) return hip_fail((hipError_t)rc, "simulation kernel launch");
  if ((e = hipEventRecord(p->ev1, s)) != hipSuccess) return hip_fail(e, "hipEventRecord");
  p->executed = true;
  return PST_OK;
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
  st->kernel_ms = (double)ms;
  return PST_OK;
}

static int run_host(const pst_network* net, const pst_config* cfg, const double* rates, const int32_t* x0,
                    int32_t n_sets, int32_t* O, int32_t* flags, pst_stats* stats) {
  if (!O) return fail(PST_EINVAL, "null trajectory array");
  pst_plan* p = nullptr;
  int rc = pst_plan_create(net, cfg, rates, x0, n_sets, &p);
  if (rc != PST_OK) return rc;
  const size_t n_out = (size_t)cfg->kappa * cfg->ita * (size_t)cfg->n_sims;
  int32_t* dO = nullptr;
  int32_t* dF = nullptr;
  hipError_t e;
  if (n_out > 0 && (e = hipMalloc(&dO, n_out * sizeof(int32_t))) != hipSuccess) {
    pst_plan_destroy(p);
    return hip_fail(e, "hipMalloc(O)");
  }
  if (flags && cfg->n_sims > 0 && (e = hipMalloc(&dF, (size_t)cfg->n_sims * sizeof(int32_t))) != hipSuccess) {
    (void)hipFree(dO);
    pst_plan_destroy(p);
    return hip_fail(e, "hipMalloc(flags)");
  }
  pst_stats st = {};
  if (cfg->n_sims > 0) {
    rc = pst_plan_execute(p, dO, cfg->n_sims, dF, nullptr);
    if (rc == PST_OK) rc = pst_plan_read_stats(p, &st);
    if (rc == PST_OK && (e = hipMemcpy(O, dO, n_out * sizeof(int32_t), hipMemcpyDeviceToHost)) != hipSuccess)
      rc = hip_fail(e, "hipMemcpy(O)");
    if (rc == PST_OK && flags &&
        (e = hipMemcpy(flags, dF, (size_t)cfg->n_sims * sizeof(int32_t), hipMemcpyDeviceToHost)) != hipSuccess)
      rc = hip_fail(e, "hipMemcpy(flags)");
  }
  if (dO) (void)hipFree(dO);
  if (dF) (void)hipFree(dF);
  pst_plan_destroy(p);
  if (rc == PST_OK && stats) *stats = st;
  return rc;
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

}  // extern "C"
