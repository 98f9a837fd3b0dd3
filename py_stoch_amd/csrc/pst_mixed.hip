// pst_mixed.hip -- mixed FP32/FP64 simulation kernels for gfx950.
//
// Same method as the FP64 kernels (P1_P2.py:54-297, P3.py:29-163,
// gillespie.py:27-149): same epsilon bound, same n_c critical split, same
// SSA fallback burst and sampling rules, same wavefront-uniform loop.
// Precision:
//   FP32: propensities, a0 for the leap decision, mu / sigma^2, the Cao tau
//         bound, Poisson means and the whole Poisson sampler (PTRS log-density
//         test in a cancellation-free Stirling form), uniforms (24-bit, one
//         Philox word each)
//   FP64: the simulation clock t, and the reaction choice, done as "first j
//         with sum_{r<=j} a_r >= rho2 * a0" accumulated in FP64 so that FP32
//         round-off cannot bias which reaction fires.  Poisson means of 2^22
//         and above use the FP64 sampler (FP32 cannot resolve the integers).
// Transcendentals use the hardware v_log_f32 / v_exp_f32 / v_rcp_f32.
#include "pst_device.hpp"
#include "pst_args.hpp"

namespace pst {

struct CountersM {
  unsigned long long events, leaps, rejected, ssa;
  Work w;
};

template <int NP, int MP>
__device__ __forceinline__ void propensities_f32(const TabView<NP, MP>& tab, const float (&k)[MP],
                                                 const int32_t (&x)[NP], float (&a)[MP]) {
#pragma unroll
  for (int j = 0; j < MP; ++j) {
    float v = k[j];
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const int p = tab.Ord(i, j);
      const int st = tab.Stp(i, j);
      for (int q = 0; q < p; ++q) v = v * (float)(x[i] - st * q);
    }
    a[j] = v;
  }
}

// CalculateG (P1_P2.py:447-486) in FP32
__device__ __forceinline__ float calc_g_f32(int h, int ht, int xi) {
  const float xf = (float)xi;
  if (h == 1) return 1.0f;
  if (h == 2) return (ht == 2) ? 2.0f + fast_rcp(xf - 1.0f) : 2.0f;
  if (h == 3) {
    if (ht == 3) return 3.0f + fast_rcp(xf - 1.0f) + 2.0f * fast_rcp(xf - 2.0f);
    if (ht == 2) return 1.5f * (2.0f + fast_rcp(xf - 1.0f));
    return 3.0f;
  }
  return 0.0f;
}

template <int NP, typename T>
__device__ __forceinline__ T pick_m(const T (&x)[NP], int idx) {
  T r = x[0];
#pragma unroll
  for (int i = 1; i < NP; ++i)
    if (idx == i) r = x[i];
  return r;
}

// log(1 + d) with the rounding of 1 + d corrected (Goldberg's method)
__device__ __forceinline__ float log1p_f32(float d) {
  const float u = 1.0f + d;
  const float m = u - 1.0f;
  return (m == 0.0f) ? d : fast_log(u) * (d / m);
}

struct PtrsF32 {
  float lam, slam, b, a, vr, twoa;
};

__device__ __forceinline__ PtrsF32 ptrs_setup_f32(float lam) {
  PtrsF32 P;
  P.lam = lam;
  P.slam = __builtin_sqrtf(lam);
  P.b = 0.931f + 2.53f * P.slam;
  P.a = -0.059f + 0.02483f * P.b;
  P.vr = 0.9277f - 3.6224f * fast_rcp(P.b - 2.0f);
  P.twoa = 2.0f * P.a;
  return P;
}

__device__ __forceinline__ bool ptrs_fast_f32(const PtrsF32& P, float U, float V, float& kf, float& us) {
  us = 0.5f - __builtin_fabsf(U);
  kf = __builtin_floorf((P.twoa * fast_rcp(us) + P.b) * U + P.lam + 0.43f);
  return us >= 0.07f && V <= P.vr;
}

// -lambda + k ln lambda - ln k!, the right-hand side of PTRS's acceptance test.
// For k >= 10, with x = k + 1 and Stirling's series for ln Gamma(x):
//   (x - lambda) - (x - 1/2) log1p((x - lambda)/lambda) - ln(lambda)/2
//   - ln(2 pi)/2 - 1/(12x) + 1/(360x^3)
// which keeps full FP32 accuracy where the direct form cancels.
__device__ __forceinline__ float log_pmf_f32(float kf, float lam) {
  if (kf < 10.0f) {
    const int k = (int)kf;
    const float lf = k < 2 ? 0.0f : (float)logfact_small(k);
    return -lam + kf * fast_log(lam) - lf;
  }
  const float x = kf + 1.0f;
  const float d = x - lam;
  const float r = fast_rcp(x);
  const float ser = r * (1.0f / 12.0f - r * r * (1.0f / 360.0f));
  return d - (x - 0.5f) * log1p_f32(d * fast_rcp(lam)) - 0.5f * fast_log(lam) - 0.91893853320467274f - ser;
}

__device__ __noinline__ bool ptrs_slow_f32(const PtrsF32& P, float V, float kf, float us) {
  if (kf < 0.0f || (us < 0.013f && V > us)) return false;
  const float invalpha = 1.1239f + 1.1328f * fast_rcp(P.b - 3.4f);
  const float lhs = fast_log(V) + fast_log(invalpha) - fast_log(P.a * fast_rcp(us * us) + P.b);
  return lhs <= log_pmf_f32(kf, P.lam);
}

// Full PTRS draw from block c, words of half h (iterations it = 0, 1, ...)
__device__ __noinline__ int64_t ptrs_full_f32(uint32_t k0, uint32_t k1, uint32_t s_lo, uint32_t s_hi, uint64_t c,
                                              int h, float lam, unsigned long long* wptrs) {
  Key key;
  key.k0 = k0;
  key.k1 = k1;
  key.s_lo = s_lo;
  key.s_hi = s_hi;
  const PtrsF32 P = ptrs_setup_f32(lam);
  for (int it = 0; it < PTRS_MAX_ITERS; ++it) {
    const Words4 bw = key.block(c, (uint32_t)it);
    if (it > 0) *wptrs += 1;
    const float U = u24(h ? bw.w2 : bw.w0) - 0.5f;
    const float V = u24(h ? bw.w3 : bw.w1);
    float kf, us;
    if (ptrs_fast_f32(P, U, V, kf, us)) return (int64_t)kf;
    if (ptrs_slow_f32(P, V, kf, us)) return (int64_t)kf;
  }
  return -1;
}

// Poisson means too large for FP32 integer resolution: PTRS in FP64
__device__ __noinline__ int64_t ptrs_big_f64(uint32_t k0, uint32_t k1, uint32_t s_lo, uint32_t s_hi, uint64_t c,
                                             int h, double lam, unsigned long long* wptrs) {
  Key key;
  key.k0 = k0;
  key.k1 = k1;
  key.s_lo = s_lo;
  key.s_hi = s_hi;
  const double loglam = log_f64(lam);
  const double b = 0.931 + 2.53 * __builtin_sqrt(lam);
  const double a = -0.059 + 0.02483 * b;
  const double vr = 0.9277 - 3.6224 / (b - 2.0);
  const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
  for (int it = 0; it < PTRS_MAX_ITERS; ++it) {
    const Words4 bw = key.block(c, (uint32_t)it);
    *wptrs += 1;
    const double U = u53(h ? bw.w2 : bw.w0, h ? bw.w3 : bw.w1) - 0.5;
    const double V = u53(h ? bw.w3 : bw.w1, h ? bw.w2 : bw.w0);
    const double us = 0.5 - fabs(U);
    const double kf = floor((2.0 * a / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) return (int64_t)kf;
    if (kf < 0.0 || (us < 0.013 && V > us)) continue;
    const double lhs = log_f64(V) + log_f64(invalpha) - log_f64(a / (us * us) + b);
    const double rhs = -lam + kf * loglam - lgamma_int_f64(kf + 1.0);
    if (lhs <= rhs) return (int64_t)kf;
  }
  return -1;
}

__device__ __forceinline__ int64_t inversion_f32(float u, float lam, Work& w) {
  float p = fast_exp(-lam);
  float F = p;
  int k = 0;
  while (u > F && k < INV_MAX_K) {
    k = k + 1;
    p = p * lam * fast_rcp((float)k);
    F = F + p;
  }
  w.inv += (unsigned long long)k;
  return k;
}

enum ModeM { MMODE_DECIDE = 0, MMODE_SSA = 1, MMODE_DONE = 2 };

__device__ __forceinline__ bool wave_any_m(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0ull; }

__device__ __forceinline__ unsigned long long wave_sum_m(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

constexpr float BIG_LAMBDA_F32 = 4194304.0f;  // 2^22

template <int NP, int MP>
__global__ void __launch_bounds__(256) sim_mixed_kernel(const NetTab<NP, MP> ntab, const RunArgs ra) {
  TabView<NP, MP> tab;
  tab.load(ntab);
  const int64_t sim = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = sim < ra.n_sims;
  const int M = ra.M, N = ra.N;
  const int kappa = ra.kappa, ita = ra.ita;
  const int64_t stride = ra.o_stride;
  const int64_t simc = active ? sim : 0;
  int32_t* Ob = ra.O + simc;
  const int64_t set = simc / ra.sims_per_set;
  const float eta = (float)ra.eta;

  CountersM cnt = {};
  int flag = 0;
  float k[MP];
  int32_t x[NP];
#pragma unroll
  for (int j = 0; j < MP; ++j) k[j] = (j < M) ? (float)ra.rates[set * M + j] : 0.0f;
#pragma unroll
  for (int i = 0; i < NP; ++i) x[i] = (i < N) ? ra.x0[set * N + i] : 0;

  Key key;
  key.init(ra.seed, ra.sim_offset + (uint64_t)simc);
  uint64_t bc = 0;
  double t = 0.0;
  int F = 0;
  double tnext = (double)ra.I[0];
  int mode = !active ? MMODE_DONE : (ra.method == 0 ? MMODE_DECIDE : MMODE_SSA);
  int burst = 0;
  const int nblk = (M + 1) >> 1;  // Philox blocks per Poisson attempt

  auto save = [&](int f, const int32_t (&xs)[NP]) {
    cnt.w.samples += 1;
#pragma unroll
    for (int q = 0; q < NP; ++q)
      if (q < kappa) Ob[((int64_t)q * ita + f) * stride] = pick_m<NP>(xs, tab.Ev(q));
  };
  auto fill_rest = [&](const int32_t (&xs)[NP], bool zero) {
    cnt.w.samples += (unsigned long long)(ita - F);
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      if (q < kappa) {
        const int32_t val = zero ? 0 : pick_m<NP>(xs, tab.Ev(q));
        for (int f = F; f < ita; ++f) Ob[((int64_t)q * ita + f) * stride] = val;
      }
    }
  };

  float a[MP];

  auto ssa_step = [&]() -> int {
    propensities_f32<NP, MP>(tab, k, x, a);
    double a0d = 0.0;
#pragma unroll
    for (int j = 0; j < MP; ++j) a0d += (double)a[j];
    if (a0d == 0.0) {
      fill_rest(x, false);
      return 1;
    }
    const Words4 bw = key.block(bc, 0);
    bc += 1;
    cnt.w.uniforms += 2;
    const float tau = -fast_log(u24(bw.w0)) * fast_rcp((float)a0d);
    const double target = (double)u24(bw.w1) * a0d;
    double cum = 0.0;
    int sel = -1, last = -1;
#pragma unroll
    for (int j = 0; j < MP; ++j) {
      if (j < M && sel < 0) {
        if (a[j] > 0.0f) last = j;
        cum += (double)a[j];
        if (cum >= target) sel = j;
      }
    }
    if (sel < 0) sel = last;
    t = t + (double)tau;
    while (t >= tnext) {
      save(F, x);
      F = F + 1;
      if (F == ita) return 1;
      tnext = (double)ra.I[F];
    }
    bool over = false;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      int dx = 0;
#pragma unroll
      for (int j = 0; j < MP; ++j)
        if (sel == j) dx = tab.V(i, j);
      x[i] += dx;
      over = over || (x[i] >= COUNT_LIMIT);
    }
    cnt.ssa += 1;
    cnt.events += 1;
    if (over) {
      fill_rest(x, true);
      return 2;
    }
    return 0;
  };

  auto decide = [&]() -> int {
    propensities_f32<NP, MP>(tab, k, x, a);
    float a0 = 0.0f;
#pragma unroll
    for (int j = 0; j < MP; ++j) a0 += a[j];
    if (a0 == 0.0f) {
      fill_rest(x, false);
      return MMODE_DONE;
    }
    cnt.w.decisions += 1;
    bool crit[MP];
#pragma unroll
    for (int j = 0; j < MP; ++j) {
      bool c = false;
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        const int v = tab.V(i, j);
        if (v < 0) c = c || ((-v) * ra.n_c > x[i]);
      }
      crit[j] = c;
    }
    float tau1 = __builtin_inff();
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      if (tab.Irs(i)) {
        float mu = 0.0f, s2 = 0.0f;
        bool any = false;
#pragma unroll
        for (int j = 0; j < MP; ++j) {
          const int v = tab.V(i, j);
          if (v != 0 && !crit[j]) {
            mu += (float)v * a[j];
            s2 += (float)(v * v) * a[j];
            any = true;
          }
        }
        if (any) {
          const float g_i = calc_g_f32(tab.Hv(i), tab.Htv(i), x[i]);
          const float num = fmaxf(eta * (float)x[i] * fast_rcp(g_i), 1.0f);
          const float lhs = num * fast_rcp(fabsf(mu));
          const float rhs = num * num * fast_rcp(s2);
          tau1 = fminf(tau1, fminf(lhs, rhs));
        }
      }
    }
    if (tau1 < 10.0f * fast_rcp(a0)) return MMODE_SSA;

    int32_t K[MP];
#pragma unroll
    for (int j = 0; j < MP; ++j) K[j] = 0;
    float a0c = 0.0f;
#pragma unroll
    for (int j = 0; j < MP; ++j)
      if (crit[j]) a0c += a[j];
    float tau2 = __builtin_inff();
    uint32_t cw1 = 0;
    if (a0c > 0.0f) {
      const Words4 bw = key.block(bc, 0);
      bc += 1;
      cnt.w.uniforms += 1;
      cw1 = bw.w1;
      tau2 = -fast_log(u24(bw.w0)) * fast_rcp(a0c);
    }
    float tau = fminf(tau1, tau2);
    if (tau1 > tau2) {
      const double target = (double)u24(cw1) * (double)a0c;
      cnt.w.uniforms += 1;
      double cum = 0.0;
      int sel = -1, last = -1;
#pragma unroll
      for (int j = 0; j < MP; ++j) {
        if (crit[j] && sel < 0) {
          if (a[j] > 0.0f) last = j;
          cum += (double)a[j];
          if (cum >= target) sel = j;
        }
      }
      if (sel < 0) sel = last;
#pragma unroll
      for (int j = 0; j < MP; ++j)
        if (sel == j) K[j] = 1;
    }
    if (!(tau < __builtin_inff())) {
      fill_rest(x, true);
      flag = FLAG_TAU_INF;
      return MMODE_DONE;
    }
    int64_t xp[NP];
    bool valid = false;
    for (int h = 0; h < MAX_HALVINGS; ++h) {
      const uint64_t grp = bc;
      bc += (uint64_t)nblk;
      int bad = 0;
      unsigned pending = 0;
      // first draw of every non-critical reaction: reactions 2r and 2r+1 share a block
#pragma unroll
      for (int r = 0; r < (MP + 1) / 2; ++r) {
        if (2 * r < M) {
          const Words4 bw = key.block(grp + (uint64_t)r, 0);
#pragma unroll
          for (int hh = 0; hh < 2; ++hh) {
            const int j = 2 * r + hh;
            if (j < MP && j < M && !crit[j]) {
              const float lam = a[j] * tau;
              const uint32_t wu = hh ? bw.w2 : bw.w0;
              const uint32_t wv = hh ? bw.w3 : bw.w1;
              if (!(lam > 0.0f)) {
                K[j] = 0;
              } else if (lam >= (float)LAMBDA_LIMIT) {
                bad = FLAG_OVERFLOW;
              } else if (lam < 10.0f) {
                cnt.w.uniforms += 1;
                K[j] = (int32_t)inversion_f32(u24(wu), lam, cnt.w);
              } else if (lam >= BIG_LAMBDA_F32) {
                pending |= 1u << j;
              } else {
                cnt.w.uniforms += 2;
                cnt.w.ptrs += 1;
                const PtrsF32 P = ptrs_setup_f32(lam);
                float kf, us;
                if (ptrs_fast_f32(P, u24(wu) - 0.5f, u24(wv), kf, us)) K[j] = (int32_t)kf;
                else pending |= 1u << j;
              }
            }
          }
        }
      }
      while (pending != 0u && !bad) {
        const int j = __builtin_ctz(pending);
        pending &= pending - 1u;
        const float lam = pick_m<MP>(a, j) * tau;
        unsigned long long extra = 0;
        const int64_t kk = (lam >= BIG_LAMBDA_F32)
                               ? ptrs_big_f64(key.k0, key.k1, key.s_lo, key.s_hi, grp + (uint64_t)(j >> 1), j & 1,
                                              (double)lam, &extra)
                               : ptrs_full_f32(key.k0, key.k1, key.s_lo, key.s_hi, grp + (uint64_t)(j >> 1), j & 1,
                                               lam, &extra);
        cnt.w.ptrs += extra;
        cnt.w.uniforms += 2 * extra;
        if (kk < 0) {
          bad = FLAG_PTRS;
        } else {
#pragma unroll
          for (int r = 0; r < MP; ++r)
            if (r == j) K[r] = (int32_t)kk;
        }
      }
      if (bad) {
        fill_rest(x, true);
        flag = bad;
        return MMODE_DONE;
      }
      bool ok = true;
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        int64_t s = x[i];
#pragma unroll
        for (int j = 0; j < MP; ++j) {
          const int v = tab.V(i, j);
          if (v != 0) s += (int64_t)K[j] * v;
        }
        xp[i] = s;
        ok = ok && (s >= 0);
      }
      if (ok) {
        valid = true;
        break;
      }
      cnt.rejected += 1;
      tau = tau * 0.5f;
    }
    if (!valid) {
      fill_rest(x, true);
      flag = FLAG_HALVING;
      return MMODE_DONE;
    }
    bool big = false;
#pragma unroll
    for (int i = 0; i < NP; ++i) big = big || (xp[i] >= COUNT_LIMIT);
    if (big) {
      fill_rest(x, true);
      flag = FLAG_OVERFLOW;
      return MMODE_DONE;
    }
    unsigned long long ev = 0;
#pragma unroll
    for (int j = 0; j < MP; ++j) ev += (unsigned long long)K[j];
    cnt.events += ev;
    cnt.leaps += 1;
    int32_t xn[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) xn[i] = (int32_t)xp[i];
    const double t0 = t;
    t = t + (double)tau;
    bool done = false;
    if (t >= tnext) {
      const float inv_gap = fast_rcp((float)(t - t0));
      while (t >= tnext) {
        const float frac = (float)(tnext - t0) * inv_gap;
#pragma unroll
        for (int q = 0; q < NP; ++q) {
          if (q < kappa) {
            const int s = tab.Ev(q);
            const int32_t xs = pick_m<NP>(x, s);
            const int diff = pick_m<NP>(xn, s) - xs;
            Ob[((int64_t)q * ita + F) * stride] = xs + rn_f32_to_int((float)diff * frac);
          }
        }
        cnt.w.samples += 1;
        F = F + 1;
        if (F == ita) {
          done = true;
          break;
        }
        tnext = (double)ra.I[F];
      }
    }
#pragma unroll
    for (int i = 0; i < NP; ++i) x[i] = xn[i];
    return done ? MMODE_DONE : MMODE_DECIDE;
  };

  const int slice = ra.ssa_slice > 0 ? ra.ssa_slice : 1;
  while (wave_any_m(mode != MMODE_DONE)) {
    tab.refresh();
    if (wave_any_m(mode == MMODE_DECIDE)) {
      if (mode == MMODE_DECIDE) {
        mode = decide();
        if (mode == MMODE_SSA) burst = ra.ssa_burst;
      }
    }
    const int limit = wave_any_m(mode == MMODE_DECIDE) ? slice : 0x7fffffff;
    for (int s = 0; s < limit; ++s) {
      if (!wave_any_m(mode == MMODE_SSA)) break;
      tab.refresh();
      if (mode == MMODE_SSA) {
        if (ra.method == 2 && !(t < ra.t_max)) {
          fill_rest(x, true);
          mode = MMODE_DONE;
        } else {
          const int r = ssa_step();
          if (r != 0) {
            mode = MMODE_DONE;
            if (r == 2) flag = FLAG_OVERFLOW;
          } else if (ra.method == 0) {
            burst -= 1;
            if (burst == 0) mode = MMODE_DECIDE;
          }
        }
      }
    }
  }

  if (active && ra.flags) ra.flags[sim] = flag;
  const unsigned long long flagged = (active && flag != 0) ? 1ull : 0ull;
  const unsigned long long vals[ST_SAMPLES + 1] = {cnt.events, cnt.leaps, cnt.rejected, cnt.ssa, flagged,
                                                   cnt.w.uniforms, cnt.w.ptrs, cnt.w.inv, cnt.w.decisions,
                                                   cnt.w.samples};
#pragma unroll
  for (int i = 0; i <= ST_SAMPLES; ++i) {
    const unsigned long long sum = wave_sum_m(vals[i]);
    if ((threadIdx.x & 63) == 0) atomicAdd(&ra.stats[i], sum);
  }
}

template <int NP, int MP>
static int launch_class_m(const HostNet& net, const RunArgs& ra, hipStream_t stream) {
  const NetTab<NP, MP> tab = make_tab<NP, MP>(net);
  const int block = 256;
  const int grid = (int)((ra.n_sims + block - 1) / block);
  if (grid == 0) return 0;
  hipLaunchKernelGGL((sim_mixed_kernel<NP, MP>), dim3(grid), dim3(block), 0, stream, tab, ra);
  return (int)hipGetLastError();
}

int launch_mixed(const HostNet& net, const RunArgs& ra, hipStream_t stream) {
  if (net.N <= 4 && net.M <= 4 && net.kappa <= 4) return launch_class_m<4, 4>(net, ra, stream);
  if (net.N <= 8 && net.M <= 8 && net.kappa <= 8) return launch_class_m<8, 8>(net, ra, stream);
  if (net.N <= 16 && net.M <= 16 && net.kappa <= 16) return launch_class_m<16, 16>(net, ra, stream);
  return -1;
}

}  // namespace pst
