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
// Transcendentals use the hardware v_log_f32 / v_exp_f32 / v_rcp_f32.  The
// launch loop, work queue and lane modes are in pst_driver.hpp.
#include "pst_driver.hpp"

namespace pst {

template <int NP, int MP>
__device__ __forceinline__ void propensities_f32(const JobRef<NP, MP>& tab, const float (&k)[MP],
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
                                              int h, float lam, unsigned int* wptrs) {
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
                                             int h, double lam, unsigned int* wptrs) {
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
  w.inv += (unsigned int)k;
  return k;
}

constexpr float BIG_LAMBDA_F32 = 4194304.0f;  // 2^22

// Per-lane state and steps of one mixed-precision simulation.
template <int NP, int MP>
struct EngineF32 {
  float k[MP];
  int32_t x[NP];
  Key key;
  uint64_t bc;
  double t, tnext;
  int F;
  int flag;
  int32_t* Ob;

  __device__ __forceinline__ void save(const JobRef<NP, MP>& tab, const JobParams& P, int f,
                                       const int32_t (&xs)[NP]) const {
#pragma unroll
    for (int q = 0; q < NP; ++q)
      if (q < P.kappa) Ob[((int64_t)q * P.ita + f) * P.o_stride] = pick_i<NP>(xs, tab.Ev(q));
  }
  __device__ __forceinline__ void fill_rest(const JobRef<NP, MP>& tab, const JobParams& P, bool zero,
                                            Counters& cnt) const {
    cnt.w.samples += (unsigned int)(P.ita - F);
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      if (q < P.kappa) {
        const int32_t val = zero ? 0 : pick_i<NP>(x, tab.Ev(q));
        for (int f = F; f < P.ita; ++f) Ob[((int64_t)q * P.ita + f) * P.o_stride] = val;
      }
    }
  }

  __device__ __forceinline__ int start(const JobRef<NP, MP>& jr, int64_t sim) {
    const JobParams& P = jr.p();
    const int64_t set = sim / P.sims_per_set;
#pragma unroll
    for (int j = 0; j < MP; ++j) k[j] = (j < P.M) ? (float)P.rates[set * P.M + j] : 0.0f;
#pragma unroll
    for (int i = 0; i < NP; ++i) x[i] = (i < P.N) ? P.x0[set * P.N + i] : 0;
    key.init(P.seed, P.sim_offset + (uint64_t)sim);
    bc = 0;
    t = 0.0;
    F = 0;
    tnext = (double)P.I[0];
    flag = 0;
    Ob = P.O + sim;
    return P.method == 0 ? MODE_DECIDE : MODE_SSA;
  }

  __device__ __forceinline__ bool before_tmax(const JobRef<NP, MP>& jr) const { return t < jr.p().t_max; }

  __device__ __forceinline__ void stop_at_tmax(const JobRef<NP, MP>& tab, Counters& cnt) const {
    fill_rest(tab, tab.p(), true, cnt);
  }

  __device__ __forceinline__ int ssa_step(const JobRef<NP, MP>& tab, Counters& cnt) {
    const JobParams& P = tab.p();
    const int M = P.M;
    float a[MP];
    propensities_f32<NP, MP>(tab, k, x, a);
    double a0d = 0.0;
#pragma unroll
    for (int j = 0; j < MP; ++j) a0d += (double)a[j];
    if (a0d == 0.0) {
      fill_rest(tab, P, false, cnt);
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
      save(tab, P, F, x);
      cnt.w.samples += 1;
      F = F + 1;
      if (F == P.ita) return 1;
      tnext = (double)P.I[F];
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
      fill_rest(tab, P, true, cnt);
      return 2;
    }
    return 0;
  }

  __device__ __forceinline__ int decide(const JobRef<NP, MP>& tab, Counters& cnt) {
    const JobParams& P = tab.p();
    const int M = P.M;
    const float eta = (float)P.eta;
    float a[MP];
    propensities_f32<NP, MP>(tab, k, x, a);
    float a0 = 0.0f;
#pragma unroll
    for (int j = 0; j < MP; ++j) a0 += a[j];
    if (a0 == 0.0f) {
      fill_rest(tab, P, false, cnt);
      return MODE_DONE;
    }
    cnt.w.decisions += 1;
    // critical reactions (P1_P2.py:314-327): per-lane conditions as selects,
    // branches only on table entries (uniform)
    bool crit[MP];
#pragma unroll
    for (int j = 0; j < MP; ++j) {
      bool c = false;
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        const int v = tab.V(i, j);
        if (v < 0) c = c | ((-v) * P.n_c > x[i]);
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
          if (v != 0) {
            const bool nc = !crit[j];
            mu += nc ? (float)v * a[j] : 0.0f;
            s2 += nc ? (float)(v * v) * a[j] : 0.0f;
            any = any | nc;
          }
        }
        const float g_i = calc_g_f32(tab.Hv(i), tab.Htv(i), x[i]);
        const float num = fmaxf(eta * (float)x[i] * fast_rcp(g_i), 1.0f);
        const float cand = fminf(num * fast_rcp(fabsf(mu)), num * num * fast_rcp(s2));
        tau1 = any ? fminf(tau1, cand) : tau1;
      }
    }
    if (tau1 < 10.0f * fast_rcp(a0)) return MODE_SSA;

    int32_t K[MP];
    float a0c = 0.0f;
#pragma unroll
    for (int j = 0; j < MP; ++j) {
      K[j] = 0;
      a0c += crit[j] ? a[j] : 0.0f;
    }
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
      for (int j = 0; j < MP; ++j) K[j] = (sel == j) ? 1 : K[j];
    }
    if (!(tau < __builtin_inff())) {
      fill_rest(tab, P, true, cnt);
      flag = FLAG_TAU_INF;
      return MODE_DONE;
    }
    constexpr int NB = (MP + 1) / 2;  // Philox blocks per Poisson attempt (two reactions per block)
    const int nblk = (M + 1) >> 1;
    int64_t xp[NP];
    bool valid = false;
    for (int h = 0; h < MAX_HALVINGS; ++h) {
      const uint64_t grp = bc;
      bc += (uint64_t)nblk;
      // every reaction's first draw, straight-line: independent Philox blocks and
      // PTRS squeezes that the scheduler interleaves
      Words4 W[NB];
#pragma unroll
      for (int r = 0; r < NB; ++r) W[r] = key.block(grp + (uint64_t)r, 0);
      float lam[MP];
      unsigned pend_ptrs = 0, pend_inv = 0, drawn = 0;
      bool bad = false;
#pragma unroll
      for (int j = 0; j < MP; ++j) {
        const bool live = (j < M) && !crit[j];
        const float l = live ? a[j] * tau : 0.0f;
        lam[j] = l;
        const uint32_t wu = (j & 1) ? W[j >> 1].w2 : W[j >> 1].w0;
        const uint32_t wv = (j & 1) ? W[j >> 1].w3 : W[j >> 1].w1;
        const PtrsF32 Pt = ptrs_setup_f32(fmaxf(l, 10.0f));
        float kf, us;
        const bool fast = ptrs_fast_f32(Pt, u24(wu) - 0.5f, u24(wv), kf, us);
        const bool use_ptrs = (l >= 10.0f) & (l < BIG_LAMBDA_F32);
        const bool ok = use_ptrs & fast;
        if (!crit[j]) K[j] = ok ? (int32_t)kf : 0;
        pend_ptrs |= ((l >= 10.0f) & !ok) ? (1u << j) : 0u;
        drawn |= use_ptrs ? (1u << j) : 0u;
        pend_inv |= ((l > 0.0f) & (l < 10.0f)) ? (1u << j) : 0u;
        bad = bad | (l >= (float)LAMBDA_LIMIT);
      }
      cnt.w.ptrs += (unsigned int)__builtin_popcount(drawn);
      cnt.w.uniforms += 2u * (unsigned int)__builtin_popcount(drawn);
      if (bad) {
        fill_rest(tab, P, true, cnt);
        flag = FLAG_OVERFLOW;
        return MODE_DONE;
      }
      // small means: sequential-search inversion for all such reactions together
      if (pend_inv != 0u) {
        float pj[MP], Fj[MP], uj[MP];
        int kj[MP];
        unsigned act = 0;
#pragma unroll
        for (int j = 0; j < MP; ++j) {
          const bool on = (pend_inv >> j) & 1u;
          uj[j] = u24((j & 1) ? W[j >> 1].w2 : W[j >> 1].w0);
          pj[j] = fast_exp(-lam[j]);
          Fj[j] = pj[j];
          kj[j] = 0;
          act |= (on && uj[j] > Fj[j]) ? (1u << j) : 0u;
        }
        cnt.w.uniforms += (unsigned int)__builtin_popcount(pend_inv);
        int it = 0;
        while (act != 0u && it < INV_MAX_K) {
          it += 1;
#pragma unroll
          for (int j = 0; j < MP; ++j) {
            const bool on = (act >> j) & 1u;
            const int kn = kj[j] + 1;
            const float pn = pj[j] * lam[j] * fast_rcp((float)kn);
            const float Fn = Fj[j] + pn;
            kj[j] = on ? kn : kj[j];
            pj[j] = on ? pn : pj[j];
            Fj[j] = on ? Fn : Fj[j];
            act = (on && !(uj[j] > Fn)) ? (act & ~(1u << j)) : act;
          }
        }
#pragma unroll
        for (int j = 0; j < MP; ++j) {
          if ((pend_inv >> j) & 1u) {
            K[j] = kj[j];
            cnt.w.inv += (unsigned int)kj[j];
          }
        }
      }
      // squeeze failures and very large means: the full PTRS draw, one shared loop
      bool ptrs_fail = false;
      while (pend_ptrs != 0u && !ptrs_fail) {
        const int j = __builtin_ctz(pend_ptrs);
        pend_ptrs &= pend_ptrs - 1u;
        const float l = pick_m<MP>(lam, j);
        unsigned int extra = 0;
        const int64_t kk = (l >= BIG_LAMBDA_F32)
                               ? ptrs_big_f64(key.k0, key.k1, key.s_lo, key.s_hi, grp + (uint64_t)(j >> 1), j & 1,
                                              (double)l, &extra)
                               : ptrs_full_f32(key.k0, key.k1, key.s_lo, key.s_hi, grp + (uint64_t)(j >> 1), j & 1,
                                               l, &extra);
        cnt.w.ptrs += extra;
        cnt.w.uniforms += 2 * extra;
        if (kk < 0) {
          ptrs_fail = true;
        } else {
#pragma unroll
          for (int r = 0; r < MP; ++r) K[r] = (r == j) ? (int32_t)kk : K[r];
        }
      }
      if (ptrs_fail) {
        fill_rest(tab, P, true, cnt);
        flag = FLAG_PTRS;
        return MODE_DONE;
      }
      // TentativeUpdatedState / ValidState (P1_P2.py:509-540)
      bool ok = true;
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        int64_t sacc = x[i];
#pragma unroll
        for (int j = 0; j < MP; ++j) {
          const int v = tab.V(i, j);
          if (v != 0) sacc += (int64_t)K[j] * v;
        }
        xp[i] = sacc;
        ok = ok & (sacc >= 0);
      }
      if (ok) {
        valid = true;
        break;
      }
      cnt.rejected += 1;
      tau = tau * 0.5f;
    }
    if (!valid) {
      fill_rest(tab, P, true, cnt);
      flag = FLAG_HALVING;
      return MODE_DONE;
    }
    bool big = false;
#pragma unroll
    for (int i = 0; i < NP; ++i) big = big | (xp[i] >= COUNT_LIMIT);
    if (big) {
      fill_rest(tab, P, true, cnt);
      flag = FLAG_OVERFLOW;
      return MODE_DONE;
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
          if (q < P.kappa) {
            const int s = tab.Ev(q);
            const int32_t xs = pick_i<NP>(x, s);
            const int diff = pick_i<NP>(xn, s) - xs;
            Ob[((int64_t)q * P.ita + F) * P.o_stride] = xs + rn_f32_to_int((float)diff * frac);
          }
        }
        cnt.w.samples += 1;
        F = F + 1;
        if (F == P.ita) {
          done = true;
          break;
        }
        tnext = (double)P.I[F];
      }
    }
#pragma unroll
    for (int i = 0; i < NP; ++i) x[i] = xn[i];
    return done ? MODE_DONE : MODE_DECIDE;
  }
};

template <int NP, int MP>
__global__ void __launch_bounds__(256) sim_mixed_kernel(const JobDev<NP, MP>* __restrict__ jobs, int n_jobs,
                                                        unsigned int* __restrict__ queue,
                                                        unsigned long long* __restrict__ stats) {
  run_driver<EngineF32<NP, MP>, NP, MP>(jobs, n_jobs, queue, stats);
}

template <int NP, int MP>
static int launch_cls_m(const void* jobs, int n_jobs, unsigned int* queue, unsigned long long* stats, int waves,
                        hipStream_t stream) {
  const int block = 256;
  const int grid = (waves + 3) / 4;
  if (grid <= 0) return 0;
  hipLaunchKernelGGL((sim_mixed_kernel<NP, MP>), dim3(grid), dim3(block), 0, stream,
                     (const JobDev<NP, MP>*)jobs, n_jobs, queue, stats);
  return (int)hipGetLastError();
}

template <int NP, int MP>
static int occupancy_cls_m() {
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, sim_mixed_kernel<NP, MP>, 256, 0) != hipSuccess) return 1;
  return blocks > 0 ? blocks : 1;
}

int launch_mixed(int cls, const void* jobs, int n_jobs, unsigned int* queue, unsigned long long* stats, int waves,
                 hipStream_t stream) {
  switch (cls) {
    case 0: return launch_cls_m<4, 4>(jobs, n_jobs, queue, stats, waves, stream);
    case 1: return launch_cls_m<8, 8>(jobs, n_jobs, queue, stats, waves, stream);
    case 2: return launch_cls_m<16, 16>(jobs, n_jobs, queue, stats, waves, stream);
    default: return -1;
  }
}

int blocks_per_cu_mixed(int cls) {
  switch (cls) {
    case 0: return occupancy_cls_m<4, 4>();
    case 1: return occupancy_cls_m<8, 8>();
    case 2: return occupancy_cls_m<16, 16>();
    default: return 1;
  }
}

}  // namespace pst
