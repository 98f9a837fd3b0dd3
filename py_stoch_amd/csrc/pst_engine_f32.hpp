// pst_engine_f32.hpp -- mixed FP32/FP64 simulation engine for gfx950.
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
#pragma once
#include "pst_driver.hpp"

namespace pst {

template <class Ref, int NP = Ref::kNP, int MP = Ref::kMP>
__device__ __forceinline__ void propensities_f32(const Ref& tab, const float (&k)[Ref::kKP],
                                                 const int32_t (&x)[NP], float (&a)[MP]) {
  if constexpr (Ref::kCustom) {
    Ref::template props<float>(x, k, a);  // compiled hazard expressions
  } else {
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

// log(1 + d) with the rounding of 1 + d corrected (Goldberg's method)
__device__ __forceinline__ float log1p_f32(float d) {
  const float u = 1.0f + d;
  const float m = u - 1.0f;
  return (m == 0.0f) ? d : fast_log(u) * (d * fast_rcp(m));
}

struct PtrsF32 {
  float lam, slam, b, a, vr, twoa;
};

__device__ __forceinline__ PtrsF32 ptrs_setup_f32(float lam) {
  PtrsF32 P;
  P.lam = lam;
  P.slam = __builtin_amdgcn_sqrtf(lam);  // hardware v_sqrt_f32 (1 ulp)
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

__device__ __forceinline__ bool ptrs_slow_f32(const PtrsF32& P, float V, float kf, float us) {
  if (kf < 0.0f || (us < 0.013f && V > us)) return false;
  const float invalpha = 1.1239f + 1.1328f * fast_rcp(P.b - 3.4f);
  const float lhs = fast_log(V) + fast_log(invalpha) - fast_log(P.a * fast_rcp(us * us) + P.b);
  return lhs <= log_pmf_f32(kf, P.lam);
}

// PTRS for Poisson means of 2^16 and above (rare: the inversion of
// poisson_asym_f32 covers the means below): iteration i of the draw in word
// slot `slot` of block c uses block (c, 1 + 4 i + slot), U from its word 0
// and V from word 1, so no two draws share a block.  FP32 below 2^22, FP64
// above (FP32 cannot resolve the integers there).  k = -1 at the iteration
// limit.
constexpr float BIG_LAMBDA_F32 = 4194304.0f;  // 2^22

__device__ __noinline__ Draw ptrs_wide(uint32_t k0, uint32_t k1, uint32_t s_lo, uint32_t s_hi, uint64_t c, int slot,
                                       float lamf) {
  Key key;
  key.k0 = k0;
  key.k1 = k1;
  key.s_lo = s_lo;
  key.s_hi = s_hi;
  Draw d = {-1, 0u};
  if (lamf < BIG_LAMBDA_F32) {
    const PtrsF32 P = ptrs_setup_f32(lamf);
    for (int i = 0; i < PTRS_MAX_ITERS; ++i) {
      const Words4 bw = key.block(c, 1u + 4u * (uint32_t)i + (uint32_t)slot);
      d.iters += 1;
      const float U = u24(bw.w0) - 0.5f;
      const float V = u24(bw.w1);
      float kf, us;
      if (ptrs_fast_f32(P, U, V, kf, us) || ptrs_slow_f32(P, V, kf, us)) {
        d.k = (int64_t)kf;
        break;
      }
    }
    return d;
  }
  const double lam = (double)lamf;
  const double loglam = log_f64(lam);
  const double b = 0.931 + 2.53 * __builtin_sqrt(lam);
  const double a = -0.059 + 0.02483 * b;
  const double vr = 0.9277 - 3.6224 / (b - 2.0);
  const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
  for (int i = 0; i < PTRS_MAX_ITERS; ++i) {
    const Words4 bw = key.block(c, 1u + 4u * (uint32_t)i + (uint32_t)slot);
    d.iters += 1;
    const double U = u53(bw.w0, bw.w1) - 0.5;
    const double V = u53(bw.w1, bw.w0);
    const double us = 0.5 - __builtin_fabs(U);
    const double kf = __builtin_floor((2.0 * a / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) {
      d.k = (int64_t)kf;
      break;
    }
    if (kf < 0.0 || (us < 0.013 && V > us)) continue;
    const double lhs = log_f64(V) + log_f64(invalpha) - log_f64(a / (us * us) + b);
    const double rhs = -lam + kf * loglam - lgamma_int_f64(kf + 1.0);
    if (lhs <= rhs) {
      d.k = (int64_t)kf;
      break;
    }
  }
  return d;
}

// Per-lane state and steps of one mixed-precision simulation.
template <class Ref>
struct EngineF32 {
  static constexpr int NP = Ref::kNP, MP = Ref::kMP;
  float k[Ref::kKP];  // rates row: per-reaction constants, or hazard parameters
  int32_t x[NP];
  Key key;
  uint64_t bc;
  double t, tnext;
  int F;
  int flag;
  int burst;
  uint32_t nsteps;  // leap decisions + SSA steps so far (step budget)
  int32_t* Ob;

  using Out = SampleOut<Ref::kOutLds>;

  __device__ __forceinline__ void fill_rest(const Ref& tab, bool zero, Counters& cnt) const {
    const auto& P = tab.p();
    cnt.w.samples += (unsigned int)(P.ita - F);
    if (P.layout != LAYOUT_BLOCKS) {
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        if (q < P.kappa) {
          const int32_t val = zero ? 0 : pick_i<NP>(x, tab.Ev(q));
          for (int f = F; f < P.ita; ++f) glb(Ob)[((int64_t)q * P.ita + f) * P.o_stride] = val;
        }
      }
      return;
    }
    for (int f = F; f < P.ita; ++f) {
#pragma unroll
      for (int q = 0; q < NP; ++q)
        if (q < P.kappa) Out::put(P, Ob, q, f, zero ? 0 : pick_i<NP>(x, tab.Ev(q)));
      Out::row_done(P, Ob, P.kappa, f);
    }
  }

  __device__ __forceinline__ int start(const Ref& jr, int64_t sim) {
    const auto& P = jr.p();
    const int64_t set = sim / P.sims_per_set;
#pragma unroll
    for (int j = 0; j < Ref::kKP; ++j) k[j] = (j < P.n_rates) ? (float)glb(P.rates)[set * P.n_rates + j] : 0.0f;
#pragma unroll
    for (int i = 0; i < NP; ++i) x[i] = (i < P.N) ? glb(P.x0)[set * P.N + i] : 0;
    key.init(P.seed, P.sim_offset + (uint64_t)sim);
    bc = 0;
    t = 0.0;
    F = 0;
    tnext = (double)glb(P.I)[0];
    flag = 0;
    burst = 0;
    nsteps = 0;
    Ob = Out::base(P, sim);
    return P.method == 0 ? MODE_DECIDE : MODE_SSA;
  }

  // One step: leap decision and leap (P1_P2.py:96-294), or one SSA step
  // (P3.py:73-149, gillespie.py:70-149) -- see EngineF64::step for the
  // structure; the arithmetic is FP32 except the clock and the cumulative
  // sum of the reaction choice.
  template <bool SSA_ONLY>
  __device__ __forceinline__ int step(const Ref& tab, Counters& cnt, int mode) {
    const auto& P = tab.p();
    const int M = P.M;
    bool ssa = SSA_ONLY || (mode == MODE_SSA);
    if (ssa && P.method == 2 && !(t < P.t_max)) {
      fill_rest(tab, true, cnt);
      return MODE_DONE;
    }
    // step budget, checked before every leap decision and every SSA step
    if (P.max_steps != 0u && nsteps >= P.max_steps) {
      fill_rest(tab, true, cnt);
      flag = FLAG_STEPS;
      return MODE_DONE;
    }
    nsteps += 1;
    // Philox blocks bc .. bc + NB - 1, generated together as independent
    // chains: block bc serves the one-block draw (SSA step, or tau2 and the
    // critical reaction) and the first Poisson attempt takes NB consecutive
    // blocks after it, or from bc when there is no one-block draw; block
    // bc + NB is generated at the leap, only if a lane of the wavefront made
    // a one-block draw.
    constexpr int NB = (MP + 3) / 4;
    const uint64_t bc0 = bc;
    Words4 B[NB + 1];
    B[0] = key.block(bc, 0);
    B[NB] = Words4{0u, 0u, 0u, 0u};
    if (!SSA_ONLY) {
#pragma unroll
      for (int r = 1; r < NB; ++r)
        if (4 * (r - 1) < M) B[r] = key.block(bc + (uint64_t)r, 0);
    }
    float a[MP];
    propensities_f32<Ref>(tab, k, x, a);
    float a0 = 0.0f;
#pragma unroll
    for (int j = 0; j < MP; ++j) a0 += a[j];
    if (a0 == 0.0f) {
      fill_rest(tab, false, cnt);
      return MODE_DONE;
    }
    bool crit[MP];
#pragma unroll
    for (int j = 0; j < MP; ++j) crit[j] = false;
    float tau1 = __builtin_inff();
    if (!SSA_ONLY && !ssa) {
      cnt.w.decisions += 1;
      // critical reactions (P1_P2.py:314-327)
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
      const float eta = (float)P.eta;
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
      if (tau1 < 10.0f * fast_rcp(a0)) {
        ssa = true;
        burst = P.ssa_burst;
        if (P.max_steps != 0u && nsteps >= P.max_steps) {
          fill_rest(tab, true, cnt);
          flag = FLAG_STEPS;
          return MODE_DONE;
        }
        nsteps += 1;
      }
    }
    bool msk[MP];
    double asum = 0.0;
#pragma unroll
    for (int j = 0; j < MP; ++j) {
      msk[j] = ssa ? (j < M) : crit[j];
      asum += msk[j] ? (double)a[j] : 0.0;
    }
    float tauX = __builtin_inff();
    uint32_t cw1 = 0;
    const bool one = asum > 0.0;
    if (one) {
      const Words4 bw = B[0];
      bc += 1;
      cnt.w.uniforms += ssa ? 2 : 1;
      cw1 = bw.w1;
      tauX = -fast_log(u24(bw.w0)) * fast_rcp((float)asum);
    }
    const float tau_f = ssa ? tauX : fminf(tau1, tauX);
    int sel = -1;
    if (ssa || tau1 > tauX) {
      if (!ssa) cnt.w.uniforms += 1;
      // first j of the mask with sum_{r<=j} a_r >= rho2 * asum, summed in FP64
      const double target = (double)u24(cw1) * asum;
      double cum = 0.0;
      int last = -1;
      if constexpr (SSA_ONLY) {
        // The running sum never decreases and rho2 * asum > 0, so that j is
        // the number of positions whose sum is still below the target (a
        // masked-out position repeats its predecessor's sum and is never the
        // first to reach it): no branches in the SSA-only steps (Schlogl pure
        // SSA 68 -> 59 ms; the leap step keeps the loop, which measured faster
        // there).  The last positive masked reaction covers a sum that ends
        // below the target.
        int below = 0;
#pragma unroll
        for (int j = 0; j < MP; ++j) {
          cum += msk[j] ? (double)a[j] : 0.0;
          last = (msk[j] && a[j] > 0.0f) ? j : last;
          below += (cum < target) ? 1 : 0;
        }
        sel = below < MP ? below : last;
      } else {
#pragma unroll
        for (int j = 0; j < MP; ++j) {
          if (msk[j] && sel < 0) {
            if (a[j] > 0.0f) last = j;
            cum += (double)a[j];
            if (cum >= target) sel = j;
          }
        }
        if (sel < 0) sel = last;
      }
    }
    int32_t xn[NP];
    float tau = tau_f;
    if (ssa) {
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        int dx = 0;
#pragma unroll
        for (int j = 0; j < MP; ++j)
          if (sel == j) dx = tab.V(i, j);
        xn[i] = x[i] + dx;
      }
    } else if (!SSA_ONLY) {
      if (!(tau < __builtin_inff())) {
        fill_rest(tab, true, cnt);
        flag = FLAG_TAU_INF;
        return MODE_DONE;
      }
      int32_t K[MP];
#pragma unroll
      for (int j = 0; j < MP; ++j) K[j] = (sel == j) ? 1 : 0;
      if (wave_any(one)) B[NB] = key.block(bc0 + (uint64_t)NB, 0);
      const int st = leap(tab, cnt, a, crit, tau, K, xn, B, one);
      if (st != 0) {
        fill_rest(tab, true, cnt);
        flag = st;
        return MODE_DONE;
      }
      unsigned long long ev = 0;
#pragma unroll
      for (int j = 0; j < MP; ++j) ev += (unsigned long long)K[j];
      cnt.events += ev;
      cnt.leaps += 1;
    }
    const double t0 = t;
    t = t + (double)tau;
    if (t == t0 && tau > 0.0f) {  // tau == 0 (a draw of u = 1) is a valid event
      fill_rest(tab, true, cnt);
      flag = FLAG_STALL;
      return MODE_DONE;
    }
    if (t >= tnext) {
      const float inv_gap = fast_rcp((float)(t - t0));
      while (t >= tnext) {
        // sample-and-hold for an SSA step: a zero fraction keeps the pre-step state
        const float frac = ssa ? 0.0f : (float)(tnext - t0) * inv_gap;
#pragma unroll
        for (int q = 0; q < NP; ++q) {
          if (q < P.kappa) {
            const int s = tab.Ev(q);
            const int32_t xs = pick_i<NP>(x, s);
            const int diff = pick_i<NP>(xn, s) - xs;
            Out::put(P, Ob, q, F, xs + rn_f32_to_int((float)diff * frac));
          }
        }
        Out::row_done(P, Ob, P.kappa, F);
        cnt.w.samples += 1;
        F = F + 1;
        if (F == P.ita) return MODE_DONE;
        tnext = (double)glb(P.I)[F];
      }
    }
    if (ssa) {
      bool over = false;
#pragma unroll
      for (int i = 0; i < NP; ++i) over = over | (xn[i] >= COUNT_LIMIT);
      cnt.ssa += 1;
      cnt.events += 1;
#pragma unroll
      for (int i = 0; i < NP; ++i) x[i] = xn[i];
      if (over) {
        fill_rest(tab, true, cnt);
        flag = FLAG_OVERFLOW;
        return MODE_DONE;
      }
      if (P.method != 0) return MODE_SSA;
      burst -= 1;
      return burst > 0 ? MODE_SSA : MODE_DECIDE;
    }
#pragma unroll
    for (int i = 0; i < NP; ++i) x[i] = xn[i];
    return MODE_DECIDE;
  }

  // Poisson leap with tau halving (P1_P2.py:222-259,509-540).  Every draw is
  // an inversion of the Poisson CDF at one 24-bit uniform: reaction j reads
  // word j % 4 of block grp + j / 4 (so one block serves four reactions);
  // lambda < 10 by sequential search, 10 <= lambda < 2^16 by the asymptotic
  // inversion with exact FP64 decisions near integer boundaries (both are
  // inversions, so no draw is rejected and repeated); lambda >= 2^16 by
  // PTRS on blocks of its own (ptrs_wide).  tau is halved in place: the clock
  // advances by the tau of the accepted draw (P1_P2.py:258-261).  Returns 0
  // or a flag.
  __device__ __forceinline__ int leap(const Ref& tab, Counters& cnt,
                                      const float (&a)[MP], const bool (&crit)[MP], float& tau,
                                      int32_t (&K)[MP], int32_t (&xn)[NP], const Words4 (&B)[(MP + 3) / 4 + 1],
                                      bool shifted) {
    constexpr int NB = (MP + 3) / 4;
    const auto& P = tab.p();
    const int M = P.M;
    const int nblk = (M + 3) >> 2;
    for (int h = 0; h < MAX_HALVINGS; ++h) {
      const uint64_t grp = bc;
      bc += (uint64_t)nblk;
      // first attempt: the blocks generated with the step's; retries: new ones
      Words4 W[NB];
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        if (4 * r < M) {
          if (h == 0) {
            W[r].w0 = shifted ? B[r + 1].w0 : B[r].w0;
            W[r].w1 = shifted ? B[r + 1].w1 : B[r].w1;
            W[r].w2 = shifted ? B[r + 1].w2 : B[r].w2;
            W[r].w3 = shifted ? B[r + 1].w3 : B[r].w3;
          } else {
            W[r] = key.block(grp + (uint64_t)r, 0);
          }
        }
      }
      float lam[MP], u[MP];
      unsigned inv = 0, exact = 0, wide = 0, nasym = 0;
      int32_t mb[MP];
      bool bad = false;
#pragma unroll
      for (int j = 0; j < MP; ++j) {
        const float l = (j < M && !crit[j]) ? a[j] * tau : 0.0f;
        lam[j] = l;
        const Words4& wd = W[j >> 2];
        const uint32_t word = (j & 3) == 0 ? wd.w0 : (j & 3) == 1 ? wd.w1 : (j & 3) == 2 ? wd.w2 : wd.w3;
        u[j] = u24(word);
        const bool asym = (l >= POIS_ASYM_MIN) & (l < POIS_ASYM_MAX);
        bool ex = false;
        int32_t m = -1;
        int32_t k = 0;
        if (wave_any(asym)) k = poisson_asym_f32(__builtin_fmaxf(l, POIS_ASYM_MIN), u[j], ex, m);
        if (!crit[j]) K[j] = asym ? k : 0;
        mb[j] = m;
        exact |= (asym & ex) ? (1u << j) : 0u;
        nasym += asym ? 1u : 0u;
        inv |= ((l > 0.0f) & (l < POIS_ASYM_MIN)) ? (1u << j) : 0u;
        wide |= (l >= POIS_ASYM_MAX) ? (1u << j) : 0u;
        bad = bad | (l >= (float)LAMBDA_LIMIT);
      }
      if (bad) return FLAG_OVERFLOW;
      cnt.w.asym += nasym;
      cnt.w.uniforms += nasym;
      // near an integer boundary, or in a far tail: decide exactly in FP64
      if (wave_any(exact != 0u)) {
        cnt.w.exact += (unsigned int)__builtin_popcount(exact);
#pragma unroll
        for (int j = 0; j < MP; ++j)
          if ((exact >> j) & 1u) K[j] = poisson_exact_f64((double)lam[j], (double)u[j], mb[j]);
      }
      // sequential-search inversion (0 < lambda < 10) for all such reactions
      // in lock-step, two search steps per trip: a live search at k = it
      // steps to it + 1 and it + 2 and stops at the first k with u <= F(k);
      // p and F past a search's end are computed but never read
      if (wave_any(inv != 0u)) {
        float p[MP], Fc[MP];
        bool live[MP];
        bool any_live = false;
#pragma unroll
        for (int j = 0; j < MP; ++j) {
          const bool on = (inv >> j) & 1u;
          p[j] = fast_exp(-lam[j]);
          Fc[j] = p[j];
          live[j] = on && u[j] > Fc[j];
          if (on) K[j] = 0;
          any_live = any_live | live[j];
        }
        cnt.w.uniforms += (unsigned int)__builtin_popcount(inv);
        for (int it = 0; it < INV_MAX_K && wave_any(any_live); it += 2) {
          const float r1 = fast_rcp((float)(it + 1));
          const float r2 = fast_rcp((float)(it + 2));
          any_live = false;
#pragma unroll
          for (int j = 0; j < MP; ++j) {
            const float p1 = p[j] * lam[j] * r1;
            const float F1 = Fc[j] + p1;
            const float p2 = p1 * lam[j] * r2;
            const float F2 = F1 + p2;
            K[j] = live[j] ? ((u[j] > F1) ? it + 2 : it + 1) : K[j];
            // a search also ends where the FP32 sum stops growing: past the
            // mode every later term is smaller still, and a u above that sum
            // (u24 can be 1.0, the sum can end a few ulps below it) would
            // otherwise run to INV_MAX_K
            live[j] = live[j] && (u[j] > F2) && (F2 > Fc[j]);
            any_live = any_live | live[j];
            p[j] = p2;
            Fc[j] = F2;
          }
        }
#pragma unroll
        for (int j = 0; j < MP; ++j)
          if ((inv >> j) & 1u) cnt.w.inv += (unsigned int)K[j];
      }
      // means of 2^16 and above (rare): PTRS, one reaction at a time
      if (wave_any(wide != 0u)) {
#pragma unroll
        for (int j = 0; j < MP; ++j) {
          if ((wide >> j) & 1u) {
            const Draw dr = ptrs_wide(key.k0, key.k1, key.s_lo, key.s_hi, grp + (uint64_t)(j >> 2), j & 3,
                                      lam[j]);
            cnt.w.ptrs += dr.iters;
            cnt.w.uniforms += 2 * dr.iters;
            if (dr.k < 0) return FLAG_PTRS;
            K[j] = (int32_t)dr.k;
          }
        }
      }
      bool ok = true, big = false;
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        int64_t sacc = x[i];
#pragma unroll
        for (int j = 0; j < MP; ++j) {
          const int v = tab.V(i, j);
          if (v != 0) sacc += (int64_t)K[j] * v;
        }
        ok = ok & (sacc >= 0);
        big = big | (sacc >= COUNT_LIMIT);
        xn[i] = (int32_t)sacc;
      }
      if (ok) return big ? FLAG_OVERFLOW : 0;
      cnt.rejected += 1;
      tau = tau * 0.5f;
    }
    return FLAG_HALVING;
  }
};

}  // namespace pst
