// pst_engine_f64.hpp -- plain FP64 simulation engine for gfx950 (one simulation per lane).
//
// This translation unit is compiled with -ffp-contract=off: every FP64
// operation is a single IEEE-754 rounding, so the integer trajectories are the
// same bits as a plain-double CPU evaluation of the same algorithm on the same
// Philox stream.
//
// Algorithm (per lane, one launch for the whole run):
//   tau-leaping with SSA fallback   P1_P2.py:54-297 + P3.py:29-163, iterated per
//                                   simulation as the host loop
//                                   cuTauLeaping.py:48-76 iterates them
//   forced SSA (gillespie=1)        P1_P2.py:67-74 -> P3 bursts only
//   pure SSA                        gillespie.py:27-149 (stops at t >= t_max)
// The reference's two launches per iteration, the per-iteration gpuarray.sum
// and the host round trip are replaced by a per-lane loop inside one launch:
// simulations never communicate, so each lane runs its simulation to the end.
//
// Random numbers are drawn in wavefront-convergent groups (see Key in
// pst_device.hpp): the first PTRS iteration of every reaction of a leap is
// evaluated for all reactions together from one group of Philox blocks; only
// the rare squeeze failures and rejections go through a shared per-lane loop.
// The launch loop, work queue and lane modes are in pst_driver.hpp.
#pragma once
#include "pst_driver.hpp"

namespace pst {

template <class Ref, int NP = Ref::kNP, int MP = Ref::kMP>
__device__ __forceinline__ void propensities_f64(const Ref& tab, const double (&k)[Ref::kKP],
                                                 const int32_t (&x)[NP], double (&a)[MP]) {
  if constexpr (Ref::kCustom) {
    Ref::template props<double>(x, k, a);  // compiled hazard expressions
  } else {
    // UpdatePropensities (parser.py:32-55) in table form, species ascending
#pragma unroll
    for (int j = 0; j < MP; ++j) {
      double v = k[j];
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        const int p = tab.Ord(i, j);
        const int st = tab.Stp(i, j);
        for (int q = 0; q < p; ++q) v = v * (double)(x[i] - st * q);
      }
      a[j] = v;
    }
  }
}

// CalculateG, P1_P2.py:447-486
__device__ __forceinline__ double calc_g_f64(int h, int ht, int xi) {
  if (h == 1) return 1.0;
  if (h == 2) {
    if (ht == 2) return 2.0 + 1.0 / ((double)xi - 1.0);
    return 2.0;
  }
  if (h == 3) {
    if (ht == 3) return 3.0 + 1.0 / ((double)xi - 1.0) + 2.0 / ((double)xi - 2.0);
    if (ht == 2) return 3.0 / 2.0 * (2.0 + 1.0 / ((double)xi - 1.0));
    return 3.0;
  }
  return 0.0;
}

// x[idx] for a run-time idx in [0, NP) (doubles), masked like pick_i (the bits
// are copied, so the value is exact)
template <int NP>
__device__ __forceinline__ double pick(const double (&x)[NP], int idx) {
  uint64_t r = 0;
#pragma unroll
  for (int i = 0; i < NP; ++i) r |= __builtin_bit_cast(uint64_t, x[i]) & (0ull - (uint64_t)(idx == i));
  return __builtin_bit_cast(double, r);
}

// PTRS (Hoermann 1993) constants for one Poisson mean
struct PtrsF64 {
  double lam, loglam, b, a, vr;
};

__device__ __forceinline__ PtrsF64 ptrs_setup_f64(double lam) {
  PtrsF64 P;
  P.lam = lam;
  P.loglam = log_f64(lam);
  const double slam = exp_f64(0.5 * P.loglam);
  P.b = 0.931 + 2.53 * slam;
  P.a = -0.059 + 0.02483 * P.b;
  P.vr = 0.9277 - 3.6224 / (P.b - 2.0);
  return P;
}

// The squeeze of one PTRS iteration: kf and whether it is accepted without
// the log-density test.
__device__ __forceinline__ bool ptrs_fast_f64(const PtrsF64& P, double U, double V, double& kf, double& us) {
  us = 0.5 - __builtin_fabs(U);
  kf = __builtin_floor((2.0 * P.a / us + P.b) * U + P.lam + 0.43);
  return us >= 0.07 && V <= P.vr;
}

// The rest of one PTRS iteration after a failed squeeze.
__device__ __forceinline__ bool ptrs_slow_f64(const PtrsF64& P, double V, double kf, double us) {
  if (kf < 0.0 || (us < 0.013 && V > us)) return false;
  const double invalpha = 1.1239 + 1.1328 / (P.b - 3.4);
  const double lhs = log_f64(V) + log_f64(invalpha) - log_f64(P.a / (us * us) + P.b);
  const double rhs = -P.lam + kf * P.loglam - lgamma_int_f64(kf + 1.0);
  return lhs <= rhs;
}

// PTRS iterations it0, it0 + 1, ... of the draw whose iteration `it` uses
// block (c, it); -1 at the iteration limit.  Used after iteration 0 failed.
__device__ __noinline__ Draw ptrs_rest_f64(uint32_t k0, uint32_t k1, uint32_t s_lo, uint32_t s_hi, uint64_t c,
                                           int it0, double lam) {
  Key key;
  key.k0 = k0;
  key.k1 = k1;
  key.s_lo = s_lo;
  key.s_hi = s_hi;
  const PtrsF64 P = ptrs_setup_f64(lam);
  Draw d = {-1, 0u};
  for (int it = it0; it < PTRS_MAX_ITERS; ++it) {
    const Words4 bw = key.block(c, (uint32_t)it);
    d.iters += 1;
    const double U = u53(bw.w0, bw.w1) - 0.5;
    const double V = u53(bw.w2, bw.w3);
    double kf, us;
    if (ptrs_fast_f64(P, U, V, kf, us) || ptrs_slow_f64(P, V, kf, us)) {
      d.k = (int64_t)kf;
      break;
    }
  }
  return d;
}

// Per-lane state and steps of one FP64 simulation.
template <class Ref>
struct EngineF64 {
  static constexpr int NP = Ref::kNP, MP = Ref::kMP;
  double k[Ref::kKP];  // rates row: per-reaction constants, or hazard parameters
  int32_t x[NP];
  Key key;
  uint64_t bc;  // Philox block counter
  double t, tnext;
  int F;
  int flag;
  int burst;    // SSA steps left in the current fallback burst (method 0)
  uint32_t nsteps;  // leap decisions + SSA steps so far (step budget)
  int32_t* Ob;  // the simulation's base in O (SampleOut::base)

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
    for (int j = 0; j < Ref::kKP; ++j) k[j] = (j < P.n_rates) ? glb(P.rates)[set * P.n_rates + j] : 0.0;
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

  // One step of the simulation: for a lane in DECIDE mode the P1_P2 body
  // (P1_P2.py:96-294) -- leap decision, then either the leap or, when the
  // decision falls back to SSA, the first step of the P3 burst; for a lane in
  // SSA mode one direct-method step (P3.py:73-149; gillespie.py:70-149 for
  // pure SSA).  An SSA step has the shape of the critical-reaction draw of a
  // leap (one Philox block: tau from rho1, the reaction from the cumulative
  // sum against rho2), over all reactions instead of the critical ones, so
  // both share that code.  SSA_ONLY drops the leap code for wavefronts whose
  // active lanes are all in SSA mode.  Returns the next mode.
  template <bool SSA_ONLY>
  __device__ __forceinline__ int step(const Ref& tab, Counters& cnt, int mode) {
    const auto& P = tab.p();
    const int M = P.M;
    bool ssa = SSA_ONLY || (mode == MODE_SSA);
    if (ssa && P.method == 2 && !(t < P.t_max)) {
      // kernel_Gillespie stops at t >= t_max; later samples stay zero (cuGillespie.py:71)
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
    double a[MP];
    propensities_f64<Ref>(tab, k, x, a);
    double a0 = 0.0;
#pragma unroll
    for (int j = 0; j < MP; ++j) a0 = a0 + a[j];
    if (a0 == 0.0) {
      fill_rest(tab, false, cnt);
      return MODE_DONE;
    }
    bool crit[MP];
#pragma unroll
    for (int j = 0; j < MP; ++j) crit[j] = false;
    double tau1 = __builtin_inf();
    if (!SSA_ONLY && !ssa) {
      cnt.w.decisions += 1;
      // DetermineCriticalReactions on V (P1_P2.py:314-327)
#pragma unroll
      for (int j = 0; j < MP; ++j) {
        bool c = false;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
          const int v = tab.V(i, j);
          if (v < 0) c = c || ((-v) * P.n_c > x[i]);
        }
        crit[j] = c;
      }
      // CalculateMuSigma (P1_P2.py:330-383) and CalculateTau (P1_P2.py:385-445)
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        if (tab.Irs(i)) {
          double mu = 0.0, s2 = 0.0;
          bool any = false;
#pragma unroll
          for (int j = 0; j < MP; ++j) {
            const int v = tab.V(i, j);
            if (v != 0 && !crit[j]) {
              mu = mu + ((double)v * a[j]);
              s2 = s2 + ((double)(v * v) * a[j]);
              any = true;
            }
          }
          if (any) {
            // one candidate per V entry of this species; all equal, so one suffices
            const double g_i = calc_g_f64(tab.Hv(i), tab.Htv(i), x[i]);
            const double num = (double)fmaxf((float)(P.eta * (double)x[i] / g_i), 1.0f);
            const double lhs = num / (double)fabsf((float)mu);
            const double rhs = (num * num) / s2;
            const double tt = (double)fminf((float)lhs, (float)rhs);
            if (tt < tau1) tau1 = tt;
          }
        }
      }
      // SSA switch (P1_P2.py:157-173): the burst starts with this step
      if (tau1 < 10.0 / a0) {
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
    // the one-block draw: SSA step, or tau2 + SingleCriticalReaction (P1_P2.py:175-220)
    bool msk[MP];
    double asum = 0.0;
#pragma unroll
    for (int j = 0; j < MP; ++j) {
      msk[j] = ssa ? (j < M) : crit[j];
      if (msk[j]) asum = asum + a[j];
    }
    double tauX = __builtin_inf();
    uint32_t cw2 = 0, cw3 = 0;
    if (asum > 0.0) {
      const Words4 bw = key.block(bc, 0);
      bc += 1;
      cnt.w.uniforms += ssa ? 2 : 1;
      cw2 = bw.w2;
      cw3 = bw.w3;
      const double rho1 = u53(bw.w0, bw.w1);
      tauX = (1.0 / asum) * log_f64(1.0 / rho1);
    }
    double tau = ssa ? tauX : (double)fminf((float)tau1, (float)tauX);
    int sel = -1;
    if (ssa || tau1 > tauX) {
      // first j of the mask with sum_{r<=j} a_r/asum >= rho (P3.py:179-194,
      // P1_P2.py:488-507); on round-off fall-through take the last j with a_j > 0
      if (!ssa) cnt.w.uniforms += 1;
      const double rho = u53(cw2, cw3);
      double rc = 0.0;
      int last = -1;
#pragma unroll
      for (int j = 0; j < MP; ++j) {
        if (msk[j] && sel < 0) {
          if (a[j] > 0.0) last = j;
          rc = rc + a[j] / asum;
          if (rc >= rho) sel = j;
        }
      }
      if (sel < 0) sel = last;
    }
    int32_t xn[NP];
    if (ssa) {
      // UpdateState (P3.py:196-208), applied after the samples below
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        int dx = 0;
#pragma unroll
        for (int j = 0; j < MP; ++j)
          if (sel == j) dx = tab.V(i, j);
        xn[i] = x[i] + dx;
      }
    } else if (!SSA_ONLY) {
      if (!(tau < __builtin_inf())) {
        fill_rest(tab, true, cnt);
        flag = FLAG_TAU_INF;
        return MODE_DONE;
      }
      int32_t K[MP];
#pragma unroll
      for (int j = 0; j < MP; ++j) K[j] = (sel == j) ? 1 : 0;
      const int st = leap(tab, cnt, a, crit, tau, K, xn);
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
    // time advance; samples: sample-and-hold of the pre-step state for SSA
    // (SaveDynamics, P3.py:165-177), interpolation for a leap
    // (SaveInterpolatedDynamics, P1_P2.py:261-286,542-584)
    const double t_prev = t;
    t = t + tau;
    if (t == t_prev && tau > 0.0) {  // tau == 0 (a draw of u = 1) is a valid event
      fill_rest(tab, true, cnt);
      flag = FLAG_STALL;
      return MODE_DONE;
    }
    while (t >= tnext) {
      const double t0 = t - tau;
      const float record_time = glb(P.I)[F];
      const double gap = t - t0;
      const double nrt = (double)record_time - t0;
      const double frac = nrt / gap;
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        if (q < P.kappa) {
          const int s = tab.Ev(q);
          const int32_t xs = pick_i<NP>(x, s);
          const int diff = pick_i<NP>(xn, s) - xs;
          const double inc = (double)diff * frac;
          Out::put(P, Ob, q, F, ssa ? xs : xs + rn_f32_to_int((float)inc));
        }
      }
      Out::row_done(P, Ob, P.kappa, F);
      cnt.w.samples += 1;
      F = F + 1;
      if (F == P.ita) return MODE_DONE;  // an SSA step ending the run is not applied
      tnext = (double)glb(P.I)[F];
    }
    if (ssa) {
      bool over = false;
#pragma unroll
      for (int i = 0; i < NP; ++i) over = over || (xn[i] >= COUNT_LIMIT);
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

  // Poisson leap with tau halving (P1_P2.py:222-259,509-540).  K holds the
  // critical firing (if any) on entry and the accepted counts on return; tau
  // is halved in place, and the clock advances by the tau of the accepted
  // draw (P1_P2.py:258-261); returns 0 or the flag that ends the simulation.
  __device__ __forceinline__ int leap(const Ref& tab, Counters& cnt,
                                      const double (&a)[MP], const bool (&crit)[MP], double& tau,
                                      int32_t (&K)[MP], int32_t (&xn)[NP]) {
    const auto& P = tab.p();
    const int M = P.M;
    for (int h = 0; h < MAX_HALVINGS; ++h) {
      const uint64_t grp = bc;
      bc += (uint64_t)M;
      // reaction j's draw uses blocks (grp + j, it); every reaction's
      // iteration-0 block and squeeze are evaluated together
      Words4 W[MP];
      double lam[MP];
      unsigned inv = 0, pend = 0;
      bool bad = false;
#pragma unroll
      for (int j = 0; j < MP; ++j) {
        if (j < M) W[j] = key.block(grp + (uint64_t)j, 0);
        const double l = (j < M && !crit[j]) ? a[j] * tau : 0.0;
        lam[j] = l;
        if (!crit[j]) K[j] = 0;
        bad = bad || (l >= LAMBDA_LIMIT);
        inv |= (l > 0.0 && l < 10.0) ? (1u << j) : 0u;
      }
      if (bad) return FLAG_OVERFLOW;
      // PTRS (lambda >= 10), iteration 0
#pragma unroll
      for (int j = 0; j < MP; ++j) {
        const bool use = lam[j] >= 10.0;
        if (wave_any(use)) {
          if (use) {
            cnt.w.uniforms += 2;
            cnt.w.ptrs += 1;
            const PtrsF64 Pt = ptrs_setup_f64(lam[j]);
            const double U = u53(W[j].w0, W[j].w1) - 0.5;
            const double V = u53(W[j].w2, W[j].w3);
            double kf, us;
            if (ptrs_fast_f64(Pt, U, V, kf, us) || ptrs_slow_f64(Pt, V, kf, us)) K[j] = (int32_t)kf;
            else pend |= 1u << j;
          }
        }
      }
      // sequential-search inversion (0 < lambda < 10), all such reactions in
      // lock-step: after `it` trips every live search is at k = it
      if (inv != 0u) {
        double p[MP], Fc[MP], u[MP];
        unsigned act = 0;
#pragma unroll
        for (int j = 0; j < MP; ++j) {
          p[j] = 0.0;
          Fc[j] = 0.0;
          u[j] = 0.0;
          if (wave_any((inv >> j) & 1u)) {
            if ((inv >> j) & 1u) {
              cnt.w.uniforms += 1;
              u[j] = u53(W[j].w0, W[j].w1);
              p[j] = exp_f64(-lam[j]);
              Fc[j] = p[j];
              if (u[j] > Fc[j]) act |= 1u << j;
            }
          }
        }
        int it = 0;
        while (act != 0u && it < INV_MAX_K) {
          it += 1;
          const double dk = (double)it;
#pragma unroll
          for (int j = 0; j < MP; ++j) {
            if (wave_any((act >> j) & 1u)) {
              if ((act >> j) & 1u) {
                p[j] = p[j] * lam[j] / dk;
                Fc[j] = Fc[j] + p[j];
                K[j] = it;
                if (!(u[j] > Fc[j])) act &= ~(1u << j);
              }
            }
          }
        }
#pragma unroll
        for (int j = 0; j < MP; ++j)
          if ((inv >> j) & 1u) cnt.w.inv += (unsigned int)K[j];
      }
      // PTRS draws rejected at iteration 0: continue from iteration 1
      while (pend != 0u) {
        const int j = __builtin_ctz(pend);
        pend &= pend - 1u;
        const Draw dr = ptrs_rest_f64(key.k0, key.k1, key.s_lo, key.s_hi, grp + (uint64_t)j, 1, pick<MP>(lam, j));
        cnt.w.ptrs += dr.iters;
        cnt.w.uniforms += 2 * dr.iters;
        const int64_t kk = dr.k;
        if (kk < 0) return FLAG_PTRS;
#pragma unroll
        for (int r = 0; r < MP; ++r)
          if (r == j) K[r] = (int32_t)kk;
      }
      // TentativeUpdatedState / ValidState (P1_P2.py:509-540)
      bool ok = true;
      bool big = false;
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        int64_t s = x[i];
#pragma unroll
        for (int j = 0; j < MP; ++j) {
          const int v = tab.V(i, j);
          if (v != 0) s += (int64_t)K[j] * v;
        }
        ok = ok && (s >= 0);
        big = big || (s >= COUNT_LIMIT);
        xn[i] = (int32_t)s;
      }
      if (ok) return big ? FLAG_OVERFLOW : 0;
      cnt.rejected += 1;
      tau = tau / 2.0;
#pragma unroll
      for (int j = 0; j < MP; ++j)
        if (!crit[j]) K[j] = 0;
    }
    return FLAG_HALVING;
  }
};

}  // namespace pst
