// pst_fp64.hip -- plain FP64 simulation kernels for gfx950 (one simulation per lane).
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
#include "pst_driver.hpp"

namespace pst {

template <int NP, int MP>
__device__ __forceinline__ void propensities_f64(const JobRef<NP, MP>& tab, const double (&k)[MP],
                                                 const int32_t (&x)[NP], double (&a)[MP]) {
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

template <int NP, typename T>
__device__ __forceinline__ T pick(const T (&x)[NP], int idx) {
  T r = x[0];
#pragma unroll
  for (int i = 1; i < NP; ++i)
    if (idx == i) r = x[i];
  return r;
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
  us = 0.5 - fabs(U);
  kf = floor((2.0 * P.a / us + P.b) * U + P.lam + 0.43);
  return us >= 0.07 && V <= P.vr;
}

// The rest of one PTRS iteration after a failed squeeze.
__device__ __noinline__ bool ptrs_slow_f64(const PtrsF64& P, double V, double kf, double us) {
  if (kf < 0.0 || (us < 0.013 && V > us)) return false;
  const double invalpha = 1.1239 + 1.1328 / (P.b - 3.4);
  const double lhs = log_f64(V) + log_f64(invalpha) - log_f64(P.a / (us * us) + P.b);
  const double rhs = -P.lam + kf * P.loglam - lgamma_int_f64(kf + 1.0);
  return lhs <= rhs;
}

// Full PTRS draw from block counter c (iterations it = 0, 1, ...); -1 at the
// iteration limit.  Used for draws whose first squeeze failed.
__device__ __noinline__ int64_t ptrs_full_f64(const Key& key, uint64_t c, double lam, Work& w) {
  const PtrsF64 P = ptrs_setup_f64(lam);
  for (int it = 0; it < PTRS_MAX_ITERS; ++it) {
    const Words4 bw = key.block(c, (uint32_t)it);
    if (it > 0) {
      w.ptrs += 1;
      w.uniforms += 2;
    }
    const double U = u53(bw.w0, bw.w1) - 0.5;
    const double V = u53(bw.w2, bw.w3);
    double kf, us;
    if (ptrs_fast_f64(P, U, V, kf, us)) return (int64_t)kf;
    if (ptrs_slow_f64(P, V, kf, us)) return (int64_t)kf;
  }
  return -1;
}

// Sequential-search inversion (Devroye 1986, X.3) for 0 < lambda < 10
__device__ __forceinline__ int64_t inversion_f64(double u, double lam, Work& w) {
  double p = exp_f64(-lam);
  double F = p;
  int64_t k = 0;
  while (u > F && k < INV_MAX_K) {
    k = k + 1;
    p = p * lam / (double)k;
    F = F + p;
  }
  w.inv += (unsigned int)k;
  return k;
}

// Per-lane state and steps of one FP64 simulation.
template <int NP, int MP>
struct EngineF64 {
  double k[MP];
  int32_t x[NP];
  Key key;
  uint64_t bc;  // Philox block counter
  double t, tnext;
  int F;
  int flag;
  int32_t* Ob;  // &O[0][0][sim]

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
    for (int j = 0; j < MP; ++j) k[j] = (j < P.M) ? P.rates[set * P.M + j] : 0.0;
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

  // one direct-method SSA step (P3.py:73-149); 1 when the run ended
  // (absorbed or last sample written), 2 on overflow, 0 otherwise
  __device__ __forceinline__ int ssa_step(const JobRef<NP, MP>& tab, Counters& cnt) {
    const JobParams& P = tab.p();
    const int M = P.M;
    double a[MP];
    propensities_f64<NP, MP>(tab, k, x, a);
    double a0 = 0.0;
#pragma unroll
    for (int j = 0; j < MP; ++j) a0 = a0 + a[j];
    if (a0 == 0.0) {
      fill_rest(tab, P, false, cnt);
      return 1;
    }
    const Words4 bw = key.block(bc, 0);
    bc += 1;
    cnt.w.uniforms += 2;
    const double rho1 = u53(bw.w0, bw.w1);
    const double tau = (1.0 / a0) * log_f64(1.0 / rho1);
    // SingleCriticalReaction (P3.py:179-194): first j with sum_{r<=j} a_r/a0 >= rho;
    // the reference falls off the end on round-off: take the last j with a_j > 0
    const double rho = u53(bw.w2, bw.w3);
    double rc = 0.0;
    int sel = -1, last = -1;
#pragma unroll
    for (int j = 0; j < MP; ++j) {
      if (j < M && sel < 0) {
        if (a[j] > 0.0) last = j;
        rc = rc + a[j] / a0;
        if (rc >= rho) sel = j;
      }
    }
    if (sel < 0) sel = last;
    t = t + tau;
    while (t >= tnext) {
      save(tab, P, F, x);  // SaveDynamics: pre-reaction state (P3.py:165-177)
      cnt.w.samples += 1;
      F = F + 1;
      if (F == P.ita) return 1;
      tnext = (double)P.I[F];
    }
    // UpdateState (P3.py:196-208)
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

  // the P1_P2 body (P1_P2.py:96-294): returns the next mode
  __device__ __forceinline__ int decide(const JobRef<NP, MP>& tab, Counters& cnt) {
    const JobParams& P = tab.p();
    const int M = P.M;
    double a[MP];
    propensities_f64<NP, MP>(tab, k, x, a);
    double a0 = 0.0;
#pragma unroll
    for (int j = 0; j < MP; ++j) a0 = a0 + a[j];
    if (a0 == 0.0) {
      fill_rest(tab, P, false, cnt);
      return MODE_DONE;
    }
    cnt.w.decisions += 1;
    // DetermineCriticalReactions on V (P1_P2.py:314-327)
    bool crit[MP];
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
    double tau1 = __builtin_inf();
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
    // SSA switch (P1_P2.py:157-173)
    if (tau1 < 10.0 / a0) return MODE_SSA;

    int32_t K[MP];
#pragma unroll
    for (int j = 0; j < MP; ++j) K[j] = 0;
    double a0c = 0.0;
#pragma unroll
    for (int j = 0; j < MP; ++j)
      if (crit[j]) a0c = a0c + a[j];
    double tau2 = __builtin_inf();
    uint32_t cw2 = 0, cw3 = 0;
    if (a0c > 0.0) {
      const Words4 bw = key.block(bc, 0);
      bc += 1;
      cnt.w.uniforms += 1;
      cw2 = bw.w2;
      cw3 = bw.w3;
      const double rho1 = u53(bw.w0, bw.w1);
      tau2 = (1.0 / a0c) * log_f64(1.0 / rho1);
    }
    double tau = (double)fminf((float)tau1, (float)tau2);
    if (tau1 > tau2) {
      // SingleCriticalReaction (P1_P2.py:488-507)
      const double rho = u53(cw2, cw3);
      cnt.w.uniforms += 1;
      double rc = 0.0;
      int sel = -1, last = -1;
#pragma unroll
      for (int j = 0; j < MP; ++j) {
        if (crit[j] && sel < 0) {
          if (a[j] > 0.0) last = j;
          rc = rc + a[j] / a0c;
          if (rc >= rho) sel = j;
        }
      }
      if (sel < 0) sel = last;
#pragma unroll
      for (int j = 0; j < MP; ++j)
        if (sel == j) K[j] = 1;
    }
    if (!(tau < __builtin_inf())) {
      fill_rest(tab, P, true, cnt);
      flag = FLAG_TAU_INF;
      return MODE_DONE;
    }
    // Poisson leap with tau halving (P1_P2.py:228-259)
    int64_t xp[NP];
    bool valid = false;
    for (int h = 0; h < MAX_HALVINGS; ++h) {
      const uint64_t grp = bc;
      bc += (uint64_t)M;
      int bad = 0;
      unsigned pending = 0;
      // first draw of every non-critical reaction, from its own block
#pragma unroll
      for (int j = 0; j < MP; ++j) {
        if (j < M && !crit[j]) {
          const double lam = a[j] * tau;
          const Words4 bw = key.block(grp + (uint64_t)j, 0);
          if (!(lam > 0.0)) {
            K[j] = 0;
          } else if (lam >= LAMBDA_LIMIT) {
            bad = FLAG_OVERFLOW;
          } else if (lam < 10.0) {
            cnt.w.uniforms += 1;
            K[j] = (int32_t)inversion_f64(u53(bw.w0, bw.w1), lam, cnt.w);
          } else {
            cnt.w.uniforms += 2;
            cnt.w.ptrs += 1;
            const PtrsF64 Pt = ptrs_setup_f64(lam);
            double kf, us;
            if (ptrs_fast_f64(Pt, u53(bw.w0, bw.w1) - 0.5, u53(bw.w2, bw.w3), kf, us)) K[j] = (int32_t)kf;
            else pending |= 1u << j;
          }
        }
      }
      // squeeze failures: the full PTRS draw, one shared loop
      while (pending != 0u && !bad) {
        const int j = __builtin_ctz(pending);
        pending &= pending - 1u;
        const double lam = pick<MP>(a, j) * tau;
        const int64_t kk = ptrs_full_f64(key, grp + (uint64_t)j, lam, cnt.w);
        if (kk < 0) {
          bad = FLAG_PTRS;
        } else {
#pragma unroll
          for (int r = 0; r < MP; ++r)
            if (r == j) K[r] = (int32_t)kk;
        }
      }
      if (bad) {
        fill_rest(tab, P, true, cnt);
        flag = bad;
        return MODE_DONE;
      }
      // TentativeUpdatedState / ValidState (P1_P2.py:509-540)
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
      tau = tau / 2.0;
    }
    if (!valid) {
      fill_rest(tab, P, true, cnt);
      flag = FLAG_HALVING;
      return MODE_DONE;
    }
    bool big = false;
#pragma unroll
    for (int i = 0; i < NP; ++i) big = big || (xp[i] >= COUNT_LIMIT);
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
    // time advance + SaveInterpolatedDynamics (P1_P2.py:261-286,542-584)
    t = t + tau;
    bool done = false;
    while (t >= tnext) {
      const double t0 = t - tau;
      const float record_time = P.I[F];
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
          Ob[((int64_t)q * P.ita + F) * P.o_stride] = xs + rn_f32_to_int((float)inc);
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
#pragma unroll
    for (int i = 0; i < NP; ++i) x[i] = xn[i];
    return done ? MODE_DONE : MODE_DECIDE;
  }
};

template <int NP, int MP>
__global__ void __launch_bounds__(256) sim_fp64_kernel(const JobDev<NP, MP>* __restrict__ jobs, int n_jobs,
                                                       unsigned int* __restrict__ queue,
                                                       unsigned long long* __restrict__ stats) {
  run_driver<EngineF64<NP, MP>, NP, MP>(jobs, n_jobs, queue, stats);
}

template <int NP, int MP>
static int launch_cls(const void* jobs, int n_jobs, unsigned int* queue, unsigned long long* stats, int waves,
                      hipStream_t stream) {
  const int block = 256;
  const int grid = (waves + 3) / 4;
  if (grid <= 0) return 0;
  hipLaunchKernelGGL((sim_fp64_kernel<NP, MP>), dim3(grid), dim3(block), 0, stream,
                     (const JobDev<NP, MP>*)jobs, n_jobs, queue, stats);
  return (int)hipGetLastError();
}

template <int NP, int MP>
static int occupancy_cls() {
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, sim_fp64_kernel<NP, MP>, 256, 0) != hipSuccess) return 1;
  return blocks > 0 ? blocks : 1;
}

int launch_fp64(int cls, const void* jobs, int n_jobs, unsigned int* queue, unsigned long long* stats, int waves,
                hipStream_t stream) {
  switch (cls) {
    case 0: return launch_cls<4, 4>(jobs, n_jobs, queue, stats, waves, stream);
    case 1: return launch_cls<8, 8>(jobs, n_jobs, queue, stats, waves, stream);
    case 2: return launch_cls<16, 16>(jobs, n_jobs, queue, stats, waves, stream);
    default: return -1;
  }
}

int blocks_per_cu_fp64(int cls) {
  switch (cls) {
    case 0: return occupancy_cls<4, 4>();
    case 1: return occupancy_cls<8, 8>();
    case 2: return occupancy_cls<16, 16>();
    default: return 1;
  }
}

}  // namespace pst
