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
#include "pst_device.hpp"
#include "pst_args.hpp"

namespace pst {

struct Counters {
  unsigned long long events, leaps, rejected, ssa;
  Work w;
};

template <int NP, int MP>
__device__ __forceinline__ void propensities_f64(const TabView<NP, MP>& tab, const double (&k)[MP],
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
  w.inv += (unsigned long long)k;
  return k;
}

enum Mode { MODE_DECIDE = 0, MODE_SSA = 1, MODE_DONE = 2 };

__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0ull; }

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// One simulation per lane.  The loop structure is wavefront-uniform: every
// lane is in one of three modes -- DECIDE (the P1_P2 body: leap decision and,
// if chosen, the leap), SSA (inside a P3 burst of ssa_burst steps) or DONE --
// and each pass of the outer loop runs one DECIDE step for the lanes that
// need it and then a slice of SSA steps for the lanes in a burst.  The
// sequence of operations of every simulation is exactly the reference's
// (decision, burst of SSA steps, decision, ...); only the interleaving of
// different simulations within a wavefront differs.
template <int NP, int MP>
__global__ void __launch_bounds__(256) sim_fp64_kernel(const NetTab<NP, MP> ntab, const RunArgs ra) {
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

  Counters cnt = {};
  int flag = 0;
  double k[MP];
  int32_t x[NP];
#pragma unroll
  for (int j = 0; j < MP; ++j) k[j] = (j < M) ? ra.rates[set * M + j] : 0.0;
#pragma unroll
  for (int i = 0; i < NP; ++i) x[i] = (i < N) ? ra.x0[set * N + i] : 0;

  Key key;
  key.init(ra.seed, ra.sim_offset + (uint64_t)simc);
  uint64_t bc = 0;  // Philox block counter
  double t = 0.0;
  int F = 0;
  double tnext = (double)ra.I[0];
  int mode = !active ? MODE_DONE : (ra.method == 0 ? MODE_DECIDE : MODE_SSA);
  int burst = 0;

  auto save = [&](int f, const int32_t (&xs)[NP]) {
    cnt.w.samples += 1;
#pragma unroll
    for (int q = 0; q < NP; ++q)
      if (q < kappa) Ob[((int64_t)q * ita + f) * stride] = pick<NP>(xs, tab.Ev(q));
  };
  auto fill_rest = [&](const int32_t (&xs)[NP], bool zero) {
    cnt.w.samples += (unsigned long long)(ita - F);
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      if (q < kappa) {
        const int32_t val = zero ? 0 : pick<NP>(xs, tab.Ev(q));
        for (int f = F; f < ita; ++f) Ob[((int64_t)q * ita + f) * stride] = val;
      }
    }
  };

  double a[MP];

  // one direct-method SSA step (P3.py:73-149); returns 1 when the run ended
  // (absorbed or last sample written), 2 on overflow, 0 otherwise
  auto ssa_step = [&]() -> int {
    propensities_f64<NP, MP>(tab, k, x, a);
    double a0 = 0.0;
#pragma unroll
    for (int j = 0; j < MP; ++j) a0 = a0 + a[j];
    if (a0 == 0.0) {
      fill_rest(x, false);
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
      save(F, x);  // SaveDynamics: pre-reaction state (P3.py:165-177)
      F = F + 1;
      if (F == ita) return 1;
      tnext = (double)ra.I[F];
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
      fill_rest(x, true);
      return 2;
    }
    return 0;
  };

  // the P1_P2 body (P1_P2.py:96-294): returns the lane's next mode
  auto decide = [&]() -> int {
    propensities_f64<NP, MP>(tab, k, x, a);
    double a0 = 0.0;
#pragma unroll
    for (int j = 0; j < MP; ++j) a0 = a0 + a[j];
    if (a0 == 0.0) {
      fill_rest(x, false);
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
        if (v < 0) c = c || ((-v) * ra.n_c > x[i]);
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
          const double num = (double)fmaxf((float)(ra.eta * (double)x[i] / g_i), 1.0f);
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
      fill_rest(x, true);
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
            const PtrsF64 P = ptrs_setup_f64(lam);
            double kf, us;
            if (ptrs_fast_f64(P, u53(bw.w0, bw.w1) - 0.5, u53(bw.w2, bw.w3), kf, us)) K[j] = (int32_t)kf;
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
        fill_rest(x, true);
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
      fill_rest(x, true);
      flag = FLAG_HALVING;
      return MODE_DONE;
    }
    bool big = false;
#pragma unroll
    for (int i = 0; i < NP; ++i) big = big || (xp[i] >= COUNT_LIMIT);
    if (big) {
      fill_rest(x, true);
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
      const float record_time = ra.I[F];
      const double gap = t - t0;
      const double nrt = (double)record_time - t0;
      const double frac = nrt / gap;
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        if (q < kappa) {
          const int s = tab.Ev(q);
          const int32_t xs = pick<NP>(x, s);
          const int diff = pick<NP>(xn, s) - xs;
          const double inc = (double)diff * frac;
          Ob[((int64_t)q * ita + F) * stride] = xs + rn_f32_to_int((float)inc);
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
#pragma unroll
    for (int i = 0; i < NP; ++i) x[i] = xn[i];
    return done ? MODE_DONE : MODE_DECIDE;
  };

  const int slice = ra.ssa_slice > 0 ? ra.ssa_slice : 1;
  while (wave_any(mode != MODE_DONE)) {
    tab.refresh();
    if (wave_any(mode == MODE_DECIDE)) {
      if (mode == MODE_DECIDE) {
        mode = decide();
        if (mode == MODE_SSA) burst = ra.ssa_burst;
      }
    }
    // SSA steps: the whole burst when no lane waits for a decision, else a slice
    const int limit = wave_any(mode == MODE_DECIDE) ? slice : 0x7fffffff;
    for (int s = 0; s < limit; ++s) {
      if (!wave_any(mode == MODE_SSA)) break;
      tab.refresh();
      if (mode == MODE_SSA) {
        if (ra.method == 2 && !(t < ra.t_max)) {
          // kernel_Gillespie stops at t >= t_max; later samples stay zero (cuGillespie.py:71)
          fill_rest(x, true);
          mode = MODE_DONE;
        } else {
          const int r = ssa_step();
          if (r != 0) {
            mode = MODE_DONE;
            if (r == 2) flag = FLAG_OVERFLOW;
          } else if (ra.method == 0) {
            burst -= 1;
            if (burst == 0) mode = MODE_DECIDE;
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
    const unsigned long long sum = wave_sum(vals[i]);
    if ((threadIdx.x & 63) == 0) atomicAdd(&ra.stats[i], sum);
  }
}

template <int NP, int MP>
static int launch_class(const HostNet& net, const RunArgs& ra, hipStream_t stream) {
  const NetTab<NP, MP> tab = make_tab<NP, MP>(net);
  const int block = 256;
  const int grid = (int)((ra.n_sims + block - 1) / block);
  if (grid == 0) return 0;
  hipLaunchKernelGGL((sim_fp64_kernel<NP, MP>), dim3(grid), dim3(block), 0, stream, tab, ra);
  return (int)hipGetLastError();
}

int launch_fp64(const HostNet& net, const RunArgs& ra, hipStream_t stream) {
  if (net.N <= 4 && net.M <= 4 && net.kappa <= 4) return launch_class<4, 4>(net, ra, stream);
  if (net.N <= 8 && net.M <= 8 && net.kappa <= 8) return launch_class<8, 8>(net, ra, stream);
  if (net.N <= 16 && net.M <= 16 && net.kappa <= 16) return launch_class<16, 16>(net, ra, stream);
  return -1;
}

}  // namespace pst
