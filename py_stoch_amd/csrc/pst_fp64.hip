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
#include "pst_device.hpp"
#include "pst_args.hpp"

namespace pst {

struct Counters {
  unsigned long long events, leaps, rejected, ssa;
};

template <int NP, int MP>
__device__ __forceinline__ void propensities_f64(const NetTab<NP, MP>& tab, const double (&k)[MP],
                                                 const int32_t (&x)[NP], double (&a)[MP]) {
  // UpdatePropensities (parser.py:32-55) in table form, species ascending
#pragma unroll
  for (int j = 0; j < MP; ++j) {
    double v = k[j];
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const int p = tab.ord[i * MP + j];
      const int st = tab.stp[i * MP + j];
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

template <int NP>
__device__ __forceinline__ int32_t pick(const int32_t (&x)[NP], int idx) {
  int32_t r = x[0];
#pragma unroll
  for (int i = 1; i < NP; ++i)
    if (idx == i) r = x[i];
  return r;
}

template <int NP, int MP>
__device__ int run_sim_f64(const NetTab<NP, MP>& tab, const RunArgs& ra, int64_t sim, Counters& cnt) {
  const int M = ra.M, N = ra.N;
  const int kappa = ra.kappa, ita = ra.ita;
  const int64_t stride = ra.o_stride;
  int32_t* Ob = ra.O + sim;
  const int64_t set = sim / ra.sims_per_set;

  double k[MP];
  int32_t x[NP];
#pragma unroll
  for (int j = 0; j < MP; ++j) k[j] = (j < M) ? ra.rates[set * M + j] : 0.0;
#pragma unroll
  for (int i = 0; i < NP; ++i) x[i] = (i < N) ? ra.x0[set * N + i] : 0;

  Stream g;
  g.init(ra.seed, ra.sim_offset + (uint64_t)sim);
  double t = 0.0;
  int F = 0;
  double tnext = (double)ra.I[0];

  auto save = [&](int f, const int32_t (&xs)[NP]) {
#pragma unroll
    for (int q = 0; q < NP; ++q)
      if (q < kappa) Ob[((int64_t)q * ita + f) * stride] = pick<NP>(xs, tab.E[q]);
  };
  auto fill_rest = [&](const int32_t (&xs)[NP], bool zero) {
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      if (q < kappa) {
        const int32_t val = zero ? 0 : pick<NP>(xs, tab.E[q]);
        for (int f = F; f < ita; ++f) Ob[((int64_t)q * ita + f) * stride] = val;
      }
    }
  };
  auto over_limit = [&](const int32_t (&xs)[NP]) {
    bool o = false;
#pragma unroll
    for (int i = 0; i < NP; ++i) o = o || (xs[i] >= COUNT_LIMIT);
    return o;
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
    const double rho1 = g.uniform_f64();
    const double tau = (1.0 / a0) * log_f64(1.0 / rho1);
    // SingleCriticalReaction (P3.py:179-194): first j with sum_{r<=j} a_r/a0 >= rho;
    // the reference falls off the end on round-off: take the last j with a_j > 0
    const double rho = g.uniform_f64();
    double rc = 0.0;
    int sel = -1, last = -1;
#pragma unroll
    for (int j = 0; j < MP; ++j) {
      if (j < M) {
        if (sel < 0) {
          if (a[j] > 0.0) last = j;
          rc = rc + a[j] / a0;
          if (rc >= rho) sel = j;
        }
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
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      int dx = 0;
#pragma unroll
      for (int j = 0; j < MP; ++j)
        if (sel == j) dx = tab.v[i * MP + j];
      x[i] += dx;
    }
    cnt.ssa += 1;
    cnt.events += 1;
    if (over_limit(x)) {
      fill_rest(x, true);
      return 2;
    }
    return 0;
  };

  if (ra.method == 2) {
    // kernel_Gillespie (gillespie.py:70-146)
    while (t < ra.t_max) {
      const int r = ssa_step();
      if (r == 1) return 0;
      if (r == 2) return FLAG_OVERFLOW;
    }
    fill_rest(x, true);  // samples past the last crossing stay zero (cuGillespie.py:71)
    return 0;
  }

  for (;;) {
    bool ssa = true;
    if (ra.method == 0) {
      // ---- kernel_P1_P2 (P1_P2.py:96-294) ----
      propensities_f64<NP, MP>(tab, k, x, a);
      double a0 = 0.0;
#pragma unroll
      for (int j = 0; j < MP; ++j) a0 = a0 + a[j];
      if (a0 == 0.0) {
        fill_rest(x, false);
        return 0;
      }
      // DetermineCriticalReactions on V (P1_P2.py:314-327)
      bool crit[MP];
#pragma unroll
      for (int j = 0; j < MP; ++j) {
        bool c = false;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
          const int v = tab.v[i * MP + j];
          if (v < 0) c = c || ((-v) * ra.n_c > x[i]);
        }
        crit[j] = c;
      }
      // CalculateMuSigma (P1_P2.py:330-383) and CalculateTau (P1_P2.py:385-445)
      double tau1 = __builtin_inf();
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        if (tab.irs[i]) {
          double mu = 0.0, s2 = 0.0;
          bool any = false;
#pragma unroll
          for (int j = 0; j < MP; ++j) {
            const int v = tab.v[i * MP + j];
            if (v != 0 && !crit[j]) {
              mu = mu + ((double)v * a[j]);
              s2 = s2 + ((double)(v * v) * a[j]);
              any = true;
            }
          }
          if (any) {
            // one candidate per V entry of this species; all equal, so one suffices
            const double g_i = calc_g_f64(tab.H[i], tab.Ht[i], x[i]);
            const double num = (double)fmaxf((float)(ra.eta * (double)x[i] / g_i), 1.0f);
            const double lhs = num / (double)fabsf((float)mu);
            const double rhs = (num * num) / s2;
            const double tt = (double)fminf((float)lhs, (float)rhs);
            if (tt < tau1) tau1 = tt;
          }
        }
      }
      // SSA switch (P1_P2.py:157-173)
      if (!(tau1 < 10.0 / a0)) {
        ssa = false;
        int32_t K[MP];
#pragma unroll
        for (int j = 0; j < MP; ++j) K[j] = 0;
        double a0c = 0.0;
#pragma unroll
        for (int j = 0; j < MP; ++j)
          if (crit[j]) a0c = a0c + a[j];
        double tau2 = __builtin_inf();
        if (a0c > 0.0) {
          const double rho1 = g.uniform_f64();
          tau2 = (1.0 / a0c) * log_f64(1.0 / rho1);
        }
        double tau = (double)fminf((float)tau1, (float)tau2);
        if (tau1 > tau2) {
          // SingleCriticalReaction (P1_P2.py:488-507)
          const double rho = g.uniform_f64();
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
          return FLAG_TAU_INF;
        }
        // Poisson leap with tau halving (P1_P2.py:228-259)
        int64_t xp[NP];
        bool valid = false;
        for (int h = 0; h < MAX_HALVINGS; ++h) {
          int bad = 0;
#pragma unroll
          for (int j = 0; j < MP; ++j) {
            if (!crit[j] && !bad) {
              const double lam = a[j] * tau;
              if (lam >= LAMBDA_LIMIT) {
                bad = FLAG_OVERFLOW;
              } else {
                const int64_t kk = poisson_f64(g, lam);
                if (kk < 0) bad = FLAG_PTRS;
                else K[j] = (int32_t)kk;
              }
            }
          }
          if (bad) {
            fill_rest(x, true);
            return bad;
          }
          // TentativeUpdatedState / ValidState (P1_P2.py:509-540)
          bool ok = true;
#pragma unroll
          for (int i = 0; i < NP; ++i) {
            int64_t s = x[i];
#pragma unroll
            for (int j = 0; j < MP; ++j) {
              const int v = tab.v[i * MP + j];
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
          return FLAG_HALVING;
        }
        bool big = false;
#pragma unroll
        for (int i = 0; i < NP; ++i) big = big || (xp[i] >= COUNT_LIMIT);
        if (big) {
          fill_rest(x, true);
          return FLAG_OVERFLOW;
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
            const int s = tab.E[q];
            const int32_t xs = pick<NP>(x, s);
            const int diff = pick<NP>(xn, s) - xs;
            const double inc = (double)diff * frac;
            Ob[((int64_t)q * ita + F) * stride] = xs + rn_f32_to_int((float)inc);
            }
          }
          F = F + 1;
          if (F == ita) {
            done = true;
            break;
          }
          tnext = (double)ra.I[F];
        }
#pragma unroll
        for (int i = 0; i < NP; ++i) x[i] = xn[i];
        if (done) return 0;
      }
    }
    if (ssa) {
      // ---- kernel_P3 burst (P3.py:71-151) ----
      for (int c = 0; c < ra.ssa_burst; ++c) {
        const int r = ssa_step();
        if (r == 1) return 0;
        if (r == 2) return FLAG_OVERFLOW;
      }
    }
  }
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int NP, int MP>
__global__ void __launch_bounds__(256) sim_fp64_kernel(const NetTab<NP, MP> tab, const RunArgs ra) {
  const int64_t sim = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  Counters cnt = {0, 0, 0, 0};
  unsigned long long flagged = 0;
  if (sim < ra.n_sims) {
    const int f = run_sim_f64<NP, MP>(tab, ra, sim, cnt);
    if (ra.flags) ra.flags[sim] = f;
    flagged = (f != 0);
  }
  // all 64 lanes are reconverged here: reduce the counters per wavefront
  const unsigned long long e = wave_sum(cnt.events);
  const unsigned long long l = wave_sum(cnt.leaps);
  const unsigned long long r = wave_sum(cnt.rejected);
  const unsigned long long s = wave_sum(cnt.ssa);
  const unsigned long long fl = wave_sum(flagged);
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&ra.stats[ST_EVENTS], e);
    atomicAdd(&ra.stats[ST_LEAPS], l);
    atomicAdd(&ra.stats[ST_REJECTED], r);
    atomicAdd(&ra.stats[ST_SSA], s);
    atomicAdd(&ra.stats[ST_FLAGGED], fl);
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
