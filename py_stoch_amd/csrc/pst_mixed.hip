// pst_mixed.hip -- mixed FP32/FP64 simulation kernels for gfx950.
//
// Same method as the FP64 kernels (P1_P2.py:54-297, P3.py:29-163,
// gillespie.py:27-149): same epsilon bound, same n_c critical split, same
// SSA fallback burst and sampling rules.  Precision:
//   FP32: propensities, a0 for the leap decision, mu / sigma^2, the Cao tau
//         bound, Poisson means and the Poisson sampler's rejection loop,
//         uniforms (24-bit, one Philox word each)
//   FP64: the simulation clock t, and the SSA reaction choice, done as
//         "first j with sum_{r<=j} a_r >= rho2 * a0" accumulated in FP64 so
//         that FP32 round-off cannot bias which reaction fires.
// Transcendentals use the hardware v_log_f32 / v_exp_f32 / v_rcp_f32.
#include "pst_device.hpp"
#include "pst_args.hpp"

namespace pst {

struct CountersM {
  unsigned long long events, leaps, rejected, ssa;
};

template <int NP, int MP>
__device__ __forceinline__ void propensities_f32(const NetTab<NP, MP>& tab, const float (&k)[MP],
                                                 const int32_t (&x)[NP], float (&a)[MP]) {
#pragma unroll
  for (int j = 0; j < MP; ++j) {
    float v = k[j];
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const int p = tab.ord[i * MP + j];
      const int st = tab.stp[i * MP + j];
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

template <int NP>
__device__ __forceinline__ int32_t pick_m(const int32_t (&x)[NP], int idx) {
  int32_t r = x[0];
#pragma unroll
  for (int i = 1; i < NP; ++i)
    if (idx == i) r = x[i];
  return r;
}

template <int NP, int MP>
__device__ int run_sim_mixed(const NetTab<NP, MP>& tab, const RunArgs& ra, int64_t sim, CountersM& cnt) {
  const int M = ra.M, N = ra.N;
  const int kappa = ra.kappa, ita = ra.ita;
  const int64_t stride = ra.o_stride;
  int32_t* Ob = ra.O + sim;
  const int64_t set = sim / ra.sims_per_set;
  const float eta = (float)ra.eta;

  float k[MP];
  int32_t x[NP];
#pragma unroll
  for (int j = 0; j < MP; ++j) k[j] = (j < M) ? (float)ra.rates[set * M + j] : 0.0f;
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
      if (q < kappa) Ob[((int64_t)q * ita + f) * stride] = pick_m<NP>(xs, tab.E[q]);
  };
  auto fill_rest = [&](const int32_t (&xs)[NP], bool zero) {
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      if (q < kappa) {
        const int32_t val = zero ? 0 : pick_m<NP>(xs, tab.E[q]);
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
    const float rho1 = g.uniform_f32();
    const float tau = -fast_log(rho1) * fast_rcp((float)a0d);
    const double target = (double)g.uniform_f32() * a0d;
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
    bool o = false;
#pragma unroll
    for (int i = 0; i < NP; ++i) o = o || (x[i] >= COUNT_LIMIT);
    if (o) {
      fill_rest(x, true);
      return 2;
    }
    return 0;
  };

  if (ra.method == 2) {
    while (t < ra.t_max) {
      const int r = ssa_step();
      if (r == 1) return 0;
      if (r == 2) return FLAG_OVERFLOW;
    }
    fill_rest(x, true);
    return 0;
  }

  for (;;) {
    bool ssa = true;
    if (ra.method == 0) {
      propensities_f32<NP, MP>(tab, k, x, a);
      float a0 = 0.0f;
#pragma unroll
      for (int j = 0; j < MP; ++j) a0 += a[j];
      if (a0 == 0.0f) {
        fill_rest(x, false);
        return 0;
      }
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
      float tau1 = __builtin_inff();
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        if (tab.irs[i]) {
          float mu = 0.0f, s2 = 0.0f;
          bool any = false;
#pragma unroll
          for (int j = 0; j < MP; ++j) {
            const int v = tab.v[i * MP + j];
            if (v != 0 && !crit[j]) {
              mu += (float)v * a[j];
              s2 += (float)(v * v) * a[j];
              any = true;
            }
          }
          if (any) {
            const float g_i = calc_g_f32(tab.H[i], tab.Ht[i], x[i]);
            const float num = fmaxf(eta * (float)x[i] * fast_rcp(g_i), 1.0f);
            const float lhs = num * fast_rcp(fabsf(mu));
            const float rhs = num * num * fast_rcp(s2);
            tau1 = fminf(tau1, fminf(lhs, rhs));
          }
        }
      }
      if (!(tau1 < 10.0f * fast_rcp(a0))) {
        ssa = false;
        int32_t K[MP];
#pragma unroll
        for (int j = 0; j < MP; ++j) K[j] = 0;
        float a0c = 0.0f;
#pragma unroll
        for (int j = 0; j < MP; ++j)
          if (crit[j]) a0c += a[j];
        float tau2 = __builtin_inff();
        if (a0c > 0.0f) tau2 = -fast_log(g.uniform_f32()) * fast_rcp(a0c);
        float tau = fminf(tau1, tau2);
        if (tau1 > tau2) {
          const double target = (double)g.uniform_f32() * (double)a0c;
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
          return FLAG_TAU_INF;
        }
        int64_t xp[NP];
        bool valid = false;
        for (int h = 0; h < MAX_HALVINGS; ++h) {
          int bad = 0;
#pragma unroll
          for (int j = 0; j < MP; ++j) {
            if (!crit[j] && !bad) {
              const float lam = a[j] * tau;
              if (lam >= (float)LAMBDA_LIMIT) {
                bad = FLAG_OVERFLOW;
              } else {
                const int64_t kk = poisson_mixed(g, lam);
                if (kk < 0) bad = FLAG_PTRS;
                else K[j] = (int32_t)kk;
              }
            }
          }
          if (bad) {
            fill_rest(x, true);
            return bad;
          }
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
          tau = tau * 0.5f;
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
              const int s = tab.E[q];
              const int32_t xs = pick_m<NP>(x, s);
              const int diff = pick_m<NP>(xn, s) - xs;
              Ob[((int64_t)q * ita + F) * stride] = xs + rn_f32_to_int((float)diff * frac);
            }
          }
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
        if (done) return 0;
      }
    }
    if (ssa) {
      for (int c = 0; c < ra.ssa_burst; ++c) {
        const int r = ssa_step();
        if (r == 1) return 0;
        if (r == 2) return FLAG_OVERFLOW;
      }
    }
  }
}

__device__ __forceinline__ unsigned long long wave_sum_m(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int NP, int MP>
__global__ void __launch_bounds__(256) sim_mixed_kernel(const NetTab<NP, MP> tab, const RunArgs ra) {
  const int64_t sim = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  CountersM cnt = {0, 0, 0, 0};
  unsigned long long flagged = 0;
  if (sim < ra.n_sims) {
    const int f = run_sim_mixed<NP, MP>(tab, ra, sim, cnt);
    if (ra.flags) ra.flags[sim] = f;
    flagged = (f != 0);
  }
  const unsigned long long e = wave_sum_m(cnt.events);
  const unsigned long long l = wave_sum_m(cnt.leaps);
  const unsigned long long r = wave_sum_m(cnt.rejected);
  const unsigned long long s = wave_sum_m(cnt.ssa);
  const unsigned long long fl = wave_sum_m(flagged);
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&ra.stats[ST_EVENTS], e);
    atomicAdd(&ra.stats[ST_LEAPS], l);
    atomicAdd(&ra.stats[ST_REJECTED], r);
    atomicAdd(&ra.stats[ST_SSA], s);
    atomicAdd(&ra.stats[ST_FLAGGED], fl);
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
