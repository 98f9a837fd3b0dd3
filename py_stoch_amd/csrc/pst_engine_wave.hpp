// pst_engine_wave.hpp -- mixed-precision engine for large networks: one
// wavefront per simulation.
//
// Networks up to 64 species and 256 reactions (random mass-action networks
// of 5-50 species and 10-200 reactions, for example) do not fit the
// one-simulation-per-lane kernels: their state, propensities and tables would
// not stay in registers.  Here the 64 lanes of a wavefront share one
// simulation: lane l owns reactions l, l + 64, l + 128, l + 192 and species
// l.  The method is the same as in pst_engine_f32.hpp (P1_P2.py:96-294,
// P3.py:73-149, gillespie.py:70-149) with the sums over reactions and species
// done as wavefront reductions:
//   propensities         per reaction lane, counts read from LDS
//   a0, a0c              wavefront sums (FP32 for the leap decision, FP64 for
//                        the reaction choice)
//   critical reactions   per reaction lane (P1_P2.py:314-327)
//   mu, sigma^2, tau1    per species lane over its column of V (transposed
//                        table, fixed order), then a wavefront minimum
//   reaction choice      FP64 prefix sum over reactions in index order
//   Poisson draws        every reaction lane draws its own K_j from Philox
//                        block (group + j, it) -- all draws of a leap in
//                        parallel
//   x' = x + V K         64-bit integer LDS atomics (order-independent)
// State of the simulation (time, sample cursor, Philox counter, mode) is
// wavefront-uniform.  Per wavefront LDS: x, x', propensities of non-critical
// reactions and the critical flags (about 2 KB).
#pragma once
#include "pst_engine_f32.hpp"

namespace pst {

constexpr int WAVE_MAX_N = 64;
constexpr int WAVE_MAX_M = 256;
constexpr int WAVE_R = WAVE_MAX_M / 64;  // reactions per lane
constexpr int WAVE_MAX_REACT = 4;        // species with a propensity factor, per reaction
constexpr int WAVE_MAX_V = 8;            // non-zero V entries per reaction
constexpr int WAVE_MAX_T = 2048;         // non-zero V entries in total

// Network of one job in sparse form (built by the host library).
struct WaveTab {
  int32_t N, M, kappa, nnz;
  uint8_t nr[WAVE_MAX_M];                  // propensity factors of reaction j
  uint8_t rs[WAVE_MAX_M][WAVE_MAX_REACT];  // their species
  uint8_t ro[WAVE_MAX_M][WAVE_MAX_REACT];  // order p (factors of the species)
  uint8_t rst[WAVE_MAX_M][WAVE_MAX_REACT]; // step: 1 falling factorial, 0 power
  uint8_t nv[WAVE_MAX_M];                  // V entries of reaction j
  uint8_t vs[WAVE_MAX_M][WAVE_MAX_V];      // their species
  int8_t vv[WAVE_MAX_M][WAVE_MAX_V];       // their values
  int8_t H[WAVE_MAX_N], Ht[WAVE_MAX_N], irs[WAVE_MAX_N], E[WAVE_MAX_N];
  uint16_t tstart[WAVE_MAX_N + 1];         // column of V of species i: entries tstart[i] .. tstart[i+1]-1
  uint16_t tj[WAVE_MAX_T];                 // reaction of the entry
  int8_t tv[WAVE_MAX_T];                   // value of the entry
};

struct WaveJob {
  JobParams p;
  WaveTab t;
};

struct WaveLds {
  long long xn[WAVE_MAX_N];  // tentative state (x + V K)
  int32_t x[WAVE_MAX_N];     // current state
  float anc[WAVE_MAX_M];     // propensity of non-critical reactions, 0 for critical ones
  int32_t K[WAVE_MAX_M];     // firings of the current leap attempt
  uint8_t crit[WAVE_MAX_M];
};

// LDS writes of some lanes visible to reads of others in the same wavefront
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ float wave_min_f(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ long long wave_sum_ll(long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
// inclusive prefix sum over lanes
__device__ __forceinline__ double wave_scan_d(double v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  return v;
}

// One simulation per wavefront, lanes over reactions and species.  All
// members are wavefront-uniform except the per-lane rate constants.
struct EngineWave {
  float k[WAVE_R];  // rates of the lane's reactions

  Key key;
  uint64_t bc;
  double t, tnext;
  int F, flag, burst, mode;
  uint32_t nsteps;  // leap decisions + SSA steps so far (step budget)
  int64_t sim;

  // index of sample f of recorded species q in O (rows or blocks layout,
  // pst_driver.hpp; stored sample by sample by the species lane)
  __device__ __forceinline__ int64_t out_index(const JobParams& P, int q, int f) const {
    if (P.layout != LAYOUT_BLOCKS) return ((int64_t)q * P.ita + f) * P.o_stride + sim;
    return (((int64_t)q * ((P.ita + OUT_BLOCK - 1) / OUT_BLOCK) + f / OUT_BLOCK) * P.o_stride + sim) * OUT_BLOCK +
           (f & (OUT_BLOCK - 1));
  }

  // blocks layout: zeros after the last sample in its block
  __device__ __forceinline__ void pad_last_block(const JobParams& P, int q) const {
    if (P.layout != LAYOUT_BLOCKS) return;
    const int end = (P.ita + OUT_BLOCK - 1) / OUT_BLOCK * OUT_BLOCK;
    for (int f = P.ita; f < end; ++f) glb(P.O)[out_index(P, q, f)] = 0;
  }

  // write samples F..ita-1: x of the recorded species, or zeros
  __device__ __forceinline__ void fill_rest(const WaveJob& J, const WaveLds& L, int lane, bool zero, Counters& cnt) {
    const JobParams& P = J.p;
    if (lane == 0) cnt.w.samples += (unsigned int)(P.ita - F);
    if (lane < P.kappa) {
      const int32_t val = zero ? 0 : L.x[J.t.E[lane]];
      for (int f = F; f < P.ita; ++f) glb(P.O)[out_index(P, lane, f)] = val;
      pad_last_block(P, lane);
    }
  }

  __device__ __forceinline__ void start(const WaveJob& J, WaveLds& L, int lane, int64_t s) {
    const JobParams& P = J.p;
    const int64_t set = s / P.sims_per_set;
    sim = s;
#pragma unroll
    for (int r = 0; r < WAVE_R; ++r) {
      const int j = lane + 64 * r;
      k[r] = (j < P.M) ? (float)glb(P.rates)[set * P.M + j] : 0.0f;
    }
    if (lane < P.N) L.x[lane] = glb(P.x0)[set * P.N + lane];
    wave_lds_sync();
    key.init(P.seed, P.sim_offset + (uint64_t)s);
    bc = 0;
    t = 0.0;
    F = 0;
    tnext = (double)glb(P.I)[0];
    flag = 0;
    burst = 0;
    nsteps = 0;
    mode = P.method == 0 ? MODE_DECIDE : MODE_SSA;
  }

  // one step (leap decision + leap, or one SSA step); returns true when the
  // simulation has ended
  __device__ bool step(const WaveJob& J, WaveLds& L, int lane, Counters& cnt) {
    const JobParams& P = J.p;
    const WaveTab& T = J.t;
    const int M = P.M, N = P.N;
    bool ssa = mode == MODE_SSA;
    if (ssa && P.method == 2 && !(t < P.t_max)) {
      fill_rest(J, L, lane, true, cnt);
      return true;
    }
    if (P.max_steps != 0u && nsteps >= P.max_steps) {
      fill_rest(J, L, lane, true, cnt);
      flag = FLAG_STEPS;
      return true;
    }
    nsteps += 1;
    // propensities of the lane's reactions
    float a[WAVE_R];
#pragma unroll
    for (int r = 0; r < WAVE_R; ++r) {
      const int j = lane + 64 * r;
      float v = 0.0f;
      if (64 * r < M && j < M) {  // whole slots past M are skipped by the wavefront
        v = k[r];
        const int nr = T.nr[j];
        for (int e = 0; e < nr; ++e) {
          const int xi = L.x[T.rs[j][e]];
          const int p = T.ro[j][e], st = T.rst[j][e];
          for (int q = 0; q < p; ++q) v = v * (float)(xi - st * q);
        }
      }
      a[r] = v;
    }
    float a_lane = 0.0f;
    double ad_lane = 0.0;
#pragma unroll
    for (int r = 0; r < WAVE_R; ++r) {
      a_lane += a[r];
      ad_lane += (double)a[r];
    }
    const float a0 = wave_sum_f(a_lane);
    if (a0 == 0.0f) {
      fill_rest(J, L, lane, false, cnt);
      return true;
    }
    bool crit[WAVE_R];
#pragma unroll
    for (int r = 0; r < WAVE_R; ++r) crit[r] = false;
    float tau1 = __builtin_inff();
    if (!ssa) {
      if (lane == 0) cnt.w.decisions += 1;
      // critical reactions on V (P1_P2.py:314-327)
#pragma unroll
      for (int r = 0; r < WAVE_R; ++r) {
        const int j = lane + 64 * r;
        bool c = false;
        if (64 * r < M && j < M) {
          const int nv = T.nv[j];
          for (int e = 0; e < nv; ++e) {
            const int v = T.vv[j][e];
            if (v < 0) c = c || ((-v) * P.n_c > L.x[T.vs[j][e]]);
          }
          L.anc[j] = c ? 0.0f : a[r];
          L.crit[j] = c ? 1 : 0;
        }
        crit[r] = c;
      }
      wave_lds_sync();
      // mu, sigma^2 and the Cao bound per species (P1_P2.py:330-445)
      float cand = __builtin_inff();
      if (lane < N && T.irs[lane]) {
        float mu = 0.0f, s2 = 0.0f;
        bool any = false;
        const int e1 = T.tstart[lane + 1];
        for (int e = T.tstart[lane]; e < e1; ++e) {
          const int j = T.tj[e];
          const float v = (float)T.tv[e];
          const float aj = L.anc[j];
          mu += v * aj;
          s2 += v * v * aj;
          any = any || (L.crit[j] == 0);
        }
        if (any) {
          const int xi = L.x[lane];
          const float g_i = calc_g_f32(T.H[lane], T.Ht[lane], xi);
          const float num = fmaxf((float)P.eta * (float)xi * fast_rcp(g_i), 1.0f);
          cand = fminf(num * fast_rcp(fabsf(mu)), num * num * fast_rcp(s2));
        }
      }
      tau1 = wave_min_f(cand);
      if (tau1 < 10.0f * fast_rcp(a0)) {
        ssa = true;
        mode = MODE_SSA;
        burst = P.ssa_burst;
        if (P.max_steps != 0u && nsteps >= P.max_steps) {
          fill_rest(J, L, lane, true, cnt);
          flag = FLAG_STEPS;
          return true;
        }
        nsteps += 1;
      }
    }
    // one-block draw: SSA step, or tau2 and the critical reaction
    double asum;
    if (ssa) {
      asum = wave_sum_d(ad_lane);
    } else {
      double c_lane = 0.0;
#pragma unroll
      for (int r = 0; r < WAVE_R; ++r) c_lane += crit[r] ? (double)a[r] : 0.0;
      asum = wave_sum_d(c_lane);
    }
    float tauX = __builtin_inff();
    uint32_t cw1 = 0;
    if (asum > 0.0) {
      const Words4 bw = key.block(bc, 0);
      bc += 1;
      if (lane == 0) cnt.w.uniforms += ssa ? 2 : 1;
      cw1 = bw.w1;
      tauX = -fast_log(u24(bw.w0)) * fast_rcp((float)asum);
    }
    const float tau_f = ssa ? tauX : fminf(tau1, tauX);
    int sel = -1;
    if (ssa || tau1 > tauX) {
      if (!ssa && lane == 0) cnt.w.uniforms += 1;
      // first reaction of the mask with sum_{r<=j} a_r >= rho2 * asum, in index order (FP64)
      const double target = (double)u24(cw1) * asum;
      double carry = 0.0;
      int last = -1;
#pragma unroll
      for (int r = 0; r < WAVE_R; ++r) {
        if (64 * r >= M) break;  // wavefront-uniform
        const int j = lane + 64 * r;
        const bool m = (j < M) && (ssa || crit[r]);
        const double v = m ? (double)a[r] : 0.0;
        const double cum = wave_scan_d(v, lane) + carry;
        carry = __shfl(cum, 63, 64);
        const unsigned long long hit = __builtin_amdgcn_ballot_w64(m && cum >= target);
        const unsigned long long pos = __builtin_amdgcn_ballot_w64(m && a[r] > 0.0f);
        if (sel < 0 && hit != 0ull) sel = 64 * r + __builtin_ctzll(hit);
        if (pos != 0ull) last = 64 * r + 63 - __builtin_clzll(pos);
      }
      if (sel < 0) sel = last;
    }
    // tentative state in LDS
    if (lane < N) L.xn[lane] = L.x[lane];
    wave_lds_sync();
    float tau = tau_f;
    if (ssa) {
      if (sel >= 0 && lane < T.nv[sel]) {
        const int s = T.vs[sel][lane];
        L.xn[s] = L.xn[s] + T.vv[sel][lane];  // distinct species per entry
      }
      wave_lds_sync();
    } else {
      if (!(tau < __builtin_inff())) {
        fill_rest(J, L, lane, true, cnt);
        flag = FLAG_TAU_INF;
        return true;
      }
      long long ev = 0;
      const int st = leap(J, L, lane, cnt, sel, tau, ev);
      if (st != 0) {
        fill_rest(J, L, lane, true, cnt);
        flag = st;
        return true;
      }
      if (lane == 0) {
        cnt.events += (unsigned long long)ev;
        cnt.leaps += 1;
      }
    }
    // time advance and samples (sample-and-hold for SSA, interpolation for a leap)
    const double t0 = t;
    t = t + (double)tau;
    if (t == t0 && tau > 0.0f) {  // tau == 0 (a draw of u = 1) is a valid event
      fill_rest(J, L, lane, true, cnt);
      flag = FLAG_STALL;
      return true;
    }
    if (t >= tnext) {
      const float inv_gap = fast_rcp((float)(t - t0));
      while (t >= tnext) {
        const float frac = ssa ? 0.0f : (float)(tnext - t0) * inv_gap;
        if (lane < P.kappa) {
          const int s = T.E[lane];
          const int32_t xs = L.x[s];
          const int diff = (int)(L.xn[s] - (long long)xs);
          glb(P.O)[out_index(P, lane, F)] = xs + rn_f32_to_int((float)diff * frac);
          if (F == P.ita - 1) pad_last_block(P, lane);
        }
        if (lane == 0) cnt.w.samples += 1;
        F = F + 1;
        if (F == P.ita) return true;  // an SSA step ending the run is not applied
        tnext = (double)glb(P.I)[F];
      }
    }
    const bool over = __builtin_amdgcn_ballot_w64(lane < N && L.xn[lane] >= COUNT_LIMIT) != 0ull;
    if (lane < N) L.x[lane] = (int32_t)L.xn[lane];
    wave_lds_sync();
    if (ssa) {
      if (lane == 0) {
        cnt.ssa += 1;
        cnt.events += 1;
      }
      if (over) {
        fill_rest(J, L, lane, true, cnt);
        flag = FLAG_OVERFLOW;
        return true;
      }
      if (P.method == 0) {
        burst -= 1;
        if (burst <= 0) mode = MODE_DECIDE;
      }
    }
    return false;
  }

  // Poisson leap with halving; x' left in L.xn, tau halved in place (the
  // clock advances by the tau of the accepted draw, P1_P2.py:258-261).
  // Returns 0 or a flag.  The
  // lane's reaction slots are walked in a rolled loop with the counts in LDS,
  // so the sampler's code (and its registers) exists once.
  __device__ int leap(const WaveJob& J, WaveLds& L, int lane, Counters& cnt, int sel, float& tau, long long& ev) {
    const JobParams& P = J.p;
    const WaveTab& T = J.t;
    const int M = P.M, N = P.N;
    for (int h = 0; h < MAX_HALVINGS; ++h) {
      const uint64_t grp = bc;
      bc += (uint64_t)M;
      bool bad = false, fail = false;
#pragma unroll 1
      for (int r = 0; r < WAVE_R; ++r) {
        const int j = lane + 64 * r;
        if (j >= M) continue;
        int32_t kk = (j == sel) ? 1 : 0;
        const float l = L.anc[j] * tau;  // 0 for critical reactions
        bad = bad || (l >= (float)LAMBDA_LIMIT);
        if (l > 0.0f && !bad) {
          const Words4 bw = key.block(grp + (uint64_t)j, 0);
          if (l < POIS_ASYM_MIN) {
            // sequential-search inversion
            cnt.w.uniforms += 1;
            const float u = u24(bw.w0);
            float p = fast_exp(-l), Fc = p;
            int n = 0;
            while (u > Fc && n < INV_MAX_K) {
              n += 1;
              p = p * l * fast_rcp((float)n);
              const float Fn = Fc + p;
              if (!(Fn > Fc)) break;  // the FP32 sum stopped growing (far tail; u may be 1.0)
              Fc = Fn;
            }
            cnt.w.inv += (unsigned int)n;
            kk = n;
          } else if (l < POIS_ASYM_MAX) {
            // asymptotic inversion, decided in FP64 near an integer boundary
            cnt.w.asym += 1;
            cnt.w.uniforms += 1;
            const float u = u24(bw.w0);
            bool ex = false;
            int32_t m = -1;
            kk = poisson_asym_f32(l, u, ex, m);
            if (ex) {
              cnt.w.exact += 1;
              kk = poisson_exact_f64((double)l, (double)u, m);
            }
          } else {
            const Draw d = ptrs_wide(key.k0, key.k1, key.s_lo, key.s_hi, grp + (uint64_t)j, 0, l);
            cnt.w.ptrs += d.iters;
            cnt.w.uniforms += 2 * d.iters;
            if (d.k < 0) fail = true;
            else kk = (int32_t)d.k;
          }
        }
        L.K[j] = kk;
      }
      if (__builtin_amdgcn_ballot_w64(bad) != 0ull) return FLAG_OVERFLOW;
      if (__builtin_amdgcn_ballot_w64(fail) != 0ull) return FLAG_PTRS;
      // x' = x + V K (TentativeUpdatedState, P1_P2.py:509-524)
      if (lane < N) L.xn[lane] = L.x[lane];
      wave_lds_sync();
      long long ev_lane = 0;
#pragma unroll
      for (int r = 0; r < WAVE_R; ++r) {
        if (64 * r >= M) break;  // wavefront-uniform
        const int j = lane + 64 * r;
        if (j >= M) continue;
        const int32_t kk = L.K[j];
        if (kk != 0) {
          ev_lane += kk;
          const int nv = T.nv[j];
          for (int e = 0; e < nv; ++e)
            atomicAdd((unsigned long long*)&L.xn[T.vs[j][e]], (unsigned long long)((long long)kk * T.vv[j][e]));
        }
      }
      wave_lds_sync();
      // ValidState (P1_P2.py:526-540)
      const bool neg = lane < N && L.xn[lane] < 0;
      if (__builtin_amdgcn_ballot_w64(neg) == 0ull) {
        const bool big = lane < N && L.xn[lane] >= COUNT_LIMIT;
        if (__builtin_amdgcn_ballot_w64(big) != 0ull) return FLAG_OVERFLOW;
        ev = wave_sum_ll(ev_lane);
        return 0;
      }
      if (lane == 0) cnt.rejected += 1;
      tau = tau * 0.5f;
    }
    return FLAG_HALVING;
  }
};

}  // namespace pst
