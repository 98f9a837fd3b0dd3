/*
 * pst_oracle.c -- CPU oracle for the simulator hot path (TEST INFRASTRUCTURE).
 *
 * This file is a plain-C restatement of py_stoch's GPU simulation algorithm:
 *   - the tau-leaping step           mod/simulator/cuTauLeaping/P1_P2.py:54-584
 *   - the 100-step SSA burst         mod/simulator/cuTauLeaping/P3.py:29-208
 *   - the host iteration loop        mod/simulator/cuTauLeaping/cuTauLeaping.py:48-84
 *   - the single-launch pure SSA     mod/simulator/cuGillespie/gillespie.py:27-193
 * with cuRAND (closed source) replaced by a counter-based Philox4x32-10 stream
 * and curand_poisson replaced by Devroye inversion / Hoermann PTRS.
 *
 * It is the checker for the HIP kernels, never the product: only tests/, the
 * smoke check and bench.py's cpu_baseline leg may load it.  It follows the
 * reference line by line (sparse species-major V table, reference summation
 * orders, the reference's FP32 rounding points) so that the FP64 GPU path can
 * be compared with it bit for bit on the same Philox stream.
 *
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off -fopenmp, no fast-math).
 *
 * Pinning (tests/test_oracle_*.py, tests/test_golden.py): Philox known-answer
 * vectors (Random123); log/exp against numpy to <= 2 ulp; the Poisson sampler
 * against the scipy pmf (a planted off-by-one fails); SSA against exact laws
 * (birth-death Poisson law, first-order chain binomials, the Schlogl master
 * equation); the leap decision against known answers and every CalculateG
 * branch; the critical-reaction branch (tau_2, SingleCriticalReaction)
 * against exponential and competing-decay laws (planted mistakes fail); tau
 * halving; the step budget; the infinite-tau flag against an extinction law;
 * committed golden trajectories of three seeded workloads.
 */
#include <stdint.h>
#include <string.h>
#include <math.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_OK 0
#define ORC_EINVAL -1

/* per-simulation flag bits (identical meaning in the HIP library) */
#define ORC_FLAG_OVERFLOW   1  /* a count or a Poisson mean exceeded 2^30   */
#define ORC_FLAG_HALVING    2  /* tau halved ORC_MAX_HALVINGS times         */
#define ORC_FLAG_STEPS     16  /* max_steps reached (not in the reference)  */
#define ORC_FLAG_STALL     32  /* tau > 0 but t + tau == t (clock stuck)    */
#define ORC_FLAG_TAU_INF    4  /* leap with a non-finite tau                */
#define ORC_FLAG_PTRS       8  /* PTRS rejection loop limit                 */

#define ORC_MAX_HALVINGS 2000
#define ORC_COUNT_LIMIT 1073741824     /* 2^30 */
#define ORC_LAMBDA_LIMIT 1.0e9
#define ORC_INV_MAX_K 100
#define ORC_PTRS_MAX_ITERS 1000

/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)                          */
/* ------------------------------------------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/*
 * The per-simulation stream.  Every random draw comes from a Philox block
 *   Philox(ctr = (c_lo, c_hi | it << 16, s_lo, s_hi), key = (seed_lo, seed_hi))
 * of simulation s, where c is a per-simulation block counter and it a retry
 * index (0 except for PTRS rejection iterations).  Blocks are taken in groups
 * at fixed points of the algorithm, each draw at a fixed word position:
 *   SSA step:            1 block  (c = b++): rho1 = U(w0,w1), rho2 = U(w2,w3)
 *   critical tau2/pick:  1 block  (c = b++): rho1 = U(w0,w1), rho  = U(w2,w3)
 *   Poisson attempt:     M blocks (c = b .. b+M-1, b += M): reaction j draws
 *                        from block b+j (inversion u = U(w0,w1); PTRS
 *                        iteration it: U = U(w0,w1), V = U(w2,w3) of block
 *                        (b+j, it))
 * An FP64 uniform from two words:
 *   m = (w_a >> 5) * 2^26 + (w_b >> 6),   u = (m + 1) * 2^-53  in (0, 1]
 * which keeps curand_uniform's (0, 1] convention (P1_P2.py:201, P3.py:114).
 */
typedef struct {
    uint32_t key[2];
    uint64_t sim;
    uint64_t b;            /* next block counter */
} orc_rng;

static void rng_init(orc_rng* g, uint64_t seed, uint64_t sim)
{
    g->key[0] = (uint32_t)seed;
    g->key[1] = (uint32_t)(seed >> 32);
    g->sim = sim;
    g->b = 0;
}

static void rng_block(const orc_rng* g, uint64_t c, uint32_t it, uint32_t w[4])
{
    uint32_t ctr[4] = {(uint32_t)c, (uint32_t)(c >> 32) | (it << 16),
                       (uint32_t)g->sim, (uint32_t)(g->sim >> 32)};
    orc_philox4x32_10(ctr, g->key, w);
}

static double u53(uint32_t wa, uint32_t wb)
{
    uint64_t m = ((uint64_t)(wa >> 5) << 26) | (uint64_t)(wb >> 6);
    return (double)(m + 1u) * 0x1p-53;
}

/* exported for the known-answer tests: 2 uniforms per block, blocks 0, 1, ... */
void orc_uniform_stream(uint64_t seed, uint64_t sim, int64_t n, double* out)
{
    orc_rng g;
    rng_init(&g, seed, sim);
    uint32_t w[4];
    for (int64_t i = 0; i < n; ++i) {
        if ((i & 1) == 0) rng_block(&g, (uint64_t)(i >> 1), 0, w);
        out[i] = (i & 1) ? u53(w[2], w[3]) : u53(w[0], w[1]);
    }
}

/* ------------------------------------------------------------------------- */
/* log / exp from published methods, built only from + - * / and bit access   */
/* so that host and device give identical bits.                               */
/* ------------------------------------------------------------------------- */
static const double ORC_LN2_HI = 0x1.62e42p-1;           /* 32 trailing zero bits */
static const double ORC_LN2_LO = 0x1.fdf473de6af28p-22;
static const double ORC_INV_LN2 = 0x1.71547652b82fep+0;
static const double ORC_SQRT2 = 1.4142135623730951;

static inline uint64_t d2u(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }
static inline double u2d(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }

/*
 * Natural log: x = 2^e * m with m in [sqrt(2)/2, sqrt(2)), then
 * ln m = 2 atanh(s), s = (m-1)/(m+1), summed as the odd series
 * 2s (1 + z/3 + z^2/5 + ... + z^10/21), z = s^2, |s| <= 0.1716.
 */
double orc_log(double x)
{
    if (!(x > 0.0)) return (x == 0.0) ? -INFINITY : NAN;
    if (x == INFINITY) return INFINITY;
    int e_adj = 0;
    if (x < 0x1p-1022) { x = x * 0x1p54; e_adj = -54; }
    uint64_t b = d2u(x);
    int e = (int)((b >> 52) & 0x7ff) - 1023 + e_adj;
    double m = u2d((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > ORC_SQRT2) { m = m * 0.5; e = e + 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double p = 1.0 / 21.0;
    p = p * z + 1.0 / 19.0;
    p = p * z + 1.0 / 17.0;
    p = p * z + 1.0 / 15.0;
    p = p * z + 1.0 / 13.0;
    p = p * z + 1.0 / 11.0;
    p = p * z + 1.0 / 9.0;
    p = p * z + 1.0 / 7.0;
    p = p * z + 1.0 / 5.0;
    p = p * z + 1.0 / 3.0;
    double s2 = s + s;
    double lm = s2 + s2 * (z * p);
    double de = (double)e;
    return de * ORC_LN2_HI + (lm + de * ORC_LN2_LO);
}

/*
 * exp: x = k ln2 + r (Cody-Waite split of ln2), |r| <= ln2/2, exp(r) by its
 * Taylor polynomial to r^13/13!, then scaled by 2^k through the exponent bits.
 */
double orc_exp(double x)
{
    if (x != x) return x;
    if (x > 709.0) return INFINITY;
    if (x < -708.0) return 0.0;
    double k = floor(x * ORC_INV_LN2 + 0.5);
    double r = (x - k * ORC_LN2_HI) - k * ORC_LN2_LO;
    double p = 1.0 / 6227020800.0;          /* 1/13! */
    p = p * r + 1.0 / 479001600.0;          /* 1/12! */
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    int ki = (int)k;
    return p * u2d((uint64_t)(ki + 1023) << 52);
}

/* ln(k!) for k = 0..9, then Stirling's series for ln Gamma(x), x = k+1 >= 11 */
static const double ORC_LOGFACT[10] = {
    0.0, 0.0,
    0.6931471805599453094172321, 1.791759469228055000812477,
    3.178053830347945619646942, 4.787491742782045994247701,
    6.579251212010100995060178, 8.525161361065414300165531,
    10.60460290274525022841723, 12.80182748008146961120772};

double orc_lgamma_int(double x)   /* x = k + 1, k >= 0 integral */
{
    if (x < 10.5) return ORC_LOGFACT[(int)x - 1];
    double r = 1.0 / x;
    double r2 = r * r;
    double ser = r * (1.0 / 12.0 - r2 * (1.0 / 360.0 - r2 * (1.0 / 1260.0 - r2 * (1.0 / 1680.0))));
    return (x - 0.5) * orc_log(x) - x + 0.9189385332046727417803297 + ser;
}

/*
 * Poisson(lambda) sampler replacing curand_poisson (P1_P2.py:239), drawing
 * from block c of the stream (retry index it for PTRS iteration it).
 *   lambda <= 0      -> 0
 *   lambda < 10      -> sequential-search inversion (Devroye 1986, X.3)
 *   lambda >= 10     -> PTRS transformed rejection (Hoermann 1993)
 * Returns -1 when the PTRS loop limit is reached.
 */
static int64_t orc_poisson(const orc_rng* g, uint64_t c, double lambda)
{
    uint32_t w[4];
    if (!(lambda > 0.0)) return 0;
    if (lambda < 10.0) {
        rng_block(g, c, 0, w);
        double u = u53(w[0], w[1]);
        double p = orc_exp(-lambda);
        double F = p;
        int64_t k = 0;
        while (u > F && k < ORC_INV_MAX_K) {
            k = k + 1;
            p = p * lambda / (double)k;
            F = F + p;
        }
        return k;
    }
    double loglam = orc_log(lambda);
    double slam = orc_exp(0.5 * loglam);
    double b = 0.931 + 2.53 * slam;
    double a = -0.059 + 0.02483 * b;
    double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (int it = 0; it < ORC_PTRS_MAX_ITERS; ++it) {
        rng_block(g, c, (uint32_t)it, w);
        double U = u53(w[0], w[1]) - 0.5;
        double V = u53(w[2], w[3]);
        double us = 0.5 - fabs(U);
        double kf = floor((2.0 * a / us + b) * U + lambda + 0.43);
        if (us >= 0.07 && V <= vr) return (int64_t)kf;
        if (kf < 0.0 || (us < 0.013 && V > us)) continue;
        double invalpha = 1.1239 + 1.1328 / (b - 3.4);
        double lhs = orc_log(V) + orc_log(invalpha) - orc_log(a / (us * us) + b);
        double rhs = -lambda + kf * loglam - orc_lgamma_int(kf + 1.0);
        if (lhs <= rhs) return (int64_t)kf;
    }
    return -1;
}

/* exported for the distribution tests: sample i uses block counter i */
void orc_poisson_stream(uint64_t seed, uint64_t sim, double lambda, int64_t n, int64_t* out)
{
    orc_rng g;
    rng_init(&g, seed, sim);
    for (int64_t i = 0; i < n; ++i) out[i] = orc_poisson(&g, (uint64_t)i, lambda);
}

/* ------------------------------------------------------------------------- */
/* Network tables (reference TlArgs fields, tl_parser.py:50-78)              */
/* ------------------------------------------------------------------------- */
typedef struct {
    int32_t n_species;       /* N */
    int32_t n_reactions;     /* M */
    int32_t a_size;          /* nnz(Pre) */
    const int32_t* A;        /* a_size x 4: (species, reaction, value, 0), species-major */
    int32_t v_size;          /* nnz(Post - Pre) */
    const int32_t* V;        /* v_size x 4 */
    const int32_t* H;        /* N: highest order of reaction (tl_parser.py:97-120) */
    const int32_t* H_type;   /* N */
    const int32_t* prop_order;  /* N*M (species-major): factors of species i in a_j */
    const int32_t* prop_step;   /* N*M: 1 -> x(x-1)(x-2).., 0 -> x^p */
} orc_network;

typedef struct {
    int32_t method;       /* 0 tau-leaping, 1 forced SSA (gillespie=1), 2 pure SSA */
    int32_t n_c;          /* tl_parser.py:43 */
    double eta;           /* tl_parser.py:44 */
    int32_t ssa_burst;    /* P3.py:71 (100) */
    double t_max;         /* gillespie.py:70 */
    int32_t ita;
    const float* I;       /* ita sample times (float32, tl_parser.py:76-77) */
    int32_t kappa;
    const int32_t* E;     /* kappa output species */
    int32_t n_sims;       /* U */
    int32_t sims_per_set; /* sims sharing one (rates, x0) row */
    uint64_t seed;
    uint64_t sim_offset;
    int64_t max_steps;    /* 0: no limit; else leap decisions + SSA steps per simulation */
} orc_config;

typedef struct {
    int64_t events;
    int64_t leaps;
    int64_t rejected_leaps;
    int64_t ssa_steps;
    int64_t flagged;
    int64_t flagged_events;  /* events of the simulations counted in flagged */
} orc_stats;

#define ORC_NMAX 64
#define ORC_MMAX 256

/* UpdatePropensities (generated by parser.py:32-55) in table form:
 * a_j = k_j * prod_i prod_{q<p_ij} (x_i - step_ij*q), species ascending. */
static void propensities(const orc_network* net, const double* k, const int32_t* x, double* a)
{
    const int N = net->n_species, M = net->n_reactions;
    for (int j = 0; j < M; ++j) {
        double v = k[j];
        for (int i = 0; i < N; ++i) {
            int p = net->prop_order[i * M + j];
            int st = net->prop_step[i * M + j];
            for (int q = 0; q < p; ++q) v = v * (double)(x[i] - st * q);
        }
        a[j] = v;
    }
}

/* CalculateG, P1_P2.py:447-486 */
static double calc_g(const orc_network* net, int x_i, int i)
{
    int h = net->H[i], ht = net->H_type[i];
    if (h == 1) return 1.0;
    if (h == 2) {
        if (ht == 2) return 2.0 + 1.0 / ((double)x_i - 1.0);
        return 2.0;
    }
    if (h == 3) {
        if (ht == 3) return 3.0 + 1.0 / ((double)x_i - 1.0) + 2.0 / ((double)x_i - 2.0);
        if (ht == 2) return 3.0 / 2.0 * (2.0 + 1.0 / ((double)x_i - 1.0));
        return 3.0;
    }
    return 0.0;
}

/* __float2int_rn (P1_P2.py:572): round half to even; NaN -> 0, clamped */
static int32_t rn_f32_to_int(float f)
{
    if (f != f) return 0;
    if (f >= 2147483648.0f) return 2147483647;  /* 2^31 and above saturate */
    if (f <= -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)nearbyintf(f);
}

/* fill samples F..ita-1 with the current state (P1_P2.py:111-133) */
static void fill_rest(const orc_config* c, int32_t* O, int64_t U, int64_t sim, int F, const int32_t* x, int zero)
{
    for (int k = 0; k < c->kappa; ++k)
        for (int f = F; f < c->ita; ++f)
            O[((int64_t)k * c->ita + f) * U + sim] = zero ? 0 : x[c->E[k]];
}

/*
 * One simulation through the reference's per-thread control flow.
 * Returns the flag bits.
 */
/* Planted mistakes, for the tests that show the exact-law checks catch them
 * (set only from tests; 0 everywhere else):
 *   1: tau2 drawn with the total propensity a0 instead of a0c (P1_P2.py:200-204)
 *   2: the critical pick normalised by a0 instead of a0c (P1_P2.py:500)      */
static int orc_plant = 0;
void orc_set_plant(int p) { orc_plant = p; }

/* The critical part of a leap (P1_P2.py:175-220): a0c = sum of the critical
 * propensities; if a0c > 0, tau2 = (1/a0c) ln(1/rho1) from the next block;
 * tau = fminf(tau1, tau2); if tau1 > tau2, SingleCriticalReaction
 * (P1_P2.py:488-507) picks the first critical j with sum a_r / a0c >= rho
 * (rho from the same block) and stores it in *jc (else -1).  The reference
 * falls off the end if round-off leaves the sum below rho: the last critical
 * reaction with a_j > 0 is taken instead.  Returns tau. */
static double critical_draw(orc_rng* g, const double* a, const int32_t* Xeta, int M, double a0, double tau1,
                            int* jc)
{
    double a0c = 0.0;
    for (int j = 0; j < M; ++j) if (Xeta[j] == 1) a0c = a0c + a[j];
    double tau2 = INFINITY;
    uint32_t wc[4] = {0, 0, 0, 0};
    if (a0c > 0.0) {
        rng_block(g, g->b, 0, wc);
        g->b += 1;
        double rho1 = u53(wc[0], wc[1]);
        tau2 = (1.0 / (orc_plant == 1 ? a0 : a0c)) * orc_log(1.0 / rho1);
    }
    double tau = (double)fminf((float)tau1, (float)tau2);
    *jc = -1;
    if (tau1 > tau2) {
        double rho = u53(wc[2], wc[3]);
        double rc = 0.0;
        int last = -1;
        for (int r = 0; r < M; ++r) {
            if (Xeta[r] == 1) {
                if (a[r] > 0.0) last = r;
                rc = rc + a[r] / (orc_plant == 2 ? a0 : a0c);
                if (rc >= rho) { *jc = r; break; }
            }
        }
        if (*jc < 0) *jc = last;
    }
    return tau;
}

/* n independent critical draws (test hook): draw i uses simulation id
 * `sim` + i of stream `seed` from block 0, as the first leap of that
 * simulation would.  tau_out[i] = tau, j_out[i] = picked reaction or -1. */
void orc_critical_draws(const double* a, const int32_t* Xeta, int M, double tau1, uint64_t seed, uint64_t sim,
                        int64_t n, double* tau_out, int32_t* j_out)
{
    double a0 = 0.0;
    for (int j = 0; j < M; ++j) a0 = a0 + a[j];
    for (int64_t i = 0; i < n; ++i) {
        orc_rng g;
        rng_init(&g, seed, sim + (uint64_t)i);
        int jc = -1;
        tau_out[i] = critical_draw(&g, a, Xeta, M, a0, tau1, &jc);
        j_out[i] = jc;
    }
}

static int simulate_one(const orc_network* net, const double* k, const int32_t* x0,
                        const orc_config* c, int64_t U, int64_t sim, int32_t* O, orc_stats* st)
{
    const int N = net->n_species, M = net->n_reactions;
    int32_t x[ORC_NMAX];
    int64_t xp[ORC_NMAX];
    double a[ORC_MMAX], mu[ORC_NMAX], sigma2[ORC_NMAX];
    int32_t Xeta[ORC_MMAX], K[ORC_MMAX], Irs[ORC_NMAX];
    orc_rng g;
    rng_init(&g, c->seed, c->sim_offset + (uint64_t)sim);
    for (int i = 0; i < N; ++i) x[i] = x0[i];
    double t = 0.0;
    int F = 0;
    int flags = 0;
    int64_t steps = 0;  /* leap decisions + SSA steps, for max_steps */

    /* I_rs: species with any Pre entry (P1_P2.py:338-346) */
    for (int i = 0; i < N; ++i) Irs[i] = 0;
    for (int e = 0; e < net->a_size; ++e) Irs[net->A[4 * e]] = 1;

    if (c->method == 2) {
        /* kernel_Gillespie, gillespie.py:70-146 */
        while (t < c->t_max) {
            if (c->max_steps > 0 && steps >= c->max_steps) {
                flags |= ORC_FLAG_STEPS; fill_rest(c, O, U, sim, F, x, 1); return flags;
            }
            ++steps;
            for (int j = 0; j < M; ++j) a[j] = 0.0;
            propensities(net, k, x, a);
            double a0 = 0.0;
            for (int j = 0; j < M; ++j) a0 = a0 + a[j];
            if (a0 == 0.0) { fill_rest(c, O, U, sim, F, x, 0); return flags; }
            uint32_t w[4];
            rng_block(&g, g.b, 0, w);
            g.b += 1;
            double rho1 = u53(w[0], w[1]);
            double tau = (1.0 / a0) * orc_log(1.0 / rho1);
            double rho = u53(w[2], w[3]);
            double rc = 0.0;
            int j = -1, last = -1;
            for (int r = 0; r < M; ++r) {
                if (a[r] > 0.0) last = r;
                rc = rc + a[r] / a0;
                if (rc >= rho) { j = r; break; }
            }
            if (j < 0) j = last;
            {
                const double t_old = t;
                t = t + tau;
                if (t == t_old && tau > 0.0) { flags |= ORC_FLAG_STALL; fill_rest(c, O, U, sim, F, x, 1); return flags; }
            }
            while (t >= (double)c->I[F]) {
                for (int q = 0; q < c->kappa; ++q)
                    O[((int64_t)q * c->ita + F) * U + sim] = x[c->E[q]];
                F = F + 1;
                if (F == c->ita) return flags;
            }
            for (int e = 0; e < net->v_size; ++e)
                if (net->V[4 * e + 1] == j) x[net->V[4 * e]] += net->V[4 * e + 2];
            st->ssa_steps++;
            st->events++;
            for (int i = 0; i < N; ++i)
                if (x[i] >= ORC_COUNT_LIMIT) { flags |= ORC_FLAG_OVERFLOW; fill_rest(c, O, U, sim, F, x, 1); return flags; }
        }
        /* t >= t_max before the last sample: remaining samples stay zero
         * (cuGillespie.py:71 zero-fills O) */
        fill_rest(c, O, U, sim, F, x, 1);
        return flags;
    }

    /* tau-leaping driver: {P1_P2; P3 if Q==0} until Q == -1 (cuTauLeaping.py:49-76) */
    for (;;) {
        int Q;
        if (c->method == 1) {
            Q = 0;   /* GILLESPIE macro, P1_P2.py:67-74 */
        } else {
            /* ---- kernel_P1_P2 (P1_P2.py:96-294) ---- */
            if (c->max_steps > 0 && steps >= c->max_steps) {
                flags |= ORC_FLAG_STEPS; fill_rest(c, O, U, sim, F, x, 1); return flags;
            }
            ++steps;
            for (int j = 0; j < M; ++j) a[j] = 0.0;
            propensities(net, k, x, a);
            double a0 = 0.0;
            for (int j = 0; j < M; ++j) a0 = a0 + a[j];
            if (a0 == 0.0) { fill_rest(c, O, U, sim, F, x, 0); return flags; }

            /* DetermineCriticalReactions on the net change V (P1_P2.py:314-327) */
            for (int j = 0; j < M; ++j) Xeta[j] = 0;
            for (int e = 0; e < net->v_size; ++e) {
                int v = net->V[4 * e + 2];
                if (v < 0) {
                    int av = -v;
                    if (av * c->n_c > x[net->V[4 * e]]) Xeta[net->V[4 * e + 1]] = 1;
                }
            }
            /* CalculateMuSigma (P1_P2.py:330-383) */
            for (int i = 0; i < N; ++i) { mu[i] = 0.0; sigma2[i] = 0.0; }
            for (int e = 0; e < net->v_size; ++e) {
                int i = net->V[4 * e], j = net->V[4 * e + 1], v = net->V[4 * e + 2];
                if (Irs[i] == 1 && Xeta[j] == 0) {
                    mu[i] = mu[i] + ((double)v * a[j]);
                    sigma2[i] = sigma2[i] + ((double)(v * v) * a[j]);
                }
            }
            /* CalculateTau (P1_P2.py:385-445): fmaxf/fabsf/fminf round through FP32 */
            double tau1 = INFINITY;
            for (int e = 0; e < net->v_size; ++e) {
                int i = net->V[4 * e], j = net->V[4 * e + 1];
                if (Irs[i] == 1 && Xeta[j] == 0) {
                    double g_i = calc_g(net, x[i], i);
                    double num = (double)fmaxf((float)(c->eta * (double)x[i] / g_i), 1.0f);
                    double lhs = num / (double)fabsf((float)mu[i]);
                    double rhs = (num * num) / sigma2[i];
                    double tt = (double)fminf((float)lhs, (float)rhs);
                    if (tt < tau1) tau1 = tt;
                }
            }
            /* SSA switch (P1_P2.py:157-173) */
            if (tau1 < 10.0 / a0) {
                Q = 0;
            } else {
                Q = 1;
                for (int j = 0; j < M; ++j) K[j] = 0;
                int jc = -1;
                double tau = critical_draw(&g, a, Xeta, M, a0, tau1, &jc);
                if (jc >= 0) K[jc] = 1;
                if (!(tau < INFINITY)) {
                    flags |= ORC_FLAG_TAU_INF;
                    fill_rest(c, O, U, sim, F, x, 1);
                    return flags;
                }
                /* Poisson leap with tau halving (P1_P2.py:228-259) */
                int valid = 0;
                for (int h = 0; h < ORC_MAX_HALVINGS; ++h) {
                    int bad = 0;
                    uint64_t grp = g.b;
                    g.b += (uint64_t)M;
                    for (int j = 0; j < M; ++j) {
                        if (Xeta[j] == 0) {
                            double lam = a[j] * tau;
                            if (lam >= ORC_LAMBDA_LIMIT) { bad = ORC_FLAG_OVERFLOW; break; }
                            int64_t kk = orc_poisson(&g, grp + (uint64_t)j, lam);
                            if (kk < 0) { bad = ORC_FLAG_PTRS; break; }
                            K[j] = (int32_t)kk;
                        }
                    }
                    if (bad) { flags |= bad; fill_rest(c, O, U, sim, F, x, 1); return flags; }
                    /* TentativeUpdatedState (P1_P2.py:509-528) */
                    for (int i = 0; i < N; ++i) xp[i] = x[i];
                    for (int e = 0; e < net->v_size; ++e)
                        xp[net->V[4 * e]] = xp[net->V[4 * e]] + (int64_t)K[net->V[4 * e + 1]] * net->V[4 * e + 2];
                    /* ValidState (P1_P2.py:530-540) */
                    int ok = 1;
                    for (int i = 0; i < N; ++i) if (xp[i] < 0) { ok = 0; break; }
                    if (ok) { valid = 1; break; }
                    st->rejected_leaps++;
                    tau = tau / 2.0;
                }
                if (!valid) { flags |= ORC_FLAG_HALVING; fill_rest(c, O, U, sim, F, x, 1); return flags; }
                for (int i = 0; i < N; ++i)
                    if (xp[i] >= ORC_COUNT_LIMIT) { flags |= ORC_FLAG_OVERFLOW; fill_rest(c, O, U, sim, F, x, 1); return flags; }
                int64_t ev = 0;
                for (int j = 0; j < M; ++j) ev += K[j];
                st->events += ev;
                st->leaps++;
                /* time advance and SaveInterpolatedDynamics (P1_P2.py:261-286,542-584) */
                {
                    const double t_old = t;
                    t = t + tau;
                    if (t == t_old && tau > 0.0) { flags |= ORC_FLAG_STALL; fill_rest(c, O, U, sim, F, x, 1); return flags; }
                }
                int done = 0;
                while (t >= (double)c->I[F]) {
                    double t0 = t - tau;
                    float record_time = c->I[F];
                    double gap = t - t0;
                    double nrt = (double)record_time - t0;
                    double frac = nrt / gap;
                    for (int q = 0; q < c->kappa; ++q) {
                        int s = c->E[q];
                        int diff = (int)xp[s] - x[s];
                        double inc = (double)diff * frac;
                        O[((int64_t)q * c->ita + F) * U + sim] = x[s] + rn_f32_to_int((float)inc);
                    }
                    F = F + 1;
                    if (F == c->ita) { done = 1; break; }
                }
                for (int i = 0; i < N; ++i) x[i] = (int32_t)xp[i];
                if (done) return flags;
            }
        }
        if (Q == 0) {
            /* ---- kernel_P3 (P3.py:55-161) ---- */
            for (int cnt = 0; cnt < c->ssa_burst; ++cnt) {
                if (c->max_steps > 0 && steps >= c->max_steps) {
                    flags |= ORC_FLAG_STEPS; fill_rest(c, O, U, sim, F, x, 1); return flags;
                }
                ++steps;
                for (int j = 0; j < M; ++j) a[j] = 0.0;
                propensities(net, k, x, a);
                double a0 = 0.0;
                for (int j = 0; j < M; ++j) a0 = a0 + a[j];
                if (a0 == 0.0) { fill_rest(c, O, U, sim, F, x, 0); return flags; }
                uint32_t w[4];
                rng_block(&g, g.b, 0, w);
                g.b += 1;
                double rho1 = u53(w[0], w[1]);
                double tau = (1.0 / a0) * orc_log(1.0 / rho1);
                /* SingleCriticalReaction, P3.py:179-194 (divides each term by a0) */
                double rho = u53(w[2], w[3]);
                double rc = 0.0;
                int j = -1, last = -1;
                for (int r = 0; r < M; ++r) {
                    if (a[r] > 0.0) last = r;
                    rc = rc + a[r] / a0;
                    if (rc >= rho) { j = r; break; }
                }
                if (j < 0) j = last;
                {
                    const double t_old = t;
                    t = t + tau;
                    if (t == t_old && tau > 0.0) { flags |= ORC_FLAG_STALL; fill_rest(c, O, U, sim, F, x, 1); return flags; }
                }
                while (t >= (double)c->I[F]) {
                    /* SaveDynamics: the pre-reaction state (P3.py:165-177) */
                    for (int q = 0; q < c->kappa; ++q)
                        O[((int64_t)q * c->ita + F) * U + sim] = x[c->E[q]];
                    F = F + 1;
                    if (F == c->ita) return flags;
                }
                /* UpdateState, P3.py:196-208 */
                for (int e = 0; e < net->v_size; ++e)
                    if (net->V[4 * e + 1] == j) x[net->V[4 * e]] += net->V[4 * e + 2];
                st->ssa_steps++;
                st->events++;
                for (int i = 0; i < N; ++i)
                    if (x[i] >= ORC_COUNT_LIMIT) { flags |= ORC_FLAG_OVERFLOW; fill_rest(c, O, U, sim, F, x, 1); return flags; }
            }
        }
    }
}

/*
 * Run n_sims simulations (sims [0, n_sims) of the config, global ids offset by
 * sim_offset).  rates is [n_sets][M], x0 is [n_sets][N]; simulation s uses row
 * s / sims_per_set.  O is [kappa][ita][U] with U = o_stride (>= n_sims).
 * flags (optional) receives per-simulation flag bits.  threads <= 0 uses the
 * OpenMP default.
 */
int orc_simulate(const orc_network* net, const double* rates, const int32_t* x0,
                 const orc_config* c, int32_t* O, int64_t o_stride, int32_t* flags,
                 orc_stats* stats, int threads)
{
    if (!net || !c || !O || !rates || !x0) return ORC_EINVAL;
    if (net->n_species < 1 || net->n_species > ORC_NMAX) return ORC_EINVAL;
    if (net->n_reactions < 1 || net->n_reactions > ORC_MMAX) return ORC_EINVAL;
    if (c->ita < 1 || c->kappa < 1 || c->n_sims < 0 || o_stride < c->n_sims) return ORC_EINVAL;
    if (c->sims_per_set < 1) return ORC_EINVAL;
    int64_t ev = 0, lp = 0, rj = 0, ss = 0, fl = 0, fe = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : ev, lp, rj, ss, fl, fe)
    for (int64_t s = 0; s < c->n_sims; ++s) {
        orc_stats st = {0, 0, 0, 0, 0, 0};
        int64_t set = s / c->sims_per_set;
        int f = simulate_one(net, rates + set * net->n_reactions, x0 + set * net->n_species,
                             c, o_stride, s, O, &st);
        if (flags) flags[s] = f;
        ev += st.events; lp += st.leaps; rj += st.rejected_leaps; ss += st.ssa_steps;
        fl += (f != 0);
        fe += (f != 0) ? st.events : 0;
    }
    if (stats) {
        stats->events = ev; stats->leaps = lp; stats->rejected_leaps = rj;
        stats->ssa_steps = ss; stats->flagged = fl; stats->flagged_events = fe;
    }
    return ORC_OK;
}

/*
 * Test hooks: the leap-decision quantities of P1_P2.py:96-173 at one state
 * (propensities, a0, critical flags, tau1, and the SSA switch), and the
 * interpolation rule of SaveInterpolatedDynamics (P1_P2.py:542-584).
 */
int orc_leap_decision(const orc_network* net, const double* k, const int32_t* x, int32_t n_c, double eta,
                      double* a_out, double* a0_out, int32_t* crit_out, double* tau1_out, int32_t* use_ssa)
{
    const int N = net->n_species, M = net->n_reactions;
    if (N < 1 || N > ORC_NMAX || M < 1 || M > ORC_MMAX) return ORC_EINVAL;
    double a[ORC_MMAX], mu[ORC_NMAX], sigma2[ORC_NMAX];
    int32_t Xeta[ORC_MMAX], Irs[ORC_NMAX];
    for (int i = 0; i < N; ++i) Irs[i] = 0;
    for (int e = 0; e < net->a_size; ++e) Irs[net->A[4 * e]] = 1;
    propensities(net, k, x, a);
    double a0 = 0.0;
    for (int j = 0; j < M; ++j) a0 = a0 + a[j];
    for (int j = 0; j < M; ++j) Xeta[j] = 0;
    for (int e = 0; e < net->v_size; ++e) {
        int v = net->V[4 * e + 2];
        if (v < 0 && (-v) * n_c > x[net->V[4 * e]]) Xeta[net->V[4 * e + 1]] = 1;
    }
    for (int i = 0; i < N; ++i) { mu[i] = 0.0; sigma2[i] = 0.0; }
    for (int e = 0; e < net->v_size; ++e) {
        int i = net->V[4 * e], j = net->V[4 * e + 1], v = net->V[4 * e + 2];
        if (Irs[i] == 1 && Xeta[j] == 0) {
            mu[i] = mu[i] + ((double)v * a[j]);
            sigma2[i] = sigma2[i] + ((double)(v * v) * a[j]);
        }
    }
    double tau1 = INFINITY;
    for (int e = 0; e < net->v_size; ++e) {
        int i = net->V[4 * e], j = net->V[4 * e + 1];
        if (Irs[i] == 1 && Xeta[j] == 0) {
            double g_i = calc_g(net, x[i], i);
            double num = (double)fmaxf((float)(eta * (double)x[i] / g_i), 1.0f);
            double lhs = num / (double)fabsf((float)mu[i]);
            double rhs = (num * num) / sigma2[i];
            double tt = (double)fminf((float)lhs, (float)rhs);
            if (tt < tau1) tau1 = tt;
        }
    }
    for (int j = 0; j < M; ++j) { a_out[j] = a[j]; crit_out[j] = Xeta[j]; }
    *a0_out = a0;
    *tau1_out = tau1;
    *use_ssa = (tau1 < 10.0 / a0) ? 1 : 0;
    return ORC_OK;
}

double orc_calc_g(int32_t h, int32_t h_type, int32_t x_i)
{
    int32_t H[1] = {h}, Ht[1] = {h_type};
    orc_network net;
    memset(&net, 0, sizeof(net));
    net.H = H;
    net.H_type = Ht;
    return calc_g(&net, x_i, 0);
}

int32_t orc_interpolate(int32_t x, int32_t x_prime, double t_old, double tau, float record_time)
{
    double t = t_old + tau;
    double t0 = t - tau;
    double gap = t - t0;
    double nrt = (double)record_time - t0;
    double frac = nrt / gap;
    int diff = x_prime - x;
    double inc = (double)diff * frac;
    return x + rn_f32_to_int((float)inc);
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
