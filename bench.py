"""Benchmark: reaction events/s of the simulator hot path on one MI355X.

Workload (one "step" = one full run of the workload, trajectories written to
device memory): Lotka-Volterra tau-leaping with SSA fallback,
X -> 2X (10), X + Y -> 2Y (0.01), Y -> 0 (10), X0 = Y0 = 1000,
epsilon = 0.03, n_c = 10, SSA burst 100, 65536 simulations, t_end = 10,
1000 sample times, both species recorded.

``value`` is the mixed-precision path (FP32 propensities / tau bound, FP64
clock and reaction choice) with the network compiled into the kernel at plan
creation (untimed).  The same workload is also timed on the plain FP64 path
(``fp64_plain``: the library's table-driven kernel) and on FP64 with the
network compiled in (``fp64_specialized``); the mixed result is checked
against FP64 (relative error of per-species means and variances and the KS
distance at every sample time), and the CPU oracle (the FP64 method in C,
OpenMP over simulations) is timed on a bounded sample of the same workload
(``cpu_baseline``).

Events are counted over the simulations that end unflagged: a simulation
flagged for a count above 2^30, a step budget or a stuck clock has invalid
output, and on this workload a handful of explosive ones (Y extinct, X
growing like e^(10 t)) would otherwise contribute ~10^9 events each, a share
that depends on the random stream.  The two precisions draw different
streams, so the speedups compare the kernel time of the same simulations.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "reaction events/s, 1 GPU: mixed-prec batched vs FP64 plain at ≤1% moment err"
VALU_PEAK = 256 * 4 * 16 * 2.4e9      # lane-instructions/s: 256 CU x 4 SIMD x 16 lanes x 2.4 GHz
HBM_PEAK = 8.0e12                      # bytes/s

U = 65536
T_END = 10.0
ITA = 1000
SEED = 2
X0 = 1000
CAP_FACTOR = 20
COPY_BYTES = 256 << 20
# launch shapes are timed on the whole horizon: on Lotka-Volterra the second
# half of the samples costs ~3x the first (the oscillations widen and
# simulations pass through low populations), and a shape timed on the first
# half only (15 SSA steps per pass: 4 % faster there) was 7 % slower overall
TUNE_FRACTION = 1.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_lv65536_mixed.txt")
ISSUE_FILE = os.path.join(ROOT, "profiles", "pmc_lv65536_mixed_instruction_mix.txt")
SIMDS = 256 * 4
CLOCK_HZ = 2.4e9


def read_traffic(path=TRAFFIC_FILE):
    """HBM bytes per launch of the mixed kernel from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by the reading of a copy of known size)."""
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 2 and parts[0] == "bytes_per_launch":
                    return float(parts[1])
    except OSError:
        pass
    return None


def read_issue(kernel_ms, path=ISSUE_FILE):
    """Issue-slot use of the mixed kernel: (VALU + SALU + branch instructions
    per SIMD, from the committed SQ counter summary of one launch) over the
    SIMD's quad-cycles in this run's kernel time at 2.4 GHz.  A wavefront
    issues at most one instruction per quad-cycle, so this is the fraction
    of the issue bound the kernel reaches."""
    try:
        cnt = {}
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 2 and parts[0].startswith("SQ_"):
                    cnt[parts[0]] = float(parts[1])
        ins = (cnt["SQ_INSTS_VALU"] + cnt["SQ_INSTS_SALU"] + cnt["SQ_INSTS_BRANCH"]) / SIMDS
    except (OSError, KeyError):
        return None
    quad = kernel_ms * 1e-3 * CLOCK_HZ / 4.0
    return {"frac": ins / quad, "instructions_per_simd": ins, "quad_cycles": quad,
            "counters": os.path.relpath(path, ROOT)}


def work_model(tabs, kappa, st):
    """Algorithmic VALU lane-operations of one mixed-precision run from the
    kernel's counters.

    Weights per counted unit (add/mul/compare/convert = 1, divide/log/exp/
    sqrt = 4, Philox4x32-10 block = 140):
      uniform              37   one Philox word (a quarter block) + conversion
      PTRS iteration       62   beyond its two uniforms (means >= 2^16)
      inversion step        3   sequential search (means < 10)
      asymptotic inversion 95   normal quantile (log, sqrt, 10-term
                                polynomial), rsqrt, three polynomials of
                                degree 10/10/6, floor, band test
      FP64 decision       300   pmf, log/exp/ln Gamma and the first terms of
                                the CDF sum (rare)
      leap decision        P + 2M + 23 nnz(V) + 5   propensities, a0, critical
                                                   test, mu/sigma^2, Cao bound
      accepted leap        2 nnz(V) + N + 1         state update and checks
      rejected leap        2 nnz(V) + N
      SSA step             P + 4M + 13 + nnz(V)/M   propensities, a0, tau,
                                                   selection, update
      sample row           8 kappa
    with P = sum_j (1 + 2 order_j).
    """
    po = np.asarray(tabs["prop_order"])
    N, M = int(tabs["N"]), int(tabs["M"])
    nnz = int(np.asarray(tabs["V"]).reshape(-1, 4).shape[0])
    P = int(sum(1 + 2 * int(po[:, j].sum()) for j in range(M)))
    w = (37 * st["uniforms"] + 62 * st["ptrs_iterations"] + 3 * st["inversion_steps"]
         + 95 * st["asym_draws"] + 300 * st["exact_draws"]
         + (P + 2 * M + 23 * nnz + 5) * st["leap_decisions"]
         + (2 * nnz + N + 1) * st["leaps"] + (2 * nnz + N) * st["rejected_leaps"]
         + (P + 4 * M + 13 + nnz / M) * st["ssa_steps"] + 8 * kappa * st["sample_rows"])
    return float(w)


class Moments:
    """Per-(species, sample) mean and variance over simulations, accumulated on
    the GPU set by set and merged with Chan's pairwise update (no sum-of-squares
    cancellation at near-deterministic early samples).

    Simulations that ended with a flag, in which a recorded species is
    extinct at the last sample, or in which a recorded species exceeds `cap`
    at some sample are left out, identically for every path compared: while
    the predator is at or near zero the prey grows like e^(10 t), and those
    rare trajectories (about 0.1% of the workload) would dominate and
    randomise every late-time mean and variance."""

    def __init__(self, cap):
        self.cap = cap
        self.n = 0
        self.mean = None
        self.m2 = None
        self.kept = 0
        self.total = 0

    def add(self, O, flags):
        keep = (flags == 0) & (O[:, -1, :] > 0).all(dim=0) & (O.amax(dim=(0, 1)) <= self.cap)
        k = int(keep.sum().item())
        self.total += O.shape[-1]
        if k == 0:
            return
        x = O[:, :, keep].double()
        mb = x.mean(-1)
        m2b = ((x - mb.unsqueeze(-1)) ** 2).sum(-1)
        if self.n == 0:
            self.mean, self.m2, self.n = mb, m2b, k
        else:
            n = self.n + k
            d = mb - self.mean
            self.mean = self.mean + d * (k / n)
            self.m2 = self.m2 + m2b + d * d * (self.n * k / n)
            self.n = n
        self.kept += k

    def mean_var(self):
        return self.mean, self.m2 / (self.n - 1)


def compare(A, B):
    """Max over species and sample times of the relative mean / variance error,
    and the 99th percentile of the variance error."""
    import torch
    ma, va = A.mean_var()
    mb, vb = B.mean_var()
    mm = mb.abs() > 0
    # sample times at which every kept simulation has the same count (t = 0)
    # have a variance that is rounding residue; they are not compared
    vm = vb > 1e-9 * (mb * mb + 1.0)
    me = ((ma - mb).abs() / mb.abs().clamp_min(1e-300))[mm].max().item() if mm.any() else 0.0
    vr = ((va - vb).abs() / vb.clamp_min(1e-300))[vm]
    ve = vr.max().item() if vm.any() else 0.0
    v99 = torch.quantile(vr, 0.99).item() if vm.any() else 0.0
    return float(me), float(ve), float(v99)


def ks_max(O_a, O_b):
    """Max over species and sample times of the two-sample KS distance."""
    import torch
    k, f, n = O_a.shape
    sa = torch.sort(O_a.reshape(k * f, n), dim=-1).values.contiguous()
    sb = torch.sort(O_b.reshape(k * f, n), dim=-1).values.contiguous()
    vals = torch.cat([sa, sb], dim=-1).contiguous()
    fa = torch.searchsorted(sa, vals, right=True).double() / n
    fb = torch.searchsorted(sb, vals, right=True).double() / n
    return float((fa - fb).abs().max().item())


def accuracy(p_mixed, p_fp64, O_blk, flags, stream, sets):
    """Mixed vs FP64 on `sets` independent sets of U simulations each, next to
    the same statistics between two independent FP64 sample sets (the noise
    floor of a comparison of this size).  The plans write blocks into O_blk;
    each run is converted to rows for the statistics."""
    import torch
    from py_stoch_amd import _native
    cap = CAP_FACTOR * X0
    mix, ref, ref2 = Moments(cap), Moments(cap), Moments(cap)
    kappa = O_blk.shape[0]
    O = torch.empty((kappa, ITA, U), dtype=torch.int32, device=O_blk.device)
    O_first = {}
    for r in range(sets):
        for name, plan, acc, off in (("mixed", p_mixed, mix, r), ("fp64", p_fp64, ref, r),
                                     ("fp64b", p_fp64, ref2, sets + r)):
            plan.set_stream(SEED, off * U)
            plan.execute(O_blk.data_ptr(), U, flags.data_ptr(), stream.cuda_stream)
            _native.blocks_to_rows(O_blk.data_ptr(), U, O.data_ptr(), U, kappa, ITA, U, stream.cuda_stream)
            torch.cuda.synchronize()
            acc.add(O, flags)
            if r == 0:
                O_first[name] = O.clone()
    p_mixed.set_stream(SEED, 0)
    p_fp64.set_stream(SEED, 0)
    me, ve, v99 = compare(mix, ref)
    ne, nv, n99 = compare(ref2, ref)
    out = {"simulations_per_path": sets * U,
           "excluded": "flagged, extinct at t_end, or a species above %d at some sample" % cap,
           "kept_fraction": {"mixed": mix.kept / mix.total, "fp64": ref.kept / ref.total},
           "mean_rel_err_max": me, "var_rel_err_max": ve, "var_rel_err_p99": v99,
           "ks_max": ks_max(O_first["mixed"], O_first["fp64"]),
           "noise_floor": {"mean_rel_err_max": ne, "var_rel_err_max": nv, "var_rel_err_p99": n99,
                           "ks_max": ks_max(O_first["fp64b"], O_first["fp64"])}}
    out["pass"] = bool(me <= 0.01 and ve <= 0.01 and out["ks_max"] <= 0.02)
    return out


def time_plan(plan, O, K, W, stream):
    import torch
    for _ in range(W):
        plan.execute(O.data_ptr(), U, None, stream.cuda_stream)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(K):
        plan.execute(O.data_ptr(), U, None, stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = plan.stats()
    return (t1 - t0), e0.elapsed_time(e1) / K, st


def parse_launch(text):
    """--launch SHIFT,SLICE[,TRIM[,BURST_WAVES]] -> the four launch-shape values."""
    v = [int(x) for x in text.split(",")]
    if not 2 <= len(v) <= 4:
        raise SystemExit("--launch takes SHIFT,SLICE[,TRIM[,BURST_WAVES]]")
    return tuple((v + [0, 0])[:4])


def valid_events(st):
    """Reaction events of the simulations that ended unflagged."""
    return st["events"] - st["flagged_events"]


def host_threads():
    """Cores this process may run on (its CPU affinity), at most the
    OMP_NUM_THREADS share when one is set (the GPU box grants a GPU job 16)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_model_name():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def cpu_baseline(tabs, rates, x0, I, E, budget_s=8.0, repeats=3):
    """The FP64 oracle on the host cores: the first n simulations of the same
    workload, n sized so one run takes about budget_s; the median of
    `repeats` runs."""
    from oracle import oracle
    threads = host_threads()
    probe = 256
    t0 = time.perf_counter()
    oracle.simulate(tabs, rates, x0, method=0, t_max=T_END, I=I, E=E, n_sims=probe, seed=SEED,
                    threads=threads)
    dt = time.perf_counter() - t0
    n = int(min(U, max(probe, (budget_s / max(dt, 1e-6)) * probe)))
    n = max(threads, (n // threads) * threads)
    rates_s = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        _, _, st = oracle.simulate(tabs, rates, x0, method=0, t_max=T_END, I=I, E=E, n_sims=n, seed=SEED,
                                   threads=threads)
        dt = time.perf_counter() - t0
        rates_s.append((valid_events(st) / dt, dt))
    rates_s.sort()
    value, dt = rates_s[len(rates_s) // 2]
    return {"value": value, "unit": "reaction events/s", "cores": threads, "kind": "port",
            "sample": "first %d of %d simulations of the same workload, FP64 oracle (C, OpenMP, %d threads), "
                      "median of %d runs (%.2f s each), %s; host reports %d CPUs"
                      % (n, U, threads, repeats, dt, cpu_model_name(), os.cpu_count() or 0)}


def _time_batch(plan, O_list, K, W, stream):
    import torch
    ptrs = [O.data_ptr() for O in O_list]
    strides = [O.shape[-1] for O in O_list]
    for _ in range(W):
        plan.execute(ptrs, strides, None, stream.cuda_stream)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(K):
        plan.execute(ptrs, strides, None, stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return t1 - t0, e0.elapsed_time(e1) / K, plan.stats()


def _ks_rows(A, B):
    """Two-sample KS distance per row of [rows][n] int tensors (same n)."""
    import torch
    sa = torch.sort(A, dim=-1).values.contiguous()
    sb = torch.sort(B, dim=-1).values.contiguous()
    vals = torch.cat([sa, sb], dim=-1).contiguous()
    n = A.shape[-1]
    fa = torch.searchsorted(sa, vals, right=True).double() / n
    fb = torch.searchsorted(sb, vals, right=True).double() / n
    return (fa - fb).abs().max(dim=-1).values


def other_workload(args):
    """Throughput of the mixed path and the FP64 path (both with the network
    compiled in) on one of the other configurations, with the accuracy
    statistic that fits that configuration."""
    import torch
    from py_stoch_amd import _native, workloads
    torch.cuda.init()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    wl = args.workload
    if wl == "birth-death":
        specs = workloads.birth_death()
    elif wl in ("schlogl", "schlogl-ssa"):
        specs = workloads.schlogl(method=2 if wl == "schlogl-ssa" else 0)
    else:
        specs = workloads.sweep()
    O_list = [torch.empty((s.kappa, s.ita, s.n_sims), dtype=torch.int32, device=dev) for s in specs]
    print("bench: %s, compiling kernels" % wl, file=sys.stderr, flush=True)
    p_mix = _native.BatchPlan(workloads.jobs(specs, _native.PRECISION_MIXED))
    p_f64 = _native.BatchPlan(workloads.jobs(specs, _native.PRECISION_FP64))
    tune = None
    if args.launch:
        shift, slc, trim, bw = parse_launch(args.launch)
        p_mix.set_launch(shift, slc, trim, bw)
        tune = {"mixed": {"lanes_shift": shift, "ssa_slice": slc, "waves_trim": trim, "burst_waves": bw, "set": True}}
    elif not args.no_tune:
        print("bench: %s, autotuning launch shapes" % wl, file=sys.stderr, flush=True)
        ptrs, strides = [O.data_ptr() for O in O_list], [s.n_sims for s in specs]
        tune = {"mixed": p_mix.autotune(ptrs, strides, None, stream.cuda_stream, fraction=TUNE_FRACTION),
                "fp64": p_f64.autotune(ptrs, strides, None, stream.cuda_stream, fraction=TUNE_FRACTION)}
    print("bench: %s, mixed" % wl, file=sys.stderr, flush=True)
    wall_m, kms_m, st_m = _time_batch(p_mix, O_list, args.steps, args.warmup, stream)
    O_mix = [O.clone() for O in O_list]
    print("bench: %s, FP64" % wl, file=sys.stderr, flush=True)
    k_f = max(1, args.steps // 2)
    wall_f, kms_f, st_f = _time_batch(p_f64, O_list, k_f, 1, stream)
    O_f64 = [O.clone() for O in O_list]
    value = valid_events(st_m) * args.steps / wall_m
    fp64_value = valid_events(st_f) * k_f / wall_f
    acc = {}
    if wl == "birth-death":
        # exact law: X(t) ~ Poisson(100 (1 - e^{-t/10}))
        t = torch.tensor(np.asarray(specs[0].kw["I"], np.float64), device=dev)
        lam = 100.0 * (1.0 - torch.exp(-0.1 * t))
        for name, O in (("mixed", O_mix[0]), ("fp64", O_f64[0])):
            x = O[0].double()
            m = x.mean(-1)
            se = (lam / x.shape[-1]).sqrt().clamp_min(1e-12)
            acc[name] = {"mean_max_abs_z_vs_exact": float(((m - lam).abs() / se)[1:].max().item())}
    elif wl in ("schlogl", "schlogl-ssa"):
        ks = _ks_rows(O_mix[0][0], O_f64[0][0])
        # noise floor: FP64 with another block of simulation ids
        p_f64.set_stream(specs[0].kw["seed"], specs[0].n_sims)
        p_f64.execute([O_list[0].data_ptr()], [specs[0].n_sims], None, stream.cuda_stream)
        torch.cuda.synchronize()
        ks_floor = _ks_rows(O_list[0][0], O_f64[0][0])
        x_m, x_f = O_mix[0][0, -1].double(), O_f64[0][0, -1].double()
        acc = {"ks_max_X": float(ks[1:].max().item()), "ks_max_X_noise_floor": float(ks_floor[1:].max().item()),
               "P(X(t_end)<300)": {"mixed": float((x_m < 300).double().mean().item()),
                                     "fp64": float((x_f < 300).double().mean().item())},
               "mean_X(t_end)": {"mixed": float(x_m.mean().item()), "fp64": float(x_f.mean().item())}}
    else:
        # per parameter set: final-time mean of each recorded species, mixed vs FP64,
        # as z-scores (flagged simulations left out on both sides)
        zs = []
        for s, Om, Of in zip(specs, O_mix, O_f64):
            per = s.kw["sims_per_set"]
            n_sets = s.n_sims // per
            for k in range(s.kappa):
                a_ = Om[k, -1].double().reshape(n_sets, per)
                b_ = Of[k, -1].double().reshape(n_sets, per)
                se = ((a_.var(-1) + b_.var(-1)) / per).sqrt()
                ok = se > 0
                zs.append(((a_.mean(-1) - b_.mean(-1)).abs() / se.clamp_min(1e-12))[ok])
        z = torch.cat(zs)
        acc = {"sets_compared": int(z.numel()), "frac_abs_z_above_4": float((z > 4).double().mean().item()),
               "median_abs_z": float(z.median().item()),
               "flagged": {"mixed": int(st_m["flagged"]), "fp64": int(st_f["flagged"])}}
    return {
        "metric": "reaction events/s, 1 GPU (%s)" % wl, "value": value, "unit": "reaction events/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall_m / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": wl, "jobs": [{"name": s.name, "simulations": s.n_sims, "N": int(s.tables["N"]),
                                             "M": int(s.tables["M"]), "samples": s.ita} for s in specs]},
        "steps_per_s": (st_m["leaps"] + st_m["ssa_steps"]) * args.steps / wall_m,
        "simulations_per_s": sum(sp.n_sims for sp in specs) * args.steps / wall_m,
        "kernel_ms": kms_m, "fp64_specialized": {"value": fp64_value, "kernel_ms": kms_f},
        "speedup_vs_fp64_specialized": (wall_f / k_f) / (wall_m / args.steps),
        "events_per_s_ratio_vs_fp64_specialized": value / fp64_value,
        "counters": {k: v for k, v in st_m.items() if k != "kernel_ms"}, "accuracy_vs_fp64": acc,
        "launch_autotune": tune,
    }


def random_networks_workload(args, n_check=8, check_sims=512):
    """256 random mass-action networks (5-50 species, 10-200 reactions) x 1024
    simulations in one plan, mixed precision.  Networks beyond the FP64 size
    classes have no FP64 GPU kernel, so accuracy is checked against the FP64
    CPU oracle: for the first n_check networks, per species, the final-time
    mean of the first check_sims simulations (flagged ones left out) as a
    z-score against an oracle run of check_sims simulations."""
    import torch
    from py_stoch_amd import _native, workloads
    from oracle import oracle
    torch.cuda.init()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    specs = workloads.random_networks()
    O_list = [torch.empty((s.kappa, s.ita, s.n_sims), dtype=torch.int32, device=dev) for s in specs]
    F_list = [torch.empty((s.n_sims,), dtype=torch.int32, device=dev) for s in specs]
    print("bench: random networks, plan", file=sys.stderr, flush=True)
    plan = _native.BatchPlan(workloads.jobs(specs, _native.PRECISION_MIXED, specialize=False))
    print("bench: random networks, timed", file=sys.stderr, flush=True)
    wall, kms, st = _time_batch(plan, O_list, args.steps, args.warmup, stream)
    plan.execute([O.data_ptr() for O in O_list], [O.shape[-1] for O in O_list], [F.data_ptr() for F in F_list],
                 stream.cuda_stream)
    torch.cuda.synchronize()
    print("bench: random networks, oracle check", file=sys.stderr, flush=True)
    zmax, flag_frac = [], []
    threads = host_threads()
    t_cpu, ev_cpu = 0.0, 0
    for s, O, F in list(zip(specs, O_list, F_list))[:n_check]:
        kw = dict(s.kw)
        kw["n_sims"] = check_sims
        t0 = time.perf_counter()
        Oc, Fc, sc = oracle.simulate(s.tables, s.rates, s.x0, threads=threads, **kw)
        t_cpu += time.perf_counter() - t0
        ev_cpu += valid_events(sc)
        Og = O[:, -1, :check_sims].cpu().numpy().astype(np.float64)
        Fg = F[:check_sims].cpu().numpy()
        x, y = Og[:, Fg == 0], Oc[:, -1, Fc == 0].astype(np.float64)
        flag_frac.append(float((Fg != 0).mean()))
        if x.shape[-1] > 10 and y.shape[-1] > 10:
            se = np.sqrt(x.var(axis=-1) / x.shape[-1] + y.var(axis=-1) / y.shape[-1])
            ok = se > 0
            zmax.append(float((np.abs(x.mean(axis=-1) - y.mean(axis=-1))[ok] / se[ok]).max()) if ok.any() else 0.0)
    value = valid_events(st) * args.steps / wall
    return {
        "metric": "reaction events/s, 1 GPU (random-networks)", "value": value, "unit": "reaction events/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "random-networks", "networks": len(specs),
                   "simulations": sum(s.n_sims for s in specs),
                   "species_range": [min(int(s.tables["N"]) for s in specs), max(int(s.tables["N"]) for s in specs)],
                   "reactions_range": [min(int(s.tables["M"]) for s in specs),
                                       max(int(s.tables["M"]) for s in specs)]},
        "kernel_ms": kms,
        "steps_per_s": (st["leaps"] + st["ssa_steps"]) * args.steps / wall,
        "counters": {k: v for k, v in st.items() if k != "kernel_ms"},
        "step_budget": {"max_steps_per_simulation": int(specs[0].kw["max_steps"]),
                        "flagged_simulations": int(st["flagged"])},
        "accuracy_vs_oracle": {"networks_checked": len(zmax), "simulations_each": check_sims,
                               "max_abs_z_per_network": zmax, "flagged_fraction": flag_frac},
        "cpu_baseline": {"value": ev_cpu / t_cpu, "unit": "reaction events/s", "cores": threads, "kind": "port",
                         "sample": "first %d networks x %d simulations, FP64 oracle (C, OpenMP), %.2f s"
                                   % (n_check, check_sims, t_cpu)},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the library's default launch shapes (skip the launch autotuner)")
    ap.add_argument("--launch", default=None, metavar="SHIFT,SLICE[,TRIM[,BURST_WAVES]]",
                    help="launch shape of the mixed plan (lanes_shift,ssa_slice,waves_trim,burst_waves of an "
                         "autotune); implies --no-tune for it (to profile the shape a default run chose)")
    ap.add_argument("--acc-sets", type=int, default=16,
                    help="independent sets of 65536 simulations per path for the accuracy check")
    ap.add_argument("--workload", default="lotka-volterra",
                    choices=["lotka-volterra", "birth-death", "schlogl", "schlogl-ssa", "sweep",
                             "random-networks"],
                    help="lotka-volterra (default) is the headline measurement; the others measure "
                         "the remaining configurations with the same JSON fields")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run one mixed-precision launch of the workload and one 256 MiB device copy "
                         "(for HBM byte counters under rocprofv3 --pmc), then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: simulations are independent, so N GPUs would each run "
              "their own range of simulation ids; multi-GPU runs are not part of this build.",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "random-networks":
        print(json.dumps(random_networks_workload(args)))
        return
    if args.workload != "lotka-volterra":
        print(json.dumps(other_workload(args)))
        return

    import torch
    from py_stoch_amd import _native, models
    from py_stoch_amd.mod.parser.tl_parser import tables_of

    a = models.make_args(models.lotka_volterra(), E=[0, 1], t_max=T_END, ita=ITA, U=U)
    tabs = tables_of(a)
    rates = a.rate_fn(a.c)
    kappa = 2
    # every plan writes the blocked trajectory layout (32-byte sectors of
    # eight samples of one simulation, each written once); the accuracy
    # check converts to rows
    common = dict(method=_native.METHOD_TAU_LEAPING, n_c=10, eta=0.03, ssa_burst=100, t_max=T_END,
                  I=a.I, E=a.E, n_sims=U, seed=SEED, layout=_native.LAYOUT_BLOCKS)
    torch.cuda.init()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    print("bench: building plans", file=sys.stderr, flush=True)
    # network-specific kernels are compiled here, before any timing
    p_mixed = _native.Plan(tabs, rates, a.x_0, precision=_native.PRECISION_MIXED, specialize=True, **common)
    p_fp64 = _native.Plan(tabs, rates, a.x_0, precision=_native.PRECISION_FP64, specialize=True, **common)
    p_plain = _native.Plan(tabs, rates, a.x_0, precision=_native.PRECISION_FP64, **common)
    ita8 = (ITA + 7) // 8
    O_m = torch.empty((kappa, ita8, U, 8), dtype=torch.int32, device=dev)
    O_f = torch.empty((kappa, ita8, U, 8), dtype=torch.int32, device=dev)
    if args.traffic_probe:
        # reference copy of known size: FETCH_SIZE / WRITE_SIZE per byte moved
        src = torch.ones((COPY_BYTES // 4,), dtype=torch.int32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        if args.launch:
            p_mixed.set_launch(*parse_launch(args.launch))
        p_mixed.execute(O_m.data_ptr(), U, None, stream.cuda_stream)
        torch.cuda.synchronize()
        print("bench: traffic probe done (copy %d bytes, one mixed launch)" % COPY_BYTES, file=sys.stderr)
        return

    # launch shapes (simulations per wavefront, SSA steps per pass, resident
    # wavefronts per SIMD) of every plan, timed on the whole horizon; results do not depend on them
    tune = None
    if args.launch:
        shift, slc, trim, bw = parse_launch(args.launch)
        p_mixed.set_launch(shift, slc, trim, bw)
        tune = {"mixed": {"lanes_shift": shift, "ssa_slice": slc, "waves_trim": trim, "burst_waves": bw, "set": True}}
    elif not args.no_tune:
        print("bench: autotuning launch shapes", file=sys.stderr, flush=True)
        tune = {name: plan.autotune(O.data_ptr(), U, None, stream.cuda_stream, fraction=TUNE_FRACTION)
                for name, plan, O in (("mixed", p_mixed, O_m), ("fp64_plain", p_plain, O_f),
                                      ("fp64_specialized", p_fp64, O_f))}
    print("bench: mixed precision, %d warmup + %d timed steps" % (args.warmup, args.steps),
          file=sys.stderr, flush=True)
    wall_m, kern_ms_m, st_m = time_plan(p_mixed, O_m, args.steps, args.warmup, stream)
    k_f = max(1, args.steps // 2)
    print("bench: plain FP64 (table-driven kernel), 1 warmup + %d timed steps" % k_f, file=sys.stderr, flush=True)
    wall_f, kern_ms_f, st_f = time_plan(p_plain, O_f, k_f, 1, stream)
    print("bench: FP64 with the network compiled in, 1 warmup + %d timed steps" % k_f, file=sys.stderr, flush=True)
    wall_s, kern_ms_s, st_s = time_plan(p_fp64, O_f, k_f, 1, stream)

    ev = valid_events(st_m)
    value = ev * args.steps / wall_m
    fp64_value = valid_events(st_f) * k_f / wall_f
    fp64s_value = valid_events(st_s) * k_f / wall_s
    print("bench: accuracy, mixed vs FP64 on %d x %d simulations" % (args.acc_sets, U), file=sys.stderr, flush=True)
    flags = torch.empty((U,), dtype=torch.int32, device=dev)
    acc = accuracy(p_mixed, p_fp64, O_f, flags, stream, args.acc_sets)

    W = work_model(tabs, kappa, st_m)
    t_k = kern_ms_m / 1e3
    traj_bytes = 4.0 * kappa * ITA * U
    roof = {"bound": "valu", "achieved": W / t_k / 1e12, "peak": VALU_PEAK / 1e12, "unit": "Tlane-op/s",
            "frac": W / t_k / VALU_PEAK, "traffic": None,
            "work_lane_ops": W, "kernel_ms": kern_ms_m,
            "hbm": {"algorithmic_bytes": traj_bytes, "achieved_GBps": traj_bytes / t_k / 1e9,
                    "frac": traj_bytes / t_k / HBM_PEAK}}
    roof["traffic"] = read_traffic()
    roof["issue"] = read_issue(kern_ms_m)

    cpu = None
    if not args.no_cpu:
        print("bench: CPU oracle baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(tabs, rates, a.x_0, a.I, a.E)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "reaction events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall_m / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "lotka-volterra tau-leaping, 65536 simulations, t_end=10, 1000 samples, "
                               "eps=0.03, n_c=10, SSA burst 100, mixed precision, network compiled into "
                               "a persistent batched kernel",
                   "trajectory_layout": "blocks [kappa][ita/8][U][8] (every plan)",
                   "simulations": U, "t_end": T_END, "samples": ITA, "species_recorded": kappa,
                   "seed": SEED, "events_per_step": ev, "events_per_step_incl_flagged": st_m["events"],
                   "flagged_simulations": st_m["flagged"]},
        "fp64_plain": {"value": fp64_value, "ms_per_step": wall_f / k_f * 1e3, "kernel_ms": kern_ms_f,
                       "events_per_step": valid_events(st_f), "flagged_simulations": st_f["flagged"],
                       "kernel": "FP64, table-driven (network read from a descriptor)"},
        "fp64_specialized": {"value": fp64s_value, "ms_per_step": wall_s / k_f * 1e3, "kernel_ms": kern_ms_s,
                             "events_per_step": valid_events(st_s), "flagged_simulations": st_s["flagged"],
                             "kernel": "FP64, network compiled in (same specialisation as the mixed value)"},
        # same simulations on both paths; their event counts differ with the random stream
        "speedup_vs_fp64": (wall_f / k_f) / (wall_m / args.steps),
        "speedup_vs_fp64_specialized": (wall_s / k_f) / (wall_m / args.steps),
        "events_per_s_ratio_vs_fp64": value / fp64_value,
        "accuracy_vs_fp64": acc,
        "launch_autotune": tune,
        "counters": {k: v for k, v in st_m.items() if k != "kernel_ms"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
