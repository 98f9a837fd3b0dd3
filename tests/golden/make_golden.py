"""Writes the golden fixtures of this directory.

  philox4x32_10_kat.json   Random123 known-answer vectors for Philox4x32-10
                           (Salmon, Moraes, Dror, Shaw, SC'11; kat_vectors
                           of the Random123 distribution), counter/key/output
                           words as published
  schlogl_cme.npz          the exact law of the Schlogl network of the
                           examples (X0 = 250, A = 1e5, B = 2e5 held constant,
                           c = 3e-7, 1e-4, 1e-3, 3.5): the chemical master
                           equation on X = 0..1400 (outflow above 1400 is
                           below 1e-12 over t <= 10) propagated with
                           scipy.sparse.linalg.expm_multiply; pmf at
                           t = 1, 2, 5, 10
  oracle_trajectories.npz  FP64 oracle trajectories (oracle/pst_oracle.c) of
                           three small seeded workloads, with the oracle's
                           event / leap / SSA-step / flag counts:
                             bd  birth-death 0 -> X (10), X -> 0 (0.1), X0 = 0,
                                 pure SSA, 256 simulations, t_end = 100,
                                 101 samples, seed 0
                             lv  Lotka-Volterra X0 = Y0 = 1000, c = 10, 0.01,
                                 10, tau-leaping eps 0.03, n_c 10, burst 100,
                                 128 simulations, t_end = 10, 101 samples,
                                 seed 2
                             sc  Schlogl as above, tau-leaping, 128
                                 simulations, t_end = 10, 101 samples, seed 3
                           They pin the oracle against regressions and are
                           the bit-exact target of the FP64 GPU kernels.

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import expm_multiply

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

# Random123 kat_vectors, philox4x32_10 lines: counter (4 words), key (2 words), output (4 words)
PHILOX_KAT = [
    ([0x00000000, 0x00000000, 0x00000000, 0x00000000], [0x00000000, 0x00000000],
     [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF], [0xFFFFFFFF, 0xFFFFFFFF],
     [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]

SCHLOGL = dict(c1=3e-7, c2=1e-4, c3=1e-3, c4=3.5, A=1e5, B=2e5, x0=250, xmax=1400)
CME_TIMES = (1.0, 2.0, 5.0, 10.0)


def schlogl_cme():
    s = SCHLOGL
    x = np.arange(s["xmax"] + 1, dtype=float)
    birth = s["c1"] / 2 * s["A"] * x * (x - 1) + s["c3"] * s["B"]   # A + 2X -> 3X, B -> X
    death = s["c2"] / 6 * x * (x - 1) * (x - 2) + s["c4"] * x      # 3X -> A + 2X, X -> B
    birth[-1] = 0.0
    Q = sp.diags([-(birth + death), birth[:-1], death[1:]], [0, -1, 1], format="csc")
    p0 = np.zeros(x.size)
    p0[s["x0"]] = 1.0
    return np.array([expm_multiply(Q * t, p0) for t in CME_TIMES])


def oracle_runs():
    from oracle import oracle
    from py_stoch_amd import models
    from py_stoch_amd.mod.parser.tl_parser import tables_of
    out = {}
    cases = [
        ("bd", models.birth_death(), [0], 100.0, 101, 256, 0, oracle.METHOD_PURE_SSA),
        ("lv", models.lotka_volterra(), [0, 1], 10.0, 101, 128, 2, oracle.METHOD_TAU_LEAPING),
        ("sc", models.schlogl(), [0], 10.0, 101, 128, 3, oracle.METHOD_TAU_LEAPING),
    ]
    for name, spn, E, t_max, ita, U, seed, method in cases:
        a = models.make_args(spn, E=E, t_max=t_max, ita=ita, U=U)
        O, flags, st = oracle.simulate(tables_of(a), a.rate_fn(a.c), a.x_0, method=method, t_max=t_max, I=a.I,
                                       E=a.E, n_sims=U, seed=seed, n_c=10, eta=0.03, ssa_burst=100)
        out[name + "_O"] = O
        out[name + "_flags"] = flags
        out[name + "_stats"] = np.array([st["events"], st["leaps"], st["rejected_leaps"], st["ssa_steps"],
                                         st["flagged"], st["flagged_events"]], np.int64)
        out[name + "_settings"] = np.array([method, t_max, ita, U, seed], np.float64)
    return out


def main():
    with open(os.path.join(HERE, "philox4x32_10_kat.json"), "w") as f:
        json.dump({"source": "Random123 kat_vectors, philox4x32_10 (Salmon et al., SC'11)",
                   "vectors": [{"counter": c, "key": k, "output": o} for c, k, o in PHILOX_KAT]}, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "schlogl_cme.npz"), t=np.array(CME_TIMES), pmf=schlogl_cme(),
                        params=np.array([SCHLOGL[k] for k in ("c1", "c2", "c3", "c4", "A", "B", "x0", "xmax")]))
    np.savez_compressed(os.path.join(HERE, "oracle_trajectories.npz"), **oracle_runs())
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
