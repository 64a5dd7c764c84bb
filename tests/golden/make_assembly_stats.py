"""Writes tests/golden/assembly_stats.json: time series of energies and oligomer statistics of
the CPU oracle (oracle/liboracle.so, FP64 restatement of the reference path) for an ensemble of
seeded 20-subunit trimer_MVM systems at the small-system parameters (C = 800 uM, c = 100 mM,
E = 2.1, dt = 0.005, K = 298, ks = 50, kb = 20, -e 12), Nose-Hoover MD (q = 5, Q = 10) and
Brownian dynamics (damp 100). The GPU test compares its own ensemble statistics with these.

Run from the repository root: python tests/golden/make_assembly_stats.py  (about 10 minutes on 8 cores).
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_lib import TEMPLATE, Oracle  # noqa: E402
from souffle_states import make_state, read_template_xyz  # noqa: E402

N_SUB = 20
SEEDS = list(range(100, 116))
STEPS = 10000
EVERY = 250
PARAMS = dict(C_uM=800.0, c_mM=100.0, elj_att=2.1, dt=0.005, K=298.0, ks=50.0, kb=20.0, ecut_c=12.0, Q=10.0,
              q_chain=5, damp=100.0)


def run(seed, brownian):
    o = Oracle(N_SUB, L=0.3, B=5, brownian=brownian, **PARAMS)
    pos, vel = make_state(read_template_xyz(TEMPLATE)[0], N_SUB, o.box, seed, T=o.T)
    o.set_state(pos, vel)
    rows = []
    for k in range(STEPS // EVERY + 1):
        if k:
            o.run(EVERY)
        e = o.energies()
        _, ms, n = o.oligomers(2.0)
        largest = int(np.nonzero(ms)[0].max()) + 1
        rows.append([k * EVERY] + [float(v) for v in e] + [n, largest])
    return rows


def main():
    out = {"n_subunits": N_SUB, "seeds": SEEDS, "steps": STEPS, "every": EVERY, "params": PARAMS,
           "columns": ["step", "ke", "se", "be", "ne", "ce", "tpe", "tke", "n_oligomers", "largest"],
           "md": [], "bd": []}
    for brownian, key in [(False, "md"), (True, "bd")]:
        for seed in SEEDS:
            t0 = time.time()
            out[key].append(run(seed, brownian))
            print(key, seed, f"{time.time() - t0:.1f} s", out[key][-1][-1], flush=True)
    with open(os.path.join(HERE, "assembly_stats.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))


if __name__ == "__main__":
    main()
