"""Throughput of the SOUFFLE step loop on one MI355X.

Workload (default): 2000 trimer_MVM subunits (data/model template, 210,000 beads) at
C = 800 uM, c = 100 mM, E = 2.1, dt = 0.005, K = 298, ks = 50, kb = 20, Nose-Hoover chain
q = 5, Q = 10, Yukawa cutoff 12/kappa, Verlet skin 0.3 sigma; seeded synthetic initial state
(non-overlapping random subunit positions and orientations, Maxwell-Boltzmann velocities).
A step is one iteration of the reference's step loop body (src/md.cpp:400-514) over all beads.
The metric is bead-steps/s = beads x timed steps / wall seconds of the timed steps, inputs
resident in HBM. Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "capsid-souffle_amd"))
sys.path.insert(0, os.path.join(REPO, "tests"))

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector


def force_kernel_bytes_per_bead(n_list, k_rebuild):
    """Algorithmic HBM bytes per bead of one k_force launch: own position 24, velocity read
    and write 48, force write 24, per-bead parameters 8, 4 bytes per list entry read, and the
    amortised list build (24 + 4 n_list) / k_rebuild. Neighbour positions are counted once per
    bead (as the bead's own position read) — re-reads are what `traffic` exposes."""
    return 104.0 + 4.0 * n_list + (24.0 + 4.0 * n_list) / max(k_rebuild, 1.0)


def run_gpu(args, brownian, steps, warmup, fp32=False):
    import souffle_gpu as sg
    from oracle_lib import TEMPLATE
    t = sg.Template(TEMPLATE)
    p = t.params(C_uM=800.0, c_mM=100.0, elj_att=2.1, dt=0.005, K=298.0, n_subunits=args.subunits,
                 brownian=1 if brownian else 0, q_chain=5, Q=10.0, damp=100.0)
    p.skin = args.skin
    p.pair_fp32 = 1 if fp32 else 0
    pos, vel = t.make_state(args.subunits, p.box, p.T, 2000)
    s = sg.System(t, args.subunits, p)
    s.set_state(pos, vel)
    s.step(warmup)
    st0 = s.stats()
    s.set_timing(True)
    t0 = time.perf_counter()
    s.step(steps)
    t1 = time.perf_counter()
    st1 = s.stats()
    s.set_timing(False)
    wall = t1 - t0
    nb = s.n_beads
    rebuilds = st1["rebuilds"] - st0["rebuilds"]
    n_list = (st1["lj_entries"] + st1["q_entries"]) / nb
    k_rebuild = steps / max(rebuilds, 1)
    return dict(nb=nb, wall=wall, value=nb * steps / wall, ms_per_step=1e3 * wall / steps,
                n_list=n_list, k_rebuild=k_rebuild, rebuilds=rebuilds, stats=st1)


def run_cpu(args):
    """The oracle (the reference algorithm restated: full per-bead lists with the subunit
    centre prefilter, OpenMP dynamic over subunits, serial integrator) on the host cores, on a
    bounded sample of the same workload: `--cpu-steps` steps after one warm-up step."""
    import souffle_gpu as sg
    from oracle_lib import TEMPLATE, Oracle
    t = sg.Template(TEMPLATE)
    p = t.params(C_uM=800.0, c_mM=100.0, elj_att=2.1, dt=0.005, K=298.0, n_subunits=args.subunits)
    pos, vel = t.make_state(args.subunits, p.box, p.T, 2000)
    threads = min(os.cpu_count() or 1, args.cpu_threads) if args.cpu_threads > 0 else (os.cpu_count() or 1)
    o = Oracle(args.subunits, C_uM=800.0, c_mM=100.0, elj_att=2.1, dt=0.005, K=298.0, L=0.3, B=5)
    o.set_state(pos, vel)
    o.time_steps(1, threads)
    secs = o.time_steps(args.cpu_steps, threads)
    return dict(value=o.N * args.cpu_steps / secs, threads=threads, secs=secs)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--subunits", type=int, default=2000)
    ap.add_argument("--skin", type=float, default=0.3)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the BD and FP32-pair runs")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the system would shard by "
              "subunit-owning spatial domains with a halo exchange; see the design notes).", file=sys.stderr)
        sys.exit(2)
    import souffle_gpu as sg
    if not os.path.exists(sg.LIB_PATH):
        import __graft_entry__
        __graft_entry__.build()

    md = run_gpu(args, brownian=False, steps=args.steps, warmup=args.warmup)
    st = md["stats"]
    t_force = st["k_force_ms"]
    b_force = force_kernel_bytes_per_bead(md["n_list"], md["k_rebuild"])
    achieved = b_force * md["nb"] / (t_force * 1e-3) / 1e9 if t_force > 0 else 0.0
    out = {
        "metric": "bead-steps/s",
        "value": md["value"],
        "unit": "bead-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": md["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"trimer_MVM x{args.subunits} plain FP64 Nose-Hoover MD (q=5), "
                               f"C=800uM c=100mM E=2.1 dt=0.005 K=298, skin {args.skin}",
                   "subunits": args.subunits, "beads": md["nb"], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_force", "kernel_ms": t_force, "bytes_per_bead": b_force,
                     "n_list": md["n_list"], "k_rebuild": md["k_rebuild"]},
        "kernels_ms": {"k_integrate": st["k_integrate_ms"], "k_force": t_force, "k_finalize": st["k_reduce_ms"]},
    }
    if not args.no_extra:
        bd = run_gpu(args, brownian=True, steps=args.steps, warmup=args.warmup)
        f32 = run_gpu(args, brownian=False, steps=args.steps, warmup=args.warmup, fp32=True)
        out["extra"] = {"bd_bead_steps_per_s": bd["value"], "bd_ms_per_step": bd["ms_per_step"],
                        "fp32_pair_bead_steps_per_s": f32["value"], "fp32_pair_ms_per_step": f32["ms_per_step"]}
    if not args.no_cpu:
        cpu = run_cpu(args)
        out["cpu_baseline"] = {"value": cpu["value"], "unit": "bead-steps/s", "cores": cpu["threads"],
                               "kind": "port",
                               "sample": f"{args.cpu_steps} NH steps of the same {args.subunits}-subunit state "
                                         f"(oracle, FP64, reference list algorithm at -L 0.3 -B 5), "
                                         f"{cpu['secs']:.1f} s"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
