"""Throughput of the SOUFFLE step loop on one MI355X.

Workload (default): 2000 trimer_MVM subunits (data/model template, 210,000 beads) at
C = 800 uM, c = 100 mM, E = 2.1, dt = 0.005, K = 298, ks = 50, kb = 20, Nose-Hoover chain
q = 5, Q = 10, Yukawa cutoff 12/kappa, Verlet skin 0.3 sigma; seeded synthetic initial state
(non-overlapping random subunit positions and orientations, Maxwell-Boltzmann velocities).
A step is one iteration of the reference's step loop body (src/md.cpp:400-514) over all beads.
The metric is bead-steps/s = beads x timed steps / wall seconds of the timed steps, with the
state resident in HBM; steps run as HIP-graph chunks. A second pass of the same length runs
with HIP events around every 10th step's kernels (on the library's stream) for the per-kernel
durations the roofline uses. Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
# CPUs this process may use at start-up, before any GPU runtime is loaded (the CPU leg restores
# this mask if the process's mask was narrowed later)
START_AFFINITY = sorted(os.sched_getaffinity(0))
sys.path.insert(0, os.path.join(REPO, "capsid-souffle_amd"))
sys.path.insert(0, os.path.join(REPO, "tests"))

HBM_PEAK_GBS = 8000.0    # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector
# data/model trimer_MVM (charged), as shipped with the reference and copied under tests/golden
TEMPLATE = os.path.join(REPO, "tests", "golden", "infiles", "trimer_MVM")
CFG = dict(C_uM=800.0, c_mM=100.0, elj_att=2.1, dt=0.005, K=298.0, q_chain=5, Q=10.0, damp=100.0)


def force_kernel_bytes_per_bead(n_list):
    """Algorithmic HBM bytes per bead of one k_force launch: own position 24, velocity read and
    write 48, force write 24, per-bead parameters 8, and 4 bytes per list entry. Neighbour
    positions are not counted (ideally served from cache); `traffic` measures what is fetched."""
    return 104.0 + 4.0 * n_list


def step_bytes_per_bead(n_list, k_rebuild):
    """Whole-step algorithmic bytes per bead-step: x r/w 48, v r/w 48, F w+r 48, parameters 8,
    list entries 4 n_list, list build (24 + 4 n_list) / k_rebuild."""
    return 152.0 + 4.0 * n_list + (24.0 + 4.0 * n_list) / max(k_rebuild, 1.0)


def measured_traffic(args):
    """HBM bytes per k_force launch from the committed rocprofv3 PMC summary, when it was
    taken on this workload; None otherwise."""
    path = os.path.join(REPO, "profiles", "k_force_trimer2000_nh_fp64.traffic.txt")
    if args.subunits != 2000 or abs(args.skin - 0.3) > 1e-12 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("k_force_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def make_system(args, brownian=False, fp32=False, skin=None):
    import souffle_gpu as sg
    t = sg.Template(TEMPLATE)
    p = t.params(n_subunits=args.subunits, brownian=1 if brownian else 0, **CFG)
    p.skin = args.skin if skin is None else skin
    p.pair_fp32 = 1 if fp32 else 0
    pos, vel = t.make_state(args.subunits, p.box, p.T, 2000)
    s = sg.System(t, args.subunits, p)
    s.set_state(pos, vel)
    return s, pos, vel


def extended_energy(s):
    """Sum of the energy file's terms (kinetic, stretch, bend, LJ, Yukawa, thermostat potential
    and kinetic) per bead: the quantity a Nose-Hoover chain conserves."""
    return float(s.energies()[0].sum()) / s.n_beads


def run_gpu(args, brownian=False, fp32=False, skin=None, timed_pass=True):
    s, pos, vel = make_system(args, brownian, fp32, skin)
    if args.no_graph:
        # direct kernel launches throughout (the library launches event-timed chunks without a
        # graph); for kernel traces of the step kernels
        s.set_timing(True)
    e_start = extended_energy(s)
    tune = None
    s.set_graph_steps(args.graph_steps)
    s.step(args.warmup)
    if args.tune_tile:
        # the force kernel's pair tile by the library's autotuner at the fixed skin, after the
        # warm-up so that the trials see clusters forming (trials run from a saved state that is
        # restored; forces do not depend on the tile)
        tune, _ = s.autotune(skins=[s.skin(0)], tiles=[128, 256, 384, 512], trial_steps=200)
    # one untimed pass with the timed run's launch lengths (graph_steps-step launches and the
    # remainder): graph capture for those lengths and the list build the tuner's last setting
    # requests happen here, not in the timed window; its rate is reported as the first-window figure
    g = s.stats()["graph_steps"]
    n_pre = args.steps % g + (g if args.steps >= g else 0)
    t0 = time.perf_counter()
    s.step(n_pre)
    first = time.perf_counter() - t0
    st0 = s.stats()
    t0 = time.perf_counter()
    s.step(args.steps)
    t1 = time.perf_counter()
    st1 = s.stats()
    wall = t1 - t0
    nb = s.n_beads
    out = dict(nb=nb, wall=wall, value=nb * args.steps / wall, ms_per_step=1e3 * wall / args.steps,
               rebuilds=st1["rebuilds"] - st0["rebuilds"],
               graph=dict(steps_per_launch=g, launches_timed=st1["graph_launches"] - st0["graph_launches"]),
               n_list=(st1["lj_entries"] + st1["q_entries"]) / nb,
               first_window=dict(steps=n_pre, bead_steps_per_s=nb * n_pre / first, ms_per_step=1e3 * first / n_pre),
               steps_run=args.warmup + n_pre + args.steps)
    out["k_rebuild"] = args.steps / max(out["rebuilds"], 1)
    out["e_start"], out["e_end"] = e_start, extended_energy(s)
    out["tile"] = tune["tile"] if tune else None
    if timed_pass:
        # per-kernel means: HIP events around every 10th step's kernels (direct launches)
        s.set_timing(True)
        s.step(min(args.steps, args.event_steps))
        if not args.no_graph:
            s.set_timing(False)
        out["stats"] = s.stats()
    out["system"] = s
    return out


def host_cpu():
    """Host CPU facts for the baseline: lscpu model and physical cores, the CPUs this process may
    run on (at start-up and now), and the OpenMP thread count the CPU leg uses (physical cores
    within this process's start-up CPU set, capped by OMP_NUM_THREADS when the environment sets
    it). If the process's CPU mask was narrowed since start-up, the start-up mask is restored so
    that the oracle's threads are not confined to it."""
    info = {}
    try:
        import subprocess
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except Exception:
        pass
    def num(k):
        try:
            return int(info.get(k, ""))
        except ValueError:
            return None
    cps, sockets, tpc = num("Core(s) per socket"), num("Socket(s)"), num("Thread(s) per core")
    physical = cps * sockets if cps and sockets else None
    now = len(os.sched_getaffinity(0))
    if now < len(START_AFFINITY):
        os.sched_setaffinity(0, START_AFFINITY)
    allowed = len(START_AFFINITY)
    threads = allowed // (tpc or 1) if physical is None else min(allowed // (tpc or 1), physical)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        threads = min(threads, int(env))
    return dict(model=info.get("Model name"), physical_cores=physical, threads_per_core=tpc,
                cpus_allowed_at_start=allowed, cpus_allowed_before_leg=now, threads=max(1, threads))


def _omp_binding():
    # the oracle's OpenMP threads are left unbound (the default): binding them to places derived
    # from a narrowed mask would confine them to a few cores
    return {k: os.environ.get(k, "unset") for k in ("OMP_PROC_BIND", "OMP_PLACES", "OMP_NUM_THREADS")}


def run_cpu(args):
    """The oracle (the reference algorithm restated: full per-bead lists with the subunit
    centre prefilter, OpenMP dynamic over subunits, serial integrator) on the host cores, on a
    bounded sample of the same workload: `--cpu-steps` steps after one warm-up step, timed as
    three consecutive blocks of a third each; the median block rate is reported."""
    binding = _omp_binding()
    import souffle_gpu as sg
    from oracle_lib import Oracle  # the CPU leg only
    t = sg.Template(TEMPLATE)
    p = t.params(n_subunits=args.subunits, **CFG)
    pos, vel = t.make_state(args.subunits, p.box, p.T, 2000)
    host = host_cpu()
    threads = host["threads"]
    if args.cpu_threads > 0:
        threads = min(threads, args.cpu_threads)
    o = Oracle(args.subunits, C_uM=CFG["C_uM"], c_mM=CFG["c_mM"], elj_att=CFG["elj_att"], dt=CFG["dt"],
               K=CFG["K"], Q=CFG["Q"], q_chain=CFG["q_chain"], L=0.3, B=5)
    o.set_state(pos, vel)
    o.time_steps(1, threads)
    n = max(1, args.cpu_steps // 3)
    rates, secs = [], 0.0
    for _ in range(3):
        t = o.time_steps(n, threads)
        rates.append(o.N * n / t)
        secs += t
    return dict(value=sorted(rates)[1], threads=threads, secs=secs, steps=3 * n, host=host, binding=binding)


def probe_copy(mb):
    """A plain device-to-device copy of known size (for scaling FETCH_SIZE / WRITE_SIZE counters)."""
    import torch
    n = mb * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["md", "replicas", "surrogate"], default="md",
                    help="md: one 2000-subunit system (default); replicas: the 256-replica sweep; "
                         "surrogate: 5000 subunits, uncut Yukawa, far-field network vs exact")
    ap.add_argument("--steps", type=int, default=None, help="timed steps (md 20000, replicas 3000, surrogate 1000)")
    ap.add_argument("--sur-subunits", type=int, default=5000)
    ap.add_argument("--sur-salt", type=float, default=1.0,
                    help="salt (mM) of the surrogate workload: the largest far field in the network's range")
    ap.add_argument("--sur-c-lo", type=float, default=1.0, help="salt range (mM) the network is trained on")
    ap.add_argument("--sur-c-hi", type=float, default=20.0)
    ap.add_argument("--sur-samples", type=int, default=1000000)
    ap.add_argument("--sur-epochs", type=int, default=20)
    ap.add_argument("--sur-steps", type=int, default=500, help="timed steps of the surrogate summary in the md line")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (md 1000, others 200)")
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--rep-steps", type=int, default=500, help="timed steps of the replica summary in the md line")
    ap.add_argument("--tune-steps", type=int, default=100)
    ap.add_argument("--graph-steps", type=int, default=100, help="steps per graph launch of the md workload")
    ap.add_argument("--integrator", choices=["md", "bd"], default="md",
                    help="integrator of the md workload's main line (bd: Brownian dynamics, for its kernel traces)")
    ap.add_argument("--rep-one-launch", type=int, default=1,
                    help="replica batch: the timed steps as one graph launch (0: 100-step launches)")
    ap.add_argument("--cpu-replicas", type=int, default=8)
    ap.add_argument("--cpu-replica-steps", type=int, default=400)
    ap.add_argument("--subunits", type=int, default=2000)
    ap.add_argument("--skin", type=float, default=0.3)
    ap.add_argument("--event-steps", type=int, default=2000, help="steps of the event-timed pass (per-kernel means)")
    ap.add_argument("--tune-tile", type=int, default=1, help="autotune the force kernel's pair tile (0: library default)")
    ap.add_argument("--cpu-steps", type=int, default=150)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the BD and FP32-pair runs")
    ap.add_argument("--probe-copy-mb", type=int, default=0, help="run a known-size copy kernel first")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the step kernels directly instead of replaying HIP graphs (for kernel traces)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"replicas": 3000, "surrogate": 1000}.get(args.workload, 20000)
    if args.warmup is None:
        args.warmup = 1000 if args.workload == "md" else 200
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one large system would shard "
              "into subunit-owning spatial domains with one halo exchange per step).", file=sys.stderr)
        sys.exit(2)
    import souffle_gpu as sg
    if not os.path.exists(sg.LIB_PATH):
        import __graft_entry__
        __graft_entry__.build()
    if args.probe_copy_mb > 0:
        probe_copy(args.probe_copy_mb)

    if args.workload == "replicas":
        out = bench_replicas(args)
    elif args.workload == "surrogate":
        out = bench_surrogate(args)
    else:
        out = bench_md(args)
    print(json.dumps(out), flush=True)


FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA


def flops_per_launch(kernel, workload):
    """(fp64, fp32, bf16_mfma) FLOPs per launch of `kernel` on `workload` from the committed
    rocprofv3 PMC summary, or None."""
    path = os.path.join(REPO, "profiles", "flops_per_launch.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 5 and f[0] == kernel and f[1] == workload:
            return tuple(float(x) for x in f[2:])
    return None


def flop_rate(kernel, workload, ms):
    """Counted FLOP rates of one kernel against the MI355X peaks, from its FLOPs per launch
    (PMC) and its event-timed mean launch duration."""
    f = flops_per_launch(kernel, workload)
    if f is None or not ms or ms <= 0:
        return None
    out = {"kernel_ms": ms, "workload": workload}
    for name, flops, peak in (("fp64", f[0], FP64_PEAK_TFLOPS), ("fp32", f[1], FP32_PEAK_TFLOPS),
                              ("bf16_mfma", f[2], BF16_PEAK_TFLOPS)):
        if flops > 0:
            tf = flops / (ms * 1e-3) / 1e12
            out[name] = {"flop_per_launch": flops, "tflops": tf, "peak_tflops": peak, "frac": tf / peak}
    return out


def roofline_entry(nb, n_list, k_rebuild, t_force, value, traffic):
    b_force = force_kernel_bytes_per_bead(n_list)
    achieved = b_force * nb / (t_force * 1e-3) / 1e9 if t_force > 0 else 0.0
    b_step = step_bytes_per_bead(n_list, k_rebuild)
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE+WRITE_SIZE, copy-scaled)",
            "algorithmic_bytes_per_launch": b_force * nb,
            "kernel": "k_force", "kernel_ms": t_force, "bytes_per_bead": b_force,
            "n_list": n_list, "k_rebuild": k_rebuild,
            "step_bytes_per_bead": b_step, "step_achieved_GBs": b_step * value / 1e9}


def bench_replicas(args):
    """A seeded sweep of uneven replicas in one handle: subunit counts uniform in [20, 400],
    salt 50-800 mM, LJ depth 1.5-2.5, temperature 280-320 K, C = 800 uM, FP32 pair / FP64
    integrate, NH MD; skin and gather tile picked by the library's autotuner. value is the
    aggregate bead-steps/s. A few replicas are rerun alone and compared bit for bit."""
    import numpy as np
    import souffle_gpu as sg
    import replica_sweep
    t = sg.Template(TEMPLATE)
    sizes, params, pos, vel, draws = replica_sweep.sweep(t, args.replicas, 4)
    batch = sg.System(t, sizes, params)
    batch.set_state(np.concatenate(pos), np.concatenate(vel))
    skins = getattr(args, "tune_skins", [0.2, 0.3, 0.5, 0.8, 1.2])
    tiles = getattr(args, "tune_tiles", [128, 256, 384, 512])
    tune, trial_ms = batch.autotune(skins=skins, tiles=tiles, trial_steps=args.tune_steps)
    # the timed steps as one graph launch (no host round trip inside the timed window) when they
    # fit one graph; an untimed pass of the same length captures it first (and runs the list
    # build the tuner's last setting requests)
    g = min(args.steps, 20000) if args.rep_one_launch else 100
    batch.set_graph_steps(g)
    n_pre = args.steps % g + (g if args.steps >= g else 0)
    batch.step(n_pre)
    batch.step(args.warmup)
    st0 = batch.stats()
    t0 = time.perf_counter()
    batch.step(args.steps)
    wall = time.perf_counter() - t0
    st1 = batch.stats()
    nb = batch.n_beads
    value = nb * args.steps / wall
    n_list = (st1["lj_entries"] + st1["q_entries"]) / nb
    rebuilds_per_rep = (st1["rebuilds"] - st0["rebuilds"]) / len(sizes)
    k_rebuild = args.steps / max(rebuilds_per_rep, 1)
    bp, bv = batch.get_state()
    # bit-for-bit check of the smallest, largest, lowest-salt and highest-salt replicas
    salts = [d["c_mM"] for d in draws]
    picks = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), int(np.argmin(salts)), int(np.argmax(salts))})
    offs = np.concatenate([[0], np.cumsum(sizes)]) * t.n_beads
    checked = []
    for k in picks:
        single = sg.System(t, sizes[k], params[k])
        single.set_state(pos[k], vel[k])
        single.step(n_pre + args.warmup + args.steps)
        sp, sv = single.get_state()
        same = bool(np.array_equal(bp[offs[k]:offs[k + 1]], sp) and np.array_equal(bv[offs[k]:offs[k + 1]], sv))
        checked.append({"replica": k, "subunits": sizes[k], "c_mM": round(salts[k], 1), "bitwise_equal": same})
        single.close()
    batch.set_timing(True)
    batch.step(min(args.steps, 1000))
    batch.set_timing(False)
    st = batch.stats()
    out = {
        "metric": "bead-steps/s", "value": value, "unit": "bead-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 pairs, f64 integrate", "data": "synthetic",
        "config": {"workload": f"{len(sizes)} replicas of trimer_MVM, subunits U[20,400] (total {sum(sizes)}), "
                               f"c U[50,800] mM, E U[1.5,2.5], K U[280,320], C=800uM dt=0.005, NH q=5, "
                               f"FP32 pair arithmetic, autotuned skin/tile",
                   "replicas": len(sizes), "subunits": sum(sizes), "beads": nb, "parallelism": "single GPU"},
        "roofline": roofline_entry(nb, n_list, k_rebuild, st["k_force_ms"], value, None),
        "kernels_ms": {"k_integrate": st["k_integrate_ms"], "k_rebuild": st["k_rebuild_ms"],
                       "k_force": st["k_force_ms"], "k_finalize": st["k_reduce_ms"]},
        "autotune": {"skin": tune["skin"], "tile": tune["tile"], "trial_ms_per_step": trial_ms.round(4).tolist(),
                     "skins": skins, "tiles": tiles},
        "single_run_check": checked,
        "graph": {"steps_per_launch": g, "launches_timed": st1["graph_launches"] - st0["graph_launches"]},
    }
    if not args.no_cpu:
        from oracle_lib import Oracle  # the CPU leg only
        # the oracle on a seeded sample of the replicas, one after another on all host cores
        # (the sequential parameter-sweep equivalent), each at its smallest valid -L, -B 5
        rng = np.random.default_rng(9)
        sample = sorted(rng.choice(len(sizes), size=min(args.cpu_replicas, len(sizes)), replace=False).tolist())
        binding = _omp_binding()
        host = host_cpu()
        threads = host["threads"]
        if args.cpu_threads > 0:
            threads = min(threads, args.cpu_threads)
        work = secs = 0.0
        for k in sample:
            p, d = params[k], draws[k]
            L = replica_sweep.reference_list_radius(TEMPLATE, p.ecut, p.ecut_el)
            o = Oracle(sizes[k], C_uM=800.0, c_mM=d["c_mM"], elj_att=d["elj_att"], K=d["K"], dt=0.005,
                       Q=10.0, q_chain=5, L=L, B=5)
            o.set_state(pos[k], vel[k])
            o.time_steps(1, threads)
            secs += o.time_steps(args.cpu_replica_steps, threads)
            work += o.N * args.cpu_replica_steps
            o.close()
        out["cpu_baseline"] = {"value": work / secs, "unit": "bead-steps/s", "cores": threads, "kind": "port",
                               "host": host, "omp": binding,
                               "sample": f"{len(sample)} of the replicas, {args.cpu_replica_steps} NH steps each, "
                                         f"sequentially (oracle, FP64, reference list algorithm at each replica's "
                                         f"smallest valid -L, -B 5), {secs:.1f} s"}
    return out


def run_surrogate(args, steps, warmup):
    """Uncut Yukawa at 5000 subunits (C = 800 uM, c = args.sur_salt mM, other parameters as the
    default workload): the plain FP64 path (exact far field, subunit-pair tiles) against FP32 pair
    arithmetic with the far field from the network (trained here on the GPU), from the same
    seeded state. The accuracy bar is 1e-3 relative RMS error of the total force; the library's
    audit measures the network's error relative to the far field it replaces, so its tolerance is
    set to 1e-3 / (the far field's share of the force on the plain run's warmed state), the
    far-field error that keeps the total-force error at 1e-3."""
    import numpy as np
    import souffle_gpu as sg
    t = sg.Template(TEMPLATE)
    ns = args.sur_subunits
    model = sg.FarModel(t, 20.0, args.sur_c_lo, args.sur_c_hi, seed=2)
    t0 = time.perf_counter()
    rep = model.train(n_samples=args.sur_samples, epochs=args.sur_epochs, batch=4096, lr=2e-3, seed=77)
    train_s = time.perf_counter() - t0
    cfg = dict(CFG, c_mM=args.sur_salt, ecut_c=0.0)
    out = {"train": rep, "train_wall_s": train_s}
    rel = lambda a, b: float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))
    systems, tol = {}, None
    for name, fp32, use_model in (("fp64_exact", False, False), ("fp32_model", True, True)):
        p = t.params(n_subunits=ns, **cfg)
        p.far_rf = 20.0
        p.skin = args.skin
        p.pair_fp32 = 1 if fp32 else 0
        pos, vel = t.make_state(ns, p.box, p.T, 2000)
        s = sg.System(t, ns, p)
        s.set_state(pos, vel)
        if args.no_graph:
            s.set_timing(True)  # direct launches throughout (for kernel traces and counters)
        if use_model:
            s.set_far_model(model, tol=tol)
        s.step(warmup)
        if not use_model:
            # the far field's share of the force and the network's error on the warmed state (on
            # the start state the subunits sit at their template geometry, the bonded forces are
            # near zero and the far field's share is tens of times its share in a running system)
            f64 = s.forces()[0]
            s.far_drop(True)
            fn = s.forces()[0]
            s.far_drop(False)
            s.set_far_model(model, tol=1e9)
            fm = s.forces()[0]
            s.set_far_model(None)
            share = rel(fn, f64)
            out["warmed_state"] = {"steps": warmup, "far_field_share_of_force": share,
                                   "model_err_rel_far_field": rel(fm, f64) / share, "model_err_rel_force": rel(fm, f64)}
            tol = 1e-3 / share
            out["audit_tol_rel_far_field"] = tol
        t1 = time.perf_counter()
        s.step(steps)
        wall = time.perf_counter() - t1
        s.set_timing(True)
        s.step(min(steps, 500))
        if not args.no_graph:
            s.set_timing(False)
        st = s.stats()
        out[name] = {"bead_steps_per_s": s.n_beads * steps / wall, "ms_per_step": 1e3 * wall / steps,
                     "kernels_ms": {"k_integrate": st["k_integrate_ms"], "k_rebuild": st["k_rebuild_ms"],
                                    "k_far": st["k_far_ms"], "k_force": st["k_force_ms"],
                                    "k_finalize": st["k_reduce_ms"]}}
        if use_model:
            out[name]["audits"] = s.far_stats()
        systems[name] = s
    out["speedup_vs_fp64_exact"] = out["fp32_model"]["bead_steps_per_s"] / out["fp64_exact"]["bead_steps_per_s"]
    # force error of the surrogate form on one state (the plain run's last state)
    pos, vel = systems["fp64_exact"].get_state()
    ref = systems["fp64_exact"]
    ref.set_state(pos, vel)
    f64 = ref.forces()[0]
    sur = systems["fp32_model"]
    sur.set_state(pos, vel)
    fs = sur.forces()[0]
    ref.far_drop(True)
    fn = ref.forces()[0]
    ref.far_drop(False)
    out["force_rel_rms_err"] = rel(fs, f64)
    out["far_field_share_of_force"] = rel(fn, f64)  # the error of leaving the far field out
    # the network's error relative to the far field it replaces (FP64 pairs, network far field)
    ref.set_far_model(model, tol=1e9)
    fm = ref.forces()[0]
    ref.set_far_model(None)
    out["model_err_rel_far_field"] = rel(fm, f64) / out["far_field_share_of_force"]
    out["workload"] = (f"trimer_MVM x{ns}, C=800uM c={args.sur_salt}mM, uncut Yukawa (-e 0 over the box), "
                       f"far field beyond 20 sigma centre distance, NH MD, skin {args.skin}")
    out["beads"] = systems["fp64_exact"].n_beads
    for x in systems.values():
        x.close()
    return out


def bench_surrogate(args):
    res = run_surrogate(args, args.steps, args.warmup)
    return {
        "metric": "bead-steps/s", "value": res["fp32_model"]["bead_steps_per_s"], "unit": "bead-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["fp32_model"]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 pairs, bf16 far-field network (f32 accumulate), f64 integrate", "data": "synthetic",
        "config": {"workload": res["workload"] + "; far field from the network", "subunits": args.sur_subunits,
                   "beads": res["beads"], "parallelism": "single GPU"},
        "plain_fp64": res["fp64_exact"], "surrogate_fp32": res["fp32_model"],
        "speedup_vs_fp64_exact": res["speedup_vs_fp64_exact"], "force_rel_rms_err": res["force_rel_rms_err"],
        "far_field_share_of_force": res["far_field_share_of_force"], "training": res["train"],
        "audit_tol_rel_far_field": res["audit_tol_rel_far_field"], "warmed_state": res["warmed_state"],
        "flop_rates": {k: v for k, v in (
            ("k_far_exact", flop_rate("k_far_exact", "trimer5000_uncut", res["fp64_exact"]["kernels_ms"]["k_far"])),
            ("k_far_mlp", flop_rate("k_far_mlp", "trimer5000_uncut", res["fp32_model"]["kernels_ms"]["k_far"])))
            if v is not None and args.sur_subunits == 5000},
        "training_wall_s": res["train_wall_s"],
    }


def run_drift(args, n_steps=20000, block=5000):
    """The FP32-pair form against the FP64 path from the same seeded state (config 3): both run
    n_steps; every `block` steps the force of the FP32 form on the FP64 trajectory's state is
    compared with the FP64 force on it (relative RMS over all beads), and the extended
    Nose-Hoover energy per bead is taken at the start and the end of each trajectory."""
    import numpy as np
    import souffle_gpu as sg
    s64, pos, vel = make_system(args)
    s32, _, _ = make_system(args, fp32=True)
    cmp32, _, _ = make_system(args, fp32=True)
    e0 = {"fp64": extended_energy(s64), "fp32_pair": extended_energy(s32)}
    errs, done, stopped = [], 0, None
    try:
        while done < n_steps:
            p64, v64 = s64.get_state()
            cmp32.set_state(p64, v64)
            f64 = s64.forces()[0]
            f32 = cmp32.forces()[0]
            errs.append({"step": done, "force_rel_rms_err": float(np.sqrt(np.sum((f32 - f64) ** 2) / np.sum(f64 ** 2)))})
            k = min(block, n_steps - done)
            s64.step(k)
            s32.step(k)
            done += k
    except sg.SfError as e:
        stopped = str(e)
    out = {"steps": done, "force_err_samples": errs,
           "fp64_energy_per_bead": [e0["fp64"], extended_energy(s64)],
           "fp32_pair_energy_per_bead": [e0["fp32_pair"], extended_energy(s32)]}
    out["fp64_energy_drift_per_bead"] = out["fp64_energy_per_bead"][1] - out["fp64_energy_per_bead"][0]
    out["fp32_pair_energy_drift_per_bead"] = out["fp32_pair_energy_per_bead"][1] - out["fp32_pair_energy_per_bead"][0]
    if stopped:
        out["stopped"] = stopped
    for x in (s64, s32, cmp32):
        x.close()
    return out


def replica_summary(args):
    """Config 4 in the default line: the 256-replica sweep with a reduced tuner grid and a short
    timed run; the CPU leg on a small sample of the replicas."""
    import copy
    a = copy.copy(args)
    a.steps, a.warmup, a.tune_steps = args.rep_steps, 100, 100
    a.tune_skins, a.tune_tiles = [0.3, 0.5, 0.8], [256, 384, 512]
    a.cpu_replicas, a.cpu_replica_steps = 4, 100
    r = bench_replicas(a)
    keep = {k: r[k] for k in ("value", "ms_per_step", "steps", "warmup", "autotune", "single_run_check", "kernels_ms",
                              "graph")}
    keep["workload"] = r["config"]["workload"]
    keep["beads"] = r["config"]["beads"]
    keep["roofline_frac"] = r["roofline"]["frac"]
    if "cpu_baseline" in r:
        keep["cpu_baseline"] = r["cpu_baseline"]
    return keep


def bench_md(args):
    bd_main = args.integrator == "bd"
    md = run_gpu(args, brownian=bd_main)
    st = md["stats"]
    t_force = st["k_force_ms"]
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
        "config": {"workload": f"trimer_MVM x{args.subunits} plain FP64 "
                               + ("Brownian dynamics (-D b -d 100), " if bd_main else "Nose-Hoover MD (q=5, Q=10), ") +
                               f"C=800uM c=100mM E=2.1 dt=0.005 K=298 ks=50 kb=20, skin {args.skin}",
                   "subunits": args.subunits, "beads": md["nb"], "tile": md["tile"], "parallelism": "single GPU"},
        "roofline": roofline_entry(md["nb"], md["n_list"], md["k_rebuild"], t_force, md["value"],
                                   measured_traffic(args)),
        "kernels_ms": {"k_integrate": st["k_integrate_ms"], "k_rebuild": st["k_rebuild_ms"],
                       "k_force": t_force, "k_finalize": st["k_reduce_ms"]},
        # the event-timed pass launches the step kernels directly; each mean also holds the
        # cost of its two event records (~3 us against a kernel trace of the same launches)
        "kernel_ms_note": "HIP-event means over direct launches, event cost included",
        # counted FLOP rates (rocprofv3 PMC FLOPs per launch / event-timed mean; issue-slot FLOPs)
        "flop_rates": {k: v for k, v in (
            ("k_force", flop_rate("k_force", "trimer2000_nh_fp64", t_force) if args.subunits == 2000 else None),
            ("k_integrate", flop_rate("k_integrate", "trimer2000_nh_fp64", st["k_integrate_ms"])
             if args.subunits == 2000 else None)) if v is not None},
        # the untimed pass before the timed window (graph capture of its chunk lengths, the tuner's
        # list build)
        "first_window": md["first_window"],
        "graph": md["graph"],
    }
    if not args.no_extra:
        bd = run_gpu(args, brownian=True, timed_pass=False)
        f32 = run_gpu(args, fp32=True, timed_pass=False)
        out["extra"] = {"bd_bead_steps_per_s": bd["value"], "bd_ms_per_step": bd["ms_per_step"],
                        "fp32_pair_bead_steps_per_s": f32["value"], "fp32_pair_ms_per_step": f32["ms_per_step"],
                        "first_window": {"fp64_md": md["first_window"], "bd": bd["first_window"],
                                         "fp32_pair": f32["first_window"]}}
        for r in (bd, f32):
            r["system"].close()
        md["system"].close()
        out["extra"]["fp32_pair_form"] = run_drift(args)
        out["extra"]["replicas_256"] = replica_summary(args)
        sur = run_surrogate(args, args.sur_steps, 100)
        out["extra"]["surrogate_5000"] = {
            "workload": sur["workload"],
            "fp64_exact_bead_steps_per_s": sur["fp64_exact"]["bead_steps_per_s"],
            "fp32_model_bead_steps_per_s": sur["fp32_model"]["bead_steps_per_s"],
            "speedup_vs_fp64_exact": sur["speedup_vs_fp64_exact"],
            "force_rel_rms_err": sur["force_rel_rms_err"],
            "far_field_share_of_force": sur["far_field_share_of_force"],
            "model_err_rel_far_field": sur["model_err_rel_far_field"],
            "audit_tol_rel_far_field": sur["audit_tol_rel_far_field"], "warmed_state": sur["warmed_state"],
            "k_far_ms": {"exact": sur["fp64_exact"]["kernels_ms"]["k_far"], "model": sur["fp32_model"]["kernels_ms"]["k_far"]},
            "flop_rates": {k: v for k, v in (
                ("k_far_exact", flop_rate("k_far_exact", "trimer5000_uncut", sur["fp64_exact"]["kernels_ms"]["k_far"])),
                ("k_far_mlp", flop_rate("k_far_mlp", "trimer5000_uncut", sur["fp32_model"]["kernels_ms"]["k_far"])))
                if v is not None and args.sur_subunits == 5000},
            "audits": sur["fp32_model"]["audits"], "training_wall_s": sur["train_wall_s"],
            "val_rel_err_force": sur["train"]["val_rel_err_force"]}
    for k in ("system",):
        md.pop(k, None)
    if not args.no_cpu:
        cpu = run_cpu(args)
        out["cpu_baseline"] = {"value": cpu["value"], "unit": "bead-steps/s", "cores": cpu["threads"],
                               "kind": "port", "host": cpu["host"], "omp": cpu["binding"],
                               "sample": f"{cpu['steps']} NH steps of the same {args.subunits}-subunit state "
                                         f"(oracle, FP64, the reference list algorithm at -L 0.3 -B 5), "
                                         f"{cpu['secs']:.1f} s; median of three blocks of {cpu['steps'] // 3}"}
    return out


if __name__ == "__main__":
    main()
