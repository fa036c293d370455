"""Benchmark: flip attempts/ns of the multicanonical weight iteration.

Default workload (plain method at reference scale on one GPU): L = 72 (5184
spins, 5185 D-bins), 4096 walkers, random initial spins (weight keying, seed
1234), ln W = 0 initially, recursive weight update, therm = N and NUPDATES = 1
per iteration with the reference's x1.01 schedule once the histogram spans all
bins.  One step = one weight iteration, the loop body of
muca_delta_multigpu.cu:374-477: one launch of the update kernel (therm +
NUPDATES attempts per walker) and the device weight update (hist_range, width,
D_KL, weight_recursion, normalisation, float copy, schedule), replayed from a
captured hipGraph with nothing returning to the host between iterations.

`--workload batch96`: 96 independent instances, L in {32,48,64,72,96,128} x
Delta in {1.90,1.95,1.98,2.00} x T in {0.5,0.6,0.7,0.8}, seed = instance id,
`--walkers` walkers each (default 4096, as the L = 72 workload), autotuned
tiles; one step = one weight iteration of every instance.  (The weight
problem in D = sum s^2 depends on T only; Delta enters at reweighting, so
instances that share (L, T) differ by seed.)

Prints one JSON line.  `value` is whole-step throughput (attempts / wall time
of the timed steps); `roofline` prices the update kernel alone (HIP events on
the batch's stream, over as many further iterations launched eagerly) against
HBM at 2 algorithmic bytes per attempt; `roofline_valu` prices the same
kernel rate against the binding ceiling, the Philox4x32-7 block rate measured
in this run by muca_amd/libmuca_probe.so (an attempt needs half a block);
`host_loop` is the same workload driven the reference's way (host weight
update between launches); `cpu_baseline` times the CPU oracle (the same
chain, threaded over walkers) on a bounded sample of the same workload and
compares its histograms with the timed batched path's.

`--workload seeded`: GPU seconds to a flat histogram (width == N and
D_KL < 1e-4, muca_delta_multigpu.cu:441-447) for an unseen (L, T, Delta) =
(`--L` 128, 0.55, 1.97), trivial update, `--jacks` 4 independent jacks of
`--walkers`/4 walkers each in one batch (default total: the reference's
walker count for one GPU, maxThreadsPerCU x CUs, muca_delta_multigpu.cu:100-106;
524 288 on MI355X); a mode's time runs until every jack is flat.  "seeded":
ln W from the surrogate (trained on the GPU on the committed corpus of
converged weights; the jacks alternate between two networks), stratified
start, NUPDATES from `--s0` sweeps growing x`--growth` per iteration up to
`--smax` sweeps; "nonet": the same from ln W = 0; "plain": the reference's
defaults (ln W = 0, NUPDATES = 1, x1.01 once width == N, random start).
Each mode stops at its device-time cap; one that hits it reports a lower
bound.  ln W accuracy: the per-bin jackknife error over the jacks.

`--workload c5`: large production, L = 256 (65 536 spins), 16 384 walkers,
T = 0.5, produce keying and FP64 acceptance, ln W from the surrogate, one
launch of therm = N and NUPDATES = N per walker into ten per-block sets of
exact per-D moments (count, sum |M|, sum M^2, sum M^4; 4.2 MB each) or,
with `--observables joint`, the reference's joint histograms H(D, M) (ten x
17.2 GB in HBM), then FP64 reweighting on the device to <|M|>, chi and the
Binder cumulant over a Delta grid with jackknife errors; block 0 at
Delta = 1.98 is checked against the CPU oracle's long-double sums.

    python bench.py [--steps K] [--warmup W] [--walkers 4096] [--L 72]
    python bench.py --workload batch96 [--walkers 4096]
    python bench.py --workload seeded [--L 128] [--mode all|both|seeded|plain|nonet]
    python bench.py --workload c5 [--L 256] [--walkers 16384]
    python bench.py --copy-probe BYTES   # plain copy kernel for counter scaling
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
ALGO_BYTES_PER_ATTEMPT = 2.0   # one int8 spin read + one write per site visit
METRIC = "flip attempts/ns; GPU s to flat H(M) at L=128, surrogate-seeded vs plain"
# the sources that define the team update kernel; the committed counter file
# (profiles/update_kernel_L72_W4096_traffic.json) records their hash, and its
# figures are reported only while the hash still matches
TEAM_KERNEL_SOURCES = ["team_run.h", "muca_team.hip", "kernels.h", "philox.h", "acclog.h",
                       "acclog_table.h"]


def team_kernel_source_sha():
    import hashlib
    h = hashlib.sha256()
    for f in TEAM_KERNEL_SOURCES:
        with open(os.path.join(ROOT, "muca_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def valu_ceiling(rounds=7, grid=2048, iters=20000, reps=5):
    """The Philox4x32-R block rate of muca_amd/libmuca_probe.so's
    microbenchmark (every lane of a full-occupancy grid draws the two keyed
    streams of a walker and nothing else).  An attempt needs half a block, so
    twice the block rate bounds any update kernel's attempts/ns."""
    import ctypes
    path = os.path.join(ROOT, "muca_amd", "libmuca_probe.so")
    lib = ctypes.CDLL(path)
    f = lib.muca_probe_philox_rate
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float)]
    bpn, ms = ctypes.c_double(), ctypes.c_float()
    rc = f(0, rounds, grid, iters, reps, ctypes.byref(bpn), ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(f"muca_probe_philox_rate failed with HIP error {rc}")
    return {"philox_blocks_per_ns": bpn.value, "attempts_per_ns": 2.0 * bpn.value,
            "launch_ms": ms.value, "lanes": grid * 256, "blocks_per_lane": 2 * iters,
            "rounds": rounds}


def host_cpu():
    """Threads for the CPU baseline (this process's CPUs, at most the 16 a
    one-GPU share of the host holds) and the CPU model."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return min(avail, 16), avail, model


def batch_hists(M, ins, n):
    """The timed batched path (same instance, graph-replayed iterations) run
    again from its start, returning the D histogram of each of its first n
    iterations."""
    out = []
    with M.Batch(ins) as b:
        b.init(0)
        for _ in range(n):
            b.iterate(1, use_graph=True)
            out.append(b.get_hist(0))
    return out


def copy_probe(nbytes):
    import torch
    x = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    x.fill_(1.0)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "launches": 5}))


def cpu_weight_loop(L, W, seed, recursion, budget_s, nthreads, beta=2.0, max_iters=10000):
    """The oracle's weight loop on the host: iterations until budget_s seconds
    of wall time are spent.  Returns (attempts, seconds, iterations, hists)."""
    from oracle import oracle as O
    N = L * L
    lat = O.init_lattice(O.KEYING_WEIGHT, L, W, seed=seed)
    E, D, _ = O.state(L, lat)
    lnw = np.zeros(N + 1)
    lnw32 = lnw.astype(np.float32)
    p_acc = np.zeros(N + 1)
    width, nup, attempts, k = 10, 1, 0, 0
    hists = []
    t0 = time.perf_counter()
    while True:
        if width == N:
            nup = int((nup + 1) * 1.01)
        h = O.mcs(L, lat, E, D, None, keying=0, fp64=False, lnw32=lnw32, iter=k, seed=seed,
                  therm=N, nupdates=nup, beta=beta, nthreads=nthreads)
        attempts += W * (N + nup)
        hists.append(h)
        s, e = O.hist_range(h)
        width = max(width, (e - s) & 0xFFFFFFFF)
        if recursion:
            O.weight_recursion(h, p_acc, s, e, lnw)
        else:
            O.update_trivial(h, lnw)
        lnw32 = O.normalise(lnw)
        k += 1
        el = time.perf_counter() - t0
        if el >= budget_s or k >= max_iters:
            break
    return attempts, el, k, hists


def host_loop(M, L, W, seed, recursion, steps, warmup, kernel):
    """The reference's arrangement: update launch, D2H histogram, host weight
    update, H2D weights.  Returns (attempts/ns, mean kernel ms, hists)."""
    N = L * L
    ctx = M.Context(L, W, device=0, kernel=kernel)
    ctx.init_lattice(mm=0, gpu=0, seed=seed)
    ctx.compute_state()
    lnw, p_acc = np.zeros(N + 1), np.zeros(N + 1)
    ctx.set_log_weight(lnw)
    st = dict(width=10, nup=1, k=0)
    hists = []

    def step():
        if st["width"] == N:
            st["nup"] = int((st["nup"] + 1) * 1.01)
        hist, ms = ctx.run(st["k"], seed, N, st["nup"])
        hists.append(hist)
        s, e = M.hist_range(hist)
        st["width"] = max(st["width"], (e - s) & 0xFFFFFFFF)
        M.kull_leibler(hist)
        if recursion:
            M.update_recursive(hist, s, e, p_acc, lnw)
        else:
            M.update_trivial(hist, lnw)
        M.normalise(lnw)
        ctx.set_log_weight(lnw)
        st["k"] += 1
        return W * (N + st["nup"]), ms

    for _ in range(warmup):
        step()
    att, kms = 0, []
    t0 = time.perf_counter()
    for _ in range(steps):
        a, ms = step()
        att += a
        kms.append(ms)
    wall = time.perf_counter() - t0
    ctx.close()
    return att / wall / 1e9, float(np.mean(kms)), hists


def batch96(walkers):
    Ls = [32, 48, 64, 72, 96, 128]
    deltas = [1.90, 1.95, 1.98, 2.00]
    Ts = [0.5, 0.6, 0.7, 0.8]
    ins, meta = [], []
    for L in Ls:
        for d in deltas:
            for T in Ts:
                i = len(ins)
                ins.append(dict(L=L, walkers=walkers, seed=i, beta=1.0 / T, recursion=True))
                meta.append((L, d, T))
    return ins, meta


def run_batch(M, ins, steps, warmup, autotune):
    """Times `steps` graph-replayed iterations, then as many eager iterations
    with per-kernel HIP events.  Returns a dict of measurements."""
    with M.Batch(ins) as b:
        tune = None
        if autotune:
            kind, block, ms = b.autotune(0)
            tune = {"tiles": {M.KERNEL_TEAM: "team", M.KERNEL_LANE: "lane"}[kind],
                    "block": block, "iteration_ms": ms}
        else:
            b.init(0)
        if warmup:
            b.iterate(warmup, use_graph=True)
        a0 = sum(s["attempts"] for s in b.status())
        t0 = time.perf_counter()
        gpu_ms = b.iterate(steps, use_graph=True)   # synchronises before returning
        wall = time.perf_counter() - t0
        a1 = sum(s["attempts"] for s in b.status())
        upd_ms, w_ms = b.iterate_profiled(steps)
        a2 = sum(s["attempts"] for s in b.status())
        st = b.status()
        return dict(attempts=a1 - a0, wall=wall, gpu_ms=gpu_ms, prof_attempts=a2 - a1,
                    update_ms=upd_ms, weight_ms=w_ms, tiles=b.tiles, block=b.block,
                    grid=b.grid, ntiles=b.ntiles, tune=tune, status=st)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["c2", "batch96", "seeded", "c5"], default="c2")
    ap.add_argument("--therm-mult", type=int, default=1, help="c5: therm = mult * N")
    ap.add_argument("--nup-mult", type=int, default=1, help="c5: NUPDATES = mult * N")
    ap.add_argument("--blocks", type=int, default=10, help="c5: jackknife blocks")
    ap.add_argument("--weights", default="", help="c5: .npz of converged ln W (the seeded_jack* "
                    "arrays a seeded run saves); default: the surrogate seed")
    ap.add_argument("--stratified", action="store_true",
                    help="c5: walkers start spread over D (stratified start) instead of the "
                    "reference's uniform random spins")
    ap.add_argument("--observables", choices=["moments", "joint"], default="moments",
                    help="c5: per-D moments (default) or the reference's joint H(D,M)")
    ap.add_argument("--mode", choices=["all", "both", "seeded", "plain", "nonet"],
                    default="both")
    ap.add_argument("--T", type=float, default=0.55)
    ap.add_argument("--delta", type=float, default=1.97)
    ap.add_argument("--jacks", type=int, default=4, help="seeded workload: independent jacks")
    ap.add_argument("--jack0", type=int, default=0,
                    help="seeded workload: index of the first jack (seed 777 + 1000 j)")
    ap.add_argument("--nets", type=int, default=2,
                    help="seeded workload: networks (different initial parameters) the "
                         "jacks alternate between")
    ap.add_argument("--s0", type=float, default=50.0,
                    help="seeded/nonet: NUPDATES of the first iteration, in sweeps")
    ap.add_argument("--growth", type=float, default=2.0,
                    help="seeded/nonet: NUPDATES factor per iteration once width == N")
    ap.add_argument("--smax", type=float, default=3200.0,
                    help="seeded/nonet: largest NUPDATES, in sweeps")
    ap.add_argument("--cap-seeded", type=float, default=900.0)
    ap.add_argument("--cap-plain", type=float, default=900.0)
    ap.add_argument("--cap-nonet", type=float, default=600.0)
    ap.add_argument("--save-lnw", default="", help="seeded: write every jack's final ln W "
                    "to this .npz")
    ap.add_argument("--init-lnw", default="", help="seeded: start every jack from the mean of "
                    "the seeded_jack* curves in this .npz (the reference's continuing mode, one "
                    "init-weight file for all jacks) instead of the network")
    ap.add_argument("--L", type=int, default=None, help="72 for c2, 128 for seeded")
    ap.add_argument("--walkers", type=int, default=0)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--trivial", action="store_true", help="trivial instead of recursive update")
    ap.add_argument("--kernel", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-host-loop", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0)
    args = ap.parse_args()

    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this engine runs on one GPU; multi-GPU runs are not supported "
              "(walkers are independent chains, so the path would shard by walkers "
              "with one histogram all-reduce per weight iteration)", file=sys.stderr)
        sys.exit(2)

    import muca_amd as M
    nth, cpus_avail, cpu_name = host_cpu()
    recursion = not args.trivial
    if args.workload == "seeded":
        seeded_workload(M, args)
        return
    if args.workload == "c5":
        c5_workload(M, args)
        return

    if args.L is None:
        args.L = 72
    if args.workload == "c2":
        L, seed = args.L, args.seed
        W = args.walkers or 4096
        N = L * L
        ins = [dict(L=L, walkers=W, seed=seed, beta=2.0, recursion=recursion)]
        r = run_batch(M, ins, args.steps, args.warmup, autotune=False)
        workload = f"multicanonical weight iterations, plain method, L={L}, {W} walkers"
        config = {"workload": workload, "L": L, "walkers": W, "therm": "N",
                  "nupdates_init": 1, "recursive_update": int(recursion), "seed": seed,
                  "parallelism": "1 GPU; persistent update kernel over 16-walker tiles, "
                                 "device weight update, hipGraphs of eight iterations"}
    else:
        W = args.walkers or 4096
        ins, meta = batch96(W)
        r = run_batch(M, ins, args.steps, args.warmup, autotune=not args.no_autotune)
        workload = f"batched weight iterations, 96 instances x {W} walkers"
        config = {"workload": workload, "instances": 96, "walkers_per_instance": W,
                  "L": [32, 48, 64, 72, 96, 128], "delta": [1.90, 1.95, 1.98, 2.00],
                  "T": [0.5, 0.6, 0.7, 0.8], "seed": "instance id", "therm": "N",
                  "recursive_update": 1,
                  "parallelism": "1 GPU; one persistent launch over all instances' tiles, "
                                 "device weight update, hipGraphs of eight iterations"}

    value = r["attempts"] / r["wall"] / 1e9
    avg_kernel_s = r["update_ms"] / args.steps * 1e-3
    att_per_launch = r["prof_attempts"] / args.steps
    achieved = ALGO_BYTES_PER_ATTEMPT * att_per_launch / avg_kernel_s / 1e9

    # HBM bytes and VALU instructions per launch from the committed rocprofv3
    # counter passes of this exact workload (none for other shapes), used
    # only while the team kernel's sources are the ones they were taken from
    traffic, valu, counters = None, None, None
    if args.workload == "c2" and args.L == 72 and W == 4096:
        tfile = os.path.join(ROOT, "profiles", "update_kernel_L72_W4096_traffic.json")
        if os.path.exists(tfile):
            t = json.load(open(tfile))
            fresh = t.get("source_sha") == team_kernel_source_sha()
            counters = {"file": os.path.relpath(tfile, ROOT), "source_sha": t.get("source_sha"),
                        "current_source_sha": team_kernel_source_sha(), "fresh": fresh}
            insts = t.get("valu_wave_insts_per_launch")
            if fresh:
                traffic = t.get("bytes_per_launch")
            if fresh and insts:
                peak = 256 * 2.4e9   # wave64 VALU instructions/s: 256 CUs x 1/clk x 2.4 GHz
                valu = {"achieved": insts / avg_kernel_s, "peak": peak,
                        "unit": "wave instructions/s", "frac": insts / avg_kernel_s / peak}
    kernel_rate = att_per_launch / avg_kernel_s / 1e9
    ceil = valu_ceiling()
    roofline_valu = {"bound": "valu: Philox4x32-7 block rate (microbenchmark, this run)",
                     "achieved": kernel_rate, "peak": ceil["attempts_per_ns"],
                     "unit": "attempts/ns", "frac": kernel_rate / ceil["attempts_per_ns"],
                     "probe": ceil}

    host = None
    cpu = None
    if args.workload == "c2":
        if not args.no_host_loop:
            hv, hk, hhists = host_loop(M, args.L, W, args.seed, recursion, args.steps,
                                       args.warmup, args.kernel)
            host = {"value": hv, "kernel_ms_avg": hk,
                    "note": "same workload, weight update on the host between launches"}
        if not args.no_cpu_baseline:
            att, el, iters, chists = cpu_weight_loop(args.L, W, args.seed, recursion,
                                                     args.cpu_budget, nth)
            # the timed path itself, re-run from its start for as many iterations
            ghists = batch_hists(M, ins, iters)
            same = 0
            while same < iters and np.array_equal(ghists[same], chists[same]):
                same += 1
            cpu = {"value": att / el / 1e9, "unit": "attempts/ns", "cores": nth, "kind": "port",
                   "cpu": cpu_name, "cpus_available": cpus_avail,
                   "sample": f"CPU oracle (the same chain; {nth} threads over walkers; host "
                             f"weight update) on the first {iters} weight iterations of this "
                             f"workload ({att} attempts, {el:.1f} s); the D histograms of the "
                             f"timed batched GPU path (re-run from its start) are identical "
                             f"to the oracle's for the first {same} of those {iters} "
                             f"iterations"}
    elif not args.no_cpu_baseline:
        # a bounded sample: first weight iteration of instances 0, 1, ... (L = 32) until
        # the budget is spent
        from oracle import oracle as O
        att, t0, done = 0, time.perf_counter(), 0
        for q in ins:
            L = q["L"]
            N = L * L
            lat = O.init_lattice(O.KEYING_WEIGHT, L, q["walkers"], seed=q["seed"])
            E, D, _ = O.state(L, lat)
            O.mcs(L, lat, E, D, None, keying=0, fp64=False,
                  lnw32=np.zeros(N + 1, np.float32), iter=0, seed=q["seed"], therm=N,
                  nupdates=1, beta=q["beta"], nthreads=nth)
            att += q["walkers"] * (N + 1)
            done += 1
            if time.perf_counter() - t0 >= args.cpu_budget:
                break
        el = time.perf_counter() - t0
        cpu = {"value": att / el / 1e9, "unit": "attempts/ns", "cores": nth, "kind": "port",
               "cpu": cpu_name, "cpus_available": cpus_avail,
               "sample": f"CPU oracle (same chain, {nth} threads over walkers): first weight "
                         f"iteration (NUPDATES = 1) of instances 0..{done - 1}, {att} attempts, "
                         f"{el:.1f} s"}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "attempts/ns",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 spins, u32 Philox, f32 acceptance, f64 weight update",
        "data": "synthetic (random initial spins from the reference's Philox keying)",
        "config": config,
        "kernel_ms_avg": avg_kernel_s * 1e3,
        "weight_update_ms_avg": r["weight_ms"] / args.steps,
        "tiles": {"kind": {M.KERNEL_TEAM: "team", M.KERNEL_LANE: "lane"}[r["tiles"]],
                  "block": r["block"], "grid": r["grid"], "ntiles": r["ntiles"],
                  "autotune": r["tune"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "roofline_valu": roofline_valu,
        "valu_issue": valu,
        "counter_file": counters,
        "host_loop": host,
        "cpu_baseline": cpu,
    }
    if args.workload == "batch96":
        st = r["status"]
        out["instances"] = {"converged": int(sum(s["converged"] for s in st)),
                            "width_eq_N": int(sum(s["reach_iter"] != 0xFFFFFFFF for s in st))}
    print(json.dumps(out))


def seeded_workload(M, args):
    """GPU seconds to a flat histogram for an unseen (L, T, Delta), surrogate
    seeded against the method without the network and the reference's plain
    method, each as `--jacks` independent jacks in one batch."""
    from muca_amd import surrogate as S
    L = args.L or 128
    Wtot = args.walkers or M.default_walkers(0)
    J = args.jacks
    W = Wtot // J
    beta = 1.0 / args.T
    N = L * L
    modes = ["seeded", "nonet", "plain"] if args.mode == "all" else (
        ["seeded", "plain"] if args.mode == "both" else [args.mode])
    caps = {"seeded": args.cap_seeded, "nonet": args.cap_nonet, "plain": args.cap_plain}

    def log(tag):
        def f(gpu, st):
            print(f"{tag}: gpu {gpu:.1f}s " + " ".join(
                f"[it {x['iter']} w {x['width']} dkl {x['dkl']:.2e} nup {x['nup']} "
                f"{'flat' if x['converged'] else ''}]" for x in st), file=sys.stderr, flush=True)
        return f

    train_info, seeds = [], None
    if "seeded" in modes and args.init_lnw:
        d = np.load(args.init_lnw)
        curves = [d[k] for k in sorted(d.files) if k.startswith("seeded_jack")]
        if not curves or len(curves[0]) != N + 1:
            raise SystemExit(f"{args.init_lnw}: no seeded_jack* arrays of N+1 = {N + 1} values")
        init = np.mean(curves, axis=0)
        init = init - init[0]
        seeds = [init] * J
        train_info = [{"init_weight": f"mean of {len(curves)} curves of "
                                      f"{os.path.basename(args.init_lnw)}"}]
        print(f"seeded: continuing from the mean of {len(curves)} curves", file=sys.stderr,
              flush=True)
    elif "seeded" in modes:
        # jacks alternate between networks trained from different initial
        # parameters, so a flat ln W that depended on its seed would show as a
        # difference between the two groups
        X, t, w = S.corpus(S.load_corpus())
        nets = []
        for k in range(args.nets):
            net, info = S.train(X, t, w, seed=k)
            nets.append(S.seed_log_weight(net, L, beta, args.delta))
            net.close()
            train_info.append({kk: v for kk, v in info.items() if kk != "loss_curve"})
            print(f"seeded: network {k} trained on {info['samples']} samples in "
                  f"{info['gpu_s']:.2f} GPU s", file=sys.stderr, flush=True)
        seeds = [nets[j % args.nets] for j in range(args.jack0, args.jack0 + J)]
    res = {}
    for mode in modes:
        ins = S.protocol_instances(L, beta, J, W, mode, s0=args.s0, growth=args.growth,
                                   smax=args.smax, jack0=args.jack0)
        out, gpu = S.converge(ins, lnw_init=seeds if mode == "seeded" else None,
                              stratified=mode != "plain", max_gpu_s=caps[mode],
                              chunk=1 if mode != "plain" else 50, log=log(mode))
        res[mode] = dict(gpu_s=gpu, out=out)
        print(f"{mode}: {sum(o['status']['converged'] for o in out)}/{J} jacks flat in "
              f"{gpu:.1f} GPU s", file=sys.stderr, flush=True)

    def summary(r):
        st = [o["status"] for o in r["out"]]
        return {"flat": all(x["converged"] for x in st), "gpu_s": r["gpu_s"],
                "jacks_flat": int(sum(x["converged"] for x in st)),
                "iterations": [x["iter"] for x in st],
                "width_reached_N_at": [None if x["reach_iter"] == 0xFFFFFFFF else x["reach_iter"]
                                       for x in st],
                "last_D_KL": [x["dkl"] for x in st],
                "attempts": int(sum(x["attempts"] for x in st))}

    sm = {k: summary(v) for k, v in res.items()}
    if args.save_lnw:
        np.savez_compressed(args.save_lnw, **{f"{k}_jack{args.jack0 + j}": o["lnw"]
                                              for k, v in res.items()
                                              for j, o in enumerate(v["out"])})
    speed = {}
    if "seeded" in sm and sm["seeded"]["flat"]:
        for other in ("plain", "nonet"):
            if other in sm:
                speed[f"{other}_over_seeded"] = {
                    "value": sm[other]["gpu_s"] / sm["seeded"]["gpu_s"],
                    "lower_bound": not sm[other]["flat"]}
    acc = {}
    for k, v in res.items():
        if not sm[k]["flat"] or J < 2:
            continue
        lnws = [o["lnw"] for o in v["out"]]
        mean, err = S.jackknife_lnw(lnws)
        rec = {"jackknife_err_median": float(np.median(err[1:])),
               "jackknife_err_max": float(np.max(err[1:]))}
        if k == "seeded" and args.nets == 2 and J >= 4:
            ga, gb = np.mean(lnws[0::2], axis=0), np.mean(lnws[1::2], axis=0)
            ea = np.std(lnws[0::2], axis=0, ddof=1) / np.sqrt(len(lnws[0::2]))
            eb = np.std(lnws[1::2], axis=0, ddof=1) / np.sqrt(len(lnws[1::2]))
            z = (ga - gb)[1:] / np.sqrt(ea[1:] ** 2 + eb[1:] ** 2)
            rec["network_groups_rms_z"] = float(np.sqrt(np.mean(z * z)))
        acc[k] = rec
        acc[k + "_mean"] = mean
    if "seeded" in acc and "plain" in acc:
        d = (acc["seeded_mean"] - acc["plain_mean"])[1:]
        e = np.sqrt(S.jackknife_lnw([o["lnw"] for o in res["seeded"]["out"]])[1] ** 2 +
                    S.jackknife_lnw([o["lnw"] for o in res["plain"]["out"]])[1] ** 2)[1:]
        acc["seeded_vs_plain"] = {"rms_diff": float(np.sqrt(np.mean(d * d))),
                                  "rms_z": float(np.sqrt(np.mean((d / e) ** 2)))}
    for k in list(acc):
        if k.endswith("_mean"):
            del acc[k]
    main_run = sm.get("seeded") or sm[modes[0]]
    out = {
        "metric": METRIC,
        "value": main_run["gpu_s"],
        "unit": "GPU s to flat H",
        "n_gpus": 1,
        "steps": int(max(main_run["iterations"])),
        "warmup": 0,
        "ms_per_step": main_run["gpu_s"] * 1e3 / max(1, max(main_run["iterations"])),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 spins, f32 acceptance, f64 weight update; surrogate bf16 MFMA / f32",
        "data": "synthetic (Philox-keyed initial spins)",
        "config": {"workload": f"weight iteration to flat H, L={L}, T={args.T}, "
                               f"Delta={args.delta}, {J} jacks x {W} walkers, trivial update",
                   "L": L, "T": args.T, "delta": args.delta, "jacks": J,
                   "walkers_per_jack": W,
                   "criterion": "every jack: width == N and D_KL < 1e-4",
                   "seeded_schedule": f"NUPDATES {args.s0} N, x{args.growth} per iteration, "
                                      f"up to {args.smax} N; stratified start",
                   "plain": "ln W = 0, NUPDATES = 1, x1.01 once width == N, random start",
                   "nonet": "ln W = 0 with the seeded schedule and start",
                   "caps_gpu_s": {m: caps[m] for m in modes}},
        "runs": sm,
        "speedup": speed,
        "lnw_accuracy": acc,
        "surrogate_training": train_info,
    }
    print(json.dumps(out))


def c5_workload(M, args):
    from muca_amd import surrogate as S
    from oracle import oracle as O
    L = args.L or 256
    W = args.walkers or 16384
    N = L * L
    beta, delta0, seed = 2.0, 1.98, args.seed
    B = args.blocks
    if args.weights:
        # converged weights: the per-bin mean of the jacks a seeded run saved
        d = np.load(args.weights)
        keys = sorted(k for k in d.files if k.startswith("seeded_jack"))
        lnw_jacks = [d[k] for k in keys]
        if not lnw_jacks or len(lnw_jacks[0]) != N + 1:
            raise SystemExit(f"{args.weights}: no seeded_jack* arrays of N+1 = {N + 1} values")
        lnw, lnw_err = S.jackknife_lnw(lnw_jacks) if len(lnw_jacks) > 1 else (lnw_jacks[0], None)
        lnw = lnw - lnw[0]
        weights_note = (f"ln W from {os.path.basename(args.weights)}: mean of "
                        f"{len(lnw_jacks)} seeded jacks")
    else:
        X, t, w = S.corpus(S.load_corpus())
        net, _ = S.train(X, t, w)
        lnw = S.seed_log_weight(net, L, beta, delta0)
        net.close()
        weights_note = "surrogate seed (not converged)"
    therm, nup = args.therm_mult * N, args.nup_mult * N
    print(f"c5: L={L} walkers={W} blocks={B} therm={therm} NUPDATES={nup}", file=sys.stderr,
          flush=True)
    joint = args.observables == "joint"
    obs_mode = M.OBS_JOINT if joint else M.OBS_MOMENTS
    with M.Context(L, W, keying=M.KEYING_PRODUCE, precision=M.FP64, joint_blocks=B,
                   observables=obs_mode, kernel=args.kernel) as ctx:
        ctx.init_lattice(mm=0, gpu=0, seed=seed, stratified=args.stratified)
        ctx.compute_state()
        ctx.set_log_weight(lnw)
        # untimed warm-up: the first launch of a kernel in a process carries
        # its one-time setup (about 13 ms against a 52 ms launch in the
        # kernel trace); 64 production attempts per walker
        _, warm_ms = ctx.run(0xFFFF, seed, 0, 64)
        _, ms = ctx.run(0, seed, therm, nup)
        kernel_used, block_used = ctx.info.kernel, ctx.info.block
        print(f"c5: production launch {ms:.1f} ms", file=sys.stderr, flush=True)
        deltas = np.round(np.linspace(1.90, 2.00, 11), 4)
        t0 = time.perf_counter()
        absM, chi, binder = ctx.reweight(deltas)
        rw_ms = (time.perf_counter() - t0) * 1e3
        if joint:
            joint0 = ctx.get_joint(0)
        else:
            mom0 = ctx.get_moments(0)
        # the same number of attempts without histogram accumulation: what
        # the accumulation costs (runs after the observables were taken)
        _, ms_therm = ctx.run(1, seed, therm + nup, 0)
    i0 = int(np.argmin(np.abs(deltas - delta0)))
    wsh = lnw.copy()
    O.weight_shift(beta, delta0, wsh)
    bz_all = O.boltz(beta, delta0, wsh)
    if joint:
        S0 = O.joint_moments(N, joint0)[0]
    else:
        S0 = np.array([float(x) for x in mom0[0]])
    # the long-double sums with the shift over the block's visited bins
    Dg = np.arange(N + 1, dtype=np.float64)
    expo = -beta * delta0 * Dg - lnw
    bz = np.where(S0 > 0, np.exp(np.minimum(expo - np.max(expo[S0 > 0]), 0.0)), 0.0)
    if joint:
        # the reference's evaluation: shift over all bins (produce2_multigpu.cu:391-400)
        a_refall, c_refall = O.observables(beta, N, joint0, bz_all)
        a_ref, c_ref = O.observables(beta, N, joint0, bz)
        del joint0
    else:
        a_refall, c_refall = O.observables_moments(beta, N, mom0[0], mom0[1], mom0[2], bz_all)
        a_ref, c_ref = O.observables_moments(beta, N, mom0[0], mom0[1], mom0[2], bz)
    obs = []
    for i, d in enumerate(deltas):
        am, ae = M.jackknife(absM[i] / N)
        cm, ce = M.jackknife(chi[i])
        bm, be = M.jackknife(binder[i])
        obs.append({"delta": float(d), "abs_m": [am, ae], "chi": [cm, ce], "binder": [bm, be]})
    attempts = W * (therm + nup)
    out = {
        "metric": METRIC,
        "value": attempts / (ms * 1e-3) / 1e9,
        "unit": "attempts/ns",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int8 spins, f64 acceptance, " +
                 ("u64 joint histograms" if joint else "exact per-D moments (u64/u128)") +
                 ", f64 reweighting",
        "data": "synthetic (" + ("stratified start over D" if args.stratified else
                                 "random initial spins from the reference's Philox keying") + ")",
        "config": {"workload": f"large production, L={L}, {W} walkers, T=0.5, "
                               f"{B} jackknife blocks", "L": L, "walkers": W,
                   "therm": f"{args.therm_mult} N", "nupdates": f"{args.nup_mult} N",
                   "observables": args.observables,
                   "accumulator_words_per_block": (N + 1) * (N + 2) // 2 if joint
                   else 8 * (N + 1),
                   "weights": weights_note},
        "kernel": {1: "generic", 2: "team", 3: "lane"}.get(kernel_used, "?"),
        "block": block_used,
        "reweight_ms": rw_ms,
        "warmup_launch": {"ms": warm_ms, "nupdates": 64,
                          "note": "untimed, same kernel, before the timed launch"},
        "therm_only": {"ms": ms_therm, "attempts_per_ns": attempts / (ms_therm * 1e-3) / 1e9,
                       "note": "same attempts per walker, no histogram accumulation"},
        "observables": obs,
        "visited_D_block0": [int(np.argmax(S0 > 0)), int(N - np.argmax(S0[::-1] > 0))],
        "oracle_block0": {"delta": delta0, "abs_M_device": float(absM[i0, 0]),
                          "abs_M_oracle": a_ref, "chi_device": float(chi[i0, 0]),
                          "chi_oracle": c_ref,
                          "reference_all_bin_shift": {"abs_M": a_refall, "chi": c_refall},
                          "rel_diff_abs_M": abs(absM[i0, 0] - a_ref) / abs(a_ref),
                          "rel_diff_chi": abs(chi[i0, 0] - c_ref) / abs(c_ref)},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
