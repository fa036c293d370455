"""Benchmark: flip attempts/ns of the multicanonical weight iteration.

Workload (plain method at reference scale on one GPU): L = 72 (5184 spins,
5185 D-bins), 4096 walkers, random initial spins (weight keying, seed 1234),
ln W = 0 initially, recursive weight update, therm = N and NUPDATES = 1 per
iteration with the reference's x1.01 schedule once the histogram spans all
bins.  One step = one weight iteration: H2D of ln W, one launch of the update
kernel (therm + NUPDATES attempts per walker), D2H of the histogram, width /
D_KL bookkeeping and the O(N) weight update on the host -- the loop body of
muca_delta_multigpu.cu:374-477.

Prints one JSON line.  `value` is whole-step throughput (attempts / wall
time); `roofline` prices the update kernel alone (HIP-event durations on its
own stream) against HBM at 2 algorithmic bytes per attempt; `cpu_baseline`
times the CPU oracle (the same chain, threaded over walkers) on a bounded
sample of the same workload.

    python bench.py [--steps K] [--warmup W] [--walkers 4096] [--L 72]
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


def cpu_baseline(L, W, seed, recursion, budget_s, nthreads):
    """The oracle's weight loop on the host for the same workload: iterations
    run until budget_s seconds of wall time are spent."""
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
                  therm=N, nupdates=nup, nthreads=nthreads)
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
        if el >= budget_s or k >= 10000:
            break
    return attempts, el, k, hists


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--L", type=int, default=72)
    ap.add_argument("--walkers", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--trivial", action="store_true", help="trivial instead of recursive update")
    ap.add_argument("--kernel", type=int, default=0)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0)
    args = ap.parse_args()

    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this engine runs on one GPU; multi-GPU runs are not part of this "
              "build (walkers are independent chains: see DESIGN.md for how it would shard)",
              file=sys.stderr)
        sys.exit(2)

    import muca_amd as M
    L, W, seed = args.L, args.walkers, args.seed
    N = L * L
    recursion = not args.trivial
    ctx = M.Context(L, W, device=0, kernel=args.kernel)
    ctx.init_lattice(mm=0, gpu=0, seed=seed)
    ctx.compute_state()
    lnw = np.zeros(N + 1)
    p_acc = np.zeros(N + 1)
    ctx.set_log_weight(lnw)
    state = dict(width=10, nup=1, k=0)
    gpu_hists = []

    def step():
        st = state
        if st["width"] == N:
            st["nup"] = int((st["nup"] + 1) * 1.01)
        hist, ms = ctx.run(st["k"], seed, N, st["nup"])
        gpu_hists.append(hist)
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

    for _ in range(args.warmup):
        step()
    attempts, kernel_ms = 0, []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a, ms = step()
        attempts += a
        kernel_ms.append(ms)
    wall = time.perf_counter() - t0
    value = attempts / wall / 1e9
    avg_kernel_s = float(np.mean(kernel_ms)) * 1e-3
    attempts_per_launch = attempts / args.steps
    achieved = ALGO_BYTES_PER_ATTEMPT * attempts_per_launch / avg_kernel_s / 1e9

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "update_kernel_L72_W4096_traffic.json")
    if os.path.exists(tfile) and L == 72 and W == 4096:
        traffic = json.load(open(tfile)).get("bytes_per_launch")

    cpu = None
    if not args.no_cpu_baseline:
        nth = min(16, os.cpu_count() or 1)
        att, el, iters, chists = cpu_baseline(L, W, seed, recursion, args.cpu_budget, nth)
        same = all(np.array_equal(a, b) for a, b in zip(gpu_hists, chists))
        cpu = {"value": att / el / 1e9, "unit": "attempts/ns", "cores": nth, "kind": "port",
               "sample": f"CPU oracle (same chain, {nth} threads over walkers) on the first "
                         f"{iters} weight iterations of this workload ({att} attempts, "
                         f"{el:.1f} s); histograms identical to the GPU's: {same}"}

    out = {
        "metric": "flip attempts/ns; GPU s to flat H(M) at L=128, surrogate-seeded vs plain",
        "value": value,
        "unit": "attempts/ns",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 spins, u32 Philox, f32 acceptance",
        "data": "synthetic (random initial spins from the reference's Philox keying)",
        "config": {"workload": f"multicanonical weight iterations, plain method, L={L}, "
                               f"{W} walkers", "L": L, "walkers": W, "therm": "N",
                   "nupdates_init": 1, "recursive_update": int(recursion), "seed": seed,
                   "parallelism": "1 GPU, one lane per walker"},
        "kernel_ms_avg": avg_kernel_s * 1e3,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
