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
HBM at 2 algorithmic bytes per attempt; `host_loop` is the same workload
driven the reference's way (host weight update between launches);
`cpu_baseline` times the CPU oracle (the same chain, threaded over walkers) on
a bounded sample of the same workload.

`--workload seeded`: GPU seconds to a flat histogram (width == N and
D_KL < 1e-4, muca_delta_multigpu.cu:441-447) for an unseen (L, T, Delta) =
(`--L`, 0.55, 1.97), 4096 walkers, trivial update.  "seeded": ln W from the
surrogate (trained on the GPU on the committed corpus of converged weights,
L = 16 .. 64, T = 0.5 .. 0.8), NUPDATES from the start = 24 N^2 (`--nup-factor`; the
reference's starting_M argument); "plain": ln W = 0, NUPDATES = 1 (the
reference's defaults).  Each run stops at its device-time cap; a run that
hits it reports a lower bound.

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
    python bench.py --workload seeded [--L 128] [--mode both|seeded|plain]
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
    ap.add_argument("--observables", choices=["moments", "joint"], default="moments",
                    help="c5: per-D moments (default) or the reference's joint H(D,M)")
    ap.add_argument("--mode", choices=["both", "seeded", "plain"], default="both")
    ap.add_argument("--T", type=float, default=0.55)
    ap.add_argument("--delta", type=float, default=1.97)
    ap.add_argument("--cap-seeded", type=float, default=600.0)
    ap.add_argument("--cap-plain", type=float, default=600.0)
    ap.add_argument("--recursive", action="store_true",
                    help="seeded workload: weight_recursion instead of the trivial update")
    ap.add_argument("--save-lnw", default="", help="seeded: write the final ln W of each run "
                    "to <path>_<run>.npy")
    ap.add_argument("--nup-factor", type=float, default=24.0,
                    help="seeded run: NUPDATES from the start = factor * N^2")
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
    nth = min(16, os.cpu_count() or 1)
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
                                 "device weight update, one hipGraph per iteration"}
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
                                 "device weight update, one hipGraph per iteration"}

    value = r["attempts"] / r["wall"] / 1e9
    avg_kernel_s = r["update_ms"] / args.steps * 1e-3
    att_per_launch = r["prof_attempts"] / args.steps
    achieved = ALGO_BYTES_PER_ATTEMPT * att_per_launch / avg_kernel_s / 1e9

    # HBM bytes and VALU instructions per launch from the committed rocprofv3
    # counter passes of this exact workload (none for other shapes)
    traffic, valu = None, None
    if args.workload == "c2" and args.L == 72 and W == 4096:
        tfile = os.path.join(ROOT, "profiles", "update_kernel_L72_W4096_traffic.json")
        if os.path.exists(tfile):
            t = json.load(open(tfile))
            traffic = t.get("bytes_per_launch")
            insts = t.get("valu_wave_insts_per_launch")
            if insts:
                peak = 256 * 2.4e9   # wave64 VALU instructions/s: 256 CUs x 1/clk x 2.4 GHz
                valu = {"achieved": insts / avg_kernel_s, "peak": peak,
                        "unit": "wave instructions/s", "frac": insts / avg_kernel_s / peak}

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
            same = None
            if host is not None:
                n = min(len(hhists), len(chists))
                same = all(np.array_equal(a, b) for a, b in zip(hhists[:n], chists[:n]))
            cpu = {"value": att / el / 1e9, "unit": "attempts/ns", "cores": nth, "kind": "port",
                   "sample": f"CPU oracle (same chain, {nth} threads over walkers) on the first "
                             f"{iters} weight iterations of this workload ({att} attempts, "
                             f"{el:.1f} s); histograms identical to the GPU host-loop run's "
                             f"for those iterations: {same}"}
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
        "valu_issue": valu,
        "host_loop": host,
        "cpu_baseline": cpu,
    }
    if args.workload == "batch96":
        st = r["status"]
        out["instances"] = {"converged": int(sum(s["converged"] for s in st)),
                            "width_eq_N": int(sum(s["reach_iter"] != 0xFFFFFFFF for s in st))}
    print(json.dumps(out))


def seeded_workload(M, args):
    from muca_amd import surrogate as S
    L = args.L or 128
    W = args.walkers or 4096
    beta = 1.0 / args.T
    N = L * L

    def log(tag):
        def f(gpu, st):
            s = st[0]
            print(f"{tag}: gpu {gpu:.1f}s iter {s['iter']} width {s['width']} D_KL "
                  f"{s['dkl']:.2e} NUPDATES {s['nup']}", file=sys.stderr, flush=True)
        return f

    res = {}
    train_info = None
    if args.mode in ("both", "seeded"):
        X, t, w = S.corpus(S.load_corpus())
        net, train_info = S.train(X, t, w)
        lnw0 = S.seed_log_weight(net, L, beta, args.delta)
        net.close()
        nup0 = S.equilibrium_nupdates(L, args.nup_factor)
        print(f"seeded: trained on {len(X)} samples in {train_info['gpu_s']:.2f} GPU s, "
              f"NUPDATES from the start {nup0}", file=sys.stderr, flush=True)
        out, _ = S.converge([dict(L=L, walkers=W, seed=777, beta=beta,
                                  recursion=args.recursive, init_nup=nup0)],
                            lnw_init=[lnw0], stratified=False,
                            max_gpu_s=args.cap_seeded, chunk=1, log=log("seeded"))
        res["seeded"] = out[0]
    if args.mode in ("both", "plain"):
        out, _ = S.converge([dict(L=L, walkers=W, seed=777, beta=beta,
                                  recursion=args.recursive, init_nup=1)], stratified=False,
                            max_gpu_s=args.cap_plain,
                            chunk=50, log=log("plain"))
        res["plain"] = out[0]

    def summary(o):
        st = o["status"]
        return {"flat": bool(st["converged"]), "gpu_s": o["gpu_s"], "iterations": st["iter"],
                "width_reached_N_at": None if st["reach_iter"] == 0xFFFFFFFF else st["reach_iter"],
                "last_D_KL": st["dkl"], "attempts": st["attempts"]}

    sm = {k: summary(v) for k, v in res.items()}
    if args.save_lnw:
        for k, v in res.items():
            np.save(f"{args.save_lnw}_{k}.npy", v["lnw"])
    ratio = None
    if "seeded" in sm and "plain" in sm and sm["seeded"]["flat"]:
        r = sm["plain"]["gpu_s"] / sm["seeded"]["gpu_s"]
        ratio = {"value": r, "lower_bound": not sm["plain"]["flat"]}
    agree = None
    if "seeded" in res and "plain" in res and sm["seeded"]["flat"] and sm["plain"]["flat"]:
        d = res["seeded"]["lnw"] - res["plain"]["lnw"]
        agree = {"max_abs_lnw_diff": float(np.max(np.abs(d))),
                 "rms_lnw_diff": float(np.sqrt(np.mean(d * d)))}
    main_run = sm.get("seeded") or sm.get("plain")
    out = {
        "metric": METRIC,
        "value": main_run["gpu_s"],
        "unit": "GPU s to flat H",
        "n_gpus": 1,
        "steps": main_run["iterations"],
        "warmup": 0,
        "ms_per_step": main_run["gpu_s"] * 1e3 / max(1, main_run["iterations"]),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 spins, f32 acceptance, f64 weight update; surrogate bf16 MFMA / f32",
        "data": "synthetic (random initial spins from the reference's Philox keying)",
        "config": {"workload": f"weight iteration to flat H, L={L}, T={args.T}, "
                               f"Delta={args.delta}, {W} walkers, "
                               f"{'recursive' if args.recursive else 'trivial'} update",
                   "seeded_nupdates": f"{args.nup_factor} N^2",
                   "L": L, "T": args.T, "delta": args.delta, "walkers": W,
                   "criterion": "width == N and D_KL < 1e-4",
                   "caps_gpu_s": {"seeded": args.cap_seeded, "plain": args.cap_plain}},
        "runs": sm,
        "speedup_plain_over_seeded": ratio,
        "lnw_agreement": agree,
        "surrogate_training": None if train_info is None else
        {k: v for k, v in train_info.items() if k != "loss_curve"},
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
    X, t, w = S.corpus(S.load_corpus())
    net, _ = S.train(X, t, w)
    lnw = S.seed_log_weight(net, L, beta, delta0)
    net.close()
    therm, nup = args.therm_mult * N, args.nup_mult * N
    print(f"c5: L={L} walkers={W} blocks={B} therm={therm} NUPDATES={nup}", file=sys.stderr,
          flush=True)
    joint = args.observables == "joint"
    obs_mode = M.OBS_JOINT if joint else M.OBS_MOMENTS
    with M.Context(L, W, keying=M.KEYING_PRODUCE, precision=M.FP64, joint_blocks=B,
                   observables=obs_mode) as ctx:
        ctx.init_lattice(mm=0, gpu=0, seed=seed)
        ctx.compute_state()
        ctx.set_log_weight(lnw)
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
        "data": "synthetic (random initial spins from the reference's Philox keying)",
        "config": {"workload": f"large production, L={L}, {W} walkers, T=0.5, "
                               f"{B} jackknife blocks", "L": L, "walkers": W,
                   "therm": f"{args.therm_mult} N", "nupdates": f"{args.nup_mult} N",
                   "observables": args.observables,
                   "accumulator_words_per_block": (N + 1) * (N + 2) // 2 if joint
                   else 8 * (N + 1),
                   "weights": "surrogate seed (not converged)"},
        "kernel": {1: "generic", 2: "team", 3: "lane"}.get(kernel_used, "?"),
        "block": block_used,
        "reweight_ms": rw_ms,
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
