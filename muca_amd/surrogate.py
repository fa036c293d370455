"""Surrogate seeding of the multicanonical weights.

The reference seeds hard weight runs from an extrapolated initial weight file
(`continuing` mode, muca_delta_multigpu.cu:223-235, `initweight_L128_T050.txt`:
N+1 whitespace-separated values of ln W).  This module replaces the
extrapolation by a small network trained on weights that the batched engine
itself has converged:

    corpus  = converged ln W(D) of many (L, T) instances      (muca_amd.Batch)
    net     = MLP 4 -> 128 -> 128 -> 128 -> 1, SiLU            (muca_amd.Mlp, MFMA kernels)
    seed    = ln W for an unseen (L, T, Delta), written as the init-weight file

Representation.  ln W(D) is extensive: ln W(D)/N -> f(D/N; beta) + O(1/L).  Its
dominant part is the ground-state line: the best D-spin configuration has
E_J = -2D, so ln W(D) ~ -2 beta D - S(D).  The network predicts the residual
r = 10 (ln W/N + 2 beta m), m = D/N, which is small and smooth, from the
features (m, 16/L, 2 (beta - 1.5), 10 (Delta - 1.95)).  Delta does not enter
the weight problem in D (it enters at reweighting), so the network learns to
ignore it; it is kept as an input so one model serves the whole (T, Delta)
grid.  The seed never touches the acceptance arithmetic: it only replaces the
initial ln W, and the weight iteration with its flatness test still decides
when a run is done.
"""
import os
import time

import numpy as np

from . import Batch, Mlp, KERNEL_AUTO, REF32

RESIDUAL_SCALE = 10.0
DELTAS = (1.90, 1.95, 1.98, 2.00)
CORPUS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                           "weight_corpus.npz")


def features(m, L, beta, delta):
    """delta: one value or one per sample."""
    m = np.asarray(m, np.float64)
    x = np.empty((m.size, 4), np.float32)
    x[:, 0] = m
    x[:, 1] = 16.0 / L
    x[:, 2] = 2.0 * (beta - 1.5)
    x[:, 3] = 10.0 * (np.asarray(delta, np.float64) - 1.95)
    return x


def targets(lnw, L, beta):
    N = L * L
    m = np.arange(N + 1) / N
    return (RESIDUAL_SCALE * (np.asarray(lnw, np.float64) / N + 2.0 * beta * m)).astype(np.float32)


def lnw_from_output(y, L, beta):
    """ln W(D), normalised to ln W(0) = 0, from network outputs at m = D/N."""
    N = L * L
    m = np.arange(N + 1) / N
    lnw = N * (np.asarray(y, np.float64) / RESIDUAL_SCALE - 2.0 * beta * m)
    return lnw - lnw[0]


def corpus(runs):
    """runs: iterable of dicts with L, beta, delta, lnw (converged, N+1 values).
    Returns features X [n][4], targets t [n] and sample weights w [n]: an
    error e in the target is an error N e / 10 in ln W, so samples are
    weighted by N (normalised to mean 1)."""
    xs, ts, ws = [], [], []
    for r in runs:
        L = r["L"]
        N = L * L
        xs.append(features(np.arange(N + 1) / N, L, r["beta"], r["delta"]))
        ts.append(targets(r["lnw"], L, r["beta"]))
        ws.append(np.full(N + 1, float(N)))
    w = np.concatenate(ws)
    return np.concatenate(xs), np.concatenate(ts), (w / w.mean()).astype(np.float32)


def init_params(seed=0):
    """Uniform(-1/sqrt(fan_in), +1/sqrt(fan_in)) weights (input layer +-1/2),
    zero biases, in the packed layout of include/muca.h."""
    H = 128
    rng = np.random.default_rng(seed)
    parts = [rng.uniform(-0.5, 0.5, H * 4), np.zeros(H),
             rng.uniform(-1, 1, H * H) / np.sqrt(H), np.zeros(H),
             rng.uniform(-1, 1, H * H) / np.sqrt(H), np.zeros(H),
             rng.uniform(-1, 1, H) / np.sqrt(H), np.zeros(1)]
    return np.concatenate(parts).astype(np.float32)


def train(X, t, w=None, *, device=0, steps=20000, batch=4096, lr=2e-3, lr_final=1e-5, seed=0):
    """Trains a fresh network on the GPU.  Samples are shuffled once per epoch
    on the host and made resident; returns (net, info)."""
    rng = np.random.default_rng(seed)
    n = len(X)
    batch = min(batch, (n // 64) * 64)
    net = Mlp(device)
    net.set_params(init_params(seed))
    nb = n // batch
    per_epoch = max(1, nb)
    done, gpu_ms, losses = 0, 0.0, []
    # one resident shuffle per epoch; the cosine schedule spans all steps, so the
    # device trains `steps` steps in chunks of one epoch with the matching
    # slice of the schedule
    while done < steps:
        perm = rng.permutation(n)
        net.set_data(X[perm], t[perm], None if w is None else w[perm])
        k = min(per_epoch, steps - done)
        a = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * done / steps))
        b = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * (done + k) / steps))
        loss, ms = net.train(k, batch, float(a), float(b))
        gpu_ms += ms
        losses.append(loss)
        done += k
    _, full = net.gradient(X, t)
    return net, dict(steps=steps, batch=batch, samples=n, gpu_s=gpu_ms / 1e3,
                     final_loss=full, loss_curve=losses)


def generate_corpus(path, Ls=(16, 24, 32, 40, 48), Ts=(0.5, 0.6, 0.7, 0.8), walkers=16384,
                    max_gpu_s=300.0, log=None):
    """Converges one batched instance per (L, T) with the reference's weight
    loop (trivial update, therm = N, NUPDATES = 1 rising x1.01 once the
    histogram spans all bins; stratified start; seed 100 + instance index)
    and stores the converged ln W as L{L}_T{T} in an .npz file.  Instances
    not flat within max_gpu_s of device time are left out.  The committed
    corpus (muca_amd/data/weight_corpus.npz) was made with these defaults."""
    ins, meta = [], []
    for L in Ls:
        for T in Ts:
            ins.append(dict(L=L, walkers=walkers, seed=100 + len(ins), beta=1.0 / T,
                            recursion=False))
            meta.append((L, T))
    out, gpu = converge(ins, stratified=True, max_gpu_s=max_gpu_s, log=log)
    keep = {f"L{L}_T{T}": o["lnw"] for (L, T), o in zip(meta, out) if o["status"]["converged"]}
    np.savez_compressed(path, **keep)
    return keep, gpu


def extend_corpus(path, Ls, Ts, *, walkers=131072, s0=50.0, growth=2.0, smax=3200.0,
                  max_gpu_s=600.0, log=None):
    """Closes the loop: a network trained on the corpus at `path` seeds new
    (L, T) instances, which the batched engine converges in one batch with
    the seeded protocol of protocol_instances (trivial update, stratified
    start, NUPDATES from s0 sweeps growing x growth per iteration up to smax
    sweeps); the flat ones join the corpus.  Returns the keys added, the
    device seconds spent and the per-instance results."""
    X, t, w = corpus(load_corpus(path))
    net, _ = train(X, t, w)
    ins, seeds, meta = [], [], []
    for L in Ls:
        for T in Ts:
            ins += protocol_instances(L, 1.0 / T, 1, walkers, "seeded", s0=s0, growth=growth,
                                      smax=smax, seed0=200 + len(ins))
            seeds.append(seed_log_weight(net, L, 1.0 / T, 1.95))
            meta.append((L, T))
    net.close()
    out, gpu = converge(ins, lnw_init=seeds, stratified=True, max_gpu_s=max_gpu_s, chunk=1,
                        log=log)
    d = dict(np.load(path))
    added = []
    for (L, T), o in zip(meta, out):
        if o["status"]["converged"]:
            d[f"L{L}_T{T}"] = o["lnw"]
            added.append(f"L{L}_T{T}")
    np.savez_compressed(path, **d)
    return added, gpu, out


def load_corpus(path=CORPUS_PATH, exclude=()):
    """Runs for corpus(): every stored (L, T) curve once.  ln W(D) of the
    D-axis weight problem does not depend on Delta (Delta enters only at
    reweighting, produce2_multigpu.cu:297), so each sample gets a Delta drawn
    uniformly from [1.90, 2.00] (fixed generator): the network sees the whole
    Delta range and learns that the target ignores it, without repeating the
    curve once per Delta."""
    d = np.load(path)
    runs = []
    for k in sorted(d.files):
        if k in exclude:
            continue
        L = int(k.split("_")[0][1:])
        T = float(k.split("_")[1][1:])
        runs.append(curve_run(L, T, d[k]))
    return runs


def curve_run(L, T, lnw):
    """One converged curve as a corpus run, each D-sample with its own Delta
    drawn uniformly from [1.90, 2.00] by a generator fixed by (L, T)."""
    rng = np.random.default_rng(1000 * L + int(round(100 * T)))
    delta = rng.uniform(DELTAS[0], DELTAS[-1], size=len(lnw))
    return dict(L=L, beta=1.0 / T, delta=delta, lnw=lnw)


def seed_log_weight(net, L, beta, delta):
    N = L * L
    y = net.predict(features(np.arange(N + 1) / N, L, beta, delta))
    return lnw_from_output(y, L, beta)


def write_init_weight(path, lnw):
    """The init-weight file read by multigpu_weight's continuing mode
    (muca_delta_multigpu.cu:223-235): N+1 whitespace-separated values."""
    with open(path, "w") as f:
        for v in lnw:
            f.write(repr(float(v)) + "\n")


def equilibrium_nupdates(L, factor=4):
    """NUPDATES of the seeded protocol: factor * N^2 attempts per walker.  A
    flat-histogram walker needs ~N^2 attempts to cross D = 0..N, so with a
    factor of a few one iteration's histogram reflects the current weights and
    the trivial update corrects a seed's errors in a few iterations instead of
    the ~100-fold NUPDATES ramp of a run from ln W = 0.  Larger lattices need a
    larger factor for D_KL < 1e-4 (statistics per bin)."""
    N = L * L
    return int(factor * N * N)


def protocol_instances(L, beta, jacks, walkers, mode, *, s0=50.0, growth=2.0, smax=3200.0,
                       seed0=777, jack0=0):
    """Instances of one weight problem, one per jack (independent chains:
    Philox seed seed0 + 1000 j for j = jack0 .. jack0 + jacks - 1, walker ids
    0..walkers-1 in each).

    mode "plain": the reference's defaults (ln W = 0, NUPDATES = 1, the
    x1.01 schedule once the histogram spans all bins, random start).
    mode "seeded" / "nonet": NUPDATES = s0 sweeps (s0 N attempts per walker)
    from the start, multiplied by `growth` each iteration once the histogram
    spans all bins, up to smax sweeps; "seeded" starts from the surrogate's
    ln W and "nonet" from ln W = 0 (the same schedule without the network).
    Both use the stratified start (walkers spread over D)."""
    N = L * L
    ins = []
    for j in range(jack0, jack0 + jacks):
        q = dict(L=L, walkers=walkers, seed=seed0 + 1000 * j, beta=beta, recursion=False)
        if mode == "plain":
            q["init_nup"] = 1
        else:
            q.update(init_nup=int(s0 * N), nup_growth=float(growth), nup_max=int(smax * N))
        ins.append(q)
    return ins


def jackknife_lnw(lnws):
    """Per-bin jackknife mean and error of J independent ln W curves (each
    normalised to ln W(0) = 0), as the reference's jakknife
    (muca_delta_multigpu.cu:620-635) applied bin by bin."""
    a = np.asarray(lnws, np.float64)
    J = a.shape[0]
    tot = a.sum(axis=0)
    loo = (tot[None, :] - a) / (J - 1)
    mean = tot / J
    err = np.sqrt((J - 1) / J * ((loo - mean[None, :]) ** 2).sum(axis=0))
    return mean, err


def converge(instances, *, lnw_init=None, stratified=True, max_gpu_s=60.0, chunk=20,
             device=0, tiles=KERNEL_AUTO, log=None):
    """Runs the batched weight iteration until every instance meets the
    reference's flatness test (width == N and D_KL < 1e-4) or max_gpu_s of
    device time is spent (checked between calls of `chunk` iterations; the
    chunk adapts to about 2-20 s of device time per call).  Returns per-instance dicts with the converged (or
    last) ln W, the status, and the device seconds spent while it was active."""
    b = Batch(instances, device=device, precision=REF32, tiles=tiles)
    try:
        b.init(0, stratified=stratified)
        if lnw_init is not None:
            for i, lw in enumerate(lnw_init):
                if lw is not None:
                    b.set_log_weight(i, lw)
        n = len(instances)
        gpu = 0.0
        active_s = [0.0] * n
        t0 = time.time()
        while True:
            st0 = b.status()
            ms = b.iterate(chunk)
            st = b.status()
            gpu += ms / 1e3
            # keep each call near 10 s of device time: once an instance's
            # histogram spans all bins its NUPDATES grows 1 % per iteration, so
            # the chunk is then kept short
            ramp = any(x["reach_iter"] != 0xFFFFFFFF and not x["converged"] for x in st)
            if ms > 20e3 and chunk > 1:
                chunk = max(1, chunk // 4)
            elif ms < 2e3 and chunk < 1000:
                chunk *= 2
            if ramp:
                chunk = min(chunk, 8)
            # device time is charged to the instances that were active in the chunk
            act = [i for i in range(n) if not st0[i]["converged"]]
            for i in act:
                active_s[i] += ms / 1e3
            if log and (time.time() - t0) > 10:
                t0 = time.time()
                log(gpu, st)
            if all(s["converged"] for s in st) or gpu >= max_gpu_s:
                break
        out = []
        for i in range(n):
            out.append(dict(status=st[i], lnw=b.get_log_weight(i), hist=b.get_hist(i),
                            gpu_s=active_s[i]))
        return out, gpu
    finally:
        b.close()


def main(argv=None):
    """python -m muca_amd.surrogate extend --L 72 96 --T 0.5 0.6 0.7 0.8
    [--corpus PATH] [--walkers 131072] [--s0 50] [--growth 2] [--smax 3200]
    [--cap 600]: converge new (L, T) curves from surrogate seeds and add the
    flat ones to the corpus file."""
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(prog="python -m muca_amd.surrogate")
    ap.add_argument("command", choices=["extend"])
    ap.add_argument("--L", type=int, nargs="+", required=True)
    ap.add_argument("--T", type=float, nargs="+", required=True)
    ap.add_argument("--corpus", default=CORPUS_PATH)
    ap.add_argument("--walkers", type=int, default=131072)
    ap.add_argument("--s0", type=float, default=50.0)
    ap.add_argument("--growth", type=float, default=2.0)
    ap.add_argument("--smax", type=float, default=3200.0)
    ap.add_argument("--cap", type=float, default=600.0)
    a = ap.parse_args(argv)

    def log(gpu, st):
        print(f"gpu {gpu:.1f}s " + " ".join(
            f"[it {x['iter']} w {x['width']} dkl {x['dkl']:.1e} c {x['converged']}]" for x in st),
            file=sys.stderr, flush=True)

    added, gpu, out = extend_corpus(a.corpus, a.L, a.T, walkers=a.walkers, s0=a.s0,
                                    growth=a.growth, smax=a.smax, max_gpu_s=a.cap, log=log)
    print(json.dumps({"added": added, "gpu_s": gpu,
                      "runs": [{"L": q, "T": T, "flat": bool(o["status"]["converged"]),
                                "iterations": o["status"]["iter"], "gpu_s": o["gpu_s"],
                                "last_dkl": o["status"]["dkl"]}
                               for (q, T), o in zip([(L, T) for L in a.L for T in a.T], out)]}))


if __name__ == "__main__":
    main()
