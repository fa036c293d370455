"""The drop-in CLIs: argv, stdout and files of multigpu_weight / produce_gpu.

GPU tests compare every file the CLIs write with a replica of the
reference's host loop (muca_delta_multigpu.cu:329-524,
produce2_multigpu.cu:283-358) driven by the CPU oracle.  Numbers are
formatted as C++ iostreams do (precision 6 or 10, %g style); timing columns
are only checked for being numbers.
"""
import os
import subprocess

import numpy as np
import pytest

import muca_amd as M

BIN = M.BIN_DIR
WEIGHT = os.path.join(BIN, "multigpu_weight")
PRODUCE = os.path.join(BIN, "produce_gpu")
NTHREADS = min(16, os.cpu_count() or 4)


def g(x, p=6):
    """C++ ostream formatting of a double at precision p (default floatfield)."""
    if np.isnan(x):
        return "-nan" if np.signbit(x) else "nan"
    return "%.*g" % (p, x)


def run(args, **kw):
    return subprocess.run(args, capture_output=True, text=True, timeout=300, **kw)


# ---- CPU: argument and file errors -----------------------------------------------
def test_weight_usage_error():
    p = run([WEIGHT, "16", "0"])
    assert p.returncode == 1 and "usage" in p.stderr


def test_weight_missing_init_file(tmp_path):
    p = run([WEIGHT, "8", "0", "0", "1", "1", "0", "1", f"--outdir={tmp_path}"])
    assert p.returncode == 1 and "init weight file" in p.stderr


def test_produce_missing_weight_file(tmp_path):
    p = run([PRODUCE, "8", "0", "0", "2", "1", "1", "1", "1.9", "0.1", f"--outdir={tmp_path}"])
    assert p.returncode == 1 and p.stdout.startswith("!error!")


def test_produce_short_weight_file(tmp_path):
    d = tmp_path / "fss" / "L4"
    d.mkdir(parents=True)
    (d / "weight_mpi_T050c16_M200_JACK0_DKL1e-4.txt").write_text("0 0\n1 -1\n")
    p = run([PRODUCE, "4", "0", "0", "2", "1", "1", "1", "1.9", "0.1", f"--outdir={tmp_path}"])
    assert p.returncode == 1 and "fewer than N+1" in p.stderr


def test_produce_moments_excludes_host_observables(tmp_path):
    p = run([PRODUCE, "4", "0", "0", "2", "1", "1", "1", "1.9", "0.1", f"--outdir={tmp_path}",
             "--moments", "--host-observables"])
    assert p.returncode == 1 and "--moments" in p.stderr


# ---- replica of the reference weight loop (all jacks) ------------------------------
def weight_replica(orc, L, W, jacks, init_nup, seed, recursion, lnw0=None, max_iter=0):
    N = L * L
    lnw = np.zeros(N + 1) if lnw0 is None else lnw0.copy()
    lnw32 = lnw.astype(np.float32)
    p_acc = np.zeros(N + 1)
    width = 10
    out = []
    for mm in range(jacks):
        lat = orc.init_lattice(orc.KEYING_WEIGHT, L, W, mm=mm, gpu=0, seed=seed)
        E, D, _ = orc.state(L, lat)
        nup, total_run, reach, rows, conv = init_nup, 0, False, [], False
        k = 0
        while not (max_iter and k >= max_iter):
            if width == N:
                nup = int((nup + 1) * 1.01) & 0xFFFFFFFF
            total_run += (nup * W) & 0xFFFFFFFF
            h = orc.mcs(L, lat, E, D, None, keying=0, fp64=False, lnw32=lnw32, iter=k, seed=seed,
                        therm=N, nupdates=nup, nthreads=NTHREADS)
            s, e = orc.hist_range(h)
            wt = (e - s) & 0xFFFFFFFF
            width = max(width, wt)
            dkl = orc.kull_leibler(h)
            rows.append((k, wt, dkl, nup, total_run))
            if width == N and not reach:
                reach = True
            if width == N and dkl < 1e-4:
                conv = True
                break
            if recursion:
                orc.weight_recursion(h, p_acc, s, e, lnw)
            else:
                orc.update_trivial(h, lnw)
            lnw32 = orc.normalise(lnw)
            k += 1
        out.append(dict(rows=rows, lnw=lnw.copy(), D0=int(D[0]), converged=conv,
                        D_init=int(orc.state(L, orc.init_lattice(0, L, 1, mm=mm, seed=seed))[1][0])))
    return out


@pytest.mark.gpu
def test_weight_cli_files_match_replica(orc, tmp_path):
    """Trivial update, two jacks (ln W and width carry over between jacks, as
    in the reference), run to convergence."""
    L, W, jacks, seed, recursion = 4, 64, 2, 1234, 0
    N = L * L
    p = run([WEIGHT, str(L), "0", "0", str(jacks), "1", str(recursion), "0", f"--walkers={W}",
             f"--seed={seed}", "--max-iter=5000", f"--outdir={tmp_path}"])
    assert p.returncode == 0, p.stderr
    rep = weight_replica(orc, L, W, jacks, 1, seed, recursion)
    assert all(r["converged"] for r in rep)
    lines = p.stdout.split("\n")
    assert lines[0] == f"E_DELTA : {rep[0]['D_init']}"
    assert lines[1] == "---------------------------------------- "
    assert p.stdout.count("D_KL is below 1e-4 !! ") == jacks
    d = tmp_path / "fss" / f"L{L}"
    for mm in range(jacks):
        wf = d / f"weight_gpu1_T050c{W}_M1_therm1N_JACK{mm}_DKL1e-4_iterupdate1.01_rand.txt"
        want = "".join(f"{i} {g(v, 10)}\n" for i, v in enumerate(rep[mm]["lnw"]))
        assert wf.read_text() == want, mm
        sf = d / f"shortiter_gpu1_T050c{W}_M1_therm1N_Jack{mm}_iterupdate1.01_rand.txt"
        rows = sf.read_text().split("\n")
        assert rows[0] == "#iter width D_KL mcs_time memcpy_time accumulated_time NUPDATES acc_NUP"
        body = [r.split(" ") for r in rows[1:] if r]
        assert len(body) == len(rep[mm]["rows"])
        for got, (k, wt, dkl, nup, tr) in zip(body, rep[mm]["rows"]):
            assert got[0] == str(k) and got[1] == str(wt) and got[2] == g(dkl)
            assert got[6] == str(nup) and got[7] == str(tr)
            float(got[3]), float(got[4]), float(got[5])
    cr = d / f"converge_result_gpu1_T050c{W}M1_therm1N_iterupdate1.01_rand.txt"
    rows = cr.read_text().split("\n")
    assert rows[0].startswith("#total_time_mean total_time_std")
    assert rows[2] == "#total_mcstime total_memtime Nwidth_time total_run Nwidth_step"
    for mm in range(jacks):
        cols = rows[3 + mm].split(" ")
        assert cols[3] == str(rep[mm]["rows"][-1][4])


@pytest.mark.gpu
def test_weight_cli_recursive_rows_match_replica(orc, tmp_path):
    """recursive_update=1 (weight_recursion): per-iteration rows of the first
    300 iterations equal the replica's."""
    L, W, seed, cap = 4, 64, 99, 300
    p = run([WEIGHT, str(L), "0", "0", "1", "1", "1", "0", f"--walkers={W}", f"--seed={seed}",
             f"--max-iter={cap}", f"--outdir={tmp_path}"])
    assert p.returncode == 0, p.stderr
    rep = weight_replica(orc, L, W, 1, 1, seed, 1, max_iter=cap)
    sf = tmp_path / "fss" / f"L{L}" / f"shortiter_gpu1_T050c{W}_M1_therm1N_Jack0_iterupdate1.01_rand.txt"
    body = [r.split(" ") for r in sf.read_text().split("\n")[1:] if r]
    assert len(body) == len(rep[0]["rows"])
    for got, (k, wt, dkl, nup, tr) in zip(body, rep[0]["rows"]):
        assert got[0] == str(k) and got[1] == str(wt) and got[2] == g(dkl)
        assert got[6] == str(nup) and got[7] == str(tr)


@pytest.mark.gpu
def test_weight_cli_max_iter_not_converged(tmp_path):
    p = run([WEIGHT, "8", "0", "0", "1", "1", "0", "0", "--walkers=32", "--max-iter=3",
             f"--outdir={tmp_path}"])
    assert p.returncode == 0 and "Noooo...not converged" in p.stdout
    d = tmp_path / "fss" / "L8"
    assert len((d / "shortiter_gpu1_T050c32_M1_therm1N_Jack0_iterupdate1.01_rand.txt")
               .read_text().strip().split("\n")) == 4


@pytest.mark.gpu
def test_weight_cli_continuing_reads_init_weights(orc, tmp_path):
    """continuing=1 seeds ln W from initweight_L128_T050.txt (whitespace-separated
    N+1 values, muca_delta_multigpu.cu:223-235)."""
    L, W, seed = 4, 64, 7
    N = L * L
    lnw0 = -0.5 * np.arange(N + 1)
    d = tmp_path / "fss" / f"L{L}"
    d.mkdir(parents=True)
    (d / "initweight_L128_T050.txt").write_text(" ".join(repr(float(x)) for x in lnw0))
    p = run([WEIGHT, str(L), "0", "0", "1", "1", "0", "1", f"--walkers={W}", f"--seed={seed}",
             "--max-iter=4000", f"--outdir={tmp_path}"])
    assert p.returncode == 0, p.stderr
    rep = weight_replica(orc, L, W, 1, 1, seed, 0, lnw0=lnw0)
    wf = d / f"weight_gpu1_T050c{W}_M1_therm1N_JACK0_DKL1e-4_fromextrapolatedinit1.txt"
    assert wf.read_text() == "".join(f"{i} {g(v, 10)}\n" for i, v in enumerate(rep[0]["lnw"]))


def produce_replica(orc, L, W, B, rng_, therm_m, nup_m, d0, gap, lnw, seed=100):
    N = L * L
    lat = orc.init_lattice(orc.KEYING_PRODUCE, L, W, seed=seed)
    E, D, Mg = orc.state(L, lat)
    lw = lnw.copy()
    delta = d0
    res = []
    bw = (W + B - 1) // B
    for l in range(rng_):
        delta = delta + gap / rng_
        orc.weight_shift(2.0, delta, lw)
        bz = orc.boltz(2.0, delta, lw)
        orc.mcs(L, lat, E, D, Mg, keying=1, fp64=True, lnw64=lnw, iter=0, seed=seed + l,
                therm=N * therm_m, nupdates=0, nthreads=NTHREADS)
        Ms, ss = [], []
        for b in range(B):
            sl = slice(b * bw, min(W, (b + 1) * bw))
            sub = np.ascontiguousarray(lat[sl])
            e, dd, m = E[sl].copy(), D[sl].copy(), Mg[sl].copy()
            j = np.zeros((N + 1) * (N + 2) // 2, np.uint64)
            orc.mcs(L, sub, e, dd, m, keying=1, fp64=True, lnw64=lnw, iter=l, seed=seed,
                    therm=0, nupdates=N * nup_m, w0=b * bw, joint=j)
            lat[sl], E[sl], D[sl], Mg[sl] = sub, e, dd, m
            a, c = orc.observables(2.0, N, j, bz)
            Ms.append(a)
            ss.append(c)
        res.append((d0 + gap * (l + 1) / rng_, orc.jackknife(Ms), orc.jackknife(ss)))
    return res


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["host", "device", "moments"])
def test_produce_cli_observables(orc, tmp_path, mode):
    host_obs = mode == "host"
    L, W, B, R = 4, 64, 4, 3
    N = L * L
    lnw = -0.8 * np.arange(N + 1) + 0.01 * np.arange(N + 1) ** 2
    d = tmp_path / "fss" / f"L{L}"
    d.mkdir(parents=True)
    (d / "weight_mpi_T050c16_M200_JACK0_DKL1e-4.txt").write_text(
        "".join(f"{i} {g(v, 10)}\n" for i, v in enumerate(lnw)))
    lnw_file = np.array([float(g(v, 10)) for v in lnw])
    args = [PRODUCE, str(L), "0", "0", str(R), "1", "20", "50", "1.9", "0.1", f"--walkers={W}",
            f"--jack-blocks={B}", f"--outdir={tmp_path}"]
    if host_obs:
        args.append("--host-observables")
    if mode == "moments":
        args.append("--moments")
    p = run(args)
    assert p.returncode == 0, p.stderr
    assert f" TOTAL UPDATES : {N * 50 * W} updates by each core : {N * 50} Therm run : {N * 20}" \
        in p.stdout
    rep = produce_replica(orc, L, W, B, R, 20, 50, 1.9, 0.1, lnw_file)
    text = (d / "observables_gpu1_T050_JACKgpu_THERM20N_NUP50N_delta1.9to0.1noreset.txt") \
        .read_text().strip().split("\n")
    assert len(text) == R
    for i, (line, (dl, (am, ae), (cm, ce))) in enumerate(zip(text, rep)):
        cols = line.split(" ")
        assert cols[0] == g(dl, 6 if i == 0 else 10)
        got = [float(c) for c in cols[1:]]
        if host_obs:
            assert cols[1:] == [g(am, 10), g(ae, 10), g(cm, 10), g(ce, 10)]
        else:
            assert got == pytest.approx([am, ae, cm, ce], rel=1e-8)
