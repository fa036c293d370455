"""Host-side pieces of the surrogate pipeline (no GPU): the target transform
round-trips, the corpus loader expands the Delta grid, the init-weight file
holds exactly N+1 values, and the committed corpus is a set of converged
weight curves (ln W(0) = 0, finite, ending near the ground-state line)."""
import numpy as np

from muca_amd import surrogate as S


def test_target_round_trip():
    L, beta = 12, 1.7
    N = L * L
    rng = np.random.default_rng(0)
    lnw = -2 * beta * np.arange(N + 1) + rng.normal(0, 3, N + 1)
    lnw -= lnw[0]
    t = S.targets(lnw, L, beta).astype(np.float64)
    back = S.lnw_from_output(t, L, beta)
    # float32 targets: relative 2^-24 of |t| * N / 10
    assert np.max(np.abs(back - lnw)) < 1e-5 * N


def test_features_ranges():
    x = S.features(np.linspace(0, 1, 5), 32, 2.0, 1.98)
    assert x.shape == (5, 4) and x.dtype == np.float32
    assert np.allclose(x[:, 1], 0.5) and np.allclose(x[:, 2], 1.0)
    assert np.allclose(x[:, 3], 0.3, atol=1e-6)


def test_committed_corpus():
    runs = S.load_corpus()
    keys = {(r["L"], round(1 / r["beta"], 3)) for r in runs}
    assert len(runs) == len(keys) and len(keys) >= 12
    for r in runs:
        # one Delta per sample, spread over the grid's range
        assert r["delta"].shape == r["lnw"].shape
        assert 1.90 <= r["delta"].min() and r["delta"].max() <= 2.00
    for r in runs:
        N = r["L"] ** 2
        lnw = r["lnw"]
        assert lnw.shape == (N + 1,) and lnw[0] == 0.0 and np.all(np.isfinite(lnw))
        # the ground-state line -2 beta D carries most of ln W at D = N
        assert abs(lnw[-1] / N + 2 * r["beta"]) < 0.1
    X, t, w = S.corpus(runs)
    assert len(X) == len(t) == len(w) == sum(r["L"] ** 2 + 1 for r in runs)
    assert abs(w.mean() - 1.0) < 1e-6


def test_init_weight_file(tmp_path):
    lnw = np.array([0.0, -1.25, -2.5000000000000004, -3.0])
    p = tmp_path / "initweight.txt"
    S.write_init_weight(p, lnw)
    assert np.array_equal(np.loadtxt(p), lnw)
    assert len(p.read_text().split()) == 4


def test_equilibrium_nupdates():
    assert S.equilibrium_nupdates(8) == 4 * 64 * 64
    assert S.equilibrium_nupdates(8, 16) == 16 * 64 * 64


def test_jackknife_lnw_matches_scalar_jackknife():
    """Bin by bin, the same numbers as the reference's jakknife
    (muca_delta_multigpu.cu:620-635) on each bin's J values: the standard
    error of the mean sqrt(var / J) (sample variance), its closed form."""
    rng = np.random.default_rng(3)
    a = rng.normal(size=(5, 7))
    mean, err = S.jackknife_lnw(a)
    assert np.allclose(mean, a.mean(axis=0))
    assert np.allclose(err, a.std(axis=0, ddof=1) / np.sqrt(5))


def test_protocol_instances():
    ins = S.protocol_instances(128, 1 / 0.55, 4, 131072, "seeded", s0=50, growth=2.0,
                               smax=3200)
    assert [q["seed"] for q in ins] == [777, 1777, 2777, 3777]
    assert all(q["init_nup"] == 50 * 16384 and q["nup_max"] == 3200 * 16384 for q in ins)
    plain = S.protocol_instances(128, 1 / 0.55, 4, 131072, "plain")
    assert all(q["init_nup"] == 1 and "nup_growth" not in q for q in plain)
