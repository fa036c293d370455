"""Pins the CPU restatement of the surrogate network (oracle/mlp_oracle.py):
its float64 backward pass against central finite differences, and its BF16
rounding against the definition (round to nearest, ties to even)."""
import numpy as np

from oracle import mlp_oracle as MO


def test_bf16_rounding_ties_to_even():
    f = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9,
                  -2.5, 3.0e-5], np.float32)
    r = MO.bf16(f)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 (even mantissa wins: 1);
    # 1 + 3 * 2^-8 a tie between 1 + 2^-7 and 1 + 2^-6 (even: 1 + 2^-6)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6
    assert r[3] == 1.0 and r[4] == 1.0 + 2 ** -7
    assert r[5] == -2.5
    assert abs(r[6] - 3.0e-5) <= 3.0e-5 * 2 ** -8


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = MO.init_params(0).astype(np.float64)
    x = rng.uniform(-1, 1, (7, 4))
    t = rng.uniform(-1, 1, 7)
    _, _, g = MO.forward_backward(p, x, t, bf16_round=False)
    idx = np.concatenate([rng.choice(np.arange(a, b), min(6, b - a), replace=False)
                          for a, b in [(MO.OW1, MO.OB1), (MO.OB1, MO.OW2), (MO.OW2, MO.OB2),
                                       (MO.OB2, MO.OW3), (MO.OW3, MO.OB3), (MO.OB3, MO.OW4),
                                       (MO.OW4, MO.OB4), (MO.OB4, MO.NPARAMS)]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        _, lp, _ = MO.forward_backward(pp, x, t, bf16_round=False)
        _, lm, _ = MO.forward_backward(pm, x, t, bf16_round=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_pack_unpack_roundtrip():
    p = MO.init_params(2).astype(np.float64)
    assert np.array_equal(MO.pack(MO.unpack(p)), p)
    assert p.size == MO.NPARAMS == 33793
