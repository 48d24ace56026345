"""Pins the CPU FP64 reference of the warm-start network (tests/mlp_ref.py): its gradient equals
central finite differences of its loss, and its bfloat16 rounding is round-to-nearest-even."""
import numpy as np

import mlp_ref as R


def test_fp64_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = rng.normal(0.0, 0.3, R.NPARAM)
    x = rng.normal(0.0, 1.0, (96, R.NIN))
    t = rng.normal(0.0, 1.0, 96)
    g, _ = R.gradient(p, x, t, bf=False)
    idx = np.concatenate([rng.choice(np.arange(a, b), 12, replace=False) for a, b in
                          ((R.OW1, R.OB1), (R.OB1, R.OW2), (R.OW2, R.OB2), (R.OB2, R.OW3), (R.OW3, R.OB3))]
                         + [np.array([R.OB3])])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = np.mean((R.forward(pp, x, bf=False)[0] - t) ** 2)
        lm = np.mean((R.forward(pm, x, bf=False)[0] - t) ** 2)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    # 1 + 2^-8 is halfway between two bfloat16 neighbours (1 and 1 + 2^-7): ties to even -> 1
    assert R.bf16(np.float32(1.0 + 2 ** -8)) == 1.0
    assert R.bf16(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6
    assert R.bf16(np.float32(1.0 + 2 ** -8 + 2 ** -20)) == 1.0 + 2 ** -7
    x = np.random.default_rng(1).normal(0, 10, 1000).astype(np.float32)
    b = R.bf16(x)
    assert np.all(np.abs(b - x) <= np.abs(x) * 2.0 ** -8)
