"""Pin the float64 seed-network oracle: its gradient against central finite
differences of its own loss (pure float64, no BF16 emulation), and the BF16
rounding helper against exactly representable values."""
import numpy as np

from oracle import mlp


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 0.0])
    got = mlp.bf16(vals)
    # BF16 keeps 7 fraction bits: 1 + 2^-9 rounds down, 1 + 3*2^-9 up
    want = np.array([1.0, 1.0 + 2 ** -7, 1.0, 1.0 + 2 ** -7, -2.5, 0.0])
    np.testing.assert_array_equal(got, want)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    p = rng.normal(scale=0.3, size=mlp.NPARAM)
    x = rng.normal(size=(37, 6))
    t = rng.normal(size=37)
    g, loss = mlp.grad(p, x, t, 1.0 / 37, emulate_bf16=False)
    idx = np.concatenate([rng.choice(np.arange(lo, hi), 12, replace=False)
                          for lo, hi in ((0, 384), (384, 448), (448, 4544), (4544, 4608),
                                         (4608, 4672))] + [np.array([4672])])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        lp = mlp.grad(pp, x, t, 1.0 / 37, emulate_bf16=False)[1]
        lm = mlp.grad(pm, x, t, 1.0 / 37, emulate_bf16=False)[1]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
