"""Pins the CPU MLP reference: its analytic gradients (without BF16
rounding) against central finite differences of its loss in float64."""
import numpy as np

import mlp_ref


def test_reference_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    n = sum(int(np.prod(s)) for _, s in mlp_ref.SIZES)
    p = (rng.standard_normal(n) * 0.08).astype(np.float64)
    x = rng.uniform(-1, 1, (8, mlp_ref.IN))
    labels = rng.integers(0, 2**32, 8, dtype=np.uint64).astype(np.uint32)
    loss, g = mlp_ref.forward_backward(p, x, labels, rounding=False)
    assert abs(loss - mlp_ref.loss_only(p, x, labels)) < 1e-12
    idx = rng.choice(n, 60, replace=False)
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp_ref.loss_only(pp, x, labels) - mlp_ref.loss_only(pm, x, labels)) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    v = np.float32([1.0, 1.00390625, 1.005859375, -2.5, 3.140625, 1e-3])
    r = mlp_ref.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0          # tie rounds to even
    assert r[2] == np.float32(1.0078125)         # above the tie rounds up
    assert r[3] == -2.5
    assert np.all(np.abs(r - v) <= np.abs(v) * 2.0 ** -8)


def test_spread_top4_picks_far_bins():
    z = np.zeros(32)
    z[10], z[11], z[12], z[20], z[3] = 9.0, 8.0, 7.0, 6.0, 5.0
    w = mlp_ref.spread_top4(z)
    bins = [(w >> (8 * k)) & 0xff for k in range(4)]
    # 10 first; 11 and 12 are within 8 of it, so 20 (>= 8 away) comes next,
    # then the best bin >= 8 from both (3 is 7 from 10: too close -> 28..31
    # or 0..2 at zero; ties go to the lowest bin, 0)
    assert bins[:2] == [10, 20]
    assert bins[2] == 0
    # plain top-4 is the spread-0 rule
    plain = mlp_ref.spread_top4(z, spread=0)
    assert [(plain >> (8 * k)) & 0xff for k in range(4)] == [10, 11, 12, 20]
    # every word names four distinct bins
    rng = np.random.default_rng(0)
    for _ in range(200):
        w = mlp_ref.spread_top4(rng.standard_normal(32))
        assert len({(w >> (8 * k)) & 0xff for k in range(4)}) == 4
