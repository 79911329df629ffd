"""Pins the CPU MLP reference: its analytic gradients (without BF16
rounding) against central finite differences of its loss in float64."""
import numpy as np
import pytest

import mlp_ref


def test_hidden_width_follows_the_parameter_count():
    for h in (32, 64, 96, 128, 256):
        n = mlp_ref.num_params(h)
        assert n == 96 * h + h + h * h + h + h * 32 + 32
        assert mlp_ref.hidden_of(n) == h
    with pytest.raises(ValueError):
        mlp_ref.hidden_of(mlp_ref.num_params(64) + 1)


@pytest.mark.parametrize("hid", [mlp_ref.HID, 32])
def test_reference_gradients_match_finite_differences(hid):
    rng = np.random.default_rng(0)
    n = mlp_ref.num_params(hid)
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


def test_spread_top4_from_probed_bins():
    """The second probe's rule: bins already probed are never picked, and
    bins near them are picked only when no far bin is left."""
    z = np.zeros(32)
    z[8], z[9], z[23], z[12], z[0] = 9.0, 8.5, 8.0, 7.0, 1.0
    # nothing near the probed bins 8 and 24 is excluded: 9 leads, then 23
    # would be 14 from 9 (far), so it follows
    w = mlp_ref.spread_top4(z, taken0=(8, 24))
    bins = [(w >> (8 * k)) & 0xff for k in range(4)]
    assert 8 not in bins and 24 not in bins
    assert bins[:2] == [9, 23]
    # radius 4 around 8 and 24: 9 and 23 are near, 12 (4 away) is the best far bin,
    # then 0 (12 from 12, 8 from 8)
    far0 = 0xffffffff
    for b in (8, 24):
        for d in range(-3, 4):
            far0 &= ~(1 << (b + d))
    w = mlp_ref.spread_top4(z, taken0=(8, 24), far0=far0)
    assert [(w >> (8 * k)) & 0xff for k in (0, 1)] == [12, 0]
    rng = np.random.default_rng(1)
    for _ in range(200):
        w = mlp_ref.spread_top4(rng.standard_normal(32), taken0=(8, 24), far0=far0)
        b = [(w >> (8 * k)) & 0xff for k in range(4)]
        assert len(set(b)) == 4 and 8 not in b and 24 not in b

