"""Pins of the oracle's normal(0,1) fill (CPU only).

The reference has no normal fill; the synthetic workloads define one
(Box-Muller on the splitmix64 stream of tenet/src/typed.rs:2086-2092).  What
fixes it independently of the oracle's own code:

* splitmix64's published known-answer sequence for seed 1234567 (the first
  five outputs of Vigna's reference generator);
* the first Box-Muller pairs computed here at 50 digits (mpmath) from those
  known draws, with the mapping u = ((z >> 11) + 1) 2^-53, element 2p = r cos,
  2p + 1 = r sin -- a swapped sin/cos, a dropped +1, 2 pi replaced by pi, or a
  draw-order slip fails this;
* ln and sin / cos of the oracle's own arithmetic against 50-digit values
  over 10^5 points and at the special points (u = 2^-53, 1; t = 1/4, 1/2, 3/4, 1);
* the distribution: moments and a Kolmogorov-Smirnov test over 10^6 draws.

The device fill (tnt_buf_fill_random, normal != 0) evaluates the same
operation sequence and is compared bit for bit with the oracle in
tests/test_gpu_parity.py.
"""
import mpmath
import numpy as np
import pytest
from scipy import stats

from oracle.tensor import log_unit, sincos_2pi, splitmix_normal, splitmix_raw

mpmath.mp.dps = 50

# splitmix64, seed 1234567: the published first outputs of the reference generator
SPLITMIX_KAT_1234567 = [6457827717110365317, 3203168211198807973, 9817491932198370423,
                        4593380528125082431, 16408922859458223821]


def ulp_err(got, want_mp):
    want = float(want_mp)
    return abs(mpmath.mpf(got) - want_mp) / mpmath.mpf(np.spacing(abs(want)) if want != 0 else 2.0 ** -1074)


def test_splitmix64_known_answers():
    assert [int(x) for x in splitmix_raw(1234567, 5)] == SPLITMIX_KAT_1234567


def test_first_pairs_by_hand():
    z = SPLITMIX_KAT_1234567[:4]
    u = [mpmath.mpf((v >> 11) + 1) / mpmath.mpf(2) ** 53 for v in z]
    want = []
    for p in range(2):
        r = mpmath.sqrt(-2 * mpmath.log(u[2 * p]))
        a = 2 * mpmath.pi * u[2 * p + 1]
        want += [r * mpmath.cos(a), r * mpmath.sin(a)]
    got = splitmix_normal(1234567, 4)
    for g, w in zip(got, want):
        assert abs(mpmath.mpf(g) - w) <= 8 * mpmath.mpf(np.spacing(abs(float(w)))), (g, w)
    # odd length: the last pair's sine is dropped, the cosine kept
    assert splitmix_normal(1234567, 3).tolist() == got[:3].tolist()


def test_unit_mapping_has_no_zero():
    # u = ((z >> 11) + 1) 2^-53 lies in [2^-53, 1]: ln u is finite, r = 0 at u = 1
    lo, hi = log_unit(np.array([2.0 ** -53, 1.0]))
    assert lo == pytest.approx(float(mpmath.log(mpmath.mpf(2) ** -53)), rel=0, abs=4e-15)
    assert hi == 0.0


def test_log_unit_accuracy():
    rng = np.random.default_rng(11)
    u = np.concatenate([rng.random(20000), 2.0 ** -rng.uniform(0, 53, 20000),
                        [2.0 ** -53, 0.5, 0.70710678118654746, 0.70710678118654757, 1 - 2.0 ** -53, 1.0]])
    u = np.clip(u, 2.0 ** -53, 1.0)
    got = log_unit(u)
    worst = max(float(ulp_err(g, mpmath.log(mpmath.mpf(float(x))))) for g, x in zip(got, u) if x != 1.0)
    assert worst <= 2.0, worst


def test_sincos_2pi_accuracy_and_special_points():
    rng = np.random.default_rng(12)
    t = np.concatenate([rng.random(20000), [2.0 ** -53, 0.125, 0.375, 1 - 2.0 ** -53]])
    s, c = sincos_2pi(t)
    for x, sg, cg in zip(t, s, c):
        a = 2 * mpmath.pi * mpmath.mpf(float(x))
        assert abs(mpmath.mpf(sg) - mpmath.sin(a)) < 4e-16
        assert abs(mpmath.mpf(cg) - mpmath.cos(a)) < 4e-16
    s, c = sincos_2pi(np.array([0.25, 0.5, 0.75, 1.0]))
    assert s.tolist() == [1.0, 0.0, -1.0, 0.0]
    assert np.array_equal(np.abs(c), [0.0, 1.0, 0.0, 1.0])
    assert c[1] == -1.0 and c[3] == 1.0
    # sin^2 + cos^2 = 1 to rounding everywhere
    s, c = sincos_2pi(rng.random(100000))
    assert np.max(np.abs(s * s + c * c - 1.0)) < 1e-15


def test_normal_distribution():
    x = splitmix_normal(7, 1_000_000)
    n = x.size
    assert np.all(np.isfinite(x))
    assert abs(x.mean()) < 5 / np.sqrt(n)
    assert abs(x.var() - 1.0) < 5 * np.sqrt(2.0 / n)
    assert abs(stats.skew(x)) < 5 * np.sqrt(6.0 / n)
    assert abs(stats.kurtosis(x)) < 5 * np.sqrt(24.0 / n)
    assert stats.kstest(x, "norm").pvalue > 1e-3
    # pairs are independent: the correlation of elements 2p and 2p+1 vanishes
    assert abs(np.corrcoef(x[0::2], x[1::2])[0, 1]) < 5 / np.sqrt(n / 2)


def test_seed_determinism_and_prefix():
    a = splitmix_normal(2 ** 63 + 5, 1001)
    b = splitmix_normal(2 ** 63 + 5, 2000)
    assert np.array_equal(a, b[:1001])
    assert not np.array_equal(a, splitmix_normal(2 ** 63 + 6, 1001))
