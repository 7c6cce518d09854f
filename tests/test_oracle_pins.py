"""Pins of the oracle's blockTransform, GLS shrinkage and sliding-window
functions against independent numpy restatements (tests/np_ref.py) and
against what the mathematics fixes.  CPU only.

* blockTransform (IFCS:9217-9373): AR(1) series whose decorrelation time
  puts the blocking optimum above q - 1, and runs at q = 3 so that the
  clamp ind = max(ind, q - 1) hides nothing: the index AND the success flag
  equal exactly; the cases cover several indices and both flag values.  The
  Java caller's zero-prepended, last-frame-dropped input (IFCS:9259-9262) is
  pinned by series with a spike in the last frame, where it decides the
  result.
* GLS (IFCS:9524-9613): the mean curve mcov and the regularized covariance
  against the numpy Schaefer-Strimmer restatement with the Java quirks; the
  quirks themselves are shown to matter; the limits of the shrinkage
  (weights 0: sample covariance; 1: the median target over 73 slots with
  the unused 0) hold for the restatement.
* Sliding windows (IFCS:9146-9172, 7331-7341): the windows' correlations
  and blocked variances from the numpy multi-tau averaged as term /
  numofsw; the correlator depth formula at worked values; one window equals
  the plain correlation.
"""
import math

import numpy as np
import pytest

import np_ref
import oracle


def _ar1(num, tau, seed, mean=100.0, sd=5.0):
    rng = np.random.default_rng(seed)
    phi = math.exp(-1.0 / tau)
    e = rng.normal(0, sd * math.sqrt(1 - phi * phi), num)
    x = np.empty(num)
    x[0] = rng.normal(0, sd)
    for t in range(1, num):
        x[t] = phi * x[t - 1] + e[t]
    return x + mean + rng.poisson(mean * 0.2, num)


CASES = [(2 ** 14 + 37, tau, seed) for tau in (1, 8, 40, 150, 600, 2500) for seed in range(5)]


@pytest.mark.parametrize("q", [3, 8])
def test_block_transform_index_and_flag_match_restatement(q):
    seen_ind, seen_flag = set(), set()
    for num, tau, seed in CASES:
        a = _ar1(num, tau, seed)
        b = _ar1(num, tau, seed) if seed % 2 else a     # ACF and unrelated-pair CCF
        a1, b1 = np.r_[0.0, a], np.r_[0.0, b]
        got = oracle.block_transform(a1, b1, q=q)
        ref = np_ref.block_transform(a1, b1, q=q)
        assert got == ref, (num, tau, seed, got, ref)
        seen_ind.add(got[0])
        seen_flag.add(got[1])
    if q == 3:
        assert len([i for i in seen_ind if i > q - 1]) >= 3, seen_ind
        assert seen_flag == {0, 1}, seen_flag


def test_block_transform_takes_the_zero_prepended_series():
    """The Java caller copies the 1-based trace 0-based: the blocking series
    is [0, a_1 .. a_{num-1}].  A huge spike in the last frame is then never
    seen, while it dominates every level of the unshifted blocking curve."""
    differs = 0
    for seed in range(6):
        a = _ar1(2 ** 14, 30, 100 + seed)
        a[-1] += 5e6
        a1 = np.r_[0.0, a]
        got = oracle.block_transform(a1, a1, q=3)
        assert got == np_ref.block_transform(a1, a1, q=3, java_shift=True)
        differs += got != np_ref.block_transform(a1, a1, q=3, java_shift=False)
    assert differs >= 1


def _gls_inputs(seed, num=8192 + 5, ind=7):
    rng = np.random.default_rng(seed)
    a = rng.poisson(60, num).astype(float) + np.convolve(rng.normal(0, 8, num), np.ones(6) / 6, "same")
    return a, ind


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_gls_mean_curve_and_shrunk_covariance_match_restatement(seed):
    a, ind = _gls_inputs(seed)
    mcov, _, _, cov = oracle.calculate_cf(np.r_[0, a], np.r_[0, a], ind, gls=True)
    _, _, Y = np_ref.multitau(a, a, ind)
    np.testing.assert_allclose(mcov[1:], Y[:, 1:].mean(0), rtol=1e-9, atol=1e-15)
    ref, lam_v, lam_c, med = np_ref.shrink_covariance(Y[:, 1:])
    np.testing.assert_allclose(cov, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())
    assert 0.0 <= lam_v <= 1.0 and 0.0 < lam_c < 1.0
    # both Java quirks change the result measurably
    alt1, *_ = np_ref.shrink_covariance(Y[:, 1:], median_zero_slot=False)
    alt2, *_ = np_ref.shrink_covariance(Y[:, 1:], running_rescale=False)
    assert np.abs(alt1 - ref).max() > 1e-9 * np.abs(ref).max()
    assert np.abs(alt2 - ref).max() > 0


def test_gls_shrinkage_limits():
    """The shrinkage interpolates between the sample covariance (weights 0)
    and the diagonal median target (weights 1); the estimated weights of a
    correlated series lie strictly inside."""
    a, ind = _gls_inputs(4)
    _, _, Y = np_ref.multitau(a, a, ind)
    Yc = Y[:, 1:]
    n = Yc.shape[0]
    S = np.cov(Yc, rowvar=False)
    zero, *_ = np_ref.shrink_covariance(Yc, lam=(0.0, 0.0))
    np.testing.assert_allclose(zero, S / n, rtol=1e-10)
    one, _, _, med = np_ref.shrink_covariance(Yc, lam=(1.0, 1.0))
    np.testing.assert_allclose(one, np.diag(np.full(72, med / n)), rtol=1e-12, atol=0)
    assert med == pytest.approx(np.sort(np.r_[0.0, np.diag(S)])[36], rel=1e-12)
    _, lv, lc, _ = np_ref.shrink_covariance(Yc)
    assert 0 < lv < 1 and 0 < lc < 1


def test_sliding_window_depth_formula():
    """lagnum = floor((ln(L / (20 + p)) + 1) / ln 2) (IFCS:7331-7341)."""
    assert oracle.sliding_window_q(2000) == 7
    assert oracle.sliding_window_q(1000) == 6
    assert oracle.sliding_window_q(10 ** 6) == 8      # capped at q
    assert oracle.sliding_window_q(150) == 3


def test_sliding_window_averages_the_windows():
    rng = np.random.default_rng(5)
    T, W, H = 6000, 3, 2
    stack = (100 + rng.poisson(30, (T, H, W))).astype(np.uint16)
    L, bg = 2000, 100
    for (px, py, px2, py2) in [(0, 0, None, None), (1, 1, 2, 1)]:
        acf, var, sd, ind, bis, qq = oracle.sliding_window_cf(stack, px, py, 1, 1, bg, 1, T, L, px2=px2, py2=py2)
        assert qq == 7
        G = V = S = 0.0
        n = T // L
        for i in range(n):
            a = stack[i * L:(i + 1) * L, py, px].astype(float) - bg
            b = a if px2 is None else stack[i * L:(i + 1) * L, py2, px2].astype(float) - bg
            k, flag = np_ref.block_transform(np.r_[0, a], np.r_[0, b], q=qq)
            g, v, _ = np_ref.multitau(a, b, k, q=qq)
            G, V, S = G + g / n, V + v / n, S + np.sqrt(v) / n
        np.testing.assert_allclose(acf, G, rtol=1e-9, atol=1e-13)
        np.testing.assert_allclose(var[1:], V[1:], rtol=1e-9)
        np.testing.assert_allclose(sd[1:], S[1:], rtol=1e-9)
        assert (ind, bis) == (k, flag)
    # a single window over the whole range is the plain correlation
    a = stack[:, 0, 0].astype(float) - bg
    acf, var, _, ind, _, qq = oracle.sliding_window_cf(stack, 0, 0, 1, 1, bg, 1, T, T)
    k, _ = oracle.block_transform(np.r_[0, a], np.r_[0, a], q=qq)
    c, v, _, _ = oracle.calculate_cf(np.r_[0, a], np.r_[0, a], k, q=qq)
    np.testing.assert_array_equal(acf, c)
    np.testing.assert_array_equal(var, v)
