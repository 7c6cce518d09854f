"""Pins the oracle (C restatement of the Java CPU path) against what the
mathematics fixes: closed forms, invariants, brute force on tiny inputs and
library routines (numpy least squares, scipy MINPACK).  CPU only."""
import math

import numpy as np
import pytest

import oracle

FT = 1e-3


# ---------------------------------------------------------------- tables
def test_tables_match_manual_groups():
    """Manual (DOC p39): group 1 lags 1..16, group 2 18..32, group 3 36..64;
    p=16, q=8 gives 73 channels ending at lag 2048 (IFCS:7391-7424)."""
    lag, samp, mtab, lagtime = oracle.tables(16, 8, 50000, FT)
    assert len(lag) == 73
    assert list(lag[1:17]) == list(range(1, 17))
    assert list(lag[17:25]) == list(range(18, 33, 2))
    assert list(lag[25:33]) == list(range(36, 65, 4))
    assert lag[-1] == 2048
    assert list(samp[:17]) == [1] * 17
    for g in range(2, 9):
        assert set(samp[g * 8 + 1: g * 8 + 9]) == {2 ** (g - 1)}
    np.testing.assert_array_equal(mtab, np.floor((50000 - lag) / samp))
    assert mtab[-1] == 374  # floor((50000 - 2048) / 128)
    np.testing.assert_allclose(lagtime, lag * FT)


def test_tables_of_other_panel_p():
    """p = 32: delays 0..32 at sample time 1, then 16 channels per group
    with lags 2^(g-1) (16 + y); p = 18 (2 mod 4): Java's integer base / 4
    gives the first group y + 8, so lag 9 appears twice and lag 18 not at all
    (IFCS:7398-7408)."""
    lag, samp, mtab, _ = oracle.tables(32, 6, 50000, FT)
    assert len(lag) == 32 + 5 * 16 + 1
    assert list(lag[:33]) == list(range(33)) and list(samp[:33]) == [1] * 33
    for g in range(2, 7):
        sl = slice(16 * g + 1, 16 * g + 17)
        assert list(lag[sl]) == [2 ** (g - 1) * (16 + y) for y in range(1, 17)]
        assert set(samp[sl]) == {2 ** (g - 1)}
    lag, samp, _, _ = oracle.tables(18, 4, 50000, FT)
    assert len(lag) == 18 + 3 * 9 + 1
    assert list(lag[:19]) == list(range(10)) + list(range(9, 18))
    assert list(samp[:19]) == [1] * 19
    assert list(lag[19:28]) == [2 * (y + 8) for y in range(1, 10)]


# ---------------------------------------------------------------- intensity
def _rand_stack(T=60, H=6, W=7, seed=0, lo=90, hi=140):
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi, size=(T, H, W)).astype(np.uint16)


@pytest.mark.parametrize("b", [1, 2, 3])
def test_get_intensity_bruteforce(b):
    st = _rand_stack()
    bg = 100
    d = oracle.get_intensity(st, 2, 1, b, b, bg, 3, 50)
    ref = (st[2:50, 1:1 + b, 2:2 + b].astype(np.int64) - bg).sum(axis=(1, 2))
    assert d[0] == 0
    np.testing.assert_array_equal(d[1:], ref)


def test_intensity_trace_integer_truncation():
    """int sum / int ave truncates toward zero, also for negative sums
    (background above the pixel values)."""
    st = _rand_stack(T=100, lo=95, hi=106, seed=3)
    bg = 100
    nopit, ave = 20, 5
    tr, it = oracle.intensity_trace(st, 1, 1, 2, 2, bg, 1, 100, nopit, FT)
    for x in range(nopit):
        s = int((st[x * ave:(x + 1) * ave, 1:3, 1:3].astype(np.int64) - bg).sum())
        assert tr[x] == math.trunc(s / ave)
        assert it[x] == pytest.approx(FT * (x + 0.5) * ave)
    assert (tr < 0).any() and (tr > 0).any()


# ---------------------------------------------------------------- bleach
def test_polyfit_is_least_squares_optimum():
    rng = np.random.default_rng(5)
    t = FT * (np.arange(1000) + 0.5) * 20
    y = np.floor(500 * np.exp(-t / 15) + 80 + rng.normal(0, 3, t.size))
    for order in (1, 2, 4, 6):
        c, st = oracle.polyfit(y, t, order)
        assert st == 0
        ref = np.polyfit(t, y, order)[::-1]
        fit_o = np.polyval(c[::-1], t)
        fit_r = np.polyval(ref[::-1], t)
        np.testing.assert_allclose(fit_o, fit_r, rtol=1e-8, atol=1e-8)


def test_bleach_correction_of_the_model_curve_is_constant():
    """I = f(t) exactly gives I' = c0 for every frame (IFCS:9744-9750)."""
    c = np.array([400.0, -20.0, 1.5, -0.05, 6e-4])
    num = 5000
    x = np.arange(1, num + 1)
    t = FT * (x + 0.5)
    f = np.polyval(c[::-1], t)
    d = np.concatenate([[0.0], f])
    out = oracle.bleach_apply(d, c, FT)
    np.testing.assert_allclose(out[1:], c[0], rtol=1e-12)


def test_bleach_correction_identity_without_bleaching():
    rng = np.random.default_rng(2)
    d = np.concatenate([[0.0], rng.normal(300, 20, 2000)])
    out = oracle.bleach_apply(d, [250.0, 0, 0], FT)
    np.testing.assert_allclose(out, d)


# ---------------------------------------------------------------- correlator
def _multitau_bruteforce(a, b, ind, p=16, q=8):
    """Direct evaluation of the multi-tau definitions (pair sums per level,
    window means per channel, block averages of 2^binct products)."""
    num = len(a)
    lag, samp, mtab, _ = oracle.tables(p, q, num, FT)
    n_ch = len(lag)
    pnum = int(math.floor(mtab[-1] / 2 ** max(ind - math.log2(samp[-1]), 0)))
    G, V = np.zeros(n_ch), np.zeros(n_ch)
    for x in range(n_ch):
        c = int(round(math.log2(samp[x])))
        nb = num >> c
        A = a[:nb << c].reshape(nb, 1 << c).sum(1)
        B = b[:nb << c].reshape(nb, 1 << c).sum(1)
        d = lag[x] >> c
        n = nb - d
        Ai, Bd = A[:n], B[d:d + n]
        am, bm = Ai.mean(), Bd.mean()
        prod = (Ai - am) * (Bd - bm)
        G[x] = prod.sum() / (n * am * bm)
        bs = 1 << (ind - c)
        nblk = n // bs
        P = prod[:nblk * bs].reshape(nblk, bs).mean(1)[:pnum]
        V[x] = (np.mean(P ** 2) - np.mean(P) ** 2) / ((pnum - 1) * (am * bm) ** 2)
    return G, V


@pytest.mark.parametrize("seed", [0, 1])
def test_calculate_cf_matches_bruteforce(seed):
    rng = np.random.default_rng(seed)
    num = 4096 + 37
    a = rng.poisson(200, num).astype(float) + np.convolve(rng.normal(0, 30, num), np.ones(20) / 20, "same")
    b = np.roll(a, 3) + rng.normal(0, 5, num)
    for ind in (7, 8):
        corav, bvar, _, _ = oracle.calculate_cf(np.r_[0, a], np.r_[0, b], ind)
        G, V = _multitau_bruteforce(a, b, ind)
        np.testing.assert_allclose(corav, G, rtol=1e-9, atol=1e-13)
        np.testing.assert_allclose(bvar[1:], V[1:], rtol=1e-9)


@pytest.mark.parametrize("p,q", [(32, 6), (18, 5), (8, 8)])
def test_calculate_cf_other_panel_p_matches_bruteforce(p, q):
    rng = np.random.default_rng(p)
    num = 8192 + 11
    a = rng.poisson(150, num).astype(float) + np.convolve(rng.normal(0, 25, num), np.ones(12) / 12, "same")
    b = np.roll(a, 2) + rng.normal(0, 4, num)
    for ind in (q - 1, q + 1):
        corav, bvar, _, _ = oracle.calculate_cf(np.r_[0, a], np.r_[0, b], ind, p=p, q=q)
        G, V = _multitau_bruteforce(a, b, ind, p=p, q=q)
        np.testing.assert_allclose(corav, G, rtol=1e-9, atol=1e-13)
        np.testing.assert_allclose(bvar[1:], V[1:], rtol=1e-9)


def test_poisson_closed_form():
    """i.i.d. Poisson: G(0) = 1/mean, G(tau > 0) = 0 within statistics."""
    rng = np.random.default_rng(4)
    m = 50.0
    a = rng.poisson(m, 2 ** 16).astype(float)
    corav, bvar, _, _ = oracle.calculate_cf(np.r_[0, a], np.r_[0, a], 7)
    assert corav[0] == pytest.approx(1 / m, rel=0.03)
    assert np.all(np.abs(corav[1:]) < 5 * np.sqrt(bvar[1:]) + 1e-4)


def test_period_two_sequence_sign_structure():
    """a_t = m + s (-1)^t: level-0 channels alternate in sign with |G| =
    (s/m)^2; pair-summed levels are constant so their G vanishes."""
    m, s, num = 100.0, 10.0, 8192
    a = m + s * (-1.0) ** np.arange(num)
    corav, _, _, _ = oracle.calculate_cf(np.r_[0, a], np.r_[0, a], 7)
    for x in range(17):
        assert corav[x] == pytest.approx((s / m) ** 2 * (-1) ** x, rel=1e-2)
    np.testing.assert_allclose(corav[17:], 0, atol=1e-12)


def test_ccf_index_direction_and_self_ccf():
    """b delayed by k frames peaks at channel lag k; a CCF of a series with
    itself equals its ACF."""
    rng = np.random.default_rng(7)
    num, k = 8192, 5
    x = rng.normal(0, 10, num + k) + 100
    a, b = x[k:], x[:-k]        # b_t = a_{t-k}: b lags a by k frames
    corav, _, _, _ = oracle.calculate_cf(np.r_[0, a], np.r_[0, b], 7)
    assert np.argmax(corav[:17]) == k
    assert corav[k] == pytest.approx(0.01, rel=0.1)
    swapped, _, _, _ = oracle.calculate_cf(np.r_[0, b], np.r_[0, a], 7)
    assert np.abs(swapped[:17]).max() < 0.1 * corav[k]
    acf1, v1, _, _ = oracle.calculate_cf(np.r_[0, a], np.r_[0, a], 7)
    ccf1, v2, _, _ = oracle.calculate_cf(np.r_[0, a], np.r_[0, a.copy()], 7)
    np.testing.assert_array_equal(acf1, ccf1)


def test_scale_invariance():
    rng = np.random.default_rng(8)
    a = rng.poisson(80, 4096).astype(float)
    c1, v1, _, _ = oracle.calculate_cf(np.r_[0, a], np.r_[0, a], 7)
    c2, v2, _, _ = oracle.calculate_cf(np.r_[0, 3 * a], np.r_[0, 3 * a], 7)
    np.testing.assert_allclose(c1, c2, rtol=1e-10)
    np.testing.assert_allclose(v1, v2, rtol=1e-9)


def test_block_transform_finds_decorrelation_time():
    """White noise blocks at the minimum index q-1; an AR(1) process with a
    long correlation time needs larger blocks."""
    rng = np.random.default_rng(9)
    num = 2 ** 15
    white = 100 + rng.normal(0, 5, num)
    ind_w, ok_w = oracle.block_transform(np.r_[0, white], np.r_[0, white])
    assert ind_w == 7
    phi = np.exp(-1 / 400)
    ar = np.zeros(num)
    e = rng.normal(0, 5, num)
    for t in range(1, num):
        ar[t] = phi * ar[t - 1] + e[t]
    ar += 100
    ind_a, _ = oracle.block_transform(np.r_[0, ar], np.r_[0, ar])
    assert ind_a >= 9


def test_gls_covariance_properties():
    rng = np.random.default_rng(10)
    a = rng.poisson(60, 8192).astype(float)
    corav, bvar, _, cov = oracle.calculate_cf(np.r_[0, a], np.r_[0, a], 7, gls=True)
    assert cov.shape == (72, 72)
    np.testing.assert_allclose(cov, cov.T, rtol=0, atol=0)
    w = np.linalg.eigvalsh(cov)
    assert w.min() > 0


# ---------------------------------------------------------------- model
def _model(ax=0.24e-6, s=0.38e-6, rx=0.0):
    return oracle.make_model(ax, ax, s, rx=rx)


def test_model_amplitude_at_zero_lag():
    """obsvol normalizes the lateral factor to 1 at tau = 0: G(0) = 1/N + G_inf.
    (Ttrip = 0 makes the triplet term 0/0 at tau = 0 in the Java expression,
    so a non-zero triplet time with Ftrip = 0 is used.)"""
    m = _model()
    for N, G in [(1.0, 0.0), (4.0, 0.01), (0.5, -0.02)]:
        pq = [N, 1e-12, 0, 0, G, 0, 0, 0, 0, 0, 1e-5]
        assert oracle.fcs_value(m, 0.0, pq) == pytest.approx(1 / N + G, rel=1e-10)


def test_model_point_detector_limit():
    """Pixels much smaller than the PSF: G(tau)/G(0) -> 1/(1 + 4 D tau / s^2)."""
    s = 0.4e-6
    m = _model(ax=s * 1e-3, s=s)
    D = 2e-12
    for tau in (1e-3, 1e-2, 0.1):
        g = oracle.fcs_value(m, tau, [1.0, D, 0, 0, 0, 0, 0, 0, 0, 0, 0])
        assert g == pytest.approx(1 / (1 + 4 * D * tau / s ** 2), rel=1e-5)


def _fd(m, tau, pq, i, h):
    p1, p2 = pq.copy(), pq.copy()
    p1[i] += h
    p2[i] -= h
    return (oracle.fcs_value(m, tau, p1) - oracle.fcs_value(m, tau, p2)) / (2 * h)


SCALE = np.array([1, 1e-12, 1e-6, 1e-6, 1e-3, 0.1, 1e-12, 0.1, 1e-12, 0.1, 1e-5])


def test_model_gradient_matches_finite_differences():
    """With equal brightness (q2 = q3 = 1) the closed-form FCS_3p gradient
    is the exact derivative of FCS_3p.value for all 11 parameters; the
    triplet factor is switched on for all entries but D (see below)."""
    m = _model(rx=0.1e-6)
    pq = np.array([2.0, 1.2e-12, 3e-6, -2e-6, 0.003, 0.2, 4e-12, 0.1, 0.3e-12, 0.2, 2e-5])
    for tau in (1e-3, 2e-2, 0.3):
        g = oracle.fcs_gradient(m, tau, pq)
        for i in range(11):
            fd = _fd(m, tau, pq, i, 1e-5 * SCALE[i])
            if i == 1:
                trip = 1 + 0.2 / 0.8 * np.exp(-tau / 2e-5)
                fd /= trip
            assert g[i] == pytest.approx(fd, rel=2e-5, abs=1e-7 / SCALE[i]), (i, tau)


def test_model_gradient_brightness_quirk():
    """For q2, q3 != 1 the Java gradient of N uses the fractions pf = x/den
    where the value has x/den^2 (IFCS:13302-13308): gradient = FD * den."""
    m = _model()
    m.q2, m.q3 = 0.6, 1.4
    pq = np.array([2.0, 1.2e-12, 0, 0, 0.0, 0.3, 4e-12, 0.2, 0.5e-12, 0.0, 1e-5])
    den = 1 - 0.3 - 0.2 + 0.6 * 0.3 + 1.4 * 0.2
    tau = 5e-3
    fd = _fd(m, tau, pq, 0, 1e-5)
    assert oracle.fcs_gradient(m, tau, pq)[0] == pytest.approx(fd * den, rel=1e-6)


def test_model_gradient_d_omits_triplet_factor():
    """FCS_3p.gradient's D entry has no triplet factor (IFCS:13310): it
    equals the true derivative divided by the triplet term."""
    m = _model()
    pq = np.array([2.0, 1.0e-12, 0, 0, 0, 0, 0, 0, 0, 0.3, 1e-4])
    tau = 2e-4
    trip = 1 + 0.3 / 0.7 * np.exp(-tau / 1e-4)
    h = 1e-17
    p1, p2 = pq.copy(), pq.copy()
    p1[1] += h
    p2[1] -= h
    fd = (oracle.fcs_value(m, tau, p1) - oracle.fcs_value(m, tau, p2)) / (2 * h)
    assert oracle.fcs_gradient(m, tau, pq)[1] * trip == pytest.approx(fd, rel=1e-5)


# ---------------------------------------------------------------- fits
def _curve(m, pq, lagtime):
    return np.array([oracle.fcs_value(m, t, pq) for t in lagtime])


def test_fit_recovers_noise_free_parameters():
    lag, samp, mtab, lagtime = oracle.tables(16, 8, 50000, FT)
    m = _model()
    true = [3.0, 1.7e-12, 0, 0, 0.002, 0, 0, 0, 0, 0, 0]
    acf = _curve(m, true, lagtime)
    var = (1e-3 * (1 + np.arange(len(lag)) / 10)) ** 2
    fr, chi2, ok, ne, ni = oracle.standard_fit(m, lagtime, acf, var)
    assert ok
    np.testing.assert_allclose(fr[[0, 1, 4]], [3.0, 1.7e-12, 0.002], rtol=1e-6)


def test_fit_matches_scipy_minpack():
    """commons-math3's LM is MINPACK lmder; scipy's least_squares(method='lm')
    wraps MINPACK itself, so both must land on the same optimum."""
    from scipy.optimize import least_squares
    rng = np.random.default_rng(12)
    lag, samp, mtab, lagtime = oracle.tables(16, 8, 50000, FT)
    m = _model()
    for trial in range(3):
        true = [rng.uniform(1, 10), rng.uniform(0.3, 5) * 1e-12, 0, 0, rng.uniform(-0.01, 0.01), 0, 0, 0, 0, 0, 0]
        acf = _curve(m, true, lagtime)
        sd = 0.02 / true[0] * (1 + np.arange(73) / 20)
        acf_n = acf + rng.normal(0, sd)
        fr, chi2, ok, ne, ni = oracle.standard_fit(m, lagtime, acf_n, sd ** 2)
        assert ok
        x, y, w = lagtime[1:], acf_n[1:], 1 / sd[1:]
        scale = np.array([1.0, 1e-12, 1.0])

        def res(q):
            pq = [q[0], q[1] * 1e-12, 0, 0, q[2], 0, 0, 0, 0, 0, 0]
            return w * (y - np.array([oracle.fcs_value(m, t, pq) for t in x]))

        def jac(q):
            pq = [q[0], q[1] * 1e-12, 0, 0, q[2], 0, 0, 0, 0, 0, 0]
            J = np.array([oracle.fcs_gradient(m, t, pq)[[0, 1, 4]] for t in x]) * scale
            return -w[:, None] * J

        sol = least_squares(res, [1.0, 1.0, 0.0], jac=jac, method="lm", xtol=1e-12, ftol=1e-12, gtol=1e-12)
        np.testing.assert_allclose(fr[[0, 1, 4]] / scale, sol.x, rtol=1e-5, atol=1e-8)
        assert chi2 == pytest.approx(np.sum(sol.fun ** 2) / (71 - 3 - 1), rel=1e-6)


def test_gls_with_diagonal_covariance_equals_weighted_fit():
    rng = np.random.default_rng(13)
    lag, samp, mtab, lagtime = oracle.tables(16, 8, 50000, FT)
    m = _model()
    acf = _curve(m, [2.5, 1.1e-12, 0, 0, 0.001, 0, 0, 0, 0, 0, 0], lagtime)
    var = (0.004 * (1 + np.arange(73) / 15)) ** 2
    acf_n = acf + rng.normal(0, np.sqrt(var))
    f1, c1, ok1, _, _ = oracle.standard_fit(m, lagtime, acf_n, var)
    f2, c2, ok2 = oracle.gls_fit(m, lagtime, acf_n, np.diag(var[1:]))
    assert ok1 and ok2
    np.testing.assert_allclose(f1[[0, 1, 4]], f2[[0, 1, 4]], rtol=1e-6)


def test_linseg_constant_trace_is_unchanged():
    """"Lin Segment" (IFCS:9767-9805): with no bleaching every segment mean
    is the same, the reference curve is flat at bcInt0 and the correction
    I/sqrt(g/c0) + c0(1 - sqrt(g/c0)) is the identity, in all three ranges
    (before the first segment centre, between centres, after the last)."""
    d = np.full(12001, 57.0)
    d[0] = 0.0
    out, seg = oracle.bleach_linseg(d, 600)
    assert len(seg) == 21
    np.testing.assert_allclose(out[1:], d[1:], rtol=1e-13)
    # segment means divide len - 1 frames by len (the reference's sum bounds)
    np.testing.assert_allclose(seg[1:], 57.0 * 599 / 600, rtol=1e-13)


def test_linseg_flattens_a_linear_decay():
    """A linear bleach between the first and the last segment centre is
    followed by the piecewise-linear curve, so the corrected trace there is
    flat up to the segment-mean offset: its slope drops by orders of
    magnitude."""
    num = 12000
    x = np.arange(num + 1, dtype=np.float64)
    d = 1000.0 - 0.03 * x
    d[0] = 0.0
    out, seg = oracle.bleach_linseg(d, 600)
    bc = len(seg) - 1
    ave = num // bc
    mid = slice(ave, (bc - 1) * ave)
    slope_in = np.polyfit(x[mid], d[mid], 1)[0]
    slope_out = np.polyfit(x[mid], out[mid], 1)[0]
    assert abs(slope_out) < 1e-3 * abs(slope_in)


def _exp_trace(nexp, seed=3):
    rng = np.random.default_rng(seed)
    t = 0.001 * (np.arange(1000) + 0.5) * 20
    if nexp == 1:
        f = 300.0 * np.exp(-t / 6.0) + 80.0
    else:
        f = 200.0 * np.exp(-t / 1.5) + 150.0 * np.exp(-t / 9.0) + 60.0
    return t, f + rng.normal(0, 2.0, t.size)


@pytest.mark.parametrize("nexp", [1, 2])
def test_expfit_reaches_the_least_squares_optimum(nexp):
    """SingleExpFit / DoubleExpFit (IFCS:10845-10950): the oracle's LM from
    the reference's initial guesses ends where MINPACK (scipy's
    least_squares 'lm', the algorithm commons-math3 translates) ends from
    the same start."""
    from scipy.optimize import least_squares
    t, y = _exp_trace(nexp)
    res, st = oracle.expfit(y, t, nexp)
    assert st == 0
    if nexp == 1:
        x0 = [y[1], t[len(t) // 2], 0.0]
        fun = lambda p: p[0] * np.exp(-t / p[1]) + p[2] - y
    else:
        x0 = [y[1] / 2, t[len(t) // 10], y[1] / 2, t[len(t) // 2], 0.0]
        fun = lambda p: p[0] * np.exp(-t / p[1]) + p[2] * np.exp(-t / p[3]) + p[4] - y
    ref = least_squares(fun, x0, method="lm", xtol=1e-14, ftol=1e-14, gtol=1e-14).x
    curve = lambda p: fun(p) + y
    np.testing.assert_allclose(curve(res), curve(ref), rtol=1e-8)
    if nexp == 1:
        np.testing.assert_allclose(res, ref, rtol=1e-6)


@pytest.mark.parametrize("nexp", [1, 2])
def test_bleach_exp_flattens_its_own_curve(nexp):
    """Correcting the exact fitted curve gives the constant f(0): I/sqrt(f/f0)
    + f0(1 - sqrt(f/f0)) = f0 when I = f (IFCS:9690, 9714); a zero amplitude
    or time skips the correction."""
    res = np.array([300.0, 6.0, 80.0]) if nexp == 1 else np.array([200.0, 1.5, 150.0, 9.0, 60.0])
    num = 5000
    x = np.arange(num + 1)
    tt = 0.001 * (x + 0.5)
    f = res[0] * np.exp(-tt / res[1]) + (res[2] if nexp == 1 else res[2] * np.exp(-tt / res[3]) + res[4])
    f[0] = 0.0
    out = oracle.bleach_exp(f, res, nexp, 0.001)
    f0 = res[0] + res[2] if nexp == 1 else res[0] + res[2] + res[4]
    np.testing.assert_allclose(out[1:], f0, rtol=1e-12)
    z = res.copy()
    z[0] = 0.0
    np.testing.assert_array_equal(oracle.bleach_exp(f, z, nexp, 0.001), f)
