"""Independent numpy restatements used to pin the oracle (test
infrastructure).  Vectorised from the definitions, not translated from the
oracle's loops:

  * multitau: per-level pair sums, window means, lag products, blocked
    products (calculateCF, Imaging_FCS_1_52.java:9425-9521);
  * block_transform: the lag-1 blocking curve and its three criteria
    (blockTransform, IFCS:9217-9373), with the zero-prepended input of the
    Java caller (IFCS:9259-9262) as an option;
  * shrink_covariance: Schaefer & Strimmer (2005, Stat. Appl. Genet. Mol.
    Biol. 4:32) shrinkage of the channel covariance towards a diagonal
    target with the median variance (Opgen-Rhein & Strimmer 2007), with the
    two Java quirks (median over 73 slots including an unused 0, running
    rescale of the covariance-weight numerator, IFCS:9544-9613) as options.
"""
import math

import numpy as np


def tables(p, q, num):
    chanum = p + (q - 1) * p // 2 + 1
    lag = np.zeros(chanum, np.int64)
    samp = np.ones(chanum)
    lag[:p // 2 + 1] = np.arange(p // 2 + 1)
    for x in range(1, q + 1):
        y = np.arange(1, p // 2 + 1)
        lag[x * (p // 2) + y] = 2 ** (x - 1) * y + (p // 4) * 2 ** x
    for x in range(2, q + 1):
        samp[x * (p // 2) + 1: x * (p // 2) + p // 2 + 1] = 2 ** (x - 1)
    mtab = np.floor((num - lag) / samp)
    return lag, samp, mtab


def _pair_average(P, times):
    for _ in range(times):
        m = len(P) // 2
        P = 0.5 * (P[0:2 * m:2] + P[1:2 * m:2])
    return P


def multitau(a, b, ind, p=16, q=8):
    """0-based series a, b (num values).  Returns G, blocked variance and
    the normalised blocked products Y[pnum][chanum] (column 0 unused) of
    calculateCF."""
    num = len(a)
    lag, samp, mtab = tables(p, q, num)
    n_ch = len(lag)
    pnum = int(math.floor(mtab[-1] / 2 ** max(ind - math.log2(samp[-1]), 0)))
    G, V = np.zeros(n_ch), np.zeros(n_ch)
    Y = np.zeros((pnum, n_ch))
    for x in range(n_ch):
        c = int(round(math.log2(samp[x])))
        nb = num >> c
        A = a[:nb << c].reshape(nb, 1 << c).sum(1)
        B = b[:nb << c].reshape(nb, 1 << c).sum(1)
        d = int(lag[x]) >> c
        n = nb - d
        Ai, Bd = A[:n], B[d:d + n]
        am, bm = Ai.mean(), Bd.mean()
        prod = (Ai - am) * (Bd - bm)
        G[x] = prod.sum() / (n * am * bm)
        P = _pair_average(prod, ind - c)[:pnum]
        V[x] = (np.mean(P ** 2) - np.mean(P) ** 2) / ((pnum - 1) * (am * bm) ** 2)
        Y[:, x] = P / (am * bm)
    return G, V, Y


def blocking_curve(a, b):
    """Lag-1 blocking curve of 0-based series: standard deviations, their
    errors and product counts per blocking level."""
    num = len(a)
    blocknum = int(math.floor(math.log(num - 1) / math.log(2))) - 2
    n = num - 1
    am, bm = a[:n].mean(), b[1:n + 1].mean()
    P = (a[:n] - am) * (b[1:n + 1] - bm)
    sd, cnt = [], []
    for y in range(blocknum):
        if y:
            P = _pair_average(P, 1)
        k = len(P)
        sd.append(math.sqrt((np.mean(P ** 2) - np.mean(P) ** 2) / (k * (am * bm) ** 2)))
        cnt.append(k)
    sd, cnt = np.array(sd), np.array(cnt, np.float64)
    return sd, sd / np.sqrt(2 * (cnt - 1)), blocknum


def block_transform(a1, b1, q=8, java_shift=True):
    """blockTransform on 1-based traces a1, b1 (index 0 unused).  With
    java_shift the series handed over is the Java caller's 0-based copy
    [0, a_1 .. a_{num-1}]; otherwise a_1 .. a_num."""
    num = len(a1) - 1
    a = a1[:num] if java_shift else a1[1:]
    b = b1[:num] if java_shift else b1[1:]
    sd, err, nb = blocking_curve(np.asarray(a, float), np.asarray(b, float))
    up, lo = sd + err, sd - err
    overlap = (up[:-1] > lo[1:]) & (up[1:] > lo[:-1])          # neighbouring error bars overlap
    triple = overlap[:-1] & overlap[1:]                          # three points overlap
    rising = np.diff(sd) > 0
    rising3 = rising[:-1] & rising[1:]                           # three points increase
    # the last rising triple whose error bars do not overlap (0 if none):
    # no blocking triple at or before it
    bad = np.flatnonzero(rising3 & ~triple)
    last0 = bad[-1] if len(bad) else 0
    cand = triple.copy()
    cand[:last0 + 1] = False
    cand[nb - 3] = cand[nb - 4] = False   # the last two points cannot start a triple
    hits = [x for x in range(1, nb - 4) if cand[x]]
    if hits:
        ind, ok = hits[0] + 1, 1
    else:
        ind, ok = (nb - 3 if nb - 3 > 0 else nb - 1), 0
    return max(ind, q - 1), ok


def shrink_covariance(Y, median_zero_slot=True, running_rescale=True, lam=None):
    """Regularised covariance covmats (m x m) of the columns of Y (n samples
    x m channels) divided by n, and the two shrinkage weights (lam = (lam_v,
    lam_c) replaces the estimated weights)."""
    n, m = Y.shape
    D = Y - Y.mean(0)
    S = D.T @ D / (n - 1)
    v = np.diag(S).copy()
    med = np.median(np.r_[0.0, v] if median_zero_slot else v)
    c = n / (n - 1.0) ** 3
    lam_v = float(np.clip(c * ((D ** 2 - v) ** 2).sum() / ((v - med) ** 2).sum(), 0, 1))
    R = S / np.sqrt(np.outer(v, v))
    Z = D / np.sqrt(v)
    Z2 = Z ** 2
    # s_xy = sum_z (Z_x Z_y - R_xy)^2
    s = Z2.T @ Z2 - 2 * R * (Z.T @ Z) + n * R ** 2
    low = np.tril_indices(m, -1)
    if running_rescale:
        num = 0.0
        for x in range(1, m):
            t = 0.0
            for y in range(x):
                t = (t + s[x, y]) * c
                num += t
    else:
        num = c * s[low].sum()
    lam_c = float(np.clip(num / (R[low] ** 2).sum(), 0, 1))
    if lam is not None:
        lam_v, lam_c = lam
    dd = lam_v * med + (1 - lam_v) * v
    out = (1 - lam_c) * R * np.sqrt(np.outer(dd, dd)) / n
    out[np.diag_indices(m)] = dd / n
    return out, lam_v, lam_c, med
