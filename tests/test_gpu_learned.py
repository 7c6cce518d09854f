"""Learned initial guesses: the BF16 perceptron's GPU kernels against the
FP64 CPU oracle (oracle/mlp.py), its training on the GPU, and what the
seeded LM gains over the plain start on simulated stacks.

Tolerances: outputs and the loss of the GPU forward pass within 1e-3 of
the mirror oracle (the same BF16 rounding points, FP64 elsewhere; the GPU
accumulates in FP32); every parameter-gradient block within 2e-3 of its
largest entry of the mirror oracle's.  The BF16 gradient's distance to the
exact FP64 gradient is stated (asserted below 5 % of the block maximum).
Seeded fits: N and D within 1e-3 relative of the plain fit where both
converge, fewer iterations on average, no more failures.
"""
import numpy as np
import pytest

import imaging_fcs_1_52_amd as imfcs
from imaging_fcs_1_52_amd.learned import InitialGuessNet
from oracle import mlp
from parity import PIX, PSF

pytestmark = pytest.mark.gpu

T_FR = 20000


def _net(engine, ncomp=1, rx=0.0, seed=3, T=T_FR, binning=1):
    geom = imfcs.make_fit_model(PIX * binning, PIX * binning, PSF, rx=rx)
    _, _, mtab, lagtime = imfcs.correlator_tables(16, 8, T, 1e-3)
    return InitialGuessNet(engine, geom, lagtime[1:], mtab[1:], ncomp=ncomp, seed=seed), geom, lagtime


def test_gradients_match_cpu_oracle(engine):
    net, _, _ = _net(engine)
    x, t = net.curves(500)                       # 500 rows: a partial 64-row tile
    assert np.all(x[:, 72:] == 0) and np.abs(x[:, :72]).max() == pytest.approx(1.0, abs=4e-3)
    p = net.params()
    g, y, loss = net.gradient(x, t)
    w = (1, 1, 0, 0, 1)
    lo, go, yo = mlp.forward_backward(p, x, t, w, mode="mirror")
    le, ge, _ = mlp.forward_backward(p, x, t, w, mode="exact")
    assert np.abs(y - yo).max() < 1e-3 * max(1.0, np.abs(yo).max())
    assert loss == pytest.approx(lo, rel=1e-3)
    blocks = [("W1", mlp.O_W1, mlp.O_B1), ("b1", mlp.O_B1, mlp.O_W2), ("W2", mlp.O_W2, mlp.O_B2),
              ("b2", mlp.O_B2, mlp.O_W3), ("W3", mlp.O_W3, mlp.O_B3), ("b3", mlp.O_B3, mlp.N_PARAMS)]
    for name, a, b in blocks:
        ref = np.abs(go[a:b]).max()
        err = np.abs(g[a:b] - go[a:b]).max() / ref
        exact_err = np.abs(g[a:b] - ge[a:b]).max() / np.abs(ge[a:b]).max()
        print("%s: |g - mirror| %.2e, |g - exact FP64| %.2e (of the block maximum)" % (name, err, exact_err))
        assert err < 2e-3, (name, err)
        assert exact_err < 5e-2, (name, exact_err)


def test_fused_inference_matches_cpu_oracle(engine):
    """The one-kernel inference (seeding and predictions) against the
    oracle's forward pass on the same BF16-normalised inputs: N, D within
    2e-3 relative and G_inf within 2e-3 of 1/N for 99 % of the curves,
    within 1e-2 for all; non-finite curves give NaN.  The hidden activations
    are rounded to BF16 between layers, and the kernel and the oracle sum
    the products in different orders: an activation that lies at a BF16
    rounding boundary rounds the other way (2^-8 relative), which moves a
    few outputs by up to a few 1e-3 (the weights are trained anew on the
    GPU each run, so which curves those are varies)."""
    net, geom, lagtime = _net(engine, seed=11)
    net.train(steps=50, batch=2048)
    cur, N, D, G = _model_curves(geom, lagtime, 333, 4)     # 333 rows: a partial 64-row tile
    cur[5, 7] = np.nan
    got = net.predict(cur)
    p = net.params()
    mu, sd = net.target_scale()
    s = np.abs(cur).max(1)
    x = np.zeros((cur.shape[0], 80))
    x[:, :72] = cur / s[:, None]
    _, _, y = mlp.forward_backward(p, x, np.zeros((cur.shape[0], 5)), mode="mirror")
    t = y * sd + mu
    ref = np.stack([np.exp(t[:, 0]) / s, np.exp(t[:, 1]), np.exp(t[:, 2]), np.clip(t[:, 3], 0, 1), t[:, 4] * s], 1)
    ok = np.ones(cur.shape[0], bool)
    ok[5] = False
    assert np.isnan(got[5]).all()
    for j in (0, 1):
        rel = np.abs(got[ok, j] - ref[ok, j]) / np.abs(ref[ok, j])
        assert np.quantile(rel, 0.99) < 2e-3, (j, np.quantile(rel, 0.99))
        assert rel.max() < 1e-2, (j, rel.max())
    eg = np.abs(got[ok, 4] - ref[ok, 4]) * ref[ok, 0]
    assert np.quantile(eg, 0.99) < 2e-3 and eg.max() < 1e-2, (np.quantile(eg, 0.99), eg.max())


def _model_curves(geom, lagtime, n, seed):
    import oracle
    om = oracle.make_model(geom.ax, geom.ay, geom.s, rx=geom.rx)
    rng = np.random.default_rng(seed)
    N = np.exp(rng.uniform(np.log(0.2), np.log(100), n))
    D = np.exp(rng.uniform(np.log(0.1), np.log(20), n))
    G = rng.uniform(-0.02, 0.02, n) / N
    cur = np.zeros((n, 72))
    for i in range(n):
        pq = [N[i], D[i] * 1e-12, 0, 0, G[i], 0, 0, 0, 0, 0, 0]
        cur[i] = [oracle.fcs_value(om, tau, pq) for tau in lagtime[1:]]
    cur += rng.normal(0, 0.01, cur.shape) / N[:, None]
    return cur, N, D, G


def test_training_lowers_the_loss_and_predicts_model_parameters(engine):
    net, geom, lagtime = _net(engine, seed=5)
    loss = net.train(steps=400, batch=4096, lr=2e-3)
    assert np.all(np.isfinite(loss))
    print("loss: first %.3f, last 20 %.4f" % (loss[:5].mean(), loss[-20:].mean()))
    assert loss[-20:].mean() < 0.1 * loss[:5].mean()
    cur, N, D, G = _model_curves(geom, lagtime, 300, 9)
    pr = net.predict(cur)
    eD = np.abs(np.log(pr[:, 1] / D))
    eN = np.abs(np.log(pr[:, 0] / N))
    print("median |ln D err| %.3f, |ln N err| %.3f" % (np.median(eD), np.median(eN)))
    assert np.median(eD) < 0.15 and np.median(eN) < 0.15
    nan = net.predict(np.full((2, 72), np.nan, np.float32))
    assert np.isnan(nan).all()


def _tiled_stack(engine, W, H, T, seed):
    """Stack of 8 x 8-pixel periodic tiles with D from 0.2 to 10 um^2/s."""
    rng = np.random.default_rng(seed)
    doms = []
    tile = 8
    for ty in range(H // tile):
        for tx in range(W // tile):
            D = float(np.exp(rng.uniform(np.log(0.2), np.log(10))))
            n = int(rng.integers(20, 120))
            doms.append(imfcs.sim_domain(tx * tile * 0.24, ty * tile * 0.24, tile * 0.24, tile * 0.24, n,
                                         D_um2s=D, counts=float(rng.uniform(10, 60))))
    return imfcs.simulate_2d_device(engine, W, H, T, doms, offset=100, tau_bleach_s=0.0, seed=seed)


def test_seeded_lm_reaches_the_same_optimum_in_fewer_iterations(engine):
    W = H = 64
    d = _tiled_stack(engine, W, H, T_FR, 17)
    acf = imfcs.make_fit_model(PIX, PIX, PSF)
    ccf = imfcs.make_fit_model(PIX, PIX, PSF, rx=PIX)
    cfg = imfcs.make_pipeline_config(W, H, T_FR, cf=(1, 0), background=100, poly_order=0, acf_model=acf,
                                     ccf_model=ccf)
    engine.set_fit_seeder(None, None)
    engine.run(cfg, d)
    plain = engine.results(cfg)
    net_a, _, _ = _net(engine, seed=7)
    net_c, _, _ = _net(engine, rx=PIX, seed=8)
    net_a.train(steps=400, batch=4096)
    net_c.train(steps=400, batch=4096)
    engine.set_fit_seeder(net_a, net_c)
    try:
        engine.refit(cfg)
        seeded = engine.results(cfg)
    finally:
        engine.set_fit_seeder(None, None)
    for pair, name in ((0, "ACF"), (1, "CCF")):
        sp, ss = plain["states"][pair], seeded["states"][pair]
        both = (sp == 0) & (ss == 0)
        fp, fs = plain["fitres"][pair], seeded["fitres"][pair]
        eN = np.abs(fs[..., 0][both] / fp[..., 0][both] - 1)
        eD = np.abs(fs[..., 1][both] / fp[..., 1][both] - 1)
        it_p, it_s = plain["iters"][pair].mean(), seeded["iters"][pair].mean()
        print("%s: converged plain %.3f seeded %.3f; iterations plain %.1f seeded %.1f; "
              "D agree (1e-3) %.4f, max N err %.1e" % (name, (sp == 0).mean(), (ss == 0).mean(), it_p, it_s,
                                                      (eD < 1e-3).mean(), eN.max()))
        assert (ss == 0).mean() >= (sp == 0).mean() - 0.002
        assert (eD < 1e-3).mean() >= 0.99 and (eN < 1e-3).mean() >= 0.99
        if pair == 0:
            assert it_s < it_p
