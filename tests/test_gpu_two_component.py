"""Learned initial guesses for the two-component fit on a 256 x 256 px x
50,000-frame stack: 256 regions of 16 x 16 pixels, each its own periodic
box with D1 uniform in 0.2-1 um^2/s, D2 uniform in 2-20 um^2/s, fraction F2
uniform in 0-1 and N uniform in 1-50 per observation area (GPU simulator).
The ACF of every pixel is fitted with N, D, G_inf, F2, D2 free (FCS_3p,
IFCS:13131-13408) twice on the same correlations: from the fit panel's
start (N = 1, D = 1, G_inf = 0, with F2 = 0.5, D2 = 10 um^2/s entered: the
panel's defaults F2 = D2 = 0 make the Hessian singular), and from the
two-component network trained on 2 M model curves of the geometry.

Reported: iterations, failure rate (not CONVERGED), fit-stage device time
and fitted pixels/s of both, and how often the two reach the same optimum
(D1 within 1e-3 relative).  Asserted: the seeded fits fail no more often
and take fewer iterations on average.
"""
import numpy as np
import pytest

import imaging_fcs_1_52_amd as imfcs
from imaging_fcs_1_52_amd.learned import InitialGuessNet
from parity import PIX, PSF

pytestmark = pytest.mark.gpu

W = H = 256
T = 50000
TILE = 16


def test_two_component_seeded_vs_plain_lm(engine):
    rng = np.random.default_rng(55)
    aeff = imfcs.obsvol(PIX, PIX, PSF) * 1e12       # um^2
    side = TILE * 0.24
    doms = []
    for ty in range(H // TILE):
        for tx in range(W // TILE):
            N = rng.uniform(1, 50)
            doms.append(imfcs.sim_domain(tx * side, ty * side, side, side, int(round(N * side * side / aeff)),
                                         D_um2s=rng.uniform(0.2, 1.0), D2_um2s=rng.uniform(2, 20),
                                         frac2=rng.uniform(0, 1), counts=20.0))
    print("simulating 256x256x50000, %d emitters" % sum(d.n_emitters for d in doms), flush=True)
    d = imfcs.simulate_2d_device(engine, W, H, T, doms, offset=100, tau_bleach_s=0.0, seed=5050)
    pf = [1, 1, 0, 0, 1, 1, 1, 0, 0, 0, 0]
    ini = [1.0, 1e-12, 0, 0, 0, 0.5, 10e-12, 0, 0, 0, 0]
    model = imfcs.make_fit_model(PIX, PIX, PSF, paramfit=pf, paraminit=ini)
    cfg = imfcs.make_pipeline_config(W, H, T, background=100, poly_order=0, acf_model=model)
    engine.set_fit_seeder(None, None)
    engine.run(cfg, d)
    plain = engine.results(cfg)
    t_plain = engine.stage_ms()[4]
    _, _, mtab, lagtime = imfcs.correlator_tables(16, 8, T, 1e-3)
    net = InitialGuessNet(engine, model, lagtime[1:], mtab[1:], ncomp=2, seed=21)
    loss = net.train(steps=500, batch=4096, lr=2e-3)     # 2.05 M curves
    print("training loss: first %.3f last %.4f" % (loss[:5].mean(), loss[-20:].mean()), flush=True)
    engine.set_fit_seeder(net, None)
    try:
        engine.refit(cfg)
        seeded = engine.results(cfg)
        t_seed = engine.stage_ms()[4]
    finally:
        engine.set_fit_seeder(None, None)
    npx = W * H
    sp, ss = plain["states"][0], seeded["states"][0]
    ip, is_ = plain["iters"][0], seeded["iters"][0]
    both = (sp == 0) & (ss == 0)
    fp, fs = plain["fitres"][0], seeded["fitres"][0]
    agree = np.abs(fs[..., 1][both] / fp[..., 1][both] - 1) < 1e-3
    chi_p, chi_s = plain["chi2"][0][both], seeded["chi2"][0][both]
    print("plain : fail %.4f, iterations mean %.1f p99 %.0f, fit stage %.2f ms, %.3g fitted px/s"
          % ((sp != 0).mean(), ip.mean(), np.percentile(ip, 99), t_plain, (sp == 0).sum() / (t_plain * 1e-3)))
    print("seeded: fail %.4f, iterations mean %.1f p99 %.0f, fit stage %.2f ms, %.3g fitted px/s"
          % ((ss != 0).mean(), is_.mean(), np.percentile(is_, 99), t_seed, (ss == 0).sum() / (t_seed * 1e-3)))
    print("same optimum (D1 within 1e-3): %.4f of %d fits converged both ways; seeded chi2 lower in %.4f"
          % (agree.mean(), both.sum(), (chi_s < chi_p * (1 - 1e-9)).mean()))
    assert (ss != 0).mean() <= (sp != 0).mean()
    assert is_.mean() < ip.mean()
    assert both.sum() > 0.5 * npx
