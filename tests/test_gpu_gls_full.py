"""GLS at full size: 256 x 256 px x 100,000 frames (13.1 GB of uint16),
emitter density 4 / um^2 (15,575 emitters over the 2-pixel periodic margin),
D = 1 um^2/s, brightness rising from 10 to 2000 photons per emitter and
frame across the field (so that long sums reach 10^7 - 10^8 counts at the
bright edge), mild bleaching (tau = 100 s) with order-4 correction; ACF and
CCF to the x + 1 neighbour with the regularized 72 x 72 channel covariance
of every pair (calculateCF GLS branch, IFCS:9524-9639) and the GLS fit with
its FP64 Cholesky factor (GLSFit, IFCS:10622-10789).  The stack comes from
the GPU simulator; crops of it go to the oracle.

Tolerances: the GLS data mcov within 1e-4 relative (floor 1e-4 of the
curve amplitude); the regularized covariance within 1e-4 of its largest
entry; blocking index equal for >= 90 % of the finite pixels of the crops;
GLS-fitted N and D within 1e-3 relative and G_inf within 1e-3 of 1/N where
both sides fitted.
"""
import numpy as np
import pytest

import imaging_fcs_1_52_amd as imfcs
import oracle
from parity import PIX, PSF, rel_err

pytestmark = pytest.mark.gpu

W = H = 256
T = 100000
MARGIN = 2


def test_256x256x100000_gls_varied_brightness_matches_oracle(engine):
    side = (W + 2 * MARGIN) * 0.24
    n_emit = int(round(4.0 * side * side))
    dom = imfcs.sim_domain(-MARGIN * 0.24, -MARGIN * 0.24, side, side, n_emit, D_um2s=1.0, counts=10.0,
                           counts_hi=2000.0)
    print("simulating 256x256x100000 on the GPU (%d emitters)" % n_emit, flush=True)
    d = imfcs.simulate_2d_device(engine, W, H, T, dom, offset=100, tau_bleach_s=100.0, seed=404)
    stack = engine.download(d, (T, H, W))
    bg = int(stack.min())
    acf_m = imfcs.make_fit_model(PIX, PIX, PSF)
    ccf_m = imfcs.make_fit_model(PIX, PIX, PSF, rx=PIX)
    cfg = imfcs.make_pipeline_config(W, H, T, binning=1, cf=(1, 0), background=bg, poly_order=4, do_acf=True,
                                     do_ccf=True, acf_model=acf_m, ccf_model=ccf_m, gls=True)
    print("running the pipeline with GLS", flush=True)
    engine.run(cfg, d)
    engine.synchronize()
    g = engine.results(cfg)
    assert g["acf"].shape == (2, H, W - 1, 73)
    counts = np.asarray(stack[:2000].mean(0), np.float64) - bg
    print("mean counts per pixel and frame: left %.1f, right %.1f" % (counts[:, :8].mean(), counts[:, -8:].mean()),
          flush=True)
    om = (oracle.make_model(PIX, PIX, PSF), oracle.make_model(PIX, PIX, PSF, rx=PIX))
    compared = fitted = 0
    for (x, y) in [(0, 0), (60, 131), (128, 7), (190, 250), (251, 120)]:
        crop = np.ascontiguousarray(stack[:, y:y + 3, x:x + 4])
        for pair, cf in ((0, (0, 0)), (1, (1, 0))):
            ocfg = oracle.make_config(4, 3, T, cf=cf, background=bg, poly_order=4, roi=(0, 0, 3, 3),
                                      model=om[pair], gls=True, nthreads=8)
            o = oracle.pipeline(crop, ocfg)
            ga = g["acf"][pair][y:y + 3, x:x + 3]
            gi = g["ind"][pair][y:y + 3, x:x + 3]
            fin = np.isfinite(o["acf"]).all(-1)
            assert np.array_equal(np.isfinite(ga).all(-1), fin)
            same = (gi == o["ind"]) & fin
            assert same[fin].mean() >= 0.9, ((x, y, pair), same[fin].mean())
            e = rel_err(ga[same], o["acf"][same], 1e-4 * np.abs(o["acf"][same][:, 1:]).max(-1, keepdims=True))
            assert e.max() < 1e-4, ((x, y, pair), e.max())
            compared += int(same.sum())
            # the regularized covariance of the crop's first pixel
            if same[0, 0]:
                ser = []
                for px in (0, cf[0]):
                    dd = oracle.get_intensity(crop, px, 0, 1, 1, bg, 1, T)
                    tr, it = oracle.intensity_trace(crop, px, 0, 1, 1, bg, 1, T, 1000, 1e-3)
                    coef, _ = oracle.polyfit(tr, it, 4)
                    ser.append(oracle.bleach_apply(dd, coef, 1e-3))
                _, _, _, ocov = oracle.calculate_cf(ser[0], ser[1], int(o["ind"][0, 0]), gls=True)
                pix_index = (pair * H + y) * (W - 1) + x
                gcov = engine.gls_covariance(pix_index)
                assert np.max(np.abs(gcov - ocov)) / np.max(np.abs(ocov)) < 1e-4, (x, y, pair)
            ok = (g["states"][pair][y:y + 3, x:x + 3] == 0) & o["fitted"] & same
            if ok.any():
                gf = g["fitres"][pair][y:y + 3, x:x + 3]
                N_o = o["fitres"][..., 0][ok]
                assert rel_err(gf[..., 0][ok], N_o, 1e-300).max() < 1e-3
                assert rel_err(gf[..., 1][ok], o["fitres"][..., 1][ok], 1e-300).max() < 1e-3
                assert (np.abs(gf[..., 4][ok] - o["fitres"][..., 4][ok]) * np.abs(N_o)).max() < 1e-3
                fitted += int(ok.sum())
    print("compared %d GLS correlations, %d GLS fits" % (compared, fitted), flush=True)
    assert compared >= 60 and fitted >= 30
