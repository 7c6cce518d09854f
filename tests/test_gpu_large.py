"""Parity at full scale: a 256 x 256 x 100,000 uint16 stack (13.1 GB; the
corrected series alone take 26 GB), binning 1, ACF + CCF to the x+1
neighbour, order-4 bleach correction, LM fits.

A pixel's correlation, blocked variance and fit depend only on its own
series and its partner's (and on the background, passed explicitly), so
the oracle run on small crops of the stack gives the exact reference for
those pixels at full size.  This checks the 64-bit indexing of every
stage at a size where 32-bit offsets overflow.  Tolerances as in the
varied-configuration tests.
"""
import numpy as np
import pytest

import imaging_fcs_1_52_amd as imfcs
import oracle
from parity import PIX, PSF, rel_err

pytestmark = pytest.mark.gpu

W = H = 256
T = 100000
CROPS = [(0, 0), (121, 77), (252, 253)]   # (x, y) of 3 x 3 output pixels; x + 3 <= W - 1 for the partner


def test_full_scale_crops_match_oracle(engine):
    print("simulating 256x256x100000", flush=True)
    stack = imfcs.simulate_2d(W, H, T, 8000, D_um2s=1.0, tau_bleach_s=100.0, seed=2024)
    bg = int(stack.min())
    acf = imfcs.make_fit_model(PIX, PIX, PSF)
    ccf = imfcs.make_fit_model(PIX, PIX, PSF, rx=PIX)
    cfg = imfcs.make_pipeline_config(W, H, T, binning=1, cf=(1, 0), background=bg, poly_order=4, do_fit=True,
                                     acf_model=acf, ccf_model=ccf)
    print("uploading", flush=True)
    engine.upload(stack)
    print("running", flush=True)
    engine.run(cfg)
    engine.synchronize()
    g = engine.results(cfg)
    assert g["acf"].shape == (2, H, W - 1, 73)
    om_a, om_c = oracle.make_model(PIX, PIX, PSF), oracle.make_model(PIX, PIX, PSF, rx=PIX)
    compared = fitted = 0
    for (x, y) in CROPS:
        print("crop", x, y, flush=True)
        xe, ye = min(x + 4, W), min(y + 3, H)
        crop = np.ascontiguousarray(stack[:, y:ye, x:xe])
        ch, cw = crop.shape[1], crop.shape[2]
        for pair, (cf, model) in enumerate((((0, 0), om_a), ((1, 0), om_c))):
            ocfg = oracle.make_config(cw, ch, T, cf=cf, background=bg, poly_order=4,
                                      roi=(0, 0, cw - 1, ch), model=model, nthreads=8)
            o = oracle.pipeline(crop, ocfg)
            ga = g["acf"][pair][y:ye, x:x + cw - 1]
            gv = g["var"][pair][y:ye, x:x + cw - 1]
            gi = g["ind"][pair][y:ye, x:x + cw - 1]
            fin = np.isfinite(o["acf"]).all(-1)
            assert np.array_equal(np.isfinite(ga).all(-1), fin), "NaN pattern differs"
            if not fin.any():
                continue
            compared += int(fin.sum())
            amp = np.abs(o["acf"][fin][:, 1:]).max(-1, keepdims=True)
            e = rel_err(ga[fin], o["acf"][fin], 1e-4 * amp).max()
            assert e < 1e-4, e
            same = (gi == o["ind"]) & fin
            assert same[fin].mean() >= 0.8
            assert rel_err(gv[same][:, 1:], o["var"][same][:, 1:], 1e-300).max() < 1e-4
            if pair == 0:
                gf = g["fitres"][0][y:ye, x:x + cw - 1]
                ok = (g["states"][0][y:ye, x:x + cw - 1] == 0) & o["fitted"] & same
                if ok.any():
                    fitted += int(ok.sum())
                    assert rel_err(gf[..., 0][ok], o["fitres"][..., 0][ok], 1e-12).max() < 1e-3
                    assert rel_err(gf[..., 1][ok], o["fitres"][..., 1][ok], 1e-30).max() < 1e-3
    print("compared %d correlations, %d fits" % (compared, fitted), flush=True)
    assert compared >= 27 and fitted >= 9
