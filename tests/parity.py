"""Helpers that run the GPU pipeline and the oracle on the same stack and
compare them.  Test infrastructure (imports the oracle)."""
import numpy as np

import imaging_fcs_1_52_amd as imfcs
import oracle

PIX = 0.24e-6      # object-space pixel size (m)
PSF = 0.38e-6      # fit PSF size s = 2 sigma of the simulated PSF (m)


def models(binning, cf):
    ax = PIX * binning
    acf = imfcs.make_fit_model(ax, ax, PSF)
    ccf = imfcs.make_fit_model(ax, ax, PSF, rx=PIX * cf[0], ry=PIX * cf[1])
    oacf = oracle.make_model(ax, ax, PSF)
    occf = oracle.make_model(ax, ax, PSF, rx=PIX * cf[0], ry=PIX * cf[1])
    return acf, ccf, oacf, occf


def run_gpu(engine, stack, binning=1, cf=(0, 0), poly_order=4, do_fit=True, tol=1e-16, maxit=2000,
            background=None, do_acf=True, p=16, q=8):
    T, H, W = stack.shape
    bg = int(stack.min()) if background is None else background
    acf, ccf, _, _ = models(binning, cf)
    cfg = imfcs.make_pipeline_config(W, H, T, binning=binning, cf=cf, background=bg, poly_order=poly_order,
                                     do_fit=do_fit, acf_model=acf, ccf_model=ccf, fit_tolerance=tol,
                                     fit_max_iterations=maxit, do_acf=do_acf, p=p, q=q)
    engine.upload(stack)
    engine.run(cfg)
    return engine.results(cfg), cfg


def run_oracle(stack, binning=1, cf=(0, 0), poly_order=4, do_fit=True, roi_cf=None, background=None,
               nthreads=8, p=16, q=8):
    """kcf=0 CPU path; roi_cf fixes the ROI to the one of the GPU run."""
    T, H, W = stack.shape
    bg = int(stack.min()) if background is None else background
    _, _, oacf, occf = models(binning, cf)
    rcf = cf if roi_cf is None else roi_cf
    cfg = oracle.make_config(W, H, T, binning=binning, cf=cf, background=bg, poly_order=poly_order,
                             roi=(0, 0, W - rcf[0], H - rcf[1]), do_fit=do_fit,
                             model=occf if cf != (0, 0) else oacf, nthreads=nthreads, p=p, q=q)
    return oracle.pipeline(stack, cfg)


def rel_err(a, b, floor):
    return np.abs(a - b) / np.maximum(np.abs(b), floor)
