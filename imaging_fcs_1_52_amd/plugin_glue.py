"""Python restatement of the plugin's GPU glue over the drop-in C ABI.

Imaging_FCS_1_52.java drives the native library through a fixed sequence
(GPU_Calculate_ACF_All, IFCS:15287-15510):

  GPU_Initialize_GPUparams (14844-14906)  -> ACFParameters geometry
  GPU_get_pixels           (14908-14986)  -> float stack, binned, background removed
  GPU_Calculate_BleachCorrection (14988-15074) -> calcDataBleachCorrection + LINEAR_1D fit
  GPU_Calculate_ACF        (14812-14842)  -> calcACF
  GPU_ACF_Fit              (15076-15285)  -> ACF_1D fit, keep CONVERGED pixels

This module reproduces that packing so a host program (or a test) sends the
native side exactly what the Java plugin would.  It is a client of the
engine, not part of the compute path.
"""
import numpy as np

from . import (ACFParameters, FitModel, MODEL_ACF_1D, MODEL_LINEAR_1D, calc_acf, calc_binning,
               calc_data_bleach_correction, correlator_tables, fit, is_binning_memory_sufficient)


def initialize_gpu_params(width, height, roi_w, roi_h, binning, cf, frametime, firstframe, lastframe, background,
                          poly_order, p=16, q=8, overlap=False):
    """GPU_Initialize_GPUparams (IFCS:14844-14906) for a ROI starting at 0."""
    pixbin = 1 if overlap else binning
    cfx, cfy = cf
    w_out = (min(roi_w + cfx, width) - cfx) // pixbin
    h_out = (min(roi_h + cfy, height) - cfy) // pixbin
    if w_out <= 0 or h_out <= 0:
        return None
    win_star = w_out * pixbin + cfx
    hin_star = h_out * pixbin + cfy
    num = lastframe - firstframe + 1
    lag, samp, mtab, _ = correlator_tables(p, q, num, frametime)
    nopit = num if num < 1000 else 1000
    a = ACFParameters()
    a.width, a.height = w_out, h_out
    a.win_star, a.hin_star = win_star, hin_star
    a.w_temp, a.h_temp = win_star - binning + 1, hin_star - binning + 1
    a.pixbinX = a.pixbinY = pixbin
    a.binningX = a.binningY = binning
    a.firstframe, a.lastframe = firstframe, lastframe
    a.framediff = num
    a.cfXDistance, a.cfYDistance = cfx, cfy
    a.correlatorp, a.correlatorq = float(p), float(q)
    a.frametime = frametime
    a.background = background
    a.mtab1 = float(mtab[1])
    a.mtabchanumminus1 = float(mtab[-1])
    a.sampchanumminus1 = float(samp[-1])
    a.chanum = len(lag)
    a.isNBcalculation = 0
    a.bleachcorr_gpu = 1 if poly_order > 0 else 0
    a.bleachcorr_order = poly_order + 1 if poly_order > 0 else 0
    a.nopit = nopit
    a.ave = num // nopit
    return a


def get_pixels(stack, params):
    """GPU_get_pixels (IFCS:14908-14986): the float[] stack the plugin sends,
    stride-1 binned to w_temp x h_temp, background * b^2 subtracted."""
    a = params
    T = a.lastframe - a.firstframe + 1
    sub = np.asarray(stack[a.firstframe - 1:a.lastframe, :a.hin_star, :a.win_star], np.float32)
    if a.binningX == 1 and a.binningY == 1:
        pixels = np.ascontiguousarray(sub[:, :a.h_temp, :a.w_temp]).reshape(-1)
    elif is_binning_memory_sufficient(a):
        pixels = calc_binning(sub.reshape(-1), a)
    else:
        pixels = np.zeros((T, a.h_temp, a.w_temp), np.float32)
        for k in range(a.binningY):
            for i in range(a.binningX):
                pixels += sub[:, k:k + a.h_temp, i:i + a.w_temp]
        pixels = pixels.reshape(-1)
    return pixels - np.float32(a.background * a.binningX * a.binningY)


def bleach_correction(pixels, params, max_iterations=2000):
    """GPU_Calculate_BleachCorrection (IFCS:14988-15074): per-position
    means of `ave` frames, then a LINEAR_1D fit started from the last trace
    value with tolerance 1e-16f.  Returns bleachcorr_params (double)."""
    a = params
    nfits, npts = a.w_temp * a.h_temp, a.nopit
    data = calc_data_bleach_correction(pixels, a)
    m = FitModel(nfits, npts, True, MODEL_LINEAR_1D, tolerance=np.float32(1e-16),
                 max_number_iterations=max_iterations, num_valid_coefs=a.bleachcorr_order,
                 parameters_to_fit=[True] + [i < a.bleachcorr_order for i in range(1, 11)],
                 user_info_size=npts * 4)
    init = np.zeros((nfits, 11), np.float32)
    init[:, 0] = data.reshape(nfits, npts)[:, -1]
    m.initial_parameters[:] = init.reshape(-1)
    m.weights[:] = 1.0
    m.user_info[:npts] = (a.frametime * (np.arange(npts) + 0.5) * a.ave).astype(np.float32)
    m.data[:] = data
    r = fit(m)
    return r.parameters.reshape(nfits, 11)[:, :a.bleachcorr_order].astype(np.float64).reshape(-1), r


def calculate_acf(pixels, bleachcorr_params, params):
    """GPU_Calculate_ACF (IFCS:14812-14842)."""
    a = params
    lag, samp, _, _ = correlator_tables(int(a.correlatorp), int(a.correlatorq), a.framediff, a.frametime)
    bc = bleachcorr_params if bleachcorr_params is not None else np.zeros(1)
    return calc_acf(pixels, bc, samp, lag, a)


def acf_fit(pixels1, blockvar, params, geom, init11, paramfit11, max_iterations=2000):
    """GPU_ACF_Fit (IFCS:15076-15285): fit channels 1..chanum-1 with weights
    1/blockvar; only CONVERGED pixels keep their parameters (others NaN)."""
    a = params
    nfits, npts = a.width * a.height, a.chanum - 1
    lagtime = correlator_tables(int(a.correlatorp), int(a.correlatorq), a.framediff, a.frametime)[3]
    true = np.array(list(init11) + [geom.ax, geom.ay, geom.s, geom.sz, geom.rx, geom.ry, geom.fitobsvol, geom.q2,
                                    geom.q3], np.float32)
    m = FitModel(nfits, npts, True, MODEL_ACF_1D, tolerance=np.float32(1e-16),
                 max_number_iterations=max_iterations, num_valid_coefs=a.bleachcorr_order,
                 parameters_to_fit=list(paramfit11) + [False] * 9, user_info_size=npts * 4)
    m.initial_parameters[:] = np.tile(true, nfits)
    m.user_info[:npts] = lagtime[1:].astype(np.float32)
    p1 = pixels1.reshape(a.chanum, nfits)
    bv = blockvar.reshape(a.chanum, nfits)
    m.data[:] = p1[1:].T.astype(np.float32).reshape(-1)
    m.weights[:] = (np.float32(1.0) / bv[1:].T.astype(np.float32)).reshape(-1)
    r = fit(m)
    params20 = r.parameters.reshape(nfits, 20).astype(np.float64)
    conv = r.states == 0
    fitres = np.where(conv[:, None], params20[:, :11], np.nan)
    nfree = int(np.sum(paramfit11))
    chi2 = r.chi_squares / ((npts - 1) - nfree - 1)
    return fitres.reshape(a.height, a.width, 11), chi2.reshape(a.height, a.width), r


def calculate_acf_all(stack, roi_w, roi_h, binning, cf, frametime, background, poly_order, geom, init11,
                      paramfit11, do_fit=True, p=16, q=8):
    """GPU_Calculate_ACF_All (IFCS:15287-15510) for a ROI at the origin."""
    T, H, W = stack.shape
    if poly_order < 0 or cf[0] < 0 or cf[1] < 0:
        raise ValueError("GPU path supports polynomial/none bleach correction and non-negative CF distances")
    a = initialize_gpu_params(W, H, roi_w, roi_h, binning, cf, frametime, 1, T, background, poly_order, p, q)
    pixels = get_pixels(stack, a)
    bc = None
    if a.bleachcorr_gpu:
        bc, _ = bleach_correction(pixels, a)
    pixels1, bvar, blocked = calculate_acf(pixels, bc, a)
    out = dict(params=a, acf=pixels1.reshape(a.chanum, a.height, a.width), var=bvar.reshape(a.chanum, a.height,
                                                                                           a.width),
               blocked=blocked.reshape(a.chanum, a.height, a.width)[:2], bleachcorr_params=bc)
    if do_fit:
        out["fitres"], out["chi2"], _ = acf_fit(pixels1, bvar, a, geom, init11, paramfit11)
    return out
