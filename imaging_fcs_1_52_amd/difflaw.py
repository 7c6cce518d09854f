"""FCS diffusion law over pixel binnings on the native pipeline.

The plugin's diffusion law (dlfit, IFCS:11007-11231, "All" mode with
non-overlapping binning, and the conversion of determineDiffusionLaw,
IFCS:8190-8215) runs the whole correlate-and-fit pipeline once per binning
u = 1..numbin and reduces each binning to one point:

    Aeff(u) = obsvolFCS_ST2D1p(2) at pixel size u * a     [um^2]
    D(u)    = mean of the fitted D over converged pixels  [um^2/s]
    tau(u)  = Aeff / D,   err(u) = Aeff / D^2 * sd(D)

fitDiffLaw (IFCS:8253-8296) then fits the line tau = t0 + Aeff / D_eff over
a range of points with weights 1 / err^2 (LineFit, IFCS:10792-10843, whose
least-squares optimum is the closed-form weighted line below).

Here every binning is one native pipeline run on the device-resident
stack, one engine (HIP stream) per binning, so the binnings run
concurrently (imfcs_run_multi: with the polynomial correction their bleach
traces come from one pass over the stack).
"""
import numpy as np

from . import Engine, make_fit_model, make_pipeline_config, obsvol


def diffusion_law(stack, numbin, pixel_size_m, psf_m, background=None, poly_order=4, frametime=1e-3, sz=None,
                  engines=None, d_stack=None):
    """Diffusion-law points of binnings 1..numbin.

    stack: uint16 [T][H][W] (host array; uploaded once unless d_stack, a
    device pointer already holding it, is given).  Returns a dict with
    'binning', 'aeff' (um^2), 'd_mean' (um^2/s), 'd_var' ((um^2/s)^2),
    'tau' (s), 'err' (s), 'n_fitted' and, per binning, the fitted D map
    ('d_maps', um^2/s, NaN where not converged)."""
    stack = np.asarray(stack)
    T, H, W = stack.shape
    if numbin < 1 or numbin > min(W, H):
        raise ValueError("numbin must be between 1 and the image side")
    bg = int(stack.min()) if background is None else int(background)
    bins = list(range(1, numbin + 1))
    own = engines is None
    if own:
        engines = [Engine(0) for _ in bins]
    if len(engines) < len(bins):
        raise ValueError("one engine per binning is needed")
    try:
        dev = d_stack if d_stack is not None else engines[0].upload(stack)
        cfgs = []
        for u in bins:
            ax = u * pixel_size_m
            model = make_fit_model(ax, ax, psf_m, sz=sz)
            cfgs.append(make_pipeline_config(W, H, T, binning=u, cf=(0, 0), background=bg, poly_order=poly_order,
                                             frametime=frametime, do_fit=True, acf_model=model))
        Engine.run_multi(engines[:len(bins)], cfgs, dev)
        out = {k: [] for k in ("binning", "aeff", "d_mean", "d_var", "tau", "err", "n_fitted", "d_maps")}
        for u, eng, cfg in zip(bins, engines, cfgs):
            eng.synchronize()
            r = eng.results(cfg)
            ok = (np.asarray(r["states"])[0] == 0) & np.isfinite(np.asarray(r["fitres"])[0][..., 1])
            d = np.asarray(r["fitres"])[0][..., 1] * 1e12
            dv = d[ok]
            n = int(ok.sum())
            aeff = obsvol(u * pixel_size_m, u * pixel_size_m, psf_m) * 1e12
            dm = float(dv.mean()) if n else float("nan")
            var = float((dv * dv).mean() - dm * dm) if n else float("nan")
            out["binning"].append(u)
            out["aeff"].append(aeff)
            out["d_mean"].append(dm)
            out["d_var"].append(var)
            out["tau"].append(aeff / dm)
            out["err"].append(aeff / dm ** 2 * np.sqrt(max(var, 0.0)))
            out["n_fitted"].append(n)
            out["d_maps"].append(np.where(ok, d, np.nan))
        for k in ("binning", "aeff", "d_mean", "d_var", "tau", "err", "n_fitted"):
            out[k] = np.asarray(out[k])
        return out
    finally:
        if own:
            for e in engines:
                e.close()


def fit_diffusion_law(aeff, tau, err, first=1, last=5):
    """fitDiffLaw (IFCS:8253-8296): weighted line tau = intercept + slope *
    Aeff over the points first..last (1-based, inclusive) with weights
    1 / err^2.  Returns (intercept [s], slope [s/um^2]); 1 / slope is the
    effective D."""
    lo, hi = first - 1, last - 1
    if lo >= hi or lo < 0 or hi >= len(aeff):
        raise ValueError("fit start/end not correctly set")
    x = np.asarray(aeff[lo:hi + 1], np.float64)
    y = np.asarray(tau[lo:hi + 1], np.float64)
    w = 1.0 / np.asarray(err[lo:hi + 1], np.float64) ** 2
    sw, sx, sy = w.sum(), (w * x).sum(), (w * y).sum()
    sxx, sxy = (w * x * x).sum(), (w * x * y).sum()
    den = sw * sxx - sx * sx
    slope = (sw * sxy - sx * sy) / den
    intercept = (sy - slope * sx) / sw
    return intercept, slope


def psf_calibration(stack, psf_values_m, bin_start, bin_end, pixel_size_m, background=None, poly_order=4,
                    frametime=1e-3, sz=None, engines=None, d_stack=None):
    """PSF calibration (determinePSF, IFCS:8118-8160): the diffusion-law
    statistics of binnings 1..numbin (numbin = bin_end - bin_start + 1; dlfit
    runs binnings 1..numbin whatever bin_start is, and the plot labels them
    u + bin_start, IFCS:8143-8146) for each trial PSF size.

    The reference reruns the whole pipeline per PSF size; only the fit
    depends on it, so every binning is correlated once and refitted
    (imfcs_refit) for the other sizes.  psf_values_m: the trial PSF sizes
    in metres (the plugin's (psfStart + x psfStep) * lambda / NA / 1e9,
    IFCS:8142), fewer than 10 (IFCS:8135).

    Returns a dict with 'label' (u + bin_start), 'd_mean' and 'd_var'
    ([npsf][numbin], um^2/s and (um^2/s)^2, over the converged pixels), 'err'
    (the reference's error bar: the variance divided by the square root of
    Java's integer pixel count ((W / b) * H) / b, IFCS:8147) and 'n_fitted'."""
    stack = np.asarray(stack)
    T, H, W = stack.shape
    psf_values_m = list(psf_values_m)
    numbin = bin_end - bin_start + 1
    if W < 20 or H < 20:
        raise ValueError("at least 20x20 pixels are required (IFCS:8123)")
    if not psf_values_m or len(psf_values_m) >= 10:
        raise ValueError("1 to 9 PSF sizes (IFCS:8135)")
    if numbin < 1 or numbin > min(W, H):
        raise ValueError("bad binning range")
    bg = int(stack.min()) if background is None else int(background)
    bins = list(range(1, numbin + 1))
    own = engines is None
    if own:
        engines = [Engine(0) for _ in bins]
    if len(engines) < len(bins):
        raise ValueError("one engine per binning is needed")

    def cfg_for(u, psf):
        ax = u * pixel_size_m
        return make_pipeline_config(W, H, T, binning=u, cf=(0, 0), background=bg, poly_order=poly_order,
                                    frametime=frametime, do_fit=True, acf_model=make_fit_model(ax, ax, psf, sz=sz))

    try:
        dev = d_stack if d_stack is not None else engines[0].upload(stack)
        out = {"label": np.asarray(bins) + (bin_start - 1),
               "d_mean": np.full((len(psf_values_m), numbin), np.nan), "d_var": np.full((len(psf_values_m), numbin),
                                                                                        np.nan),
               "err": np.full((len(psf_values_m), numbin), np.nan), "n_fitted": np.zeros((len(psf_values_m), numbin),
                                                                                         int)}
        for x, psf in enumerate(psf_values_m):
            cfgs = [cfg_for(u, psf) for u in bins]
            if x == 0:
                Engine.run_multi(engines[:numbin], cfgs, dev)
            else:
                for e, c in zip(engines, cfgs):
                    e.refit(c)
            for i, (u, e, c) in enumerate(zip(bins, engines, cfgs)):
                e.synchronize()
                r = e.results(c)
                d = np.asarray(r["fitres"])[0][..., 1] * 1e12
                ok = (np.asarray(r["states"])[0] == 0) & np.isfinite(d)
                n = int(ok.sum())
                out["n_fitted"][x, i] = n
                if n:
                    dv = d[ok]
                    m = float(dv.mean())
                    var = float((dv * dv).mean() - m * m)
                    out["d_mean"][x, i] = m
                    out["d_var"][x, i] = var
                    out["err"][x, i] = var / np.sqrt(((W // u) * H) // u)
        return out
    finally:
        if own:
            for e in engines:
                e.close()


def diffusion_law_map(stack, numbin, roi_dim, pixel_size_m, psf_m, background=None, poly_order=4, frametime=1e-3,
                      sz=None, fit_first=1, fit_last=5, engines=None, d_stack=None):
    """Diffusion-law map ("ROI" mode of dlfit, IFCS:11068-11080, 11084-11181,
    and of determineDiffusionLaw, IFCS:8218-8245): the image is cut into
    floor(W / roi_dim) x floor(H / roi_dim) square regions; for every
    binning u = 1..numbin the pixels are taken on the overlapping grid
    (stride 1, output pixel x = 0..W-u) and region (r, s) averages the fitted
    D over its pixels x = roi_dim r .. roi_dim (r+1) - u (the same in y);
    each region's diffusion law is then fitted over points fit_first..
    fit_last (fitDiffLaw, limits 1..5 in the plugin).

    One pipeline per binning over the whole image (imfcs_run_multi), the
    regions are sums over its D map.  Returns a dict with 'aeff' (numbin,
    um^2), 'd_mean', 'd_var', 'tau', 'err' ([mh][mw][numbin]), 'n_fitted'
    and 'fit' ([mh][mw][2]: intercept in s, slope in s/um^2; NaN where a
    region has no fitted pixel at some binning of the fit range)."""
    stack = np.asarray(stack)
    T, H, W = stack.shape
    if roi_dim < 1 or roi_dim > min(W, H) or numbin < 1 or numbin > roi_dim:
        raise ValueError("roi_dim must fit the image and numbin the region")
    mw, mh = W // roi_dim, H // roi_dim
    bg = int(stack.min()) if background is None else int(background)
    bins = list(range(1, numbin + 1))
    own = engines is None
    if own:
        engines = [Engine(0) for _ in bins]
    if len(engines) < len(bins):
        raise ValueError("one engine per binning is needed")
    try:
        dev = d_stack if d_stack is not None else engines[0].upload(stack)
        cfgs = []
        for u in bins:
            ax = u * pixel_size_m
            cfgs.append(make_pipeline_config(W, H, T, binning=u, cf=(0, 0), background=bg, poly_order=poly_order,
                                             frametime=frametime, do_fit=True, overlap=True,
                                             acf_model=make_fit_model(ax, ax, psf_m, sz=sz)))
        Engine.run_multi(engines[:len(bins)], cfgs, dev)
        shape = (mh, mw, numbin)
        out = {"aeff": np.zeros(numbin), "d_mean": np.full(shape, np.nan), "d_var": np.full(shape, np.nan),
               "n_fitted": np.zeros(shape, int), "d_maps": []}
        for i, (u, eng, cfg) in enumerate(zip(bins, engines, cfgs)):
            eng.synchronize()
            r = eng.results(cfg)
            d = np.asarray(r["fitres"])[0][..., 1] * 1e12
            ok = (np.asarray(r["states"])[0] == 0) & np.isfinite(d)
            out["aeff"][i] = obsvol(u * pixel_size_m, u * pixel_size_m, psf_m) * 1e12
            out["d_maps"].append(np.where(ok, d, np.nan))
            for s in range(mh):
                for q in range(mw):
                    ys = slice(roi_dim * s, roi_dim * (s + 1) - u + 1)
                    xs = slice(roi_dim * q, roi_dim * (q + 1) - u + 1)
                    sel = ok[ys, xs]
                    n = int(sel.sum())
                    out["n_fitted"][s, q, i] = n
                    if n:
                        dv = d[ys, xs][sel]
                        m = float(dv.mean())
                        out["d_mean"][s, q, i] = m
                        out["d_var"][s, q, i] = float((dv * dv).mean() - m * m)
        out["tau"] = out["aeff"] / out["d_mean"]
        out["err"] = out["aeff"] / out["d_mean"] ** 2 * np.sqrt(np.maximum(out["d_var"], 0.0))
        fit = np.full((mh, mw, 2), np.nan)
        for s in range(mh):
            for q in range(mw):
                lo, hi = fit_first - 1, fit_last - 1
                seg_ok = np.isfinite(out["tau"][s, q, lo:hi + 1]).all() and (out["err"][s, q, lo:hi + 1] > 0).all()
                if hi < numbin and seg_ok:
                    fit[s, q] = fit_diffusion_law(out["aeff"], out["tau"][s, q], out["err"][s, q], fit_first, fit_last)
        out["fit"] = fit
        return out
    finally:
        if own:
            for e in engines:
                e.close()
