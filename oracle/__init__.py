"""ctypes bindings for the oracle: a CPU restatement of the Imaging_FCS 1.52
Java CPU path (Imaging_FCS_1_52.java) in C.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker or as the timed CPU baseline.  The product (``imaging_fcs_1_52_amd``)
never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libimfcs_oracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


class FcsModel(ctypes.Structure):
    """orc_fcs_model: fixed geometry of FCS_3p (IFCS:13131-13146) plus the
    paramfit / paraminitval arrays of prepareFit (IFCS:9986-10039)."""
    _fields_ = [("ax", ctypes.c_double), ("ay", ctypes.c_double),
                ("s", ctypes.c_double), ("sz", ctypes.c_double),
                ("rx", ctypes.c_double), ("ry", ctypes.c_double),
                ("fitobsvol", ctypes.c_double), ("q2", ctypes.c_double),
                ("q3", ctypes.c_double),
                ("paramfit", ctypes.c_int * 11),
                ("paraminit", ctypes.c_double * 11)]


class Config(ctypes.Structure):
    """orc_config: the plugin settings read by correlateROI's CPU branch."""
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int),
                ("frames", ctypes.c_int), ("firstframe", ctypes.c_int),
                ("lastframe", ctypes.c_int), ("binningX", ctypes.c_int),
                ("binningY", ctypes.c_int), ("pixbinX", ctypes.c_int),
                ("pixbinY", ctypes.c_int), ("cfXDistance", ctypes.c_int),
                ("cfYDistance", ctypes.c_int), ("background", ctypes.c_int),
                ("correlatorp", ctypes.c_int), ("correlatorq", ctypes.c_int),
                ("frametime", ctypes.c_double), ("polyOrder", ctypes.c_int),
                ("roiX", ctypes.c_int), ("roiY", ctypes.c_int),
                ("roiW", ctypes.c_int), ("roiH", ctypes.c_int),
                ("doFit", ctypes.c_int), ("fitstart", ctypes.c_int),
                ("fitend", ctypes.c_int), ("gls", ctypes.c_int),
                ("nthreads", ctypes.c_int), ("model", FcsModel)]


def build(force=False):
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-B" if force else "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_tables.restype = ctypes.c_int
        L.orc_tables.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _I, _D, _D, _D]
        L.orc_get_intensity.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 10 + [_D]
        L.orc_intensity_trace.argtypes = ([ctypes.c_void_p] + [ctypes.c_int] * 11 +
                                          [ctypes.c_double, _D, _D])
        L.orc_polyfit.restype = ctypes.c_int
        L.orc_polyfit.argtypes = [_D, _D, ctypes.c_int, ctypes.c_int, _D]
        L.orc_bleach_apply.argtypes = [_D, ctypes.c_int, _D, ctypes.c_int, ctypes.c_double]
        L.orc_bleach_linseg.restype = ctypes.c_int
        L.orc_bleach_linseg.argtypes = [_D, ctypes.c_int, ctypes.c_int, _D]
        L.orc_expfit.restype = ctypes.c_int
        L.orc_expfit.argtypes = [_D, _D, ctypes.c_int, ctypes.c_int, _D]
        L.orc_bleach_exp.argtypes = [_D, ctypes.c_int, _D, ctypes.c_int, ctypes.c_double]
        L.orc_block_transform.restype = ctypes.c_int
        L.orc_block_transform.argtypes = [_D, _D, ctypes.c_int, ctypes.c_int, _I, _D, _D,
                                          ctypes.c_double, ctypes.c_int, _I]
        L.orc_calculate_cf.argtypes = [_D, _D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I, _D, _D,
                                       _D, _D, _D, _D]
        L.orc_obsvol.restype = ctypes.c_double
        L.orc_obsvol.argtypes = [ctypes.c_double] * 3
        L.orc_fcs3p_value.restype = ctypes.c_double
        L.orc_fcs3p_value.argtypes = [ctypes.POINTER(FcsModel), ctypes.c_double, _D]
        L.orc_fcs3p_gradient.argtypes = [ctypes.POINTER(FcsModel), ctypes.c_double, _D, _D]
        L.orc_standard_fit.restype = ctypes.c_int
        L.orc_standard_fit.argtypes = [ctypes.POINTER(FcsModel), ctypes.c_int, _D, _D, _D,
                                       ctypes.c_int, ctypes.c_int, _D, _D, _I, _I]
        L.orc_gls_fit.restype = ctypes.c_int
        L.orc_gls_fit.argtypes = [ctypes.POINTER(FcsModel), ctypes.c_int, _D, _D, _D, _D, _D, _I, _I]
        L.orc_roi_dims.argtypes = [ctypes.POINTER(Config), _I, _I, _I, _I]
        L.orc_pipeline.restype = ctypes.c_int
        L.orc_pipeline.argtypes = [ctypes.c_void_p, ctypes.POINTER(Config), _D, _D, _I, _I, _D, _D, _I, _D]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


def tables(p, q, num, frametime):
    """Correlator tables lag, samp, mtab, lagtime (IFCS:7391-7424)."""
    n = p + (q - 1) * p // 2 + 1
    lag = np.zeros(n, np.int32)
    samp, mtab, lagtime = (np.zeros(n) for _ in range(3))
    c = lib().orc_tables(p, q, num, frametime, _ip(lag), _dp(samp), _dp(mtab), _dp(lagtime))
    assert c == n
    return lag, samp, mtab, lagtime


def get_intensity(stack, px, py, bx, by, bckg, first, last):
    stack = np.ascontiguousarray(stack, np.uint16)
    T, H, W = stack.shape
    out = np.zeros(last - first + 2)
    lib().orc_get_intensity(stack.ctypes.data, W, H, T, px, py, bx, by, bckg, first, last, _dp(out))
    return out


def intensity_trace(stack, px, py, bx, by, bckg, first, last, nopit, frametime):
    stack = np.ascontiguousarray(stack, np.uint16)
    T, H, W = stack.shape
    tr, it = np.zeros(nopit), np.zeros(nopit)
    lib().orc_intensity_trace(stack.ctypes.data, W, H, T, px, py, bx, by, bckg, first, last, nopit,
                              frametime, _dp(tr), _dp(it))
    return tr, it


def polyfit(trace, int_time, order):
    trace = np.ascontiguousarray(trace, np.float64)
    int_time = np.ascontiguousarray(int_time, np.float64)
    coef = np.zeros(order + 1)
    st = lib().orc_polyfit(_dp(trace), _dp(int_time), len(trace), order, _dp(coef))
    return coef, st


def bleach_apply(intdat, coef, frametime):
    d = np.array(intdat, np.float64)
    coef = np.ascontiguousarray(coef, np.float64)
    lib().orc_bleach_apply(_dp(d), len(d) - 1, _dp(coef), len(coef) - 1, frametime)
    return d


def bleach_linseg(intdat, swl):
    """"Lin Segment" correction of a 1-based trace (IFCS:9767-9805).
    Returns (corrected trace, [bcInt0, segment means...])."""
    d = np.array(intdat, np.float64)
    num = len(d) - 1
    seg = np.zeros(num // swl + 1)
    n = lib().orc_bleach_linseg(_dp(d), num, swl, _dp(seg))
    if n == 0:
        raise ValueError("linear segments need two segments of >= 2 frames")
    return d, seg


def expfit(trace, int_time, nexp):
    """SingleExpFit (nexp=1) / DoubleExpFit (nexp=2) of a trace
    (IFCS:10845-10950): returns (params, status)."""
    trace = np.ascontiguousarray(trace, np.float64)
    int_time = np.ascontiguousarray(int_time, np.float64)
    res = np.zeros(3 if nexp == 1 else 5)
    st = lib().orc_expfit(_dp(trace), _dp(int_time), len(trace), nexp, _dp(res))
    return res, st


def bleach_exp(intdat, res, nexp, frametime):
    """"Single Exp" / "Double Exp" correction of a 1-based trace (IFCS:9682-9738)."""
    d = np.array(intdat, np.float64)
    res = np.ascontiguousarray(res, np.float64)
    lib().orc_bleach_exp(_dp(d), len(d) - 1, _dp(res), nexp, frametime)
    return d


def block_transform(a, b, p=16, q=8, frametime=0.001):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    num = len(a) - 1
    lag, samp, mtab, _ = tables(p, q, num, frametime)
    flag = ctypes.c_int(0)
    ind = lib().orc_block_transform(_dp(a), _dp(b), num, len(lag), _ip(lag), _dp(samp), _dp(mtab),
                                    frametime, q, ctypes.byref(flag))
    return ind, flag.value


def calculate_cf(a, b, ind, p=16, q=8, frametime=0.001, gls=False):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    num = len(a) - 1
    lag, samp, mtab, _ = tables(p, q, num, frametime)
    n = len(lag)
    corav, bvar, bsd = np.zeros(n), np.zeros(n), np.zeros(n)
    cov = np.zeros((n - 1, n - 1)) if gls else None
    lib().orc_calculate_cf(_dp(a), _dp(b), num, ind, n, _ip(lag), _dp(samp), _dp(mtab),
                           _dp(corav), _dp(bvar), _dp(bsd), _dp(cov) if gls else None)
    return corav, bvar, bsd, cov


def sliding_window_q(L, p=16, q=8):
    """Correlator depth of a sliding-window run: setParameters replaces q by
    lagnum = floor((ln(L / (swMinFrameReq + p)) + 1) / ln 2) when that is
    smaller, swMinFrameReq = 20 (IFCS:203, 7331-7341)."""
    lagnum = int(np.floor((np.log(L / (20.0 + p)) + 1.0) / np.log(2.0)))
    return min(q, lagnum)


def sliding_window_cf(stack, px, py, bx, by, bckg, first, last, L, p=16, q=8, frametime=0.001, px2=None,
                      py2=None):
    """correlate() with "Sliding Window" bleach correction (IFCS:9146-9172):
    the numofsw = floor(num / L) windows are read without correction
    (getIntensity has no branch for this mode), blocked and correlated on
    their own with the window's tables, and corav, blockvar, blocksd are
    summed as term / numofsw.  Returns (acf, var, sd, ind, blockIndS, q) with
    the last window's blocking result."""
    qq = sliding_window_q(L, p, q)
    nsw = (last - first + 1) // L
    acf = var = sd = 0.0
    ind = bis = 0
    for i in range(nsw):
        f0, f1 = i * L + first, (i + 1) * L + first - 1
        a = get_intensity(stack, px, py, bx, by, bckg, f0, f1)
        b = a if px2 is None else get_intensity(stack, px2, py2, bx, by, bckg, f0, f1)
        ind, bis = block_transform(a, b, p, qq, frametime)
        corav, bvar, bsd, _ = calculate_cf(a, b, ind, p, qq, frametime)
        acf = acf + corav / nsw
        var = var + bvar / nsw
        sd = sd + bsd / nsw
    return acf, var, sd, ind, bis, qq


def make_model(ax, ay, s, sz=None, rx=0.0, ry=0.0, q2=1.0, q3=1.0, paramfit=None, paraminit=None):
    """Build FcsModel in SI units.  Defaults follow the fit panel
    (IFCS:1959-1968): N=1, D=1 um^2/s, everything else 0; N, D, G free."""
    m = FcsModel()
    m.ax, m.ay, m.s = ax, ay, s
    m.sz = 1e6 * 515e-9 / 1.49 if sz is None else sz
    m.rx, m.ry, m.q2, m.q3 = rx, ry, q2, q3
    m.fitobsvol = obsvol(ax, ay, s)
    pf = paramfit if paramfit is not None else [1, 1, 0, 0, 1, 0, 0, 0, 0, 0, 0]
    pi = paraminit if paraminit is not None else [1.0, 1e-12, 0, 0, 0, 0, 0, 0, 0, 0, 0]
    for i in range(11):
        m.paramfit[i] = int(pf[i])
        m.paraminit[i] = float(pi[i])
    return m


def obsvol(ax, ay, s):
    return lib().orc_obsvol(ax, ay, s)


def fcs_value(model, x, pareq):
    pq = np.ascontiguousarray(pareq, np.float64)
    return lib().orc_fcs3p_value(ctypes.byref(model), float(x), _dp(pq))


def fcs_gradient(model, x, pareq):
    pq = np.ascontiguousarray(pareq, np.float64)
    g = np.zeros(11)
    lib().orc_fcs3p_gradient(ctypes.byref(model), float(x), _dp(pq), _dp(g))
    return g


def standard_fit(model, lagtime, acf, varacf, fitstart=1, fitend=None):
    lagtime = np.ascontiguousarray(lagtime, np.float64)
    acf = np.ascontiguousarray(acf, np.float64)
    varacf = np.ascontiguousarray(varacf, np.float64)
    n = len(lagtime)
    fitend = n - 1 if fitend is None else fitend
    fr, chi2 = np.zeros(11), ctypes.c_double(0)
    ne, ni = ctypes.c_int(0), ctypes.c_int(0)
    ok = lib().orc_standard_fit(ctypes.byref(model), n, _dp(lagtime), _dp(acf), _dp(varacf), fitstart,
                                fitend, _dp(fr), ctypes.byref(chi2), ctypes.byref(ne), ctypes.byref(ni))
    return fr, chi2.value, bool(ok), ne.value, ni.value


def gls_fit(model, lagtime, mcov, covmats):
    lagtime = np.ascontiguousarray(lagtime, np.float64)
    mcov = np.ascontiguousarray(mcov, np.float64)
    covmats = np.ascontiguousarray(covmats, np.float64)
    fr, chi2 = np.zeros(11), ctypes.c_double(0)
    ne, ni = ctypes.c_int(0), ctypes.c_int(0)
    ok = lib().orc_gls_fit(ctypes.byref(model), len(lagtime), _dp(lagtime), _dp(mcov), _dp(covmats),
                           _dp(fr), ctypes.byref(chi2), ctypes.byref(ne), ctypes.byref(ni))
    return fr, chi2.value, bool(ok)


def make_config(W, H, T, binning=1, cf=(0, 0), background=0, poly_order=4, frametime=1e-3,
                p=16, q=8, roi=None, do_fit=True, gls=False, nthreads=1, model=None, overlap=False,
                firstframe=1, lastframe=None):
    c = Config()
    c.width, c.height, c.frames = W, H, T
    c.firstframe = firstframe
    c.lastframe = T if lastframe is None else lastframe
    c.binningX = c.binningY = binning
    c.pixbinX = c.pixbinY = 1 if overlap else binning
    c.cfXDistance, c.cfYDistance = cf
    c.background = background
    c.correlatorp, c.correlatorq = p, q
    c.frametime = frametime
    c.polyOrder = poly_order
    rx, ry, rw, rh = roi if roi is not None else (0, 0, W - cf[0], H - cf[1])
    c.roiX, c.roiY, c.roiW, c.roiH = rx, ry, rw, rh
    c.doFit = int(do_fit)
    chanum = p + (q - 1) * p // 2 + 1
    c.fitstart, c.fitend = 1, chanum - 1
    c.gls = int(gls)
    c.nthreads = nthreads
    if model is not None:
        c.model = model
    return c


def roi_dims(cfg):
    sx, sy, ow, oh = (ctypes.c_int() for _ in range(4))
    lib().orc_roi_dims(ctypes.byref(cfg), ctypes.byref(sx), ctypes.byref(sy), ctypes.byref(ow),
                       ctypes.byref(oh))
    return sx.value, sy.value, ow.value, oh.value


def pipeline(stack, cfg):
    """Run the CPU path over the ROI.  Returns a dict of per-output-pixel
    arrays shaped (oh, ow, ...)."""
    stack = np.ascontiguousarray(stack, np.uint16)
    sx, sy, ow, oh = roi_dims(cfg)
    p, q = cfg.correlatorp, cfg.correlatorq
    n = p + (q - 1) * p // 2 + 1
    npx = ow * oh
    acf, var = np.zeros((npx, n)), np.zeros((npx, n))
    ind, bis = np.zeros(npx, np.int32), np.zeros(npx, np.int32)
    fitres, chi2 = np.full((npx, 11), np.nan), np.full(npx, np.nan)
    fitted = np.zeros(npx, np.int32)
    st = np.zeros(4)
    rc = lib().orc_pipeline(stack.ctypes.data, ctypes.byref(cfg), _dp(acf), _dp(var), _ip(ind), _ip(bis),
                            _dp(fitres), _dp(chi2), _ip(fitted), _dp(st))
    if rc < 0:
        raise ValueError("orc_pipeline failed: %d" % rc)
    sh = (oh, ow)
    return dict(acf=acf.reshape(sh + (n,)), var=var.reshape(sh + (n,)), ind=ind.reshape(sh),
                blockIndS=bis.reshape(sh), fitres=fitres.reshape(sh + (11,)), chi2=chi2.reshape(sh),
                fitted=fitted.reshape(sh).astype(bool), stage_seconds=st, origin=(sx, sy))
