"""MI355X-native imaging-FCS engine.

The hot path of Imaging_FCS 1.52 -- binning, polynomial bleach correction,
multi-tau ACF/CCF with blocked variance and the per-pixel Levenberg-Marquardt
fit -- runs as hand-written HIP kernels in ``libimfcs_hip.so``.  This module is
the host-side mirror of the plugin's JNI class ``gpufitImFCS.GpufitImFCS``
(GpufitImFCS.java): the same method names, argument meaning and error
behaviour, over the library's C ABI (include/imfcs.h).  There is no CPU
fallback: every entry point raises if the native library cannot be loaded or
no GPU is present.
"""
import ctypes
import os

import numpy as np

from . import native
from .native import (ACFParameters, FitModelGeom, PipelineConfig, SimDomain, MODEL_ACF_1D, MODEL_LINEAR_1D, ESTIMATOR_LSE,
                     STATE_CONVERGED, STATE_MAX_ITERATION, STATE_SINGULAR_HESSIAN, lib, sim_lib)

__all__ = ["ACFParameters", "FitModel", "FitResult", "Engine", "PipelineConfig", "FitModelGeom", "fit",
           "get_last_error", "is_cuda_available", "get_cuda_version", "reset_gpu", "calc_data_bleach_correction",
           "is_binning_memory_sufficient", "calc_binning", "is_acf_memory_sufficient", "calc_acf", "simulate_2d",
           "SimDomain", "sim_domain", "simulate_2d_device",
           "correlator_tables", "obsvol", "make_fit_model", "make_pipeline_config"]


class ImFCSError(RuntimeError):
    pass


def get_last_error():
    """GpufitImFCS.getLastError (GpufitImFCS.java:382)."""
    return lib().imfcs_get_last_error().decode()


def is_cuda_available():
    """GpufitImFCS.isCudaAvailable (GpufitImFCS.java:389-399)."""
    return bool(lib().imfcs_device_available())


def get_cuda_version():
    """getCudaVersionAsArray (GpufitImFCS.java:403-436): (runtime, driver)
    version strings "major.minor" of the HIP runtime."""
    rv, dv = ctypes.c_int(), ctypes.c_int()
    if lib().imfcs_get_versions(ctypes.byref(rv), ctypes.byref(dv)) != 0:
        raise ImFCSError(get_last_error())
    fmt = lambda v: "%d.%d" % (v // 1000, (v % 1000) // 10)
    return fmt(rv.value), fmt(dv.value)


def _check(rc):
    if rc != 0:
        raise ImFCSError(get_last_error())


def reset_gpu():
    """GpufitImFCS.resetGPU (GpufitImFCS.java:491)."""
    _check(lib().imfcs_reset_gpu())


class FitModel:
    """GpufitImFCS.FitModel (GpufitImFCS.java:616-733) with numpy buffers."""

    def __init__(self, number_fits, number_points, with_weights, model_id, tolerance=1e-4,
                 max_number_iterations=25, num_valid_coefs=0, parameters_to_fit=None,
                 estimator=ESTIMATOR_LSE, user_info_size=0):
        npar = {MODEL_ACF_1D: 20, MODEL_LINEAR_1D: 11, 1: 5}[model_id]
        self.number_fits, self.number_points, self.model_id = number_fits, number_points, model_id
        self.number_parameters = npar
        self.data = np.zeros(number_fits * number_points, np.float32)
        self.weights = np.zeros(number_fits * number_points, np.float32) if with_weights else None
        self.initial_parameters = np.zeros(number_fits * npar, np.float32)
        self.tolerance = np.float32(tolerance)
        self.max_number_iterations = int(max_number_iterations)
        self.num_valid_coefs = int(num_valid_coefs)
        self.parameters_to_fit = (np.ones(npar, np.int32) if parameters_to_fit is None
                                  else np.asarray([1 if p else 0 for p in parameters_to_fit], np.int32))
        self.estimator = estimator
        self.user_info = np.zeros(max(0, user_info_size), np.float32)


class FitResult:
    """GpufitImFCS.FitResult (GpufitImFCS.java:745-805)."""

    def __init__(self, number_fits, number_parameters):
        self.parameters = np.zeros(number_fits * number_parameters, np.float32)
        self.states = np.zeros(number_fits, np.int32)
        self.chi_squares = np.zeros(number_fits, np.float32)
        self.number_iterations = np.zeros(number_fits, np.int32)


def _fp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _ip(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def fit(fit_model, fit_result=None):
    """GpufitImFCS.fit (GpufitImFCS.java:340-366): runs the batched LM and
    raises on a non-zero status with the native error message.  The user
    info size passed to the native side is the buffer capacity, as in the
    reference (bytes for the native routine)."""
    m = fit_model
    r = fit_result or FitResult(m.number_fits, m.number_parameters)
    ui = np.ascontiguousarray(m.user_info, np.float32)
    status = lib().imfcs_fit(m.number_fits, m.number_points, _fp(m.data), _fp(m.weights), m.model_id,
                             _fp(m.initial_parameters), ctypes.c_float(m.tolerance), m.max_number_iterations,
                             m.num_valid_coefs, _ip(m.parameters_to_fit), m.estimator, ctypes.c_size_t(ui.size),
                             ui.ctypes.data_as(ctypes.c_void_p), _fp(r.parameters), _ip(r.states),
                             _fp(r.chi_squares), _ip(r.number_iterations))
    if status != 0:
        raise ImFCSError("status = %d, message = %s" % (status, get_last_error()))
    return r


def calc_data_bleach_correction(pixels, params):
    """calcDataBleachCorrection (GpufitImFCS.java:501)."""
    p = params
    out = np.zeros(p.w_temp * p.h_temp * p.nopit, np.float32)
    px = np.ascontiguousarray(pixels, np.float32)
    _check(lib().imfcs_calc_data_bleach_correction(_fp(px), _fp(out), ctypes.byref(p)))
    return out


def is_binning_memory_sufficient(params):
    rc = lib().imfcs_is_binning_memory_sufficient(ctypes.byref(params))
    if rc < 0:
        raise ImFCSError(get_last_error())
    return rc == 1


def calc_binning(indata, params):
    """calcBinning (GpufitImFCS.java:520)."""
    p = params
    T = p.lastframe - p.firstframe + 1
    out = np.zeros(p.w_temp * p.h_temp * T, np.float32)
    src = np.ascontiguousarray(indata, np.float32)
    _check(lib().imfcs_calc_binning(_fp(src), _fp(out), ctypes.byref(p)))
    return out


def is_acf_memory_sufficient(params):
    rc = lib().imfcs_is_acf_memory_sufficient(ctypes.byref(params))
    if rc < 0:
        raise ImFCSError(get_last_error())
    return rc == 1


def calc_acf(pixels, bleachcorr_params, samp, lag, params):
    """calcACF (GpufitImFCS.java:533-536).  Returns pixels1, blockvararray,
    blocked1D as plane-major float64 arrays of width*height*chanum."""
    p = params
    n = p.width * p.height * p.chanum
    pixels1, bvar, blocked = np.zeros(n), np.zeros(n), np.zeros(n)
    nbm, nbc = np.zeros(p.width * p.height), np.zeros(p.width * p.height)
    px = np.ascontiguousarray(pixels, np.float32)
    bc = np.ascontiguousarray(bleachcorr_params, np.float64)
    sp = np.ascontiguousarray(samp, np.float64)
    lg = np.ascontiguousarray(lag, np.int32)
    _check(lib().imfcs_calc_acf(_fp(px), _dp(pixels1), _dp(bvar), _dp(nbm), _dp(nbc), _dp(blocked), _dp(bc),
                                _dp(sp), _ip(lg), ctypes.byref(p)))
    return pixels1, bvar, blocked


def correlator_tables(p, q, num, frametime):
    """lag, samp, mtab, lagtime of setParameters (IFCS:7391-7424)."""
    base, hbase = p, p // 2
    chanum = p + (q - 1) * p // 2 + 1
    lag = np.zeros(chanum, np.int32)
    samp = np.zeros(chanum)
    for x in range(hbase + 1):
        lag[x] = x
    for x in range(1, q + 1):
        for y in range(1, hbase + 1):
            lag[x * hbase + y] = int(2 ** (x - 1) * y + (base // 4) * 2 ** x)
    samp[:base + 1] = 1
    for x in range(2, q + 1):
        for y in range(1, hbase + 1):
            samp[x * hbase + y] = 2 ** (x - 1)
    mtab = np.floor((num - lag) / samp)
    return lag, samp, mtab, lag * frametime


def obsvol(ax, ay, s):
    """obsvolFCS_ST2D1p (IFCS:13670-13702) in SI units."""
    from math import erf, exp, sqrt, pi
    pi_ = 3.14159265359
    sq = sqrt(pi_)
    px = 2 * exp(-(ax / s) ** 2) - 2
    ex = 2 * ax * erf(ax / s)
    py = 2 * exp(-(ay / s) ** 2) - 2
    ey = 2 * ay * erf(ay / s)
    return 4 * (ax * ay) ** 2 / ((s / sq * px + ex) * (s / sq * py + ey))


def make_fit_model(ax, ay, s, sz=None, rx=0.0, ry=0.0, q2=1.0, q3=1.0, paramfit=None, paraminit=None):
    """FitModelGeom in SI units with the fit panel defaults (IFCS:1959-1968):
    N=1, D=1 um^2/s, G=0; N, D, G free."""
    m = FitModelGeom()
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


def make_pipeline_config(W, H, T, binning=1, cf=(0, 0), background=0, poly_order=4, frametime=1e-3, p=16, q=8,
                         roi=None, do_acf=True, do_ccf=None, do_fit=True, acf_model=None, ccf_model=None,
                         overlap=False, firstframe=1, lastframe=None, fit_max_iterations=2000,
                         fit_tolerance=1e-16, fit_chi2_fp64=True, gls=False, sliding_window=0, lin_segment=0,
                         exp_bleach=0):
    """imfcs_config of one pipeline run.  fit_chi2_fp64 selects FP64 step
    acceptance and convergence on chi^2 (the optimum the Java CPU path
    reaches); False keeps Gpufit's FP32 chi^2 of the reference GPU fit.
    sliding_window > 0 selects the sliding-window bleach mode with windows of
    that many frames, lin_segment > 0 the linear-segment correction with
    T // lin_segment segments, exp_bleach=True the single-exponential
    correction (exp_bleach=2: the double exponential; poly_order must then
    be 0)."""
    c = PipelineConfig()
    c.width, c.height, c.frames = W, H, T
    c.firstframe = firstframe
    c.lastframe = T if lastframe is None else lastframe
    c.binning_x = c.binning_y = binning
    c.pixbin_x = c.pixbin_y = 1 if overlap else binning
    c.cf_x, c.cf_y = cf
    c.background = background
    c.correlator_p, c.correlator_q = p, q
    c.frametime = frametime
    c.poly_order = poly_order
    rx, ry, rw, rh = roi if roi is not None else (0, 0, W - cf[0], H - cf[1])
    c.roi_x, c.roi_y, c.roi_w, c.roi_h = rx, ry, rw, rh
    c.do_acf = int(do_acf)
    c.do_ccf = int((cf != (0, 0)) if do_ccf is None else do_ccf)
    c.do_fit = int(do_fit)
    c.fit_max_iterations = fit_max_iterations
    c.fit_tolerance = fit_tolerance
    c.fit_chi2_fp64 = int(bool(fit_chi2_fp64))
    c.gls = int(bool(gls))
    c.bleach_mode = 1 if sliding_window else (2 if lin_segment else ((4 if exp_bleach == 2 else 3) if exp_bleach
                                                                     else 0))
    c.sliding_window_length = int(sliding_window or lin_segment)
    if acf_model is not None:
        c.acf_model = acf_model
    if ccf_model is not None:
        c.ccf_model = ccf_model
    elif acf_model is not None:
        c.ccf_model = acf_model
    return c


class Engine:
    """A device context for the native pipeline (imfcs_create / imfcs_run)."""

    def __init__(self, device=0, _ctx=None):
        self._lib = lib()
        # _ctx: a context owned elsewhere (a batch entry), not destroyed here
        self._owned = _ctx is None
        self.ctx = self._lib.imfcs_create(device) if _ctx is None else _ctx
        if not self.ctx:
            raise ImFCSError(get_last_error())
        self._stack = None
        self._stack_bytes = 0
        self._chanum_last = 0
        self._pinned = []

    def close(self):
        if self.ctx:
            if self._stack:
                self._lib.imfcs_device_free(self.ctx, self._stack)
                self._stack = None
            for p in self._pinned:
                self._lib.imfcs_host_free(self.ctx, ctypes.c_void_p(p))
            self._pinned = []
            if self._owned:
                self._lib.imfcs_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def alloc_stack(self, nbytes):
        """Device memory of nbytes for the engine's stack (reused when large
        enough); returns its pointer."""
        if self._stack is None or self._stack_bytes < nbytes:
            if self._stack:
                self._lib.imfcs_device_free(self.ctx, self._stack)
                self._stack = None
            self._stack = self._lib.imfcs_device_alloc(self.ctx, ctypes.c_size_t(nbytes))
            if not self._stack:
                raise ImFCSError(get_last_error())
            self._stack_bytes = nbytes
        return self._stack

    def device_buffer(self, nbytes):
        """Device memory owned by the caller (free with free_buffer)."""
        p = self._lib.imfcs_device_alloc(self.ctx, ctypes.c_size_t(max(int(nbytes), 1)))
        if not p:
            raise ImFCSError(get_last_error())
        return p

    def free_buffer(self, d_ptr):
        _check(self._lib.imfcs_device_free(self.ctx, d_ptr))

    def pinned_array(self, shape, dtype=np.uint16):
        """A numpy array in pinned host memory (imfcs_host_alloc); the memory
        is freed with free_pinned(array) or when the engine closes."""
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        p = self._lib.imfcs_host_alloc(self.ctx, ctypes.c_size_t(nbytes))
        if not p:
            raise MemoryError(get_last_error())
        self._pinned.append(p)
        buf = (ctypes.c_char * nbytes).from_address(p)
        return np.frombuffer(buf, dtype=dtype).reshape(shape)

    def free_pinned(self, arr):
        p = arr.ctypes.data
        if p in self._pinned:
            self._pinned.remove(p)
            _check(self._lib.imfcs_host_free(self.ctx, ctypes.c_void_p(p)))

    def copy_in(self, d_ptr, arr):
        a = np.ascontiguousarray(arr)
        _check(self._lib.imfcs_memcpy_h2d(self.ctx, d_ptr, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes)))

    def set_fit_seeder(self, acf_net=None, ccf_net=None):
        """Start the LM fits of later runs from the networks' predictions
        (learned.InitialGuessNet, or None for the plain start)."""
        self._seeders = (acf_net, ccf_net)
        _check(self._lib.imfcs_set_fit_seeder(self.ctx, acf_net.h if acf_net else None,
                                              ccf_net.h if ccf_net else None))

    def download(self, d_ptr, shape, dtype=np.uint16):
        """Copies device memory at d_ptr into a new host array."""
        out = np.empty(shape, dtype)
        _check(self._lib.imfcs_memcpy_d2h(self.ctx, out.ctypes.data_as(ctypes.c_void_p), d_ptr,
                                          ctypes.c_size_t(out.nbytes)))
        return out

    def upload(self, stack):
        """Copies a uint16 [T][H][W] stack to device memory owned by the
        engine and returns its device pointer."""
        stack = np.ascontiguousarray(stack, np.uint16)
        self.alloc_stack(stack.nbytes)
        _check(self._lib.imfcs_memcpy_h2d(self.ctx, self._stack, stack.ctypes.data_as(ctypes.c_void_p),
                                          ctypes.c_size_t(stack.nbytes)))
        return self._stack

    def run(self, cfg, d_stack=None):
        _check(self._lib.imfcs_run(self.ctx, d_stack or self._stack, ctypes.byref(cfg)))
        self._chanum_last = cfg.correlator_p + (cfg.correlator_q - 1) * cfg.correlator_p // 2 + 1

    def synchronize(self):
        _check(self._lib.imfcs_synchronize(self.ctx))

    def refit(self, cfg):
        """imfcs_refit: the LM stage alone on the last run's correlations,
        with cfg's fit models (cfg must match the last run elsewhere)."""
        _check(self._lib.imfcs_refit(self.ctx, ctypes.byref(cfg)))

    @staticmethod
    def run_multi(engines, cfgs, d_stack):
        """imfcs_run_multi: cfgs[i] on engines[i] over one device stack (the
        binnings of a diffusion law); with the polynomial correction the
        bleach traces of all configurations come from one pass over the
        stack.  Asynchronous, as run()."""
        n = len(engines)
        if len(cfgs) != n or n == 0:
            raise ValueError("one configuration per engine is needed")
        ctxs = (ctypes.c_void_p * n)(*[e.ctx for e in engines])
        arr = (PipelineConfig * n)(*cfgs)
        _check(lib().imfcs_run_multi(ctxs, n, d_stack, arr))
        for e, c in zip(engines, cfgs):
            e._chanum_last = c.correlator_p + (c.correlator_q - 1) * c.correlator_p // 2 + 1

    def run_flags(self):
        """imfcs_get_run_flags of the last run (bit 0: its bleach pass was
        the nested pass of binnings 1, 2 and 4)."""
        f = ctypes.c_int()
        _check(self._lib.imfcs_get_run_flags(self.ctx, ctypes.byref(f)))
        return f.value

    def stage_ms(self):
        """Device times (ms) of the last run's stages: trace + polynomial fit,
        bleach correction + window means, blocking, correlation, LM fit."""
        ms = (ctypes.c_float * 5)()
        _check(self._lib.imfcs_get_stage_ms(self.ctx, ms))
        return list(ms)

    def gls_covariance(self, pair):
        """Regularized covariance (chanum-1 x chanum-1) of one pair of the
        last run with gls=True."""
        r = self._lib.imfcs_get_gls_covariance
        m = self._chanum_last - 1
        out = np.zeros((m, m))
        _check(r(self.ctx, int(pair), _dp(out)))
        return out

    def correlator_kernel_ms(self):
        """Device time (ms) of the last run's main correlator kernel."""
        ms = ctypes.c_float()
        _check(self._lib.imfcs_get_correlator_kernel_ms(self.ctx, ctypes.byref(ms)))
        return ms.value

    def results(self, cfg):
        return _results(self.ctx, cfg)


def _results(ctx, cfg):
    """imfcs_get_results of context ctx's last run of cfg, as arrays shaped
    (pair kind, out_h, out_w, ...)."""
    ow, oh, chanum = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(lib().imfcs_plan(ctypes.byref(cfg), ctypes.byref(ow), ctypes.byref(oh), ctypes.byref(chanum)))
    ow, oh, nc = ow.value, oh.value, chanum.value
    npairs = ow * oh * (int(cfg.do_acf) + int(cfg.do_ccf))
    nseries = ow * oh * (2 if cfg.do_ccf else 1)
    acf, var = np.zeros((npairs, nc)), np.zeros((npairs, nc))
    ind, bis = np.zeros(npairs, np.int32), np.zeros(npairs, np.int32)
    fitres, chi2 = np.full((npairs, 11), np.nan), np.full(npairs, np.nan)
    states, iters = np.full(npairs, -1, np.int32), np.zeros(npairs, np.int32)
    if cfg.bleach_mode == 2:   # bcInt0 and the segment means
        ncoef = (cfg.lastframe - cfg.firstframe + 1) // cfg.sliding_window_length + 1
    elif cfg.bleach_mode == 3:   # f(0), A, t1, C
        ncoef = 4
    elif cfg.bleach_mode == 4:   # f(0), A, t1, B, t2, C
        ncoef = 6
    else:
        ncoef = max(cfg.poly_order, 0) + 1
    coef = np.zeros((nseries, ncoef))
    _check(lib().imfcs_get_results(ctx, _dp(acf), _dp(var), _ip(ind), _ip(bis), _dp(fitres),
                                       _dp(chi2), _ip(states), _ip(iters), _dp(coef)))
    npair_types = int(cfg.do_acf) + int(cfg.do_ccf)
    sh = (npair_types, oh, ow)
    return dict(acf=acf.reshape(sh + (nc,)), var=var.reshape(sh + (nc,)), ind=ind.reshape(sh),
                blockIndS=bis.reshape(sh), fitres=fitres.reshape(sh + (11,)), chi2=chi2.reshape(sh),
                states=states.reshape(sh), iters=iters.reshape(sh), coef=coef)


def simulate_2d(W, H, T, n_emitters, D_um2s=1.0, pixel_um=0.24, psf_sigma_um=0.19, frametime=1e-3,
                counts_per_frame=20.0, offset=100, tau_bleach_s=15.0, margin=2, seed=1, nthreads=0):
    """Seeded 2D particle simulation -> uint16 [T][H][W] (see csrc/simulate.cpp)."""
    out = np.empty((T, H, W), np.uint16)
    d = np.asarray(D_um2s if np.ndim(D_um2s) else [D_um2s] * 4, np.float64)
    if d.size != 4:
        raise ValueError("D_um2s must be a scalar or four quadrant values")
    rc = sim_lib().imfcs_simulate_2d(out.ctypes.data_as(ctypes.c_void_p), W, H, T, margin, n_emitters, pixel_um,
                                     psf_sigma_um, frametime, counts_per_frame, offset, tau_bleach_s,
                                     d.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint64(seed), nthreads)
    if rc != 0:
        raise ValueError("imfcs_simulate_2d: bad arguments")
    return out


def sim_domain(x0_um, y0_um, w_um, h_um, n_emitters, D_um2s=1.0, D2_um2s=0.0, frac2=0.0, counts=20.0,
               counts_hi=None):
    """One periodic box of the GPU simulator: n_emitters start uniformly in
    [x0, x0 + w) x [y0, y0 + h) (um) and wrap at its edges; D_um2s is one value
    or four quadrant values of component 1; a fraction frac2 of the emitters
    diffuses with D2_um2s; brightness `counts` photons per frame, or a
    log-ramp from `counts` to `counts_hi` along x of the box."""
    d = SimDomain()
    d.x0_um, d.y0_um, d.w_um, d.h_um = x0_um, y0_um, w_um, h_um
    d.n_emitters = int(n_emitters)
    dq = np.asarray(D_um2s if np.ndim(D_um2s) else [D_um2s] * 4, np.float64)
    for q in range(4):
        d.D_um2s[q] = float(dq[q])
    d.D2_um2s, d.frac2 = float(D2_um2s), float(frac2)
    d.counts_lo = float(counts)
    d.counts_hi = float(counts if counts_hi is None else counts_hi)
    return d


def simulate_2d_device(engine, W, H, T, domains, pixel_um=0.24, psf_sigma_um=0.19, frametime=1e-3, offset=100,
                       tau_bleach_s=0.0, seed=1, d_out=None):
    """The particle simulator on the GPU into a uint16 [T][H][W] stack: the
    device buffer d_out, or one owned by `engine` (replacing its current
    stack); returns the device pointer."""
    doms = domains if isinstance(domains, (list, tuple)) else [domains]
    arr = (SimDomain * len(doms))(*doms)
    nbytes = 2 * W * H * T
    d_stack = d_out if d_out else engine.alloc_stack(nbytes)
    _check(lib().imfcs_simulate_2d_device(d_stack, W, H, T, arr, len(doms), pixel_um, psf_sigma_um, frametime,
                                          offset, tau_bleach_s, ctypes.c_uint64(seed), None))
    return d_stack


class Batch:
    """A batch of independent stacks (imfcs_batch_*): one context per stack,
    every stack's pipeline on its own stream and all fits in one persistent
    launch (own_fits=True: each pipeline fits its own pixels at the end of
    its stream, seeded fits allowed); with graph=True later runs of the same
    stacks, configurations and seeders replay a captured hipGraph;
    autotune=True times the frame chunking per stack shape first.  The same
    device stack under several configurations (binnings) shares one pass
    for the bleach traces, as Engine.run_multi."""

    def __init__(self, n, device=0, graph=True, autotune=True, own_fits=False):
        self.n = n
        self.h = lib().imfcs_batch_create(device, n, (1 if graph else 0) | (2 if autotune else 0) |
                                          (4 if own_fits else 0))
        if not self.h:
            raise ImFCSError(get_last_error())
        self._views = {}

    def engine(self, i):
        """Engine view of entry i's context (owned by the batch): results,
        refit, set_fit_seeder."""
        if i not in self._views:
            ctx = lib().imfcs_batch_context(self.h, i)
            if not ctx:
                raise ImFCSError(get_last_error())
            self._views[i] = Engine(_ctx=ctx)
        return self._views[i]

    def run(self, d_stacks, cfgs):
        if len(d_stacks) != self.n or len(cfgs) != self.n:
            raise ValueError("one stack and one configuration per batch entry")
        st = (ctypes.c_void_p * self.n)(*d_stacks)
        arr = (PipelineConfig * self.n)(*cfgs)
        _check(lib().imfcs_batch_run(self.h, st, arr))

    def synchronize(self):
        _check(lib().imfcs_batch_synchronize(self.h))

    def results(self, i, cfg):
        ctx = lib().imfcs_batch_context(self.h, i)
        if not ctx:
            raise ImFCSError(get_last_error())
        return _results(ctx, cfg)

    def slot_count(self, i):
        n = ctypes.c_int()
        _check(lib().imfcs_get_slot_count(lib().imfcs_batch_context(self.h, i), ctypes.byref(n)))
        return n.value

    def run_flags(self, i):
        f = ctypes.c_int()
        _check(lib().imfcs_get_run_flags(lib().imfcs_batch_context(self.h, i), ctypes.byref(f)))
        return f.value

    def info(self):
        g, f, t, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(lib().imfcs_batch_info(self.h, ctypes.byref(g), ctypes.byref(f), ctypes.byref(t), ctypes.byref(ms)))
        return {"graph_nodes": g.value, "fit_groups": f.value, "tuned_shapes": t.value, "autotune_ms": ms.value}

    def close(self):
        if self.h:
            for v in self._views.values():
                v.close()
            self._views = {}
            lib().imfcs_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def get_tuning(nslots, frames):
    """(apply_log2, block_log2) frame chunking recorded for a series shape."""
    a, b = ctypes.c_int(), ctypes.c_int()
    _check(lib().imfcs_get_tuning(nslots, frames, ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value
