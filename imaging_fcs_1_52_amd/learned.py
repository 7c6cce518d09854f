"""Learned initial guesses for the LM fit (include/imfcs.h section 3).

The plugin starts every pixel's LM fit from the fit panel's values
(Imaging_FCS_1_52.java:1959-1968, packed per pixel by GPU_ACF_Fit,
15090-15121).  An ``InitialGuessNet`` is a BF16 perceptron (72 -> 256 -> 256
-> 5, GELU) for one fit geometry, trained on the GPU by the library's own
MFMA forward/backward and AdamW kernels on FCS_3p model curves of that
geometry; ``Engine.set_fit_seeder`` makes the engine's LM start from its
predictions, with the plain start as the fallback for fits that do not
converge.
"""
import ctypes

import numpy as np

from .native import lib

N_IN, N_OUT = 80, 5


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class InitialGuessNet:
    """A network for the fit geometry ``geom`` (a FitModelGeom in SI units),
    lag times ``lagtime`` and product counts ``mtab`` of channels 1..72,
    ``ncomp`` = 1 (N, D, G_inf) or 2 (also D2, F2)."""

    def __init__(self, engine, geom, lagtime, mtab, ncomp=1, seed=1):
        from . import ImFCSError, get_last_error
        lt = np.ascontiguousarray(lagtime, np.float64)
        mt = np.ascontiguousarray(mtab, np.float64)
        if lt.shape != mt.shape:
            raise ValueError("lagtime and mtab must have one entry per fitted channel")
        self._engine = engine
        self.npts = lt.size
        self.ncomp = ncomp
        self.h = lib().imfcs_mlp_create(engine.ctx, ncomp, ctypes.byref(geom),
                                        lt.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), lt.size,
                                        mt.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.c_uint64(seed))
        if not self.h:
            raise ImFCSError(get_last_error())

    def _check(self, rc):
        from . import ImFCSError, get_last_error
        if rc != 0:
            raise ImFCSError(get_last_error())

    def close(self):
        if self.h:
            lib().imfcs_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def train(self, steps=600, batch=4096, lr=2e-3):
        """AdamW steps on fresh model curves; returns the loss of every step."""
        loss = np.zeros(steps, np.float32)
        self._check(lib().imfcs_mlp_train(self.h, steps, batch, ctypes.c_float(lr), _fp(loss)))
        return loss

    def params(self):
        p = np.zeros(lib().imfcs_mlp_param_count(), np.float32)
        self._check(lib().imfcs_mlp_get_params(self.h, _fp(p)))
        return p

    def target_scale(self):
        """(mu, sd) of the standardised outputs."""
        mu, sd = np.zeros(N_OUT, np.float32), np.zeros(N_OUT, np.float32)
        self._check(lib().imfcs_mlp_target_scale(self.h, _fp(mu), _fp(sd)))
        return mu, sd

    def gradient(self, inputs, targets):
        """One forward + backward pass: (parameter gradient, outputs, loss)."""
        x = np.ascontiguousarray(inputs, np.float32)
        t = np.ascontiguousarray(targets, np.float32)
        rows = x.shape[0]
        assert x.shape == (rows, N_IN) and t.shape == (rows, N_OUT)
        g = np.zeros(lib().imfcs_mlp_param_count(), np.float32)
        y = np.zeros((rows, N_OUT), np.float32)
        loss = np.zeros(1, np.float32)
        self._check(lib().imfcs_mlp_gradient(self.h, _fp(x), _fp(t), rows, _fp(g), _fp(y), _fp(loss)))
        return g, y, float(loss[0])

    def curves(self, rows):
        """The next `rows` training curves: (inputs [rows][80], standardised targets [rows][5])."""
        x = np.zeros((rows, N_IN), np.float32)
        t = np.zeros((rows, N_OUT), np.float32)
        self._check(lib().imfcs_mlp_curves(self.h, rows, _fp(x), _fp(t)))
        return x, t

    def predict(self, curves):
        """Host curves [n][72] -> [n][5]: N, D (um^2/s), D2, F2, G_inf."""
        c = np.ascontiguousarray(curves, np.float32)
        n = c.shape[0]
        e = self._engine
        d_in = e.device_buffer(c.nbytes)
        d_out = e.device_buffer(4 * N_OUT * n)
        try:
            e.copy_in(d_in, c)
            self._check(lib().imfcs_mlp_predict(self.h, d_in, n, d_out))
            return e.download(d_out, (n, N_OUT), np.float32)
        finally:
            e.free_buffer(d_in)
            e.free_buffer(d_out)
