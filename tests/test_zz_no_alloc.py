"""imfcs_run_noalloc (the allocation-free variant of imfcs_run) against
imfcs_run + imfcs_get_results on the same stacks: identical results, bit for
bit (the same kernels on the same inputs), for two calls in a row on one
context, the second with a smaller stack; a configuration beyond the scratch
array is refused with IMFCS_ERR_MEMORY and a message."""
import ctypes

import numpy as np
import pytest

import imaging_fcs_1_52_amd as imfcs
from imaging_fcs_1_52_amd import native

pytestmark = pytest.mark.gpu

PIX, PSF = 0.24e-6, 0.38e-6


def _cfg(W, H, T, bg, cf=(1, 0)):
    acf = imfcs.make_fit_model(PIX, PIX, PSF)
    ccf = imfcs.make_fit_model(PIX, PIX, PSF, rx=PIX * cf[0])
    return imfcs.make_pipeline_config(W, H, T, cf=cf, background=bg, poly_order=4, acf_model=acf, ccf_model=ccf)


def _noalloc(eng, d_stack, cfg, npairs, nc):
    L = native.lib()
    shapes = {"acf": (npairs, nc, np.float64), "var": (npairs, nc, np.float64), "ind": (npairs, 1, np.int32),
              "blocked": (npairs, 1, np.int32), "fitres": (npairs, 11, np.float64), "chi2": (npairs, 1, np.float64),
              "states": (npairs, 1, np.int32), "iters": (npairs, 1, np.int32)}
    bufs = {k: eng.device_buffer(a * b * np.dtype(t).itemsize) for k, (a, b, t) in shapes.items()}
    try:
        rc = L.imfcs_run_noalloc(eng.ctx, d_stack, ctypes.byref(cfg), *[bufs[k] for k in shapes])
        assert rc == 0, L.imfcs_get_last_error()
        eng.synchronize()
        return {k: eng.download(bufs[k], (a, b), t).reshape(a, b) if b > 1 else eng.download(bufs[k], (a,), t)
                for k, (a, b, t) in shapes.items()}
    finally:
        for p in bufs.values():
            eng.free_buffer(p)


def test_noalloc_matches_run_on_two_calls_in_a_row():
    ref_eng, eng = imfcs.Engine(0), imfcs.Engine(0)
    stacks = [imfcs.simulate_2d(16, 16, 8192, 60, seed=71), imfcs.simulate_2d(8, 8, 4096, 20, seed=72)]
    d_stacks = []
    for st in stacks:
        T, H, W = st.shape
        d = eng.device_buffer(st.nbytes)
        eng.copy_in(d, st)
        d_stacks.append(d)
    try:
        for st, d in zip(stacks, d_stacks):
            T, H, W = st.shape
            cfg = _cfg(W, H, T, int(st.min()))
            ref_eng.upload(st)
            ref_eng.run(cfg)
            r = ref_eng.results(cfg)
            npairs = r["acf"].shape[0] * r["acf"].shape[1] * r["acf"].shape[2]
            g = _noalloc(eng, d, cfg, npairs, r["acf"].shape[-1])
            np.testing.assert_array_equal(g["acf"], r["acf"].reshape(npairs, -1))
            np.testing.assert_array_equal(g["var"], r["var"].reshape(npairs, -1))
            np.testing.assert_array_equal(g["ind"], r["ind"].reshape(-1))
            np.testing.assert_array_equal(g["blocked"], r["blockIndS"].reshape(-1))
            np.testing.assert_array_equal(g["fitres"], r["fitres"].reshape(npairs, -1))
            np.testing.assert_array_equal(g["chi2"], r["chi2"].reshape(-1))
            np.testing.assert_array_equal(g["states"], r["states"].reshape(-1))
            np.testing.assert_array_equal(g["iters"], r["iters"].reshape(-1))
            assert np.isfinite(g["acf"]).any()
        # results of a no-alloc run stay in the caller's buffers
        with pytest.raises(imfcs.ImFCSError, match="noalloc"):
            eng.results(cfg)
        # beyond the scratch array (64 x 64 px x 20,000 frames: 328 MB of
        # corrected series): refused with a message; the stack given is a real
        # one of that size, so the stages enqueued before the refusal read
        # within it
        bst = imfcs.simulate_2d(64, 64, 20000, 200, seed=73)
        d_big = eng.device_buffer(bst.nbytes)
        d_stacks.append(d_big)
        eng.copy_in(d_big, bst)
        big = _cfg(64, 64, 20000, int(bst.min()))
        outs = [eng.device_buffer(8 * 64 * 64 * 2 * 73) for _ in range(8)]
        d_stacks.extend(outs)
        L = native.lib()
        rc = L.imfcs_run_noalloc(eng.ctx, d_big, ctypes.byref(big), *outs)
        assert rc == -4
        assert b"MiB of scratch" in L.imfcs_get_last_error()
        # a normal run on the same context afterwards still works
        st = stacks[1]
        T, H, W = st.shape
        cfg = _cfg(W, H, T, int(st.min()))
        eng.upload(st)
        eng.run(cfg)
        r2 = eng.results(cfg)
        ref_eng.upload(st)
        ref_eng.run(cfg)
        np.testing.assert_array_equal(r2["acf"], ref_eng.results(cfg)["acf"])
    finally:
        eng.synchronize()
        for d in d_stacks:
            eng.free_buffer(d)
        eng.close()
        ref_eng.close()
