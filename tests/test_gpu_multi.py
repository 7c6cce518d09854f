"""imfcs_run_multi against separate imfcs_run calls on the same stack: the
bleach traces taken from the shared per-pixel point sums must give
identical results (the integer window sums are exact and the moments are
summed in the same order), for several binnings, partners and background
levels; a set with a non-polynomial configuration falls back to separate
runs."""
import numpy as np
import pytest

import imaging_fcs_1_52_amd as imfcs
from parity import models

pytestmark = pytest.mark.gpu

KEYS = ("acf", "var", "ind", "blockIndS", "fitres", "chi2", "states", "iters", "coef")


def _cfg(W, H, T, b, cf, bg, order):
    acf, ccf, _, _ = models(b, cf)
    return imfcs.make_pipeline_config(W, H, T, binning=b, cf=cf, background=bg, poly_order=order, do_fit=True,
                                      acf_model=acf, ccf_model=ccf)


@pytest.mark.parametrize("T", [6000, 4321])
def test_run_multi_matches_separate_runs(engine, T):
    W, H = 24, 16
    print(f"run_multi T={T}", flush=True)
    stack = imfcs.simulate_2d(W, H, T, 150, D_um2s=1.0, tau_bleach_s=10.0, seed=7 + T)
    d = engine.upload(stack)
    bg = int(stack.min())
    cfgs = [_cfg(W, H, T, 4, (4, 0), bg, 4), _cfg(W, H, T, 2, (0, 2), bg - 3, 4), _cfg(W, H, T, 1, (1, 0), bg, 3),
            _cfg(W, H, T, 3, (0, 0), bg + 1, 4)]
    sep = []
    for c in cfgs:
        engine.run(c, d)
        engine.synchronize()
        sep.append(engine.results(c))
    engines = [imfcs.Engine(0) for _ in cfgs]
    try:
        imfcs.Engine.run_multi(engines, cfgs, d)
        for e, c, r in zip(engines, cfgs, sep):
            e.synchronize()
            m = e.results(c)
            for k in KEYS:
                np.testing.assert_array_equal(np.asarray(m[k]), np.asarray(r[k]), err_msg=k)
        # a configuration without polynomial correction: separate runs
        c0 = _cfg(W, H, T, 2, (2, 0), bg, 0)
        engine.run(c0, d)
        engine.synchronize()
        r0 = engine.results(c0)
        imfcs.Engine.run_multi(engines[:2], [c0, cfgs[0]], d)
        engines[0].synchronize()
        engines[1].synchronize()
        m0 = engines[0].results(c0)
        for k in ("acf", "var", "fitres"):
            np.testing.assert_array_equal(np.asarray(m0[k]), np.asarray(r0[k]), err_msg=k)
    finally:
        for e in engines:
            e.close()


def test_run_multi_frame_range_and_roi(engine):
    W, H, T = 20, 14, 5000
    stack = imfcs.simulate_2d(W, H, T, 120, D_um2s=1.0, tau_bleach_s=10.0, seed=41)
    d = engine.upload(stack)
    bg = int(stack.min())
    cfgs = []
    for b, cf, roi in ((1, (1, 0), (3, 2, 12, 9)), (2, (0, 2), (2, 0, 14, 12)), (3, (0, 0), (1, 1, 18, 12))):
        acf, ccf, _, _ = models(b, cf)
        cfgs.append(imfcs.make_pipeline_config(W, H, T, binning=b, cf=cf, background=bg, poly_order=4,
                                               acf_model=acf, ccf_model=ccf, roi=roi, firstframe=101,
                                               lastframe=T - 50))
    sep = []
    for c in cfgs:
        engine.run(c, d)
        engine.synchronize()
        sep.append(engine.results(c))
    engines = [imfcs.Engine(0) for _ in cfgs]
    try:
        imfcs.Engine.run_multi(engines, cfgs, d)
        for e, c, r in zip(engines, cfgs, sep):
            e.synchronize()
            m = e.results(c)
            for k in KEYS:
                np.testing.assert_array_equal(np.asarray(m[k]), np.asarray(r[k]), err_msg=k)
    finally:
        for e in engines:
            e.close()


@pytest.mark.parametrize("T", [5000, 4321])
def test_nested_binnings_share_one_bleach_pass(engine, T):
    """Binnings 1, 2 and 4 of one stack (full grids from (0, 0), each with
    its CCF partner one output pixel to the right) share one bleach pass in
    imfcs_run_multi; a binning-3 configuration of the same stack runs its
    own.  Every result equals the separate runs' bit for bit (the nested pass
    corrects, sums and stores each series as the binning's own pass does,
    and the window-mean sums are taken per 64-frame sub-chunk whatever the
    chunking); T = 4321 ends inside a sub-chunk."""
    W, H = 64, 32
    print(f"nested T={T}", flush=True)
    stack = imfcs.simulate_2d(W, H, T, 200, D_um2s=1.0, tau_bleach_s=10.0, seed=11 + T)
    d = engine.upload(stack)
    bg = int(stack.min())
    cfgs = [_cfg(W, H, T, 2, (2, 0), bg, 4), _cfg(W, H, T, 4, (4, 0), bg, 4), _cfg(W, H, T, 1, (1, 0), bg, 4),
            _cfg(W, H, T, 3, (0, 0), bg, 4)]
    sep = []
    for c in cfgs:
        engine.run(c, d)
        engine.synchronize()
        assert engine.run_flags() == 0
        sep.append(engine.results(c))
    engines = [imfcs.Engine(0) for _ in cfgs]
    try:
        imfcs.Engine.run_multi(engines, cfgs, d)
        assert [e.run_flags() for e in engines] == [1, 1, 1, 0]
        for e, c, r in zip(engines, cfgs, sep):
            e.synchronize()
            m = e.results(c)
            for k in KEYS:
                np.testing.assert_array_equal(np.asarray(m[k]), np.asarray(r[k]), err_msg=k)
    finally:
        for e in engines:
            e.close()


@pytest.mark.skip(reason="the nested bleach pass is switched off in batches: a batch run of this "
                         "triple ended in a host segmentation fault whose cause is not found yet")
def test_nested_binnings_in_a_batch_graph(engine):
    """A batch's captured graph of the nested triple, replayed, against the
    separate runs bit for bit."""


def test_nested_pass_with_vertical_partners_and_its_fallback(engine):
    """Binnings 1, 2 and 4 with partners one output pixel below (grids
    extended by a row) take the nested pass and match separate runs bit for
    bit; with a different background for binning 4 the set is not nested
    (every run does its own pass) and again matches."""
    W, H, T = 64, 32, 4500
    stack = imfcs.simulate_2d(W, H, T, 200, D_um2s=1.0, tau_bleach_s=10.0, seed=77)
    d = engine.upload(stack)
    bg = int(stack.min())
    for bg4, nested in ((bg, True), (bg - 2, False)):
        cfgs = [_cfg(W, H, T, 2, (0, 2), bg, 4), _cfg(W, H, T, 4, (0, 4), bg4, 4), _cfg(W, H, T, 1, (0, 1), bg, 4)]
        print(f"nested={nested}", flush=True)
        sep = []
        for c in cfgs:
            engine.run(c, d)
            engine.synchronize()
            sep.append(engine.results(c))
        engines = [imfcs.Engine(0) for _ in cfgs]
        try:
            imfcs.Engine.run_multi(engines, cfgs, d)
            assert [e.run_flags() for e in engines] == [int(nested)] * 3
            for e, c, r in zip(engines, cfgs, sep):
                e.synchronize()
                m = e.results(c)
                for k in KEYS:
                    np.testing.assert_array_equal(np.asarray(m[k]), np.asarray(r[k]), err_msg=k)
        finally:
            for e in engines:
                e.close()
