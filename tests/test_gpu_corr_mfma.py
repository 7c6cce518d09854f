"""GPU parity of the matrix-core correlator of levels 0..2 (k_corr_mfma,
selected with IMFCS_CORRELATOR=mfma when the context is created) against the
oracle and against the VALU correlator on the same stacks.

Cases: ACF with a CCF partner one or two output pixels to the right (partner
slot offsets 1 and 2), ACF alone, frame counts that are not multiples of 4
(the level-1 and level-2 products past the last complete value of those
levels are removed at the series end), blocks shorter than a 512-frame span
(short series), and the GLS covariance, whose block table the kernel writes
at every block end.  Tolerances as the VALU path's parity tests: correlation
1e-4 relative (channels within 1e-3 of the amplitude of zero: 1e-7 of the
amplitude absolute), blocked variance 1e-4 where the blocking index agrees,
fitted N, D 1e-3.  The FP16 hi/lo split keeps each staged value to 2^-22
relative, so the two correlators agree to about 1e-6 of the amplitude; the
results are not bit-identical (which shows that the matrix-core path ran).
"""
import os

import numpy as np
import pytest

import imaging_fcs_1_52_amd as imfcs
from parity import models, rel_err, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mfma_engine():
    old = os.environ.get("IMFCS_CORRELATOR")
    os.environ["IMFCS_CORRELATOR"] = "mfma"
    try:
        e = imfcs.Engine(0)
    finally:
        if old is None:
            del os.environ["IMFCS_CORRELATOR"]
        else:
            os.environ["IMFCS_CORRELATOR"] = old
    yield e
    e.close()


def _check(g_acf, g_var, g_ind, o):
    fin = np.isfinite(o["acf"]).all(-1)
    assert np.array_equal(np.isfinite(g_acf).all(-1), fin), "NaN pattern differs"
    amp = np.abs(o["acf"][fin][:, 1:]).max(-1, keepdims=True)
    assert rel_err(g_acf[fin], o["acf"][fin], 1e-3 * amp).max() < 1e-4
    near0 = np.abs(o["acf"][fin]) < 1e-3 * amp
    assert (np.abs(g_acf[fin] - o["acf"][fin])[near0] <= 1e-7 * np.broadcast_to(amp, near0.shape)[near0]).all()
    same = (g_ind == o["ind"]) & fin
    assert same[fin].mean() >= 0.9
    assert rel_err(g_var[same][:, 1:], o["var"][same][:, 1:], 1e-300).max() < 1e-4
    return fin


CASES = [
    # W, H, T, binning, cf, poly order, seed
    (20, 12, 5000, 1, (1, 0), 4, 201),    # partner slot offset 1
    (14, 10, 7777, 2, (4, 0), 3, 202),    # offset 2; T = 1 mod 4 (level-1 and level-2 ends)
    (20, 15, 3335, 5, (0, 0), 4, 203),    # ACF alone, binning 5; T = 3 mod 4
    (17, 9, 3002, 1, (1, 0), 2, 204),     # short series: blocks shorter than a span
]


@pytest.mark.parametrize("W,H,T,b,cf,order,seed", CASES)
def test_mfma_correlator_matches_oracle_and_valu(engine, mfma_engine, W, H, T, b, cf, order, seed):
    print(f"case W={W} H={H} T={T} b={b} cf={cf} order={order}", flush=True)
    stack = imfcs.simulate_2d(W, H, T, W * H // 2, D_um2s=1.0, tau_bleach_s=15.0, seed=seed)
    gm, _ = run_gpu(mfma_engine, stack, binning=b, cf=cf, poly_order=order)
    gv, _ = run_gpu(engine, stack, binning=b, cf=cf, poly_order=order)
    o = run_oracle(stack, binning=b, cf=(0, 0), poly_order=order, roi_cf=cf)
    fin = _check(gm["acf"][0], gm["var"][0], gm["ind"][0], o)
    ok = (gm["states"][0] == 0) & o["fitted"] & (gm["ind"][0] == o["ind"])
    assert ok.any()
    assert rel_err(gm["fitres"][0][..., 0][ok], o["fitres"][..., 0][ok], 1e-12).max() < 1e-3
    assert rel_err(gm["fitres"][0][..., 1][ok], o["fitres"][..., 1][ok], 1e-30).max() < 1e-3
    if cf != (0, 0):
        oc = run_oracle(stack, binning=b, cf=cf, poly_order=order)
        _check(gm["acf"][1], gm["var"][1], gm["ind"][1], oc)
    # matrix-core vs VALU correlator: same blocking, correlations within
    # 1e-6 of the amplitude, and not the same bits
    kinds = 2 if cf != (0, 0) else 1
    for k in range(kinds):
        f = np.isfinite(gv["acf"][k]).all(-1)
        assert np.array_equal(gm["ind"][k], gv["ind"][k])
        amp = np.abs(gv["acf"][k][f][:, 1:]).max(-1, keepdims=True)
        assert (np.abs(gm["acf"][k][f] - gv["acf"][k][f]) <= 1e-6 * amp).all()
    assert not np.array_equal(gm["acf"][0][fin], gv["acf"][0][fin])


def test_mfma_gls_block_table_matches_valu(engine, mfma_engine):
    """The GLS covariance comes from the block sums the correlator writes at
    every block end: the two correlators' covariances and GLS data agree."""
    stack = imfcs.simulate_2d(8, 6, 16384, 60, D_um2s=1.0, tau_bleach_s=15.0, seed=205)
    T, H, W = stack.shape
    acf_m, _, _, _ = models(1, (0, 0))
    cfg = imfcs.make_pipeline_config(W, H, T, binning=1, cf=(0, 0), background=int(stack.min()), poly_order=4,
                                     do_fit=True, acf_model=acf_m, gls=True)
    res = []
    for e in (mfma_engine, engine):
        e.upload(stack)
        e.run(cfg)
        r = e.results(cfg)
        r["cov"] = np.stack([e.gls_covariance(p) for p in range(W * H)])
        res.append(r)
    gm, gv = res
    fin = np.isfinite(gv["acf"][0]).all(-1)
    assert np.array_equal(np.isfinite(gm["acf"][0]).all(-1), fin)
    amp = np.abs(gv["acf"][0][fin][:, 1:]).max(-1, keepdims=True)
    assert rel_err(gm["acf"][0][fin], gv["acf"][0][fin], 1e-4 * amp).max() < 1e-4
    cf_ = fin.reshape(-1)
    cm, cv = gm["cov"][cf_], gv["cov"][cf_]
    scale = np.abs(cv).max(axis=(1, 2), keepdims=True)
    assert (np.abs(cm - cv) <= 1e-4 * scale).all()
