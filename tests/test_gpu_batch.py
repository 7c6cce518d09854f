"""Batching: 48 independent seeded stacks of uneven size -- sides uniform in
16-256 px, 10,000-100,000 frames, binning 1-5, D uniform in 0.2-10 um^2/s
(GPU simulator, 0.115 emitters per pixel, 20 photons per frame, offset 100,
bleaching tau = 200 s) -- each with order-4 correction, ACF + CCF to the
x + b neighbour and LM fits of N, D, G_inf.

The batch (imfcs_batch_*) autotunes the frame chunking per stack shape,
runs every stack's pipeline on its own stream with one persistent fit
launch over all stacks' fits, captures that into a hipGraph and replays it.
Checked: the replayed batch results equal, bit for bit, imfcs_run of every
stack alone (same kernels, same tuning); the per-stack path is itself
checked against the oracle here on crops of four stacks (ACF 1e-4 relative,
floor 1e-4 of the amplitude; fitted D 1e-3 where both converged).
Reported: graph nodes, fit groups, tuned shapes, and the wall time of a
replay against per-stack launches.
"""
import time

import numpy as np
import pytest

import imaging_fcs_1_52_amd as imfcs
import oracle
from parity import PSF, rel_err

pytestmark = pytest.mark.gpu

N = 48


def _draw():
    rng = np.random.default_rng(3030)
    side = rng.integers(16, 257, N)
    T = rng.integers(10000, 100001, N)
    b = rng.integers(1, 6, N)
    D = rng.uniform(0.2, 10.0, N)
    return side, T, b, D


def _cfg(side, T, b):
    pix = 0.24e-6 * b
    acf = imfcs.make_fit_model(pix, pix, PSF)
    ccf = imfcs.make_fit_model(pix, pix, PSF, rx=pix)
    return imfcs.make_pipeline_config(int(side), int(side), int(T), binning=int(b), cf=(int(b), 0), background=100,
                                      poly_order=4, acf_model=acf, ccf_model=ccf)


def test_48_uneven_stacks_batch_graph_equals_per_stack_runs(engine):
    side, T, b, D = _draw()
    eng = imfcs.Engine(0)
    stacks = []
    try:
        t0 = time.time()
        for i in range(N):
            s_, t_ = int(side[i]), int(T[i])
            n_emit = int(round(0.115 * (s_ + 4) ** 2))
            dom = imfcs.sim_domain(-2 * 0.24, -2 * 0.24, (s_ + 4) * 0.24, (s_ + 4) * 0.24, n_emit, D_um2s=float(D[i]),
                                   counts=20.0)
            d = eng.device_buffer(2 * s_ * s_ * t_)
            stacks.append(d)
            imfcs.simulate_2d_device(eng, s_, s_, t_, dom, offset=100, tau_bleach_s=200.0, seed=9000 + i, d_out=d)
        print("simulated 48 stacks (%.1f GB) in %.1f s" % ((2.0 * side.astype(float) ** 2 * T).sum() / 1e9,
                                                           time.time() - t0), flush=True)
        cfgs = [_cfg(side[i], T[i], b[i]) for i in range(N)]
        batch = imfcs.Batch(N, graph=True, autotune=True)
        t0 = time.time()
        batch.run(stacks, cfgs)           # autotune, eager run, capture
        batch.synchronize()
        t_first = time.time() - t0
        info = batch.info()
        print("first run (autotune + eager + capture) %.2f s: %r" % (t_first, info), flush=True)
        assert info["fit_groups"] == N and info["graph_nodes"] > 10 * N
        reps = []
        for _ in range(3):
            t0 = time.perf_counter()
            batch.run(stacks, cfgs)       # graph replay
            batch.synchronize()
            reps.append(time.perf_counter() - t0)
        res = [batch.results(i, cfgs[i]) for i in range(N)]
        nslots = [batch.slot_count(i) for i in range(N)]
        batch.close()
        tuned = {imfcs.get_tuning(nslots[i], int(T[i])) for i in range(N)}
        print("tunings used: %s" % sorted(tuned), flush=True)
        # per-stack launches, one after the other on one engine
        t0 = time.perf_counter()
        for i in range(N):
            engine.run(cfgs[i], stacks[i])
        engine.synchronize()
        t_single = time.perf_counter() - t0
        print("graph replay %.1f ms (median of 3); per-stack launches %.1f ms" % (1e3 * np.median(reps),
                                                                                 1e3 * t_single), flush=True)
        nfit = 0
        for i in range(N):
            engine.run(cfgs[i], stacks[i])
            r = engine.results(cfgs[i])
            for k in ("acf", "var", "ind", "blockIndS", "fitres", "chi2", "states", "iters"):
                assert np.array_equal(res[i][k], r[k], equal_nan=True), (i, k)
            nfit += int((r["states"] == 0).sum())
        print("bit-identical results for 48 stacks, %d converged fits" % nfit, flush=True)
        # the per-stack path against the oracle on crops of four stacks
        checked = 0
        for i in (0, 13, 29, 41):
            s_, t_, bb = int(side[i]), int(T[i]), int(b[i])
            engine.run(cfgs[i], stacks[i])
            r = engine.results(cfgs[i])
            full = eng.download(stacks[i], (t_, s_, s_))
            nx = min(3, r["acf"].shape[2])
            crop = np.ascontiguousarray(full[:, 0:3 * bb, 0:(nx + 1) * bb])
            del full
            om = oracle.make_model(0.24e-6 * bb, 0.24e-6 * bb, PSF)
            o = oracle.pipeline(crop, oracle.make_config((nx + 1) * bb, 3 * bb, t_, binning=bb, background=100,
                                                        poly_order=4, roi=(0, 0, nx * bb, 3 * bb), model=om,
                                                        nthreads=8))
            ga = r["acf"][0][0:3, 0:nx]
            fin = np.isfinite(o["acf"]).all(-1)
            assert np.array_equal(np.isfinite(ga).all(-1), fin)
            if fin.any():
                amp = np.abs(o["acf"][fin][:, 1:]).max(-1, keepdims=True)
                assert rel_err(ga[fin], o["acf"][fin], 1e-4 * amp).max() < 1e-4, i
                same = (r["ind"][0][0:3, 0:nx] == o["ind"]) & fin
                ok = same & (r["states"][0][0:3, 0:nx] == 0) & o["fitted"]
                if ok.any():
                    gd = r["fitres"][0][0:3, 0:nx][..., 1][ok]
                    assert rel_err(gd, o["fitres"][..., 1][ok], 1e-300).max() < 1e-3, i
                checked += int(fin.sum())
        assert checked >= 12, checked
    finally:
        eng.synchronize()
        for d in stacks:
            eng.free_buffer(d)
        eng.close()


def test_shared_stack_graph_with_own_and_seeded_fits_equals_run_multi(engine):
    """The binnings of a diffusion law as one batch over one device stack
    (own_fits: each pipeline fits at the end of its stream): the captured
    graph's replays give, bit for bit, what imfcs_run_multi gives, with the
    plain and with the seeded LM start (the seeders are part of what a
    replay is keyed on)."""
    from imaging_fcs_1_52_amd.learned import InitialGuessNet
    W = H = 64
    T = 20000
    stack = imfcs.simulate_2d(W, H, T, 400, D_um2s=[0.5, 1.0, 2.0, 5.0], tau_bleach_s=20.0, seed=4242)
    bins = (1, 2)
    cfgs = [_cfg(W, T, b) for b in bins]
    batch = imfcs.Batch(len(bins), graph=True, autotune=False, own_fits=True)
    ref = [imfcs.Engine(0) for _ in bins]
    try:
        views = [batch.engine(i) for i in range(len(bins))]
        d = views[0].upload(stack)
        nets = []
        for mode in ("plain", "seeded", "plain"):
            if mode == "seeded" and not nets:
                for b, cfg in zip(bins, cfgs):
                    _, _, mtab, lagtime = imfcs.correlator_tables(16, 8, T, 1e-3)
                    pix = 0.24e-6 * b
                    pair = [InitialGuessNet(views[0], imfcs.make_fit_model(pix, pix, PSF, rx=rx), lagtime[1:],
                                            mtab[1:], ncomp=1, seed=60 + 2 * b + k) for k, rx in enumerate((0, pix))]
                    for n in pair:
                        n.train(steps=60, batch=2048)
                    nets.append(pair)
            for i in range(len(bins)):
                seed = nets[i] if mode == "seeded" else (None, None)
                views[i].set_fit_seeder(*seed)
                ref[i].set_fit_seeder(*seed)
            for rep in range(3):           # eager + capture, then two replays
                batch.run([d] * len(bins), cfgs)
                batch.synchronize()
            info = batch.info()
            imfcs.Engine.run_multi(ref, cfgs, d)
            for i, cfg in enumerate(cfgs):
                a, r = batch.results(i, cfg), ref[i].results(cfg)
                for k in ("acf", "var", "ind", "blockIndS", "fitres", "chi2", "states", "iters", "coef"):
                    assert np.array_equal(a[k], r[k], equal_nan=True), (mode, bins[i], k)
            print("%s: graph of %d nodes, results equal imfcs_run_multi" % (mode, info["graph_nodes"]), flush=True)
        for pair in nets:
            for n in pair:
                n.close()
    finally:
        for e in ref:
            e.set_fit_seeder(None, None)
        batch.close()
        for e in ref:
            e.close()
