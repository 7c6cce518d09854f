#!/usr/bin/env python3
"""Headline benchmark: fitted pixels/s of the whole imaging-FCS hot path
(bin + polynomial bleach correction + multi-tau ACF/CCF with blocked variance
+ LM fit) on a 128x128x50,000-frame uint16 stack, on one MI355X, with the LM
started from the fit panel's values (plain) and from the learned initial
guesses (seeded).

One step = for each binning b in (1, 2, 4): fused b x b binning, order-4
polynomial bleach correction, ACF and CCF to the x+b neighbour with blocked
variances, and the LM fit of N, D, G_inf of every ACF and CCF pixel.  The
three binnings run as three pipelines on three HIP streams over the one
device-resident stack, enqueued each step by imfcs_run_multi: one pass over
the stack gives every pipeline its bleach trace, and one bleach-correction
pass writes the corrected series of all three binnings (--runner graph: the
three pipelines captured once into a hipGraph by imfcs_batch, each with its
own bleach-correction pass, and replayed each step).  The stack is in HBM
before timing starts.

``value``: ACF pixels fitted to state CONVERGED (finite correlation; the
plugin keeps only those, Imaging_FCS_1_52.java:15249-15256) per second of the
seeded pipeline, the layered method; ``value_plain`` the same for the plain
start, ``all_pixels_per_s`` every ACF pixel processed.  The seeded pipeline
runs the network's inference inside the timed step; its training (six
networks: three binnings x ACF/CCF geometry) happens once before timing and
is reported as ``seeder_training_s``.  ``lm_stage``: the fit stage alone
(imfcs_refit on the same correlations, HIP events), plain vs seeded.

``roofline``: the main correlator kernel at binning 1 against HBM, with the
algorithmic bytes of one pass over the stack, 2 B x W x H x T (one uint16
read), over its HIP-event duration on its stream; ``traffic`` is the
copy-calibrated FETCH_SIZE + WRITE_SIZE per launch from the committed
profile.  ``cpu_baseline``: the oracle (a C restatement of the plugin's Java
CPU path) on one host core over a pixel sample of the binning-1 ACF
workload; ``cpu_baseline_all_cores`` the same sample with OpenMP over pixels
on every host core.  ``host_to_host_per_s``: the converged fits of a step
per second with the uint16 stack's copy from pinned host memory added
(``host_to_host_pageable_per_s``: from pageable memory).
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

W = H = 128
T = 50000
N_EMITTERS = 2000
D_QUADRANTS = [0.5, 1.0, 2.0, 5.0]
BINNINGS = (1, 2, 4)
ENQUEUE_ORDER = (2, 4, 1)
PIX = 0.24e-6
PSF = 0.38e-6
HBM_PEAK_GBS = 8000.0
COPY_BYTES = 2 * W * H * T   # the whole uint16 stack
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_corr_main_128x128x50000_b1_traffic.md")


def read_traffic():
    """HBM bytes per k_corr_main launch (b = 1) measured with rocprofv3
    --pmc FETCH_SIZE / WRITE_SIZE and scaled by a plain copy, from the
    committed profile summary (None when absent)."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_dispatch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def make_stack(seed=2024):
    import imaging_fcs_1_52_amd as imfcs
    t0 = time.time()
    st = imfcs.simulate_2d(W, H, T, N_EMITTERS, D_um2s=D_QUADRANTS, pixel_um=0.24, psf_sigma_um=0.19,
                           frametime=1e-3, counts_per_frame=20.0, offset=100, tau_bleach_s=15.0, margin=2,
                           seed=seed, nthreads=min(16, os.cpu_count() or 1))
    log("simulated %dx%dx%d stack in %.1f s" % (W, H, T, time.time() - t0))
    return st


def models(b):
    import imaging_fcs_1_52_amd as imfcs
    ax = PIX * b
    return imfcs.make_fit_model(ax, ax, PSF), imfcs.make_fit_model(ax, ax, PSF, rx=PIX * b)


def configs(stack_min, p=16):
    import imaging_fcs_1_52_amd as imfcs
    out = []
    for b in BINNINGS:
        acf, ccf = models(b)
        out.append(imfcs.make_pipeline_config(W, H, T, binning=b, cf=(b, 0), background=stack_min, poly_order=4,
                                              p=p, do_acf=True, do_ccf=True, do_fit=True, acf_model=acf,
                                              ccf_model=ccf, fit_tolerance=1e-16, fit_max_iterations=2000))
    return out


def cpu_baseline(stack, stack_min, sample_side, nthreads):
    """Oracle (C restatement of the Java CPU path) on a bounded sample of the
    b=1 ACF workload: sample_side^2 pixels, all 50,000 frames."""
    import oracle
    oracle.build()
    m = oracle.make_model(PIX, PIX, PSF)
    cfg = oracle.make_config(W, H, T, binning=1, cf=(0, 0), background=stack_min, poly_order=4,
                             roi=(0, 0, sample_side, sample_side), do_fit=True, model=m, nthreads=nthreads)
    t0 = time.time()
    r = oracle.pipeline(stack, cfg)
    dt = time.time() - t0
    nfit = int(r["fitted"].sum())
    return nfit / dt, sample_side * sample_side / dt, dt


def solo(imfcs, stack, smin, args):
    """One binning's plain pipeline alone on the GPU (the standalone kernel
    times of a profile): stage times per run, HIP events."""
    b = args.solo_binning
    if b not in BINNINGS:
        raise SystemExit("--solo-binning must be one of %s" % (BINNINGS,))
    cfg = configs(smin, args.correlator_p)[BINNINGS.index(b)]
    e = imfcs.Engine(0)
    d = e.upload(stack)
    rows = []
    for i in range(args.warmup + args.steps):
        e.run(cfg, d)
        e.synchronize()
        if i >= args.warmup:
            rows.append(e.stage_ms())
    med = np.median(np.asarray(rows), 0)
    print(json.dumps({"solo_binning": b, "correlator_p": args.correlator_p, "runs": args.steps,
                      "stage_ms_median": {k: round(float(v), 3) for k, v in
                                          zip(("trace+fit", "correction+means", "blocking", "correlation", "lm"),
                                              med)}}), flush=True)
    e.close()


def fit_stats(states, iters):
    ok = states == 0
    it = iters[ok] if ok.any() else iters
    return {"converged": int(ok.sum()), "failed": int((~ok).sum()),
            "iter_mean": round(float(iters.mean()), 2), "iter_p99": float(np.percentile(iters, 99)),
            "iter_max": int(iters.max()), "iter_mean_converged": round(float(it.mean()), 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=48, help="side of the oracle's pixel sample (0 = skip)")
    ap.add_argument("--train-steps", type=int, default=400, help="AdamW steps of each seeding network")
    ap.add_argument("--runner", choices=("graph", "multi"), default="multi",
                    help="multi: imfcs_run_multi enqueued every step (binnings 1, 2, 4 share one bleach-correction "
                         "pass); graph: the three pipelines captured once into a hipGraph (imfcs_batch with one "
                         "shared stack) and replayed every step")
    ap.add_argument("--autotune", action="store_true", help="graph runner: time the frame chunking first")
    ap.add_argument("--solo-binning", type=int, default=0,
                    help="profiling aid: run only this binning's plain pipeline --steps times (after --warmup), alone "
                         "on the GPU, and print its per-run stage times instead of the benchmark line")
    ap.add_argument("--correlator-p", type=int, default=16,
                    help="with --solo-binning: correlator panel P (16 runs the fast kernels, any other even P the "
                         "table-driven correlator)")
    ap.add_argument("--enqueue-order", default=",".join(str(b) for b in ENQUEUE_ORDER),
                    help="order in which the binnings' pipelines are enqueued (captured), e.g. 1,2,4")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also copy the stack device to device twice (16-byte vector copy, 4-byte buffer-word copy; known "
                         "bytes) for PMC traffic scaling")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine has no multi-GPU path (pixels shard without "
              "communication)", file=sys.stderr)
        sys.exit(2)

    import imaging_fcs_1_52_amd as imfcs
    from imaging_fcs_1_52_amd.learned import InitialGuessNet
    stack = make_stack()
    smin = int(stack.min())
    if args.solo_binning:
        solo(imfcs, stack, smin, args)
        return
    enq = tuple(int(x) for x in args.enqueue_order.split(","))
    if sorted(enq) != sorted(BINNINGS):
        raise SystemExit("--enqueue-order must list the binnings %s" % (BINNINGS,))
    order = [BINNINGS.index(b) for b in enq]
    batch = None
    if args.runner == "graph":
        # batch entry j runs binning enq[j]; engines[] are views of its
        # contexts in BINNINGS order
        batch = imfcs.Batch(len(BINNINGS), graph=True, autotune=args.autotune, own_fits=True)
        engines = [batch.engine(enq.index(b)) for b in BINNINGS]
    else:
        engines = [imfcs.Engine(0) for _ in BINNINGS]
    # host-to-host figures: the stack copied in from pageable memory, and
    # from pinned memory (staged there before the timer starts)
    t0 = time.perf_counter()
    d_stack = engines[0].upload(stack)
    h2d_s = time.perf_counter() - t0
    pinned = engines[0].pinned_array(stack.shape)
    pinned[...] = stack
    t0 = time.perf_counter()
    engines[0].copy_in(d_stack, pinned)
    h2d_pinned_s = time.perf_counter() - t0
    engines[0].free_pinned(pinned)
    del pinned
    cfgs = configs(smin)
    n_acf_pixels = 0
    for c in cfgs:
        ow, oh, _ = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        imfcs.lib().imfcs_plan(ctypes.byref(c), ctypes.byref(ow), ctypes.byref(oh), ctypes.byref(_))
        n_acf_pixels += ow.value * oh.value

    def step():
        if batch is not None:
            batch.run([d_stack] * len(order), [cfgs[i] for i in order])
        else:
            imfcs.Engine.run_multi([engines[i] for i in order], [cfgs[i] for i in order], d_stack)
        kern = []
        for e in engines:
            e.synchronize()
            kern.append(e.correlator_kernel_ms())
        return kern

    def converged_acf():
        n, res = 0, []
        for c, e in zip(cfgs, engines):
            r = e.results(c)
            ok = (np.asarray(r["states"])[0] == 0) & np.isfinite(np.asarray(r["acf"])[0]).all(-1)
            n += int(ok.sum())
            res.append(r)
        return n, res

    def timed(label):
        for i in range(args.warmup):
            step()
        kern = np.zeros(len(BINNINGS))
        t0 = time.perf_counter()
        for i in range(args.steps):
            kern += np.asarray(step())
        dt = (time.perf_counter() - t0) / args.steps
        log("%s: %d steps, %.3f ms/step" % (label, args.steps, dt * 1e3))
        return dt, kern / args.steps

    # plain LM start
    dt_plain, kern_plain = timed("plain")
    n_conv_plain, res_plain = converged_acf()

    # seeding networks: one per binning and correlation kind (their fit geometry)
    t0 = time.perf_counter()
    nets = []
    for b, e in zip(BINNINGS, engines):
        _, _, mtab, lagtime = imfcs.correlator_tables(16, 8, T, 1e-3)
        pair = []
        for k, m in enumerate(models(b)):
            net = InitialGuessNet(e, m, lagtime[1:], mtab[1:], ncomp=1, seed=100 + 10 * b + k)
            net.train(steps=args.train_steps, batch=4096, lr=2e-3)
            pair.append(net)
        nets.append(pair)
        e.set_fit_seeder(*pair)
    train_s = time.perf_counter() - t0
    log("trained %d seeding networks in %.1f s" % (2 * len(BINNINGS), train_s))
    dt_seed, kern_seed = timed("seeded")
    n_conv_seed, res_seed = converged_acf()

    # fit stage alone on the last correlations (imfcs_refit): plain vs seeded
    lm = {}
    agree_all = []
    for b, c, e, pair, rs in zip(BINNINGS, cfgs, engines, nets, res_seed):
        row = {}
        for mode in ("plain", "seeded"):
            e.set_fit_seeder(*(pair if mode == "seeded" else (None, None)))
            ms = []
            for _ in range(max(1, args.steps)):
                e.refit(c)
                e.synchronize()
                ms.append(e.stage_ms()[4])
            r = e.results(c)
            st, it = np.asarray(r["states"]), np.asarray(r["iters"])
            row[mode] = {"fit_ms": round(float(np.median(ms)), 3),
                         "fitted_px_per_s": round(float((st == 0).sum() / (np.median(ms) * 1e-3)), 1),
                         "acf": fit_stats(st[0], it[0]), "ccf": fit_stats(st[1], it[1]), "_r": r}
        rp, rs2 = row["plain"].pop("_r"), row["seeded"].pop("_r")
        both = (np.asarray(rp["states"]) == 0) & (np.asarray(rs2["states"]) == 0)
        dD = np.abs(np.asarray(rs2["fitres"])[..., 1] / np.asarray(rp["fitres"])[..., 1] - 1)
        agree = dD[both] < 1e-3
        agree_all.append(agree)
        row["D_within_1e-3"] = round(float(agree.mean()), 5) if agree.size else None
        row["speedup"] = round(row["plain"]["fit_ms"] / row["seeded"]["fit_ms"], 3)
        lm["b%d" % b] = row
    agree_frac = float(np.concatenate(agree_all).mean())
    fit_ms_plain = sum(lm[k]["plain"]["fit_ms"] for k in lm)
    fit_ms_seed = sum(lm[k]["seeded"]["fit_ms"] for k in lm)
    for e in engines:
        e.set_fit_seeder(*([None, None]))

    value_plain = n_conv_plain / dt_plain
    value_seed = n_conv_seed / dt_seed

    # roofline of the main correlator kernel at b = 1 (HIP events on its
    # stream, averaged over the seeded timed steps), on the algorithmic bytes
    # of one pass over the stack: one uint16 read of W x H x T
    alg_bytes = 2.0 * W * H * T
    corr_ms = kern_seed[0]
    achieved = alg_bytes / (corr_ms * 1e-3) / 1e9
    # its arithmetic floor: 23 FP32 FMAs per pair-frame (17 level-0 delays on
    # every frame, 8 level-1 on every second, 8 level-2 on every fourth) for
    # the ACF and the CCF of every output pixel, at the non-packed FP32 VALU
    # peak (78.6 TFLOP/s = 39.3e12 FMA/s)
    ow1, oh1 = W - 1, H   # b = 1 output grid with the x+1 partner
    fmas = 23.0 * 2 * ow1 * oh1 * T
    floor_ms = fmas / 39.3e12 * 1e3
    roof = {"kernel": "k_corr_main (b=1, ACF+CCF)", "bound": "hbm", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": read_traffic(),
            "avg_ms": round(float(corr_ms), 3), "alg_bytes": alg_bytes,
            "valu_floor": {"fma": fmas, "floor_ms": round(floor_ms, 3), "frac": round(floor_ms / corr_ms, 4),
                           "note": "FP32 VALU FMA floor of the kernel's products; its HBM frac cannot exceed "
                                   "%.3f" % ((alg_bytes / (floor_ms * 1e-3) / 1e9) / HBM_PEAK_GBS)}}

    if args.copy_calibration:
        lib = imfcs.lib()
        dst = lib.imfcs_device_alloc(engines[0].ctx, ctypes.c_size_t(COPY_BYTES))
        if dst:
            lib.imfcs_memcpy_d2d(engines[0].ctx, dst, d_stack, ctypes.c_size_t(COPY_BYTES))
            lib.imfcs_memcpy_d2d_words(engines[0].ctx, dst, d_stack, ctypes.c_size_t(COPY_BYTES))
            lib.imfcs_device_free(engines[0].ctx, dst)
            log("copy calibration: %d bytes (16-byte vector copy, then 4-byte buffer-word copy)" % COPY_BYTES)

    iters = {}
    for b, r in zip(BINNINGS, res_plain):
        acf_ok = np.isfinite(np.asarray(r["acf"])[0]).all(-1)
        stt = np.asarray(r["states"])[0]
        iters["b%d" % b] = {"finite_acf_frac": round(float(acf_ok.mean()), 4),
                            "converged_frac": round(float((stt == 0).mean()), 4)}

    cpu = cpu_all = None
    if args.cpu_sample > 0:
        ncores = os.cpu_count() or 1
        rate, rate_px, cdt = cpu_baseline(stack, smin, args.cpu_sample, 1)
        smp = "%dx%d px ROI of the b=1 ACF workload, 50,000 frames, poly4, LM fit" % (args.cpu_sample,
                                                                                     args.cpu_sample)
        cpu = {"value": round(rate, 3), "unit": "fitted pixels/s", "cores": 1, "kind": "port",
               "sample": smp + " (%.1f s; fitted = optimizer without exception)" % cdt,
               "all_pixels_per_s": round(rate_px, 3), "cpu_model": cpu_model()}
        rate_a, rate_pa, cdt_a = cpu_baseline(stack, smin, args.cpu_sample, ncores)
        cpu_all = {"value": round(rate_a, 3), "unit": "fitted pixels/s", "cores": ncores, "kind": "port",
                   "sample": smp + " (%.1f s)" % cdt_a, "all_pixels_per_s": round(rate_pa, 3),
                   "nproc": ncores, "cpu_model": cpu_model()}

    line = {"metric": "fitted pixels/s (bin+bleach+ACF+LM), 128^2x50k; seeded vs plain LM at 1e-3 D",
            "value": round(value_seed, 1), "unit": "fitted pixels/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(dt_seed * 1e3, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u16/f32/f64 (bf16 seeding network)",
            "data": "synthetic (seeded 2D particle simulation)",
            "config": {"workload": "128x128x50000 uint16; binning 1,2,4; ACF+CCF(x+b); poly4 bleach; LM N,D,G",
                       "acf_pixels_per_step": n_acf_pixels, "converged_acf_fits_per_step": n_conv_seed,
                       "runner": ("hipGraph replay of the three pipelines (imfcs_batch)" if batch is not None
                                  else "imfcs_run_multi"),
                       "shared_bleach_pass": {str(b): bool(e.run_flags() & 1) for b, e in zip(BINNINGS, engines)}},
            "value_plain": round(value_plain, 1), "ms_per_step_plain": round(dt_plain * 1e3, 3),
            "converged_acf_fits_per_step_plain": n_conv_plain,
            "all_pixels_per_s": round(n_acf_pixels / dt_seed, 1),
            "host_to_host_per_s": round(n_conv_seed / (dt_seed + h2d_pinned_s), 1),
            "h2d_s_pinned": round(h2d_pinned_s, 4),
            "host_to_host_pageable_per_s": round(n_conv_seed / (dt_seed + h2d_s), 1), "h2d_s_pageable": round(h2d_s, 3),
            "seeded_vs_plain": {"D_within_1e-3_frac": round(agree_frac, 5),
                                "lm_stage_speedup": round(fit_ms_plain / fit_ms_seed, 3),
                                "pipeline_speedup": round(dt_plain / dt_seed, 3)},
            "lm_stage": lm, "seeder_training_s": round(train_s, 2), "fit_iterations": iters,
            "roofline": roof, "cpu_baseline": cpu, "cpu_baseline_all_cores": cpu_all}
    print(json.dumps(line), flush=True)
    for pair in nets:
        for n in pair:
            n.close()
    for e in engines[::-1]:
        e.close()
    if batch is not None:
        batch.close()


if __name__ == "__main__":
    main()
