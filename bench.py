#!/usr/bin/env python3
"""Headline benchmark: fitted pixels/s of the whole imaging-FCS hot path
(bin + polynomial bleach correction + multi-tau ACF/CCF with blocked variance
+ LM fit) on a 128x128x50,000-frame uint16 stack, on one MI355X.

One step = for each binning b in (1, 2, 4): fused b x b binning, order-4
polynomial bleach correction, ACF and CCF to the x+b neighbour with blocked
variances, and the LM fit of N, D, G_inf for every ACF and CCF pixel.  The
three binnings run as three pipelines on three HIP streams (one engine each,
all reading the one device-resident stack, started by imfcs_run_multi: one
pass over the stack sums every pixel over each bleach-trace point for all
three), so the tail of one pipeline's fit overlaps the others' work.  The
stack is device-resident before timing starts.  ``value`` counts ACF-fitted
output pixels (the CCF work is inside the timed region but not counted).

Prints one JSON line.  ``roofline``: the main correlator kernel at binning 1,
algorithmic bytes (one read of the FP32 corrected series, one write of the
level-3 sums) over its HIP-event duration on its stream, against the 8 TB/s
HBM peak, with the copy-calibrated FETCH_SIZE + WRITE_SIZE bytes per launch
read from the committed traffic profile.  ``cpu_baseline``: the oracle (a C
restatement of the plugin's Java CPU path) on one host core over a pixel
sample of the binning-1 ACF workload.
"""
import argparse
import ctypes
import json
import os
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


def make_stack(seed=2024):
    import imaging_fcs_1_52_amd as imfcs
    t0 = time.time()
    st = imfcs.simulate_2d(W, H, T, N_EMITTERS, D_um2s=D_QUADRANTS, pixel_um=0.24, psf_sigma_um=0.19,
                           frametime=1e-3, counts_per_frame=20.0, offset=100, tau_bleach_s=15.0, margin=2,
                           seed=seed, nthreads=min(16, os.cpu_count() or 1))
    log("simulated %dx%dx%d stack in %.1f s" % (W, H, T, time.time() - t0))
    return st


def configs(stack_min):
    import imaging_fcs_1_52_amd as imfcs
    out = []
    for b in BINNINGS:
        ax = PIX * b
        acf = imfcs.make_fit_model(ax, ax, PSF)
        ccf = imfcs.make_fit_model(ax, ax, PSF, rx=PIX * b)
        cfg = imfcs.make_pipeline_config(W, H, T, binning=b, cf=(b, 0), background=stack_min, poly_order=4,
                                         do_acf=True, do_ccf=True, do_fit=True, acf_model=acf, ccf_model=ccf,
                                         fit_tolerance=1e-16, fit_max_iterations=2000)
        out.append(cfg)
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
    npx = sample_side * sample_side
    return npx / dt, dt, r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=48, help="side of the oracle's pixel sample (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=1)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain device-to-device copy of the stack (known bytes) for PMC traffic scaling")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine has no multi-GPU path (pixels shard without "
              "communication)", file=sys.stderr)
        sys.exit(2)

    import imaging_fcs_1_52_amd as imfcs
    stack = make_stack()
    smin = int(stack.min())
    engines = [imfcs.Engine(0) for _ in BINNINGS]
    d_stack = engines[0].upload(stack)
    cfgs = configs(smin)
    n_fitted_pixels = 0
    for c in cfgs:
        ow, oh, _ = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        imfcs.lib().imfcs_plan(ctypes.byref(c), ctypes.byref(ow), ctypes.byref(oh), ctypes.byref(_))
        n_fitted_pixels += ow.value * oh.value

    def step():
        # the coarse binnings are enqueued first: their pipelines are short
        # and end in a few long LM fits, which then overlap the b=1 pipeline
        # (b=2, whose CCF fit tail is the longest, first of all).  One pass
        # over the stack gives every binning its bleach trace
        # (imfcs_run_multi)
        order = [BINNINGS.index(b) for b in ENQUEUE_ORDER]
        imfcs.Engine.run_multi([engines[i] for i in order], [cfgs[i] for i in order], d_stack)
        stage, kern = [], []
        for e in engines:
            e.synchronize()
            stage.append(e.stage_ms())
            kern.append(e.correlator_kernel_ms())
        return stage, kern

    for i in range(args.warmup):
        step()
        log("warmup step %d done" % i)
    per_bin = np.zeros((len(BINNINGS), 5))
    kern_ms = np.zeros(len(BINNINGS))
    t0 = time.perf_counter()
    for i in range(args.steps):
        st, km = step()
        per_bin += np.asarray(st)
        kern_ms += np.asarray(km)
        log("step %d done" % i)
    dt = time.perf_counter() - t0
    ms_step = 1e3 * dt / args.steps
    per_bin /= args.steps
    kern_ms /= args.steps
    stage_sum = per_bin.sum(0)
    value = n_fitted_pixels / (dt / args.steps)

    # roofline of the main correlator kernel at b = 1 (HIP events on its
    # stream, averaged over the timed steps): it reads the corrected FP32
    # series of every slot once (4 B per slot-frame) and writes the level-3
    # sums of every slot once (4 B per 8 frames)
    nslots_b1 = W * H   # 128 x 128 slot grid (127 output columns + the x+1 partner column)
    alg_bytes = 4.0 * nslots_b1 * T + 4.0 * nslots_b1 * (T // 8)
    corr_ms = kern_ms[0]
    achieved = alg_bytes / (corr_ms * 1e-3) / 1e9
    roof = {"kernel": "k_corr_main (b=1, ACF+CCF)", "bound": "hbm", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "avg_ms": round(float(corr_ms), 3), "alg_bytes": alg_bytes}

    traffic = read_traffic()
    if traffic is not None:
        roof["traffic"] = traffic

    if args.copy_calibration:
        # a plain copy of COPY_BYTES: its FETCH_SIZE / WRITE_SIZE under
        # rocprofv3 --pmc give the counters' bytes-per-byte scale
        lib = imfcs.lib()
        dst = lib.imfcs_device_alloc(engines[0].ctx, ctypes.c_size_t(COPY_BYTES))
        if dst:
            lib.imfcs_memcpy_d2d(engines[0].ctx, dst, d_stack, ctypes.c_size_t(COPY_BYTES))
            lib.imfcs_device_free(engines[0].ctx, dst)
            log("copy calibration: %d bytes" % COPY_BYTES)

    # LM iteration statistics of every binning (untimed re-run)
    iters = {}
    for b, c, e in zip(BINNINGS, cfgs, engines):
        e.run(c, d_stack)
        r = e.results(c)
        # ACF fits (pair 0).  Pixels whose bleach polynomial crosses zero get
        # NaN correlations (the reference's correction takes sqrt(f/c0) of a
        # negative ratio) and their fits stop at once; iteration counts are
        # over the converged fits
        acf_ok = np.isfinite(np.asarray(r["acf"])[0]).all(-1)
        stt = np.asarray(r["states"])[0]
        it = np.asarray(r["iters"])[0][stt == 0]
        iters["b%d" % b] = {"median": float(np.median(it)), "p99": float(np.percentile(it, 99)),
                            "max": int(it.max()), "finite_acf_frac": round(float(acf_ok.mean()), 4),
                            "converged_frac": round(float((stt == 0).mean()), 4)}

    cpu = None
    if args.cpu_sample > 0:
        rate, cdt, _ = cpu_baseline(stack, smin, args.cpu_sample, args.cpu_threads)
        cpu = {"value": round(rate, 3), "unit": "fitted pixels/s", "cores": args.cpu_threads, "kind": "port",
               "sample": "%dx%d px ROI of the b=1 ACF workload, 50,000 frames, poly4, LM fit (%.1f s)"
                         % (args.cpu_sample, args.cpu_sample, cdt)}

    line = {"metric": "fitted pixels/s (bin+bleach+ACF+LM), 128^2x50k", "value": round(value, 1),
            "unit": "fitted pixels/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32/f64", "data": "synthetic (seeded 2D particle simulation)",
            "config": {"workload": "128x128x50000 uint16; binning 1,2,4; ACF+CCF(x+b); poly4 bleach; LM N,D,G",
                       "fitted_pixels_per_step": n_fitted_pixels},
            "stages_ms_sum_of_streams": {"trace_polyfit": round(stage_sum[0], 3), "means": round(stage_sum[1], 3),
                          "blocking": round(stage_sum[2], 3), "correlate": round(stage_sum[3], 3),
                          "fit": round(stage_sum[4], 3)},
            "fit_iterations": iters, "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)
    for e in engines[::-1]:
        e.close()


if __name__ == "__main__":
    main()
