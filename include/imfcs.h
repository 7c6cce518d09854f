/*
 * imfcs.h -- C ABI of the MI355X imaging-FCS engine (libimfcs_hip.so).
 *
 * Two layers share one set of HIP kernels:
 *
 *  1. The drop-in layer: one function per native method of the plugin's
 *     JNI class gpufitImFCS.GpufitImFCS (GpufitImFCS.java:323-538).  Plain
 *     host pointers and sizes, the same argument meaning and the same error
 *     behaviour (status codes / last-error string).  The JNI exports in
 *     jni_shim.cpp are thin wrappers over these.
 *
 *  2. The native pipeline: a 64-bit, device-resident path for uint16 stacks
 *     (bin -> polynomial bleach correction -> multi-tau ACF/CCF with blocked
 *     variance -> LM fit) that has no Java array limits.
 *
 * Every function returns 0 (IMFCS_OK) on success and a negative code on
 * failure; imfcs_get_last_error() then describes the failure.  The last
 * error is thread-local.
 */
#ifndef IMFCS_H
#define IMFCS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IMFCS_OK 0
#define IMFCS_ERROR (-1)
#define IMFCS_ERR_ARG (-2)
#define IMFCS_ERR_DEVICE (-3)
#define IMFCS_ERR_MEMORY (-4)

/* Fit model ids: GpufitImFCS.java:879-891, agpufitjni.cu:177-182. */
#define IMFCS_MODEL_GAUSS_2D 1
#define IMFCS_MODEL_ACF_1D 2
#define IMFCS_MODEL_LINEAR_1D 3
/* Estimator ids: GpufitImFCS.java:570-587 (only LSE is implemented). */
#define IMFCS_ESTIMATOR_LSE 0
#define IMFCS_ESTIMATOR_MLE 1
/* Fit states: GpufitImFCS.java:817-866, agpufitjni.cu:187. */
#define IMFCS_STATE_CONVERGED 0
#define IMFCS_STATE_MAX_ITERATION 1
#define IMFCS_STATE_SINGULAR_HESSIAN 2
#define IMFCS_STATE_NEG_CURVATURE_MLE 3
#define IMFCS_STATE_GPU_NOT_READY 4

/* ------------------------------------------------------------------ */
/* 1. Drop-in layer                                                    */
/* ------------------------------------------------------------------ */

/* Field-for-field mirror of GpufitImFCS.ACFParameters
 * (GpufitImFCS.java:453-486); booleans are 0/1. */
typedef struct imfcs_acf_params {
    int32_t width, height;          /* output grid */
    int32_t win_star, hin_star;     /* input region */
    int32_t w_temp, h_temp;         /* stride-1 binned grid */
    int32_t pixbinX, pixbinY, binningX, binningY;
    int32_t firstframe, lastframe, framediff;
    int32_t cfXDistance, cfYDistance;
    double correlatorp, correlatorq, frametime;
    int32_t background;
    double mtab1, mtabchanumminus1, sampchanumminus1;
    int32_t chanum;
    int32_t isNBcalculation, bleachcorr_gpu;
    int32_t bleachcorr_order, nopit, ave;
} imfcs_acf_params;

/* Batched Levenberg-Marquardt fit.  Replaces gpufit() behind the JNI
 * method `fit` (GpufitImFCS.java:323, agpufitjni.cu:4614-4636, 4332).
 * data/weights: n_fits*n_points (weights may be NULL); initial_parameters
 * and output_parameters: n_fits*n_parameters(model); parameters_to_fit:
 * n_parameters(model) flags; user_info: user_info_size BYTES (lag times or
 * bleach abscissae as float).  Models ACF_1D and LINEAR_1D; estimator LSE.
 * Returns 0 or -1 (message in imfcs_get_last_error), like gpufit(). */
int imfcs_fit(int n_fits, int n_points, const float *data, const float *weights,
              int model_id, const float *initial_parameters, float tolerance,
              int max_n_iterations, int num_valid_coefs,
              const int *parameters_to_fit, int estimator_id,
              size_t user_info_size, const void *user_info,
              float *output_parameters, int *output_states,
              float *output_chi_squares, int *output_n_iterations);

/* getLastError (GpufitImFCS.java:382, agpufitjni.cu:4643-4647). */
const char *imfcs_get_last_error(void);

/* isCudaAvailableInt (GpufitImFCS.java:389, agpufitjni.cu:4656-4659):
 * 1 if a usable GPU is present, else 0. */
int imfcs_device_available(void);

/* getCudaVersionAsArray (GpufitImFCS.java:403, agpufitjni.cu:4669-4683):
 * HIP runtime and driver versions as major*1000 + minor*10. */
int imfcs_get_versions(int *runtime_version, int *driver_version);

/* resetGPU (GpufitImFCS.java:491, agpufitjni.cu:4688-4700): drops every
 * cached device allocation of this library (a full device reset would
 * destroy other users' contexts in the same process). */
int imfcs_reset_gpu(void);

/* resetDevice (GpufitImFCS.java:538): declared by the reference but never
 * implemented; here a no-op that returns 0. */
int imfcs_reset_device(void);

/* calcDataBleachCorrection (GpufitImFCS.java:501, agpufitjni.cu:4708-4809):
 * pixels float [framediff][h_temp][w_temp]; out float [h_temp][w_temp][nopit]
 * = means of `ave` consecutive frames. */
int imfcs_calc_data_bleach_correction(const float *pixels, float *outdata,
                                      const imfcs_acf_params *p);

/* isBinningMemorySufficient (GpufitImFCS.java:510, agpufitjni.cu:4816-4853):
 * 1 if (win*hin + w_temp*h_temp)*T*4 bytes fit in 90% of free memory. */
int imfcs_is_binning_memory_sufficient(const imfcs_acf_params *p);

/* calcBinning (GpufitImFCS.java:520, agpufitjni.cu:4860-4967): stride-1
 * binningX x binningY window sums of a float [T][hin_star][win_star] stack
 * into float [T][h_temp][w_temp]. */
int imfcs_calc_binning(const float *indata, float *outdata, const imfcs_acf_params *p);

/* isACFmemorySufficient (GpufitImFCS.java:531, agpufitjni.cu:4974-5092). */
int imfcs_is_acf_memory_sufficient(const imfcs_acf_params *p);

/* calcACF (GpufitImFCS.java:533-536, agpufitjni.cu:5364-5831).
 * pixels: float [framediff][h_temp][w_temp], background already removed.
 * Outputs (double, plane-major [z][height][width]): pixels1 = G(z),
 * blockvararray = blocked variance, blocked1D plane 0 = blocking index,
 * plane 1 = blocking success flag.  NBmean/NBcov ([height][width]) are
 * written when isNBcalculation.  bleachcorr_params [(y*w_temp+x)*order+i]
 * is read when bleachcorr_gpu (polynomial in seconds, as fitted by the
 * LINEAR_1D model).  samp/lag: chanum-entry correlator tables. */
int imfcs_calc_acf(const float *pixels, double *pixels1, double *blockvararray,
                   double *NBmean, double *NBcov, double *blocked1D,
                   const double *bleachcorr_params, const double *samp,
                   const int *lag, const imfcs_acf_params *p);

/* ------------------------------------------------------------------ */
/* 2. Native pipeline (device-resident uint16 stacks)                  */
/* ------------------------------------------------------------------ */

typedef struct imfcs_ctx imfcs_ctx;

/* Fit geometry of the FCS model in SI units (metres, seconds), as
 * prepareFit / GPU_ACF_Fit set it up (Imaging_FCS_1_52.java:9986-10039,
 * 15090-15110).  paramfit/paraminit follow the 11 fit parameters
 * N, D, vx, vy, G, F2, D2, F3, D3, Ftrip, Ttrip. */
typedef struct imfcs_fit_model {
    double ax, ay, s, sz, rx, ry, fitobsvol, q2, q3;
    int32_t paramfit[11];
    double paraminit[11];
} imfcs_fit_model;

typedef struct imfcs_config {
    int32_t width, height, frames;   /* stack dims W, H, T */
    int32_t firstframe, lastframe;   /* 1-based inclusive frame range */
    int32_t binning_x, binning_y;    /* b x b window sums */
    int32_t pixbin_x, pixbin_y;      /* output stride (== binning unless overlap) */
    int32_t cf_x, cf_y;              /* CCF partner offset in raw pixels */
    int32_t background;              /* subtracted from every raw pixel */
    int32_t correlator_p, correlator_q;
    double frametime;                /* seconds */
    int32_t poly_order;              /* polynomial bleach order, 0 = none */
    int32_t roi_x, roi_y, roi_w, roi_h;
    int32_t do_acf, do_ccf;          /* which correlation pairs */
    int32_t do_fit;                  /* LM fit of every computed pair */
    int32_t fit_max_iterations;      /* 2000 in the plugin */
    float fit_tolerance;             /* relative chi^2 change for convergence */
    int32_t fit_chi2_fp64;           /* 1: accept steps and test convergence on an FP64 chi^2
                                        (tighter optimum); 0: Gpufit's FP32 chi^2 */
    int32_t gls;                     /* 1: regularized channel covariance of every pair
                                        (calculateCF GLS branch, IFCS:9524-9639); the acf
                                        output becomes its mean curve and the fit is the
                                        GLS fit (GLSFit, IFCS:10622-10789) */
    imfcs_fit_model acf_model, ccf_model;
    int32_t bleach_mode;             /* IMFCS_BLEACH_*: POLYNOMIAL (of poly_order; 0 = no
                                        correction), SLIDING_WINDOW (IFCS:9146-9172),
                                        LIN_SEGMENT (IFCS:9767-9805), SINGLE_EXP
                                        (IFCS:9682-9709) or DOUBLE_EXP (IFCS:9710-9738) */
    int32_t sliding_window_length;   /* frames per window (SLIDING_WINDOW), or the length
                                        that sets the segment count (LIN_SEGMENT) */
} imfcs_config;

/* Bleach correction modes of the native pipeline (cbBleachCor, IFCS:9145). */
enum {
    IMFCS_BLEACH_POLYNOMIAL = 0,
    /* correlations of consecutive windows of sliding_window_length frames,
       uncorrected, averaged (ACF, blocked variance) over the windows; the
       correlator depth is reduced to what a window supports (IFCS:7331-7341) */
    IMFCS_BLEACH_SLIDING_WINDOW = 1,
    /* piecewise-linear bleach curve through the means of num / L segments;
       coefficients returned per series: bcInt0 then the segment means */
    IMFCS_BLEACH_LIN_SEGMENT = 2,
    /* LM fit of A e^{-t/t1} + C to the intensity trace and the correction
       with that curve (IFCS:9682-9709); coefficients returned per series:
       f(0), A, t1, C */
    IMFCS_BLEACH_SINGLE_EXP = 3,
    /* A e^{-t/t1} + B e^{-t/t2} + C (DoubleExpFit, IFCS:9710-9738,
       10898-10950) from the reference's initial guesses; coefficients per
       series: f(0), A, t1, B, t2, C.  Where the two-exponential least-squares
       problem has several minima, the LM's path decides which one is found
       (as in the reference, whose CPU and GPU fits differ the same way). */
    IMFCS_BLEACH_DOUBLE_EXP = 4
};

/* Output geometry for a config: out_w x out_h pixels, chanum channels. */
int imfcs_plan(const imfcs_config *cfg, int32_t *out_w, int32_t *out_h, int32_t *chanum);

imfcs_ctx *imfcs_create(int device);
void imfcs_destroy(imfcs_ctx *ctx);

/* Device buffer helpers (so hosts need no other GPU runtime). */
void *imfcs_device_alloc(imfcs_ctx *ctx, size_t bytes);
int imfcs_device_free(imfcs_ctx *ctx, void *ptr);
/* Pinned (page-locked) host memory, so that imfcs_memcpy_h2d of a stack
 * staged there runs at the link's full rate.  No counterpart in the
 * reference (its JNI calls copy Java arrays through pageable memory). */
void *imfcs_host_alloc(imfcs_ctx *ctx, size_t bytes);
int imfcs_host_free(imfcs_ctx *ctx, void *ptr);
int imfcs_memcpy_h2d(imfcs_ctx *ctx, void *dst, const void *src, size_t bytes);
int imfcs_memcpy_d2h(imfcs_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Device-to-device copy by a plain vector-copy kernel (also the known-size
 * reference for HBM traffic counters). */
int imfcs_memcpy_d2d(imfcs_ctx *ctx, void *dst, const void *src, size_t bytes);
/* The same copy in 4-byte words with buffer loads and stores (one word per
 * lane, the row as the scalar offset), the access pattern of the pipeline's
 * FP32 series passes: the known-size reference of the traffic counters for
 * such kernels.  No counterpart in the reference (a measurement aid);
 * bytes must be a multiple of 4. */
int imfcs_memcpy_d2d_words(imfcs_ctx *ctx, void *dst, const void *src, size_t bytes);
int imfcs_synchronize(imfcs_ctx *ctx);

/* Runs the pipeline on a device-resident uint16 stack [T][H][W].  Results
 * stay on the device inside ctx until imfcs_get_results.  Asynchronous
 * on the context's stream. */
int imfcs_run(imfcs_ctx *ctx, const uint16_t *d_stack, const imfcs_config *cfg);

/* No counterpart in the reference: imfcs_run for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Every scratch buffer of the run is
 * carved from a file-scope device array of 64 MiB per GPU (held until the
 * process ends); configurations that need more are refused with
 * IMFCS_ERR_MEMORY.  No hipMalloc or hipFree.  Results go to the caller's
 * device arrays, laid out as imfcs_get_results' host arrays (acf, var:
 * [pair][chanum]; ind, blocked, chi2, states, iters: [pair]; fitres:
 * [pair][11]; the fit arrays may be NULL when cfg->do_fit is 0).
 * Asynchronous on ctx's stream.  Every call on one GPU shares its scratch
 * array: calls that overlap on one GPU (other contexts, other threads)
 * race.  No fit seeding; imfcs_get_results and imfcs_refit do not apply to
 * such a run. */
int imfcs_run_noalloc(imfcs_ctx *ctx, const uint16_t *d_stack, const imfcs_config *cfg, double *acf, double *var,
                      int32_t *ind, int32_t *blocked, double *fitres, double *chi2, int32_t *states,
                      int32_t *iters);

/* n configurations of one device stack, cfgs[i] on ctxs[i] (distinct
 * contexts, one stream each) -- the binnings of a diffusion law
 * (determineDiffusionLaw, IFCS:8190-8215, runs one pipeline per binning).
 * No single reference call corresponds: it replaces n calls of the
 * per-binning path.  When all take the polynomial bleach correction over
 * the same frame range and frame size, one pass over the stack sums every
 * pixel over each bleach-trace point (on ctxs[0]'s stream) and every
 * pipeline takes its intensity trace from those sums instead of the stack
 * (identical traces: Java's integer division comes after the window sum);
 * the other streams wait for that pass.  Otherwise equal to n imfcs_run
 * calls.  Asynchronous; results per context as after imfcs_run. */
int imfcs_run_multi(imfcs_ctx *const *ctxs, int n, const uint16_t *d_stack, const imfcs_config *cfgs);

/* The fit stage alone on the correlations of the last imfcs_run of ctx,
 * with cfg's fit models and settings (every other field of cfg must equal
 * the last run's).  The PSF calibration (determinePSF, IFCS:8118-8160)
 * reruns dlfit -- the whole pipeline of every binning -- for each trial
 * PSF size although only the fit depends on it; no single reference call
 * corresponds.  Asynchronous; results as after imfcs_run. */
int imfcs_refit(imfcs_ctx *ctx, const imfcs_config *cfg);

/* Regularized covariance (chanum-1)^2, FP64, row-major, of one pair of the
 * last imfcs_run with cfg.gls = 1 (covmats of IFCS:9601-9613). */
int imfcs_get_gls_covariance(imfcs_ctx *ctx, int pair, double *cov);

/* Device time (ms) of the main multi-tau correlator kernel (levels 0..2,
 * the pass over the corrected series) of the last imfcs_run; measured with
 * HIP events on the context's stream. */
int imfcs_get_correlator_kernel_ms(imfcs_ctx *ctx, float *ms);

/* Copies the results of the last imfcs_run to host arrays (any may be NULL).
 * pair 0 = ACF, pair 1 = CCF (when both computed, ACF first).
 * acf/var: [pair][out_h][out_w][chanum]; ind/blocked: [pair][out_h][out_w];
 * fitres: [pair][out_h][out_w][11] (SI units); chi2: [pair][pixel];
 * states/iters: [pair][pixel]; coeffs: [series][poly_order+1] bleach
 * polynomial in seconds (series = pixels, then CCF partners). */
int imfcs_get_results(imfcs_ctx *ctx, double *acf, double *var, int32_t *ind,
                      int32_t *blocked, double *fitres, double *chi2,
                      int32_t *states, int32_t *iters, double *coeffs);

/* Per-stage device time of the last imfcs_run in milliseconds:
 * [0] trace+polyfit, [1] means, [2] blocking, [3] correlation, [4] fit. */
int imfcs_get_stage_ms(imfcs_ctx *ctx, float *stage_ms5);

/* Batches of independent stacks (no single reference call corresponds: the
 * plugin runs the whole pipeline per stack and binning,
 * Imaging_FCS_1_52.java:11146-11166, each stage a host loop of launches,
 * agpufitjni.cu:5714-5741).  A batch holds one context per stack.  Its
 * first run of a set of stacks and configurations enqueues every stack's
 * pipeline on its own stream and fits all their pixels in ONE persistent
 * launch (a work queue over the fits of every stack; used when every
 * configuration fits N, D, G_inf with the FP64 chi-square, otherwise each
 * stack fits on its own); with IMFCS_BATCH_GRAPH the same work is captured
 * into a hipGraph that later runs of the same stacks and configurations
 * replay; with IMFCS_BATCH_AUTOTUNE the frame chunking of the bleach/prefix
 * and blocking passes is first timed per stack shape and the fastest choice
 * kept for every later run of that shape (batched or not).  Results per
 * stack: imfcs_get_results(imfcs_batch_context(b, i), ...) after
 * imfcs_batch_synchronize; identical to imfcs_run of each stack with the
 * same tuning.  Where every entry is the same device stack under the
 * polynomial correction with one frame range (the binnings of a diffusion
 * law), the bleach traces of all come from one pass over it, as in
 * imfcs_run_multi, whose results the batch then reproduces. */
typedef struct imfcs_batch imfcs_batch;
#define IMFCS_BATCH_GRAPH 1
#define IMFCS_BATCH_AUTOTUNE 2
/* every pipeline fits its own pixels at the end of its own stream (no
 * shared fit launch), as imfcs_run does; seeded contexts
 * (imfcs_set_fit_seeder) need this flag */
#define IMFCS_BATCH_OWN_FITS 4
imfcs_batch *imfcs_batch_create(int device, int n, int flags);
void imfcs_batch_destroy(imfcs_batch *b);
int imfcs_batch_run(imfcs_batch *b, const uint16_t *const *d_stacks, const imfcs_config *cfgs);
int imfcs_batch_synchronize(imfcs_batch *b);
imfcs_ctx *imfcs_batch_context(imfcs_batch *b, int i);
/* graph nodes of the captured run, fit groups of the persistent launch,
 * shapes tuned and time spent tuning. */
int imfcs_batch_info(imfcs_batch *b, int *graph_nodes, int *fit_groups, int *tuned_shapes, double *autotune_ms);
/* The frame chunking recorded for a series shape (2^apply_log2 and
 * 2^block_log2 threads in flight; 19 / 19 untuned), and a context's slot
 * count (the shape key of its last run). */
int imfcs_get_tuning(int nslots, int frames, int *apply_log2, int *block_log2);
int imfcs_get_slot_count(imfcs_ctx *ctx, int *nslots);
/* How a context's last run was done (no counterpart in the reference; for
 * callers and tests that check which path ran): bit 0 (IMFCS_RUN_NESTED)
 * set when its bleach-correction pass was the one pass shared by binnings
 * 1, 2 and 4 of its stack (imfcs_run_multi). */
#define IMFCS_RUN_NESTED 1
int imfcs_get_run_flags(imfcs_ctx *ctx, int *flags);

/* ------------------------------------------------------------------ */
/* 3. Learned initial guesses for the LM fit                           */
/* ------------------------------------------------------------------ */

/* No reference call corresponds: the plugin starts every LM fit from the
 * fit panel's values (Imaging_FCS_1_52.java:1959-1968, packed for every
 * pixel by GPU_ACF_Fit, 15090-15121).  A network maps a fit's data curve
 * (channels 1..npts divided by their largest magnitude) to starting values
 * N, D, G_inf (and D2, F2 with two components), for one fit geometry. */
typedef struct imfcs_mlp imfcs_mlp;

/* A BF16 perceptron npts -> 256 -> 256 -> 5 (GELU) on ctx's device and
 * stream for the fit geometry `geom` (ax, ay, s, sz, rx, ry, fitobsvol, q2,
 * q3 used; SI units), lag times lagtime[npts] (s) and product counts
 * mtab[npts] of channels 1..npts; ncomp 1 or 2 components.  Weights are
 * He-initialised from `seed`.  NULL on failure. */
imfcs_mlp *imfcs_mlp_create(imfcs_ctx *ctx, int ncomp, const imfcs_fit_model *geom, const double *lagtime,
                            int npts, const double *mtab, uint64_t seed);
void imfcs_mlp_destroy(imfcs_mlp *net);

/* `steps` AdamW steps on batches of `batch` fresh model curves of the
 * network's geometry (FCS_3p with log-uniform N and D, noise relative to 1/N
 * log-uniform in 1e-3..0.5 and growing with the lag as sqrt(mtab_1 /
 * mtab_i)), learning rate lr with cosine decay.  loss_out[steps] (may be
 * NULL) receives the loss of every step.  Synchronous. */
int imfcs_mlp_train(imfcs_mlp *net, int steps, int batch, float lr, float *loss_out);

/* Number of FP32 parameters, and a copy of them (W1 [80][256], b1 [256],
 * W2 [256][256], b2 [256], W3 [256][16], b3 [16]; input rows >= npts and
 * output columns >= 5 are padding). */
int imfcs_mlp_param_count(void);
int imfcs_mlp_get_params(imfcs_mlp *net, float *params);

/* The standardisation of the five outputs: the network predicts (t - mu) /
 * sd of t = (ln(N s), ln D, ln D2, F2, G_inf / s), s the curve's largest
 * magnitude. */
int imfcs_mlp_target_scale(imfcs_mlp *net, float *mu, float *sd);

/* One forward and backward pass on host data (the gradient check): inputs
 * [rows][80] (rounded to BF16), standardised targets [rows][5]; grad
 * receives the parameter gradient of the weighted mean squared error,
 * outputs [rows][5] the network outputs, loss the loss. */
int imfcs_mlp_gradient(imfcs_mlp *net, const float *inputs, const float *targets, int rows, float *grad,
                       float *outputs, float *loss);

/* `rows` training curves as the trainer draws them next: inputs [rows][80]
 * (BF16 values) and standardised targets [rows][5]. */
int imfcs_mlp_curves(imfcs_mlp *net, int rows, float *inputs, float *targets);

/* Predictions for device curves [rows][npts] into device out [rows][5]:
 * N, D (um^2/s), D2 (um^2/s), F2, G_inf; NaN rows for curves that are not
 * finite.  Synchronous. */
int imfcs_mlp_predict(imfcs_mlp *net, const float *d_curves, int rows, float *d_out);

/* Seeds the LM fits of ctx's next runs: ACF fits start from acf_net's
 * prediction, CCF fits from ccf_net's (either may be NULL: those fits start
 * from the model's paraminit).  Fits that do not converge from the seed are
 * fitted again from paraminit (the plain method), their iteration counts
 * added.  The networks must outlive their use; pass NULLs to stop. */
int imfcs_set_fit_seeder(imfcs_ctx *ctx, imfcs_mlp *acf_net, imfcs_mlp *ccf_net);

/* One periodic box of the GPU simulator (object-space micrometres; detector
 * pixel (i, j) covers [i, i+1) x [j, j+1) pixel sizes).  Its emitters start
 * uniformly inside the box and wrap at its edges.  Component 1 diffuses with
 * D_um2s[q] in quadrant q of the box ((x<cx,y<cy), (x>=cx,y<cy), (x<cx,y>=cy),
 * (x>=cx,y>=cy)); an emitter is of component 2 (D2_um2s everywhere) with
 * probability frac2.  Brightness in photons per frame: counts_lo *
 * (counts_hi / counts_lo)^((x - x0) / w) at the emitter's position. */
typedef struct imfcs_sim_domain {
    double x0_um, y0_um, w_um, h_um;
    int32_t n_emitters;
    double D_um2s[4];
    double D2_um2s;
    double frac2;
    double counts_lo, counts_hi;
} imfcs_sim_domain;

/* The simulator on the GPU, into a device uint16 stack [T][H][W] (physics of
 * simulateACF2D, Imaging_FCS_1_52.java:13963-14393; shot noise drawn per
 * pixel as Poisson with the PSF-integrated mean).  stream: a hipStream_t or
 * NULL.  Synchronous; 0 or a negative code. */
int imfcs_simulate_2d_device(uint16_t *d_out, int W, int H, int T, const imfcs_sim_domain *domains,
                             int ndomains, double pixel_um, double psf_sigma_um, double frametime_s,
                             int offset, double tau_bleach_s, uint64_t seed, void *stream);

/* Seeded 2D particle simulator into a host uint16 [T][H][W] buffer
 * (physics of the plugin's simulateACF2D, Imaging_FCS_1_52.java:13963-14393).
 * Lives in libimfcs_sim.so. D_um2s: 4 quadrant diffusion coefficients. */
int imfcs_simulate_2d(uint16_t *out, int W, int H, int T, int margin, int n_emitters,
                      double pixel_um, double psf_sigma_um, double frametime_s,
                      double counts_per_frame, int offset, double tau_bleach_s,
                      const double *D_um2s, uint64_t seed, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
