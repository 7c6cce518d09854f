// Internal types shared by the HIP kernels and the host-side C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/imfcs.h"

namespace imfcs {

// The stream and device of a context (imfcs_capi.hip).
hipStream_t ctx_stream(imfcs_ctx *ctx);
int ctx_device(imfcs_ctx *ctx);

// Records msg as the calling thread's last error (imfcs_get_last_error) and
// returns code.
int set_error(int code, const std::string &msg);

constexpr int kMaxChan = 320;   // correlator channels (73 for p=16, q=8; 289 for p=64, q=8)
constexpr int kMaxLevel = 12;   // multi-tau levels (q)
// zero rows stored after each level >= 3 series array of the correlator, so
// that its readers load past the last value without a clamp
constexpr int kLevelPad = 64;
constexpr int kMaxOrder = 8;    // polynomial bleach order (BCmaxorder, IFCS:201)
constexpr int kMaxBlock = 20;   // blocking-curve levels
constexpr int kMaxPos = 2 * kMaxChan + kMaxLevel + 8;  // prefix positions

enum SourceType : int32_t { kSrcU16 = 0, kSrcF32 = 1 };

// Where the intensity series live and how one series value is formed.
// Series s < npix is output pixel s (row-major over ow x oh); series
// s >= npix is the CCF partner of pixel s - npix.  Distinct series are
// held once, in "slots": a CCF partner that coincides with another output
// pixel (partner offset a multiple of the pixel pitch) shares its slot.
// Slot k reads the window at (slot_org[2k], slot_org[2k+1]).
struct SeriesGeom {
    const void *data;
    int32_t src;          // SourceType
    int64_t plane;        // elements per frame
    int32_t stride;       // elements per row
    int32_t f0;           // first frame (0-based index into data)
    int32_t T;            // frames in each series (num)
    int32_t bx, by;       // window summed per series value (1 for F32)
    int32_t bckg_total;   // background * bx * by, subtracted per value
    int32_t ow, oh, npix;
    int32_t ox0, oy0;     // raw origin of map pixel (0,0)
    int32_t pbx, pby;     // stride between map pixels
    int32_t cfx, cfy;     // partner offset
    int32_t nseries;
    int32_t nslots;
    // slot row pitch of the pixel grid: pixel (x, y) is slot y * swp + x,
    // swp a power of two up to 64 or a multiple of 64
    int32_t swp;
    const int32_t *slot_of;    // [nseries] device
    const int32_t *slot_org;   // [nslots][2] device
    // [nseries] device: 1 for the first series of each slot, which writes
    // the slot's coarse correlator levels (stored once per slot)
    const int32_t *slot_own;
    // 2 or 4: every window row starts on a multiple of bx elements (uint16
    // source), so a row of bx = vec pixels is one 4- or 8-byte load; else 0
    int32_t vec;
};

// The bleach-corrected series of every slot, stored once per pipeline run
// as FP32 deviations from a per-slot shift mu (X[f * nslots + slot] =
// I'(f) - mu): mu is the constant term c0 of the bleach polynomial, around
// which the corrected series fluctuates (I' -> c0 when I follows the fit),
// or 0 without correction, where I' is an integer and X exact.
__host__ __device__ inline double slot_mu(const double *coef, int order, int slot) {
    return order > 0 ? coef[(int64_t)slot * (order + 1)] : 0.0;
}

// Polynomial bleach correction (per series) in a normalized abscissa
// u = (frame + t_off) * frametime / tscale.  coef[s][k] are the
// coefficients of f(u); order 0 means "no correction".
struct BleachDesc {
    const double *coef;   // [nslots][order+1]
    int32_t order;
    double frametime, tscale, t_off;
    // 1: linear segments (coef[slot] = {bcInt0, bcTrace[0..order-1]},
    // segments of seg_len frames); 2, 3: single / double exponential
    // (coef[slot] = {f(0), A, t1, C} / {f(0), A, t1, B, t2, C}); 0:
    // polynomial of the given order
    int32_t mode;
    int32_t seg_len;
};

// Multi-tau correlator layout derived from the lag/samp tables.
struct CorrTables {
    int32_t chanum, nlevels, T;
    int32_t level[kMaxChan];      // log2(samp)
    int32_t del[kMaxChan];        // lag / samp
    int32_t first_chan[kMaxLevel], nchan[kMaxLevel];
    int32_t numbin[kMaxLevel];    // floor(T / 2^level)
    double mtab_last, log2_samp_last;
    int32_t q;
    int32_t blocknum;             // blocking levels of the lag-1 curve
    // prefix-sum positions used to form every channel mean
    int32_t npos;
    int32_t pos[kMaxPos];
    int16_t pos_a_end[kMaxChan];  // index into pos of 2^c (numbin_c - del)
    int16_t pos_b_beg[kMaxChan];  // index of lag
    int16_t pos_b_end[kMaxChan];  // index of 2^c numbin_c
    int16_t pos_blk_a, pos_blk_b, pos_total;
    // 1: the p = 16 layout (level 0 delays 0..16, every other level 9..16)
    // that the fast correlator kernels implement
    int32_t p16;
};

// Pair p (< npairs) correlates series sa(p) with series sb(p).
struct PairDesc {
    int32_t npix, npairs;
    int32_t acf_first;   // 1: pairs [0,npix) are ACF
    int32_t has_ccf;     // 1: CCF pairs follow
};

__host__ __device__ inline void pair_series(const PairDesc &pd, int p, int &sa, int &sb) {
    if (pd.acf_first && p < pd.npix) {
        sa = p;
        sb = p;
    } else {
        int k = pd.acf_first ? p - pd.npix : p;
        sa = k;
        sb = pd.npix + k;
    }
}

// Fit model with fixed geometry (ACF_1D parameter slots 11..19).
struct LMModelGeom {
    float ax, ay, s, sz, rx, ry, obsvol, q2, q3;   // micrometres
};

// GLS covariance of every pair (imfcs_gls.hip).
constexpr int kGlsMaxM = 104;   // fitted channels up to q = 12 (p = 16)
struct GlsArgs {
    const CorrTables *ctp;
    SeriesGeom g;
    PairDesc pd;
    BleachDesc bd;
    const double *F;
    const int32_t *ind;
    const float *blk;     // [pair][nblk][m] block sums of the blocked variance
    int nblk;
    double *mcov;         // [pair][chanum] (channel 0 = 0), the GLS data
    double *W;            // [pair][m][m] inverse Cholesky factor (lower)
    double *yw;           // [pair][m] whitened data W mcov
    int32_t *ok;          // [pair] 1 = factorized
    double *cov;          // [pair][m][m] regularized covariance (optional)
};
void launch_gls_cov(const GlsArgs &A, int npairs, hipStream_t st);
// init: one parameter row per fit, or, with init_split >= 0, two rows (fits
// below init_split start from row 0, the others from row 1).
void launch_lm_gls(int nfits, int m, const float *x, const float *init, const int *pfit, float tol, int maxit,
                   const double *W, const double *yw, const int32_t *ok, float *out_params, int *states, float *chi2,
                   int *iters, hipStream_t st, int init_split = -1);

// Training curves of the learned initial guesses (imfcs_kernels.hip): model
// curves of FCS_3p at one fit geometry with random parameters and noise,
// normalised by their largest magnitude s, and the standardised targets
// (ln(N s), ln D, ln D2, F2, G_inf / s).
struct CurveGenArgs {
    LMModelGeom geom;          // micrometres
    const float *x;            // [npts] lag times (s)
    const float *nw;           // [npts] noise shape sqrt(mtab_1 / mtab_i)
    int npts;
    int ncomp;                 // 1: D2 = D, F2 = 0; 2: two components
    float lnN[2], lnD[2], lnD2[2];   // uniform ranges of ln N, ln D, ln D2 (um^2/s)
    float ginf;                // |G_inf| * N <= ginf
    float lg_noise[2];         // log10 range of the noise relative to 1/N
    const float *tmu, *tsd;    // [5] target standardisation
    uint64_t seed;
    int64_t base;              // index of the first curve (random stream)
    uint16_t *X;               // bf16 inputs [n][ldx] (columns >= npts zero)
    int ldx;
    float *Tg;                 // [n][5]
    int n;
};
void launch_mlp_curves(const CurveGenArgs &a, hipStream_t st);

// Learned initial guesses (imfcs_mlp.hip): writes one LM parameter row per
// fit into init ([n][20]): the template row with N, D, G_inf (and D2, F2 for
// a two-component network) replaced by the network's prediction from the
// fit's data curve ([n][npts], channels 1..npts); rows whose prediction is
// not finite and positive keep the template.
struct MlpNet;
MlpNet *mlp_of(imfcs_mlp *m);
int mlp_seed_init(MlpNet *net, const float *data, int n, const float *tmpl, float *init, hipStream_t st);
int mlp_points(const MlpNet *net);

// One group of FCS fits (N, D, G_inf free, FP64 chi-square, 65..128
// points) of one pipeline run, for the persistent fit launch over a batch.
struct LMGroup {
    int nfits, npts;
    const float *data, *weights, *x, *init;
    const int *pfit;
    float tol;
    int maxit;
    float *out_params;
    int *states;
    float *chi2;
    int *iters;
    int init_two_rows, init_split;
};
bool lm_queue_eligible(int npts, unsigned free_mask, int chi2_fp64);
// workgroups of the persistent fit kernel that are resident at once
int lm_queue_blocks();
// groups, prefix (fit offsets, ngroups + 1 entries) and counter on the device
void launch_lm_queue(const LMGroup *groups, const int64_t *prefix, int ngroups, int64_t total,
                     unsigned long long *counter, int blocks, hipStream_t st);

// Launchers (imfcs_kernels.hip)
void launch_copy_words(void *dst, const void *src, size_t bytes, hipStream_t st);
void launch_trace_moments(const SeriesGeom &g, int nopit, int ave, double tscale, double frametime,
                          int order, int integer_trace, double *part, int nchunk, hipStream_t st,
                          float *trace_out = nullptr);
// Per-pixel sums of every bleach-trace point over the whole uint16 frame,
// P[nopit][plane] (int32), and the trace moments of each slot taken from them.
void launch_point_sums(const uint16_t *data, int64_t plane, int f0, int nopit, int ave, int32_t *P,
                       hipStream_t st);
void launch_trace_points(const SeriesGeom &g, const int32_t *P, int nopit, int ave, double tscale,
                         double frametime, int order, double *part, int nchunk, hipStream_t st,
                         float *trace_out = nullptr);
// Exponential bleach fits (SingleExpFit / DoubleExpFit): initial guesses
// from the traces, the LM fit of every slot's trace (FP64 chi-square), and
// coef[slot] = {f(0), params} (f(0) = 0: no correction).
void launch_exp_init(const float *trace, const float *itime, int nslots, int nopit, int nexp, float *init,
                     hipStream_t st);
void launch_lm_exp(int nfits, int npts, const float *data, const float *x, const float *init, const int *pfit,
                   int nexp, float tol, int maxit, double *out64, int *states, float *chi2, int *iters,
                   hipStream_t st);
void launch_exp_pack(const double *out64, int nslots, int nexp, double *coef, hipStream_t st);
void launch_polyfit(const double *part, int nchunk, int nseries, int nopit, int ave, double tscale,
                    double frametime, int order, double *coef, hipStream_t st);
// l3 (level-3 values of every slot) and cmax (per-chunk max |X|) may be null.
void launch_apply(const SeriesGeom &g, const BleachDesc &bd, const CorrTables *dct, double *ctot, double *rec,
                  float *X, float *l3, float *cmax, int nchunk, int chunk_len, hipStream_t st);
// Linear-segment bleach correction: per slot the segment means bcTrace and
// the extrapolated start bcInt0, into coef[slot][0..nseg].
void launch_seg_means(const SeriesGeom &g, int seg_len, int nseg, double *coef, hipStream_t st);
// cmax / xexp (null: none): per-slot staging exponents of the matrix-core correlator.
void launch_prefix_combine(const CorrTables *dct, int nslots, const double *ctot, const double *rec,
                           int nchunk, int chunk_len, double *F, const float *cmax, int32_t *xexp, hipStream_t st);
// Sub-chunks (64 frames) of the bleach pass's prefix sums: ctot holds
// prefix_subchunks(T) totals per slot.
int prefix_subchunks(int T);

// One bleach pass for the binnings 1, 2 and 4 of one uint16 stack (the
// benchmark's and the diffusion law's nested binnings): the slot grids are
// the full pixel grids of the three binnings over the same raw rectangle
// from (0, 0) (grid b: pitch swp[b], sh[b] rows, slot (x, y) = window at
// (b x, b y)), each with its own polynomial correction, X, ctot and rec
// buffers and its own frame chunking of the prefix sums (the 64-frame
// sub-chunks, so the buffers hold the bits the binnings' own passes write).
struct NestedApply {
    const uint16_t *data;
    int64_t plane;        // elements per frame
    int32_t stride;       // elements per row
    int32_t f0, T;
    int32_t bg;           // background per raw pixel
    int32_t swp[3], sh[3];   // grids of binning 1, 2, 4
    BleachDesc bd[3];
    double *ctot[3], *rec[3];
    float *X[3];
};
void launch_apply_nested(const NestedApply &A, const CorrTables *dct, int order, int nchunk, int chunk_len,
                         hipStream_t st);
// Work split of the blocking pass: products in chunks of 256 (levels 0..7
// inside a chunk, levels >= 8 cascaded from the chunk totals).
struct BlockingPlan {
    int nsplit;            // grid.y
    int chunks_per_split;  // 256-product chunks per thread
    int nchunk_prod;       // chunks of the whole series (csum rows)
};
// threads_log2: about 2^threads_log2 threads in flight (the split of the
// frames into per-thread chunks; tuned per shape by the batch autotuner)
BlockingPlan plan_blocking(const CorrTables &ct, int npix, int threads_log2 = 19);
void launch_blocking(const SeriesGeom &g, const BleachDesc &bd, const CorrTables *dct, const PairDesc &pd,
                     const double *F, const float *X, const BlockingPlan &bp, double *bpart, double *csum,
                     hipStream_t st);
void launch_blocking_finalize(const CorrTables &ct, const CorrTables *dct, const SeriesGeom &g, const PairDesc &pd,
                              const double *F, const BlockingPlan &bp, const double *bpart, const double *csum,
                              int32_t *ind, int32_t *bis, hipStream_t st);
// Level c >= 1 of every slot (FP64) for the table-driven correlator, at
// lvg + off[c]: generic_level_values returns the values to allocate.
struct LevelOffsets {
    int64_t off[kMaxLevel];
};
size_t generic_level_values(const CorrTables &ct, int nslots, LevelOffsets &lo);
// Frames per time chunk of the table-driven correlator (part holds
// ceil(T / chunk) x npairs x chanum x 3 sums, then npairs x chanum x 2
// window means).
int generic_chunk_len(const CorrTables &ct);
void launch_correlate(const SeriesGeom &g, const BleachDesc &bd, const CorrTables &ct, const CorrTables *dct,
                      const PairDesc &pd, const double *F, const float *X, const int32_t *ind, double *corav,
                      double *bvar, double *lvg, const LevelOffsets &lo, double *part, hipStream_t st);
bool correlate_fast_supported(const CorrTables &ct);
int corr_fast_chunk_len(const CorrTables &ct, int npairs);
// xexp non-null: levels 0..2 on the matrix cores (k_corr_mfma; the level-3
// values must already be in l3), else the VALU kernel k_corr_main.
void launch_correlate_fast(const SeriesGeom &g, const BleachDesc &bd, const CorrTables &ct, const CorrTables *dct,
                           const PairDesc &pd, const double *F, const float *X, const int32_t *ind, float *part,
                           double *hpart, float *l3, float *lvd, int chunk_len, double *corav, double *bvar,
                           hipStream_t st, hipEvent_t ev_main_begin = nullptr, hipEvent_t ev_main_end = nullptr,
                           float *gls_blk = nullptr, int gls_nblk = 0, const int32_t *xexp = nullptr);
// Matrix-core correlator of levels 0..2 (imfcs_corr_mfma.hip): ACF, and
// CCF partners on the slot grid at x offset 0..4 in the same row.
bool corr_mfma_eligible(const SeriesGeom &g, const PairDesc &pd);
void launch_corr_mfma(const SeriesGeom &g, const BleachDesc &bd, const CorrTables *dct, const PairDesc &pd,
                      const double *F, const float *X, const int32_t *ind, const int32_t *xexp, int chunk_len,
                      int nchunk, float *part, float *gls, int gls_nblk, hipStream_t st);
// FP64 polish of the converged N, D, G_inf fits of the one-component model
// (up to 128 points): rewrites their parameters and chi-square in place.
void launch_lm_polish(int nfits, int npts, const float *data, const float *weights, const float *x,
                      float *out_params, int *states, float *chi2, hipStream_t st);
void launch_lm_polish_groups(const LMGroup *groups, const int64_t *prefix, int ngroups, int64_t total,
                             hipStream_t st);
void launch_lm_acf(int nfits, int npts, const float *data, const float *weights, const float *x,
                   const float *init, const int *pfit, unsigned free_mask, const LMModelGeom &geom, float tol, int maxit,
                   int chi2_fp64, float *out_params, int *states, float *chi2, int *iters, hipStream_t st,
                   int init_split = -1, int only_failed = 0);
void launch_lm_poly(int nfits, int npts, const float *data, const float *weights, const float *x,
                    const float *init, const int *pfit, unsigned free_mask, int ncoef, float tol, int maxit,
                    int chi2_fp64, float *out_params, int *states, float *chi2, int *iters, hipStream_t st);
void launch_binning_f32(const float *in, float *out, int win, int hin, int wt, int ht, int T, int bx, int by,
                        hipStream_t st);
void launch_bleach_average_f32(const float *in, float *out, int wt, int ht, int nopit, int ave,
                               hipStream_t st);
void launch_copy(void *dst, const void *src, size_t bytes, hipStream_t st);
// dst[i] = (first ? 0 : dst[i]) + src[i] / div
void launch_accum_div(double *dst, const double *src, int64_t n, double div, int first, hipStream_t st);
void launch_nb_extract(const CorrTables *dct, const SeriesGeom &g, const PairDesc &pd, const double *F,
                       const double *corav, double *nbmean, double *nbcov, hipStream_t st);
void launch_pack_fit_inputs(int chanum, int npairs, const double *corav, const double *bvar,
                            float *data, float *weights, hipStream_t st);

}  // namespace imfcs
