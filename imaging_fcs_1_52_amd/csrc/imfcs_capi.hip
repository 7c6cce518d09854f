// Host side of the C ABI declared in include/imfcs.h.
//
// The drop-in functions keep the argument meaning and error behaviour of the
// reference's JNI natives (agpufitjni.cu:4595-5831) but run on persistent
// device buffers instead of allocating, copying and resetting the device per
// call.  The native pipeline (imfcs_run) works on a device-resident uint16
// stack and never leaves the device between stages.
#include "../../include/imfcs.h"
#include "imfcs_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

using namespace imfcs;

namespace {

thread_local std::string g_last_error = "";

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(IMFCS_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Carving of device buffers from the scratch array during an
// imfcs_run_noalloc call on this thread (null otherwise).
struct Arena {
    unsigned char *base;
    size_t size, off, need;
    bool overflow;
};
thread_local Arena *g_arena = nullptr;

// A growable device buffer: hipMalloc'd, or carved from the scratch arena
// while one is active (then nothing is freed or allocated).
struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= n) return hipSuccess;
        if (g_arena) {
            const size_t a = (g_arena->off + 255) & ~(size_t)255;
            if (a + bytes > g_arena->size) {
                g_arena->overflow = true;
                g_arena->need = a + bytes;
                return hipErrorOutOfMemory;
            }
            p = g_arena->base + a;
            n = bytes;
            g_arena->off = a + bytes;
            return hipSuccess;
        }
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) n = bytes;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    template <class T>
    T *as() const { return (T *)p; }
};

// Host-to-device copy of a small per-run parameter block, skipped when the
// device buffer still holds the same bytes from an earlier run (repeated
// runs of one configuration then enqueue no copies).  The shadow is the
// copy's source and is only rewritten after the previous run's copies have
// completed (the run start waits for them).
struct CachedUpload {
    std::vector<unsigned char> host;
    void *dev = nullptr;
    void invalidate() { dev = nullptr; }
    hipError_t put(DevBuf &buf, const void *src, size_t n, hipStream_t st) {
        const bool grows = n > buf.n;
        hipError_t e = buf.ensure(n);
        if (e != hipSuccess) return e;
        if (!grows && dev == buf.p && host.size() == n && std::memcmp(host.data(), src, n) == 0) return hipSuccess;
        host.assign((const unsigned char *)src, (const unsigned char *)src + n);
        dev = buf.p;
        return hipMemcpyAsync(buf.p, host.data(), n, hipMemcpyHostToDevice, st);
    }
};

// Correlator tables exactly as setParameters builds them (IFCS:7391-7424).
struct HostTables {
    int chanum = 0;
    std::vector<int> lag;
    std::vector<double> samp, mtab, lagtime;
};

int make_host_tables(int p, int q, int num, double frametime, HostTables &t) {
    if (p < 4 || p % 2 != 0 || q < 1) return fail(IMFCS_ERR_ARG, "correlator p must be even and >= 4, q >= 1");
    int base = p, hbase = p / 2, lagnum = q;
    t.chanum = p + (q - 1) * p / 2 + 1;
    if (t.chanum > kMaxChan) return fail(IMFCS_ERR_ARG, "too many correlator channels");
    t.lag.assign(t.chanum, 0);
    t.samp.assign(t.chanum, 0);
    t.mtab.assign(t.chanum, 0);
    t.lagtime.assign(t.chanum, 0);
    for (int x = 0; x <= hbase; x++) t.lag[x] = x;
    for (int x = 1; x <= lagnum; x++)
        for (int y = 1; y <= hbase; y++)
            t.lag[x * hbase + y] = (int)(std::pow(2.0, x - 1) * y + (base / 4) * std::pow(2.0, x));
    for (int x = 0; x <= base; x++) t.samp[x] = 1;
    for (int x = 2; x <= lagnum; x++)
        for (int y = 1; y <= hbase; y++) t.samp[x * hbase + y] = std::pow(2.0, x - 1);
    for (int x = 0; x < t.chanum; x++) {
        t.mtab[x] = (int)std::floor((num - t.lag[x]) / t.samp[x]);
        t.lagtime[x] = t.lag[x] * frametime;
    }
    return 0;
}

// Derives the level layout, prefix positions and blocking constants.
int make_corr_tables(const HostTables &h, int num, int q, CorrTables &ct) {
    std::memset(&ct, 0, sizeof(ct));
    ct.chanum = h.chanum;
    ct.T = num;
    ct.q = q;
    int nlev = 0;
    for (int x = 0; x < h.chanum; x++) {
        int s = (int)h.samp[x];
        int c = 0;
        while ((1 << c) < s) c++;
        if ((1 << c) != s) return fail(IMFCS_ERR_ARG, "sample times must be powers of two");
        ct.level[x] = c;
        ct.del[x] = h.lag[x] / s;
        nlev = std::max(nlev, c + 1);
    }
    if (nlev > kMaxLevel) return fail(IMFCS_ERR_ARG, "too many correlator levels");
    ct.nlevels = nlev;
    for (int c = 0; c < nlev; c++) {
        ct.first_chan[c] = -1;
        ct.nchan[c] = 0;
        ct.numbin[c] = num >> c;
    }
    for (int x = 0; x < h.chanum; x++) {
        int c = ct.level[x];
        if (ct.first_chan[c] < 0) ct.first_chan[c] = x;
        ct.nchan[c]++;
    }
    // every level holds a contiguous channel range with ascending delays
    // (p = 2 mod 4 repeats one level-0 delay, IFCS:7398-7408 with the
    // integer base / 4); the fast kernels implement the p = 16 layout, level
    // 0 delays 0..16 and every other level delays 9..16
    ct.p16 = 1;
    for (int c = 0; c < nlev; c++) {
        const int c0 = ct.first_chan[c], n = ct.nchan[c];
        if (n <= 0) return fail(IMFCS_ERR_ARG, "correlator level without channels");
        for (int i = 0; i < n; i++) {
            if (ct.level[c0 + i] != c) return fail(IMFCS_ERR_ARG, "correlator channels out of level order");
            if (i > 0 && ct.del[c0 + i] < ct.del[c0 + i - 1]) return fail(IMFCS_ERR_ARG, "unexpected lag table");
        }
        if (ct.numbin[c] - ct.del[c0 + n - 1] < 2) return fail(IMFCS_ERR_ARG, "not enough frames for the correlator");
        const int nexp = (c == 0) ? 17 : 8, dmin = (c == 0) ? 0 : 9;
        if (n != nexp || ct.del[c0] != dmin || ct.del[c0 + n - 1] != dmin + n - 1) ct.p16 = 0;
    }
    ct.mtab_last = h.mtab[h.chanum - 1];
    ct.log2_samp_last = std::log(h.samp[h.chanum - 1]) / std::log(2.0);
    ct.blocknum = (int)std::floor(std::log(h.mtab[1]) / std::log(2.0)) - 2;
    if (ct.blocknum < 4 || ct.blocknum > kMaxBlock) return fail(IMFCS_ERR_ARG, "frame count outside the blocking range");
    // prefix positions
    std::vector<int> pos;
    for (int x = 0; x < h.chanum; x++) {
        int c = ct.level[x];
        pos.push_back((ct.numbin[c] - ct.del[x]) << c);
        pos.push_back(ct.del[x] << c);
        pos.push_back(ct.numbin[c] << c);
    }
    pos.push_back(num - 2);
    pos.push_back(num - 1);
    pos.push_back(num);
    std::sort(pos.begin(), pos.end());
    pos.erase(std::unique(pos.begin(), pos.end()), pos.end());
    if ((int)pos.size() > kMaxPos) return fail(IMFCS_ERR_ARG, "too many prefix positions");
    ct.npos = (int)pos.size();
    for (int i = 0; i < ct.npos; i++) ct.pos[i] = pos[i];
    auto idx = [&](int v) { return (int16_t)(std::lower_bound(pos.begin(), pos.end(), v) - pos.begin()); };
    for (int x = 0; x < h.chanum; x++) {
        int c = ct.level[x];
        ct.pos_a_end[x] = idx((ct.numbin[c] - ct.del[x]) << c);
        ct.pos_b_beg[x] = idx(ct.del[x] << c);
        ct.pos_b_end[x] = idx(ct.numbin[c] << c);
    }
    ct.pos_blk_a = idx(num - 2);
    ct.pos_blk_b = idx(num - 1);
    ct.pos_total = idx(num);
    return 0;
}

// The channel scheme of setParameters (IFCS:7391-7424): p + 1 delays in the
// first group, p / 2 in each of the q - 1 further groups.  The fast kernels
// implement the plugin's default p = 16; every other even p runs the
// table-driven correlator k_correlate.  Odd p leaves channels of Java's
// tables unset (chanum counts p / 2 as a fraction) and is refused, as is a
// layout of more than kMaxChan channels, before work is enqueued.
int check_correlator_layout(int p, int q) {
    if (p < 4 || p % 2 != 0)
        return fail(IMFCS_ERR_ARG, "correlator p = " + std::to_string(p) + " is not supported: p must be even and >= 4");
    if (q < 1 || q > kMaxLevel)
        return fail(IMFCS_ERR_ARG, "correlator q = " + std::to_string(q) + " outside 1.." + std::to_string(kMaxLevel));
    if (p + (int64_t)(q - 1) * p / 2 + 1 > kMaxChan)
        return fail(IMFCS_ERR_ARG, "correlator p = " + std::to_string(p) + ", q = " + std::to_string(q) + " gives more than " +
                                       std::to_string(kMaxChan) + " channels");
    return 0;
}

bool device_present() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return false;
    return n > 0;
}

}  // namespace

int imfcs::set_error(int code, const std::string &msg) { return fail(code, msg); }

// Scratch of the allocation-free entry point (imfcs_run_noalloc): every
// device buffer of one pipeline run is carved from it.  Each GPU holds its
// own copy until the process ends; its address on each device is looked up
// once, with that device current (imfcs_create), and kept in g_scratch_ptr:
// host code uses that address, never the array's name.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::vector<void *> g_scratch_ptr;   // per device

// With `device` current: records the device address of scratch.
int scratch_setup(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (g_scratch_ptr[device]) return 0;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(IMFCS_ERR_DEVICE, std::string("scratch address: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
    return 0;
}

// Frames-per-chunk choices of the bleach-correction / prefix pass and of the
// blocking pass, as the exponent of the threads kept in flight (2^e), per
// series shape (slots, frames).  The batch autotuner times candidates and
// records the fastest; every later run of that shape (batched or not) uses
// them, so batched and single runs compute identically.
struct Tune {
    int apply_log2 = 19, block_log2 = 19;
};
std::mutex g_tune_mu;
std::map<std::pair<int, int>, Tune> g_tune;

Tune tune_lookup(int nslots, int T) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find({nslots, T});
    return it == g_tune.end() ? Tune{} : it->second;
}

void tune_store(int nslots, int T, const Tune &t) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[{nslots, T}] = t;
}

void *scratch_ptr(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return device < (int)g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}
struct EngineArgs {
    SeriesGeom g;
    int order;             // polynomial order (0 = none)
    int gls;               // 1: regularized covariance of every pair (calculateCF GLS branch)
    int fit_trace;         // 1: trace + fit on device; 0: coefficients given in host_coef
    const double *host_coef;   // [nseries][order+1] in seconds (drop-in)
    int nopit, ave;
    double frametime;
    // bleach abscissa t = frametime * (frame + t_off), frame 0-based:
    // 1.5 = the Java CPU path's 1-based (x + 0.5) (IFCS:9747), 0.5 = the
    // reference GPU kernel's 0-based (i + 0.5) (agpufitjni.cu:5108)
    double t_off;
    int p, q;
    PairDesc pd;
    int seg_swl;           // > 0: "Lin Segment" correction with this sliding window length
    int exp_n;             // 1: "Single Exp" correction
    // per-pixel point sums of the stack (P[nopit][plane], this run's first
    // frame, nopit and ave) to take the trace from instead of the stack
    const int32_t *psums;
};

// A pipeline run stopped before its bleach pass (stage 1 of imfcs_run_multi
// and the batches, whose nested binnings share one pass over the stack): what
// the rest of the run (stage 2) takes up.
struct StagedRun {
    EngineArgs ea;          // with the slot grid of the run
    BleachDesc bd;
    HostTables ht;
    CorrTables ct;
    Tune tn;
    int plen = 0, npc = 0;  // frame chunks of the bleach pass
    bool mfma = false;
};
}  // namespace

// ------------------------------------------------------------------
// Context
// ------------------------------------------------------------------
struct imfcs_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};
    hipEvent_t ev_h2d = nullptr;   // after the last host-to-device copy of a run
    hipEvent_t evk[2] = {};        // around the main correlator kernel
    int evk_valid = 0;
    // host staging of a run's fit parameters and the cached uploads of the
    // per-run parameter blocks
    std::vector<float> h_xs, h_init;
    std::vector<int> h_pfit;
    CachedUpload up_ctab, up_slots, up_fx, up_finit, up_fpfit, up_eitime, up_epfit;
    std::vector<float> h_eitime;
    std::vector<int> h_epfit;
    DevBuf swacf, swvar;   // sliding-window averages
    DevBuf psums;          // per-pixel bleach-point sums shared by imfcs_run_multi
    hipEvent_t ev_psums = nullptr;
    hipEvent_t ev_stage = nullptr;   // end of a staged run's first part / of a nested bleach pass
    DevBuf etrace, eitime, epfit, einit, eout, estate, echi2, eiter;   // exponential bleach fits
    DevBuf part, coef, ctot, rec, F, bpart, ind, bis, corav, bvar, fdata, fw, fx, finit, fpfit, fout, fstate, fchi2,
        fiter, cpart, hpart, l3, lvd, ctab, csum, slots, X, glsblk, glsW, glsy, glsok, glscov, dropin_in, nbm, nbc,
        cmax, xexp, lvg;
    SeriesGeom last_g{};   // geometry (with device slot tables) of the last engine run
    std::vector<int32_t> host_slots;   // slot_of then slot origins
    std::vector<double> host_coef;     // drop-in bleach coefficients per slot (source of the async copy)   // source of the async copy into ctab
    // last run
    int npix = 0, npairs = 0, nseries = 0, nslots = 0, chanum = 0, order = 0, ow = 0, oh = 0, did_fit = 0, nfree = 0;
    int did_gls = 0;
    int force_v1 = 0;   // IMFCS_CORRELATOR=generic selects the one-thread-per-level kernel
    int want_mfma = 0;   // IMFCS_CORRELATOR=mfma: levels 0..2 on the matrix cores (k_corr_mfma)
    double chi2_norm[2] = {1.0, 1.0};   // ACF fits, CCF fits
    int fit_split = 0;                  // fits below are ACF fits
    std::vector<double> coef_scale;   // u-basis -> seconds conversion per power
    // what imfcs_refit needs of the last run: its configuration and lag times
    imfcs_config last_cfg{};
    int has_last = 0;
    std::vector<double> last_lagtime;
    // learned initial guesses for the ACF and the CCF fits (or null), and the
    // per-fit starting rows they write
    imfcs::MlpNet *seeder[2] = {nullptr, nullptr};
    DevBuf fseed;
    // 1: the last run was imfcs_run_noalloc (its results went to the
    // caller's buffers; nothing of it stays in the context)
    int last_noalloc = 0;
    // autotuner trial: chunking forced for the next runs (apply_log2 > 0)
    Tune force_tune{0, 0};
    // the chunking the last run used
    Tune used_tune{};
    // 1: run_engine stops before the bleach pass, keeping its state in
    // staged; 2: it continues from there, the pass (X, ctot, rec) done by
    // the caller; 3: it continues with the pass; 0: whole runs
    int stage = 0;
    StagedRun staged;
    int run_nested = 0;   // the last run's bleach pass was a nested pass
};

hipStream_t imfcs::ctx_stream(imfcs_ctx *ctx) { return ctx->stream; }
int imfcs::ctx_device(imfcs_ctx *ctx) { return ctx->device; }

extern "C" {

const char *imfcs_get_last_error(void) { return g_last_error.c_str(); }

int imfcs_device_available(void) { return device_present() ? 1 : 0; }

int imfcs_get_versions(int *runtime_version, int *driver_version) {
    int rv = 0, dv = 0;
    HIP_TRY(hipRuntimeGetVersion(&rv));
    HIP_TRY(hipDriverGetVersion(&dv));
    // HIP encodes major*10000000 + minor*100000 + patch
    auto enc = [](int v) { return (v / 10000000) * 1000 + ((v / 100000) % 100) * 10; };
    if (runtime_version) *runtime_version = enc(rv);
    if (driver_version) *driver_version = enc(dv);
    return 0;
}

// Correlator depth q of a run: with sliding windows, setParameters lowers it
// to what a window of L frames supports, lagnum = floor((ln(L / (20 + p)) +
// 1) / ln 2) (swMinFrameReq = 20, IFCS:203, 7331-7341).
static int effective_q(const imfcs_config *c) {
    if (c->bleach_mode != IMFCS_BLEACH_SLIDING_WINDOW) return c->correlator_q;
    const double L = (double)c->sliding_window_length;
    const int lagnum = (int)std::floor((std::log(L / (20.0 + c->correlator_p)) + 1.0) / std::log(2.0));
    return std::min(c->correlator_q, lagnum);
}

int imfcs_plan(const imfcs_config *c, int32_t *out_w, int32_t *out_h, int32_t *chanum) {
    if (!c) return fail(IMFCS_ERR_ARG, "null config");
    if (c->pixbin_x <= 0 || c->pixbin_y <= 0 || c->binning_x <= 0 || c->binning_y <= 0)
        return fail(IMFCS_ERR_ARG, "binning and pixbin must be positive");
    const int rc = check_correlator_layout(c->correlator_p, c->correlator_q);
    if (rc) return rc;
    // correlateROI ROI mapping (IFCS:8458-8466)
    int sx = (int)std::ceil((double)c->roi_x / c->pixbin_x);
    int sy = (int)std::ceil((double)c->roi_y / c->pixbin_y);
    int ex = (int)std::floor((double)(c->roi_x + c->roi_w - c->binning_x) / c->pixbin_x);
    int ey = (int)std::floor((double)(c->roi_y + c->roi_h - c->binning_y) / c->pixbin_y);
    if (out_w) *out_w = ex - sx + 1;
    if (out_h) *out_h = ey - sy + 1;
    if (chanum) *chanum = c->correlator_p + (effective_q(c) - 1) * c->correlator_p / 2 + 1;
    return 0;
}

imfcs_ctx *imfcs_create(int device) {
    if (!device_present()) {
        fail(IMFCS_ERR_DEVICE, "no HIP device available");
        return nullptr;
    }
    if (hipSetDevice(device) != hipSuccess) {
        fail(IMFCS_ERR_DEVICE, "hipSetDevice failed");
        return nullptr;
    }
    // the scratch address of this device (imfcs_run_noalloc); a failure
    // leaves the other entry points usable and is reported by that call
    (void)scratch_setup(device);
    imfcs_ctx *ctx = new imfcs_ctx();
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        fail(IMFCS_ERR_DEVICE, "hipStreamCreate failed");
        return nullptr;
    }
    for (auto &e : ctx->ev) (void)hipEventCreate(&e);
    (void)hipEventCreateWithFlags(&ctx->ev_h2d, hipEventDisableTiming);
    (void)hipEventCreateWithFlags(&ctx->ev_psums, hipEventDisableTiming);
    (void)hipEventCreateWithFlags(&ctx->ev_stage, hipEventDisableTiming);
    for (auto &e : ctx->evk) (void)hipEventCreate(&e);
    const char *sel = std::getenv("IMFCS_CORRELATOR");
    ctx->force_v1 = (sel && std::strcmp(sel, "generic") == 0) ? 1 : 0;
    ctx->want_mfma = (sel && std::strcmp(sel, "mfma") == 0) ? 1 : 0;
    return ctx;
}

static std::vector<DevBuf *> all_bufs(imfcs_ctx *ctx) {
    return {&ctx->part, &ctx->coef, &ctx->ctot, &ctx->rec, &ctx->F, &ctx->bpart, &ctx->ind,
                      &ctx->bis, &ctx->corav, &ctx->bvar, &ctx->fdata, &ctx->fw, &ctx->fx, &ctx->finit,
                      &ctx->fpfit, &ctx->fout, &ctx->fstate, &ctx->fchi2, &ctx->fiter, &ctx->cpart,
                      &ctx->hpart, &ctx->l3, &ctx->lvd, &ctx->ctab, &ctx->csum, &ctx->slots,
                      &ctx->X, &ctx->glsblk, &ctx->glsW, &ctx->glsy, &ctx->glsok, &ctx->glscov,
                      &ctx->dropin_in, &ctx->nbm, &ctx->nbc, &ctx->swacf, &ctx->swvar, &ctx->etrace,
                      &ctx->eitime, &ctx->epfit, &ctx->einit, &ctx->eout, &ctx->estate, &ctx->echi2,
                      &ctx->eiter, &ctx->psums, &ctx->fseed, &ctx->cmax, &ctx->xexp, &ctx->lvg};
}

static void invalidate_uploads(imfcs_ctx *ctx) {
    CachedUpload *ups[] = {&ctx->up_ctab, &ctx->up_slots, &ctx->up_fx, &ctx->up_finit, &ctx->up_fpfit,
                           &ctx->up_eitime, &ctx->up_epfit};
    for (CachedUpload *u : ups) u->invalidate();
}

static void release_all(imfcs_ctx *ctx) {
    for (DevBuf *b : all_bufs(ctx)) b->release();
}

void imfcs_destroy(imfcs_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    release_all(ctx);
    for (auto &e : ctx->ev) (void)hipEventDestroy(e);
    if (ctx->ev_h2d) (void)hipEventDestroy(ctx->ev_h2d);
    if (ctx->ev_psums) (void)hipEventDestroy(ctx->ev_psums);
    if (ctx->ev_stage) (void)hipEventDestroy(ctx->ev_stage);
    for (auto &e : ctx->evk) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

void *imfcs_device_alloc(imfcs_ctx *ctx, size_t bytes) {
    void *p = nullptr;
    if (!ctx) return nullptr;
    (void)hipSetDevice(ctx->device);
    if (hipMalloc(&p, bytes) != hipSuccess) {
        fail(IMFCS_ERR_MEMORY, "hipMalloc failed");
        return nullptr;
    }
    return p;
}

int imfcs_device_free(imfcs_ctx *ctx, void *ptr) {
    if (!ctx) return fail(IMFCS_ERR_ARG, "null context");
    HIP_TRY(hipFree(ptr));
    return 0;
}

void *imfcs_host_alloc(imfcs_ctx *ctx, size_t bytes) {
    void *p = nullptr;
    if (!ctx) return nullptr;
    (void)hipSetDevice(ctx->device);
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
        fail(IMFCS_ERR_MEMORY, "hipHostMalloc failed");
        return nullptr;
    }
    return p;
}

int imfcs_host_free(imfcs_ctx *ctx, void *ptr) {
    if (!ctx) return fail(IMFCS_ERR_ARG, "null context");
    HIP_TRY(hipHostFree(ptr));
    return 0;
}

int imfcs_memcpy_h2d(imfcs_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx) return fail(IMFCS_ERR_ARG, "null context");
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int imfcs_memcpy_d2h(imfcs_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx) return fail(IMFCS_ERR_ARG, "null context");
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int imfcs_memcpy_d2d(imfcs_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx) return fail(IMFCS_ERR_ARG, "null context");
    launch_copy(dst, src, bytes, ctx->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int imfcs_memcpy_d2d_words(imfcs_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx) return fail(IMFCS_ERR_ARG, "null context");
    if (bytes % 4) return fail(IMFCS_ERR_ARG, "the word copy takes a multiple of 4 bytes");
    launch_copy_words(dst, src, bytes, ctx->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int imfcs_synchronize(imfcs_ctx *ctx) {
    if (!ctx) return fail(IMFCS_ERR_ARG, "null context");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

}  // extern "C"

// ------------------------------------------------------------------
// Shared correlation engine (native pipeline and drop-in calcACF)
// ------------------------------------------------------------------
namespace {

// Distinct series ("slots").  Slots form a row-major grid over the output
// pixels extended by the CCF partner offset when that offset is a multiple
// of the pixel pitch (then every partner is a grid point and shares the
// slot of the pixel at that point); otherwise the partners get a second
// grid of their own.  Grid rows are padded to SWP slots (a multiple of 64,
// or a power of two below 64) so that the 64 lanes of a wave read whole,
// aligned row segments of the stack: unaligned uint16 rows cost ~10x in
// load throughput.  Padding slots repeat the last window of their row.
// buf = slot_of[nseries], the window origins [nslots][2], then the owner
// flags [nseries] (first series of each slot).
static int slot_pitch(const SeriesGeom &g) {
    const bool ccf = g.nseries > g.npix;
    const bool on_grid = ccf && (g.cfx % g.pbx == 0) && (g.cfy % g.pby == 0);
    const int sw = g.ow + (on_grid ? g.cfx / g.pbx : 0);
    int swp = 1;
    while (swp < sw && swp < 64) swp <<= 1;
    if (sw > 64) swp = (sw + 63) / 64 * 64;
    return swp;
}

int build_slots(const SeriesGeom &g, std::vector<int32_t> &buf) {
    const bool ccf = g.nseries > g.npix;
    const bool on_grid = ccf && (g.cfx % g.pbx == 0) && (g.cfy % g.pby == 0);
    const int dx = on_grid ? g.cfx / g.pbx : 0, dy = on_grid ? g.cfy / g.pby : 0;
    const int sw = g.ow + dx, sh = g.oh + dy;
    const int swp = slot_pitch(g);
    const int ngrid = (ccf && !on_grid) ? 2 : 1;
    const int nslots = ngrid * sh * swp;
    buf.assign(2 * (size_t)g.nseries + 2 * (size_t)nslots, 0);
    int32_t *org = buf.data() + g.nseries;
    for (int gi = 0; gi < ngrid; gi++)
        for (int y = 0; y < sh; y++)
            for (int xs = 0; xs < swp; xs++) {
                const int x = std::min(xs, sw - 1);
                const size_t k = ((size_t)gi * sh + y) * swp + xs;
                org[2 * k] = g.ox0 + x * g.pbx + (gi ? g.cfx : 0);
                org[2 * k + 1] = g.oy0 + y * g.pby + (gi ? g.cfy : 0);
            }
    for (int k = 0; k < g.npix; k++) {
        const int x = k % g.ow, y = k / g.ow;
        buf[k] = y * swp + x;
        if (ccf) buf[g.npix + k] = on_grid ? (y + dy) * swp + (x + dx) : (sh + y) * swp + x;
    }
    std::vector<char> seen(nslots, 0);
    int32_t *own = org + 2 * (size_t)nslots;
    for (int s = 0; s < g.nseries; s++) {
        own[s] = !seen[buf[s]];
        seen[buf[s]] = 1;
    }
    return nslots;
}


// The run from its bleach pass on (the pass itself unless the caller did it:
// stage 2 of a staged run), with the state its first part left in staged.
int engine_rest(imfcs_ctx *ctx, bool do_apply) {
    hipStream_t st = ctx->stream;
    StagedRun &S = ctx->staged;
    EngineArgs &ea = S.ea;
    SeriesGeom &g = ea.g;
    const BleachDesc &bd = S.bd;
    const CorrTables &ct = S.ct;
    const CorrTables *dct = ctx->ctab.as<CorrTables>();
    const Tune tn = S.tn;
    const int ns = g.nslots, np = ea.pd.npairs, plen = S.plen, npc = S.npc;
    const bool mfma = S.mfma;
    if (do_apply)
        launch_apply(g, bd, dct, ctx->ctot.as<double>(), ctx->rec.as<double>(), ctx->X.as<float>(),
                     mfma ? ctx->l3.as<float>() : nullptr, mfma ? ctx->cmax.as<float>() : nullptr, npc, plen, st);
    launch_prefix_combine(dct, ns, ctx->ctot.as<double>(), ctx->rec.as<double>(), npc, plen, ctx->F.as<double>(),
                          mfma ? ctx->cmax.as<float>() : nullptr, mfma ? ctx->xexp.as<int32_t>() : nullptr, st);
    HIP_TRY(hipEventRecord(ctx->ev[2], st));
    // blocking
    const BlockingPlan bp = plan_blocking(ct, g.npix, tn.block_log2);
    HIP_TRY(ctx->bpart.ensure(sizeof(double) * (size_t)bp.nsplit * np * 8 * 2));
    HIP_TRY(ctx->csum.ensure(sizeof(double) * (size_t)std::max(1, bp.nchunk_prod) * np));
    HIP_TRY(ctx->ind.ensure(sizeof(int32_t) * (size_t)np));
    HIP_TRY(ctx->bis.ensure(sizeof(int32_t) * (size_t)np));
    launch_blocking(g, bd, dct, ea.pd, ctx->F.as<double>(), ctx->X.as<float>(), bp, ctx->bpart.as<double>(),
                    ctx->csum.as<double>(), st);
    launch_blocking_finalize(ct, dct, g, ea.pd, ctx->F.as<double>(), bp, ctx->bpart.as<double>(),
                             ctx->csum.as<double>(), ctx->ind.as<int32_t>(), ctx->bis.as<int32_t>(), st);
    HIP_TRY(hipEventRecord(ctx->ev[3], st));
    // correlation
    HIP_TRY(ctx->corav.ensure(sizeof(double) * (size_t)np * ct.chanum));
    HIP_TRY(ctx->bvar.ensure(sizeof(double) * (size_t)np * ct.chanum));
    const int m = ct.chanum - 1, nblk = (int)ct.mtab_last;
    ctx->did_gls = 0;
    if (ea.gls) {
        if (!correlate_fast_supported(ct) || ctx->force_v1)
            return fail(IMFCS_ERR_ARG, "GLS needs the p = 16 correlator layout with blocknum >= q");
        if (m > kGlsMaxM) return fail(IMFCS_ERR_ARG, "GLS supports at most 104 fitted channels");
        HIP_TRY(ctx->glsblk.ensure(sizeof(float) * (size_t)np * nblk * m));
        HIP_TRY(ctx->glsW.ensure(sizeof(double) * (size_t)np * m * m));
        HIP_TRY(ctx->glscov.ensure(sizeof(double) * (size_t)np * m * m));
        HIP_TRY(ctx->glsy.ensure(sizeof(double) * (size_t)np * m));
        HIP_TRY(ctx->glsok.ensure(sizeof(int32_t) * (size_t)np));
    }
    if (correlate_fast_supported(ct) && !ctx->force_v1) {
        const int clen = corr_fast_chunk_len(ct, np);
        const int nch = (g.T + clen - 1) / clen;
        HIP_TRY(ctx->cpart.ensure(sizeof(float) * (size_t)nch * np * 33 * 3));
        HIP_TRY(ctx->hpart.ensure(sizeof(double) * (size_t)nch * np * (ct.chanum - 33) * 3));
        size_t nlv = 0;
        for (int c = 4; c < ct.nlevels; c++) nlv += (size_t)ct.numbin[c] + kLevelPad;
        HIP_TRY(ctx->l3.ensure(sizeof(float) * ((size_t)ct.numbin[3] + kLevelPad) * g.nslots));
        HIP_TRY(ctx->lvd.ensure(sizeof(float) * std::max<size_t>(nlv, 1) * g.nslots));
        launch_correlate_fast(g, bd, ct, dct, ea.pd, ctx->F.as<double>(), ctx->X.as<float>(), ctx->ind.as<int32_t>(),
                              ctx->cpart.as<float>(), ctx->hpart.as<double>(), ctx->l3.as<float>(),
                              ctx->lvd.as<float>(), clen,
                              ctx->corav.as<double>(), ctx->bvar.as<double>(), st, ctx->evk[0], ctx->evk[1],
                              ea.gls ? ctx->glsblk.as<float>() : nullptr, nblk,
                              mfma ? ctx->xexp.as<int32_t>() : nullptr);
        ctx->evk_valid = 1;
        if (ea.gls) {
            // mcov replaces the correlation as Java's GLS output (IFCS:9638-9639)
            GlsArgs A{dct, g, ea.pd, bd, ctx->F.as<double>(), ctx->ind.as<int32_t>(), ctx->glsblk.as<float>(), nblk,
                      ctx->corav.as<double>(), ctx->glsW.as<double>(), ctx->glsy.as<double>(),
                      ctx->glsok.as<int32_t>(), ctx->glscov.as<double>()};
            launch_gls_cov(A, np, st);
            ctx->did_gls = 1;
        }
    } else {
        LevelOffsets lo;
        const size_t nlv = generic_level_values(ct, g.nslots, lo);
        HIP_TRY(ctx->lvg.ensure(sizeof(double) * std::max<size_t>(nlv, 1)));
        const int gclen = generic_chunk_len(ct);
        const size_t gnch = (size_t)((g.T + gclen - 1) / gclen);
        HIP_TRY(ctx->hpart.ensure(sizeof(double) * (gnch * 3 + 2) * np * ct.chanum));
        HIP_TRY(hipEventRecord(ctx->evk[0], st));
        launch_correlate(g, bd, ct, dct, ea.pd, ctx->F.as<double>(), ctx->X.as<float>(), ctx->ind.as<int32_t>(),
                         ctx->corav.as<double>(), ctx->bvar.as<double>(), ctx->lvg.as<double>(), lo,
                         ctx->hpart.as<double>(), st);
        HIP_TRY(hipEventRecord(ctx->evk[1], st));
        ctx->evk_valid = 1;
    }
    HIP_TRY(hipEventRecord(ctx->ev[4], st));
    HIP_TRY(hipGetLastError());
    ctx->last_g = g;
    return 0;
}

int run_engine(imfcs_ctx *ctx, const EngineArgs &ea_in, HostTables &ht, CorrTables &ct) {
    if (ctx->stage >= 2) {   // continue a staged run (2: its bleach pass done by the caller)
        ht = ctx->staged.ht;
        ct = ctx->staged.ct;
        return engine_rest(ctx, ctx->stage == 3);
    }
    ctx->run_nested = 0;
    hipStream_t st = ctx->stream;
    EngineArgs ea = ea_in;
    SeriesGeom &g = ea.g;
    int rc = make_host_tables(ea.p, ea.q, g.T, ea.frametime, ht);
    if (rc) return rc;
    rc = make_corr_tables(ht, g.T, ea.q, ct);
    if (rc) return rc;
    // the host sources of the previous run's copies are reused below (a
    // stream being captured into a graph has drained its earlier work: the
    // batch synchronizes before capturing)
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(st, &cap));
    if (cap == hipStreamCaptureStatusNone) HIP_TRY(hipEventSynchronize(ctx->ev_h2d));
    HIP_TRY(ctx->up_ctab.put(ctx->ctab, &ct, sizeof(CorrTables), st));
    g.nslots = build_slots(g, ctx->host_slots);
    g.swp = slot_pitch(g);
    // vectorized window rows: every slot origin, the row stride and the
    // frame size are multiples of the window width (uint16 source)
    g.vec = 0;
    if (g.src == kSrcU16 && (g.bx == 2 || g.bx == 4) && g.stride % g.bx == 0 && g.plane % g.bx == 0 &&
        (uintptr_t)g.data % (2 * g.bx) == 0) {
        bool aligned = true;
        const int32_t *org = ctx->host_slots.data() + g.nseries;
        for (int k = 0; k < g.nslots && aligned; k++) aligned = (org[2 * k] % g.bx) == 0;
        if (aligned) g.vec = g.bx;
    }
    HIP_TRY(ctx->up_slots.put(ctx->slots, ctx->host_slots.data(), sizeof(int32_t) * ctx->host_slots.size(), st));
    g.slot_of = ctx->slots.as<int32_t>();
    g.slot_org = ctx->slots.as<int32_t>() + g.nseries;
    g.slot_own = g.slot_org + 2 * (size_t)g.nslots;
    ctx->nslots = g.nslots;
    const int ns = g.nslots, np = ea.pd.npairs, n = ea.order + 1;
    ctx->npairs = np;
    ctx->nseries = g.nseries;
    ctx->chanum = ct.chanum;
    ctx->order = ea.order;
    BleachDesc bd{};
    bd.order = ea.order;
    bd.frametime = ea.frametime;
    HIP_TRY(hipEventRecord(ctx->ev[0], st));
    if (ea.seg_swl > 0) {
        // "Lin Segment" (IFCS:9767-9805): bcNopit = num / L segments of
        // bcAve = num / bcNopit frames
        const int nseg = g.T / ea.seg_swl, seg_len = nseg > 0 ? g.T / nseg : 0;
        if (nseg < 2 || seg_len < 2) return fail(IMFCS_ERR_ARG, "linear segments need two segments of >= 2 frames");
        HIP_TRY(ctx->coef.ensure(sizeof(double) * (size_t)ns * (nseg + 1)));
        launch_seg_means(g, seg_len, nseg, ctx->coef.as<double>(), st);
        bd.mode = 1;
        bd.order = nseg;
        bd.seg_len = seg_len;
        bd.coef = ctx->coef.as<double>();
        bd.tscale = 1.0;
        ctx->order = nseg;
        ctx->coef_scale.assign(nseg + 1, 1.0);
    } else if (ea.exp_n > 0) {
        // "Single Exp" / "Double Exp" (IFCS:9682-9738): LM fit of the
        // nopit-point intensity trace (SingleExpFit / DoubleExpFit,
        // IFCS:10845-10950), then the correction with the fitted curve
        const int nexp = ea.exp_n, np = nexp == 1 ? 3 : 5;
        const int nopit = ea.nopit;
        const int nchunk = std::min(std::min(nopit, 128), std::max(40, (1 << 19) / std::max(1, ns)));
        HIP_TRY(ctx->part.ensure(sizeof(double) * (size_t)nchunk * ns));
        HIP_TRY(ctx->etrace.ensure(sizeof(float) * (size_t)ns * nopit));
        if (ea.psums)
            launch_trace_points(g, ea.psums, nopit, ea.ave, 1.0, ea.frametime, 0, ctx->part.as<double>(), nchunk, st,
                                ctx->etrace.as<float>());
        else
            launch_trace_moments(g, nopit, ea.ave, 1.0, ea.frametime, 0, 1, ctx->part.as<double>(), nchunk, st,
                                 ctx->etrace.as<float>());
        ctx->h_eitime.resize(nopit);
        for (int x = 0; x < nopit; x++) ctx->h_eitime[x] = (float)(ea.frametime * (x + 0.5) * ea.ave);   // IFCS:9851
        ctx->h_epfit.assign(11, 0);
        for (int k = 0; k < np; k++) ctx->h_epfit[k] = 1;
        HIP_TRY(ctx->up_eitime.put(ctx->eitime, ctx->h_eitime.data(), sizeof(float) * nopit, st));
        HIP_TRY(ctx->up_epfit.put(ctx->epfit, ctx->h_epfit.data(), sizeof(int) * 11, st));
        HIP_TRY(ctx->einit.ensure(sizeof(float) * 11 * (size_t)ns));
        HIP_TRY(ctx->eout.ensure(sizeof(double) * 5 * (size_t)ns));
        HIP_TRY(ctx->estate.ensure(sizeof(int) * (size_t)ns));
        HIP_TRY(ctx->echi2.ensure(sizeof(float) * (size_t)ns));
        HIP_TRY(ctx->eiter.ensure(sizeof(int) * (size_t)ns));
        launch_exp_init(ctx->etrace.as<float>(), ctx->eitime.as<float>(), ns, nopit, nexp, ctx->einit.as<float>(), st);
        launch_lm_exp(ns, nopit, ctx->etrace.as<float>(), ctx->eitime.as<float>(), ctx->einit.as<float>(),
                      ctx->epfit.as<int>(), nexp, 1e-16f, 2000, ctx->eout.as<double>(), ctx->estate.as<int>(),
                      ctx->echi2.as<float>(), ctx->eiter.as<int>(), st);
        HIP_TRY(ctx->coef.ensure(sizeof(double) * (size_t)ns * (np + 1)));
        launch_exp_pack(ctx->eout.as<double>(), ns, nexp, ctx->coef.as<double>(), st);
        bd.mode = nexp == 1 ? 2 : 3;
        bd.order = np;
        bd.coef = ctx->coef.as<double>();
        bd.tscale = 1.0;
        bd.t_off = ea.t_off;
        ctx->order = np;
        ctx->coef_scale.assign(np + 1, 1.0);
    } else if (ea.order > 0) {
        if (ea.order > kMaxOrder) return fail(IMFCS_ERR_ARG, "polynomial order above 8");
        HIP_TRY(ctx->coef.ensure(sizeof(double) * (size_t)ns * n));
        if (ea.fit_trace) {
            // abscissa normalized by the last trace time
            double tscale = ea.frametime * (ea.nopit - 0.5) * ea.ave;
            int nchunk = std::min(std::min(ea.nopit, 128), std::max(40, (1 << 19) / std::max(1, ns)));
            HIP_TRY(ctx->part.ensure(sizeof(double) * (size_t)nchunk * ns * n));
            if (ea.psums)
                launch_trace_points(g, ea.psums, ea.nopit, ea.ave, tscale, ea.frametime, ea.order,
                                    ctx->part.as<double>(), nchunk, st);
            else
                launch_trace_moments(g, ea.nopit, ea.ave, tscale, ea.frametime, ea.order, 1, ctx->part.as<double>(),
                                     nchunk, st);
            launch_polyfit(ctx->part.as<double>(), nchunk, ns, ea.nopit, ea.ave, tscale, ea.frametime, ea.order,
                           ctx->coef.as<double>(), st);
            bd.tscale = tscale;
            ctx->coef_scale.assign(n, 1.0);
            for (int k = 1; k < n; k++) ctx->coef_scale[k] = ctx->coef_scale[k - 1] / tscale;
        } else {
            // coefficients per series -> per slot (a slot's series share the window)
            ctx->host_coef.assign((size_t)ns * n, 0.0);
            for (int s2 = g.nseries - 1; s2 >= 0; s2--)
                for (int k = 0; k < n; k++)
                    ctx->host_coef[(size_t)ctx->host_slots[s2] * n + k] = ea.host_coef[(size_t)s2 * n + k];
            HIP_TRY(hipMemcpyAsync(ctx->coef.p, ctx->host_coef.data(), sizeof(double) * (size_t)ns * n,
                                   hipMemcpyHostToDevice, st));
            bd.tscale = 1.0;
            ctx->coef_scale.assign(n, 1.0);
        }
        bd.coef = ctx->coef.as<double>();
        bd.t_off = ea.t_off;
    } else {
        bd.tscale = 1.0;
        bd.coef = nullptr;
    }
    HIP_TRY(hipEventRecord(ctx->ev_h2d, st));
    HIP_TRY(hipEventRecord(ctx->ev[1], st));
    // chunking of the frames (autotuned per shape, default 2^19 threads in flight)
    const Tune tn = ctx->force_tune.apply_log2 > 0 ? ctx->force_tune : tune_lookup(ns, g.T);
    ctx->used_tune = tn;
    // window means
    // chunk length: about 2^apply_log2 threads in flight, 64..4096 frames
    // each (few slots, e.g. binning 4 of a 128 x 128 image: short chunks)
    int plen = 4096;
    while (plen > 64 && (int64_t)ns * ((g.T + plen - 1) / plen) < ((int64_t)1 << tn.apply_log2)) plen >>= 1;
    const int npc = (g.T + plen - 1) / plen;
    HIP_TRY(ctx->ctot.ensure(sizeof(double) * (size_t)prefix_subchunks(g.T) * ns));
    HIP_TRY(ctx->rec.ensure(sizeof(double) * (size_t)ns * ct.npos));
    HIP_TRY(ctx->F.ensure(sizeof(double) * (size_t)ns * ct.npos));
    HIP_TRY(ctx->X.ensure(sizeof(float) * (size_t)ns * g.T));
    // levels 0..2 on the matrix cores (selected with IMFCS_CORRELATOR=mfma;
    // the VALU kernel is faster on every measured shape): the bleach pass
    // also writes the level-3 values and the per-slot staging exponents
    const bool fast = correlate_fast_supported(ct) && !ctx->force_v1;
    const bool mfma = fast && ctx->want_mfma && corr_mfma_eligible(g, ea.pd) &&
                      (int64_t)g.nslots * 4 * (16 * 36) < 0x7fffffff;
    if (mfma) {
        HIP_TRY(ctx->l3.ensure(sizeof(float) * ((size_t)ct.numbin[3] + kLevelPad) * g.nslots));
        HIP_TRY(ctx->cmax.ensure(sizeof(float) * (size_t)npc * ns));
        HIP_TRY(ctx->xexp.ensure(sizeof(int32_t) * (size_t)ns));
    }
    StagedRun &S = ctx->staged;
    S.ea = ea;
    S.bd = bd;
    S.ht = ht;
    S.ct = ct;
    S.tn = tn;
    S.plen = plen;
    S.npc = npc;
    S.mfma = mfma;
    if (ctx->stage == 1) return 0;
    return engine_rest(ctx, true);
}

int check_geom(const imfcs_config *c, int ow, int oh) {
    if (c->width <= 0 || c->height <= 0 || c->frames <= 0) return fail(IMFCS_ERR_ARG, "empty stack");
    if (c->firstframe < 1 || c->lastframe > c->frames || c->lastframe - c->firstframe + 1 < 2)
        return fail(IMFCS_ERR_ARG, "frame range outside the stack");
    if (ow <= 0 || oh <= 0) return fail(IMFCS_ERR_ARG, "ROI holds no output pixel");
    if (c->cf_x < 0 || c->cf_y < 0)
        return fail(IMFCS_ERR_ARG, "negative CF distances are not supported (as in IFCS:15296-15299)");
    int sx = (int)std::ceil((double)c->roi_x / c->pixbin_x), sy = (int)std::ceil((double)c->roi_y / c->pixbin_y);
    if (sx < 0 || sy < 0) return fail(IMFCS_ERR_ARG, "ROI starts outside the stack");
    int cfx = c->do_ccf ? c->cf_x : 0, cfy = c->do_ccf ? c->cf_y : 0;
    if ((sx + ow - 1) * c->pixbin_x + c->binning_x + cfx > c->width ||
        (sy + oh - 1) * c->pixbin_y + c->binning_y + cfy > c->height)
        return fail(IMFCS_ERR_ARG, "ROI (with CCF partner) exceeds the stack");
    return 0;
}

}  // namespace

// The LM stage over the correlations of the last run held in ctx (corav,
// bvar, or the GLS-whitened data), with cfg's fit models and settings.
static int run_fit(imfcs_ctx *ctx, const imfcs_config *c, std::vector<LMGroup> *collect = nullptr) {
    {
        hipStream_t st = ctx->stream;
        const int np = ctx->npairs, npts = ctx->chanum - 1;
        HIP_TRY(ctx->fdata.ensure(sizeof(float) * (size_t)np * npts));
        HIP_TRY(ctx->fw.ensure(sizeof(float) * (size_t)np * npts));
        HIP_TRY(ctx->fout.ensure(sizeof(float) * (size_t)np * 20));
        HIP_TRY(ctx->fstate.ensure(sizeof(int) * np));
        HIP_TRY(ctx->fchi2.ensure(sizeof(float) * np));
        HIP_TRY(ctx->fiter.ensure(sizeof(int) * np));
        if (!ctx->did_gls)
            launch_pack_fit_inputs(ctx->chanum, np, ctx->corav.as<double>(), ctx->bvar.as<double>(), ctx->fdata.as<float>(),
                               ctx->fw.as<float>(), st);
        // lag times of channels 1..chanum-1 (IFCS:15122-15125) and initial
        // parameters with the fixed geometry in slots 11..19 (IFCS:15090-15110):
        // every ACF fit starts from the ACF model's row and takes its free set,
        // every CCF fit the CCF model's, so two rows of each travel
        std::vector<float> &xs = ctx->h_xs, &init = ctx->h_init;
        std::vector<int> &pfit = ctx->h_pfit;
        xs.assign(npts, 0.f);
        init.assign(2 * 20, 0.f);
        pfit.assign(2 * 20, 0);
        for (int i = 0; i < npts; i++) xs[i] = (float)ctx->last_lagtime[i + 1];
        const imfcs_fit_model *models[2] = {&c->acf_model, &c->ccf_model};
        unsigned mask[2] = {0u, 0u};
        int nfree[2] = {0, 0};
        for (int r = 0; r < 2; r++) {
            const imfcs_fit_model *m = models[r];
            float *q = &init[(size_t)r * 20];
            for (int k = 0; k < 11; k++) {
                q[k] = (float)m->paraminit[k];
                pfit[(size_t)r * 20 + k] = m->paramfit[k] ? 1 : 0;
                mask[r] |= (m->paramfit[k] ? 1u : 0u) << k;
                nfree[r] += m->paramfit[k] ? 1 : 0;
            }
            q[11] = (float)m->ax;
            q[12] = (float)m->ay;
            q[13] = (float)m->s;
            q[14] = (float)m->sz;
            q[15] = (float)m->rx;
            q[16] = (float)m->ry;
            q[17] = (float)m->fitobsvol;
            q[18] = (float)m->q2;
            q[19] = (float)m->q3;
        }
        // fits [0, n_acf) are ACF fits, the rest CCF fits
        const int n_acf = c->do_acf ? ctx->npix : 0, n_ccf = np - n_acf;
        for (int r = 0; r < 2; r++)
            if ((r == 0 ? n_acf : n_ccf) > 0 && nfree[r] == 0) return fail(IMFCS_ERR_ARG, "no free fit parameter");
        HIP_TRY(ctx->up_fx.put(ctx->fx, xs.data(), sizeof(float) * npts, st));
        HIP_TRY(ctx->up_finit.put(ctx->finit, init.data(), sizeof(float) * init.size(), st));
        HIP_TRY(ctx->up_fpfit.put(ctx->fpfit, pfit.data(), sizeof(int) * pfit.size(), st));
        HIP_TRY(hipEventRecord(ctx->ev_h2d, st));
        ctx->fit_split = n_acf;
        ctx->nfree = nfree[n_acf > 0 ? 0 : 1];
        // one launch when both kinds share the free set (the last fit first:
        // the CCF fits); else the ACF and the CCF fits as two launches, each
        // with its own parameter row
        const bool seeded = !ctx->did_gls && (ctx->seeder[0] || ctx->seeder[1]);
        const bool one = n_acf == 0 || n_ccf == 0 || (mask[0] == mask[1] && !seeded);
        struct Part { int first, count, row; };
        Part parts[2];
        int nparts = 0;
        if (one) parts[nparts++] = {0, np, n_acf > 0 ? 0 : 1};
        else {
            parts[nparts++] = {0, n_acf, 0};
            parts[nparts++] = {n_acf, n_ccf, 1};
        }
        if (ctx->did_gls)
            for (int i = 0; i < nparts; i++)
                if (mask[parts[i].row] != 0x13u) return fail(IMFCS_ERR_ARG, "the GLS fit supports the free set N, D, G_inf");
        // a batch collects the fits of its pipelines for one persistent launch
        bool collected = false;
        if (collect && !ctx->did_gls && !seeded && (n_acf == 0 || n_ccf == 0 || mask[0] == mask[1]) &&
            lm_queue_eligible(npts, mask[n_acf ? 0 : 1], c->fit_chi2_fp64 ? 1 : 0)) {
            const int row = n_acf ? 0 : 1;
            collect->push_back(LMGroup{np, npts, ctx->fdata.as<float>(), ctx->fw.as<float>(), ctx->fx.as<float>(),
                                       ctx->finit.as<float>(), ctx->fpfit.as<int>() + (size_t)row * 20,
                                       c->fit_tolerance, c->fit_max_iterations, ctx->fout.as<float>(),
                                       ctx->fstate.as<int>(), ctx->fchi2.as<float>(), ctx->fiter.as<int>(), 1,
                                       n_acf});
            nparts = 0;
            collected = true;
        }
        // learned starting rows: a seeded attempt of at most 15 % of the
        // iteration limit (a fit converging from a good start takes a few
        // tens; one still running by then is wandering along a flat valley),
        // then the plain start (the model's row) for the fits it did not
        // converge (their iterations added)
        const int seed_maxit = std::max(1, std::min(c->fit_max_iterations, (c->fit_max_iterations * 3 + 19) / 20));
        const int chi2_64 = c->fit_chi2_fp64 ? 1 : 0;
        float *fo = ctx->fout.as<float>(), *fc = ctx->fchi2.as<float>();
        int *fs = ctx->fstate.as<int>(), *fi = ctx->fiter.as<int>();
        const float *fd = ctx->fdata.as<float>(), *fw = ctx->fw.as<float>(), *fx = ctx->fx.as<float>();
        const bool all_seeded = seeded && (n_acf == 0 || (ctx->seeder[0] && mlp_points(ctx->seeder[0]) == npts)) &&
                                (n_ccf == 0 || (ctx->seeder[1] && mlp_points(ctx->seeder[1]) == npts));
        if (nparts > 0 && all_seeded && (n_acf == 0 || n_ccf == 0 || mask[0] == mask[1])) {
            // every fit seeded: one LM launch over ACF and CCF fits, one fallback launch
            HIP_TRY(ctx->fseed.ensure(sizeof(float) * 20 * (size_t)np));
            float *rows = ctx->fseed.as<float>();
            int rc = 0;
            if (n_acf) rc = mlp_seed_init(ctx->seeder[0], fd, n_acf, ctx->finit.as<float>(), rows, st);
            if (!rc && n_ccf)
                rc = mlp_seed_init(ctx->seeder[1], fd + (size_t)n_acf * npts, n_ccf, ctx->finit.as<float>() + 20,
                                   rows + (size_t)n_acf * 20, st);
            if (rc) return rc;
            const int row = n_acf ? 0 : 1;
            const int *pf = ctx->fpfit.as<int>() + (size_t)row * 20;
            launch_lm_acf(np, npts, fd, fw, fx, rows, pf, mask[row], LMModelGeom{}, c->fit_tolerance, seed_maxit,
                          chi2_64, fo, fs, fc, fi, st, -1);
            launch_lm_acf(np, npts, fd, fw, fx, ctx->finit.as<float>(), pf, mask[row], LMModelGeom{},
                          c->fit_tolerance, c->fit_max_iterations, chi2_64, fo, fs, fc, fi, st, n_acf, 1);
            nparts = 0;
        }
        for (int i = 0; i < nparts; i++) {
            const Part &pt = parts[i];
            // two init rows: with one launch, fits below n_acf take row 0 and
            // the others row 1; a split launch takes its own row for every fit
            const float *ini = ctx->finit.as<float>() + (one ? 0 : (size_t)pt.row * 20);
            const int *pf = ctx->fpfit.as<int>() + (size_t)pt.row * 20;
            const int split = one ? n_acf : pt.count;
            const size_t f0 = (size_t)pt.first;
            if (ctx->did_gls) {
                launch_lm_gls(pt.count, npts, fx, ini, pf, c->fit_tolerance, c->fit_max_iterations,
                              ctx->glsW.as<double>() + f0 * npts * npts, ctx->glsy.as<double>() + f0 * npts,
                              ctx->glsok.as<int32_t>() + f0, fo + f0 * 20, fs + f0, fc + f0, fi + f0, st, split);
                continue;
            }
            MlpNet *net = ctx->seeder[pt.row];
            if (net && mlp_points(net) == npts) {
                HIP_TRY(ctx->fseed.ensure(sizeof(float) * 20 * (size_t)np));
                float *rows = ctx->fseed.as<float>() + f0 * 20;
                const int rc = mlp_seed_init(net, fd + f0 * npts, pt.count, ini, rows, st);
                if (rc) return rc;
                launch_lm_acf(pt.count, npts, fd + f0 * npts, fw + f0 * npts, fx, rows, pf, mask[pt.row],
                              LMModelGeom{}, c->fit_tolerance, seed_maxit, chi2_64, fo + f0 * 20, fs + f0, fc + f0,
                              fi + f0, st, -1);
                launch_lm_acf(pt.count, npts, fd + f0 * npts, fw + f0 * npts, fx, ini, pf, mask[pt.row],
                              LMModelGeom{}, c->fit_tolerance, c->fit_max_iterations, chi2_64, fo + f0 * 20, fs + f0,
                              fc + f0, fi + f0, st, pt.count, 1);
            } else {
                launch_lm_acf(pt.count, npts, fd + f0 * npts, fw + f0 * npts, fx, ini, pf, mask[pt.row],
                              LMModelGeom{}, c->fit_tolerance, c->fit_max_iterations, chi2_64, fo + f0 * 20, fs + f0,
                              fc + f0, fi + f0, st, split);
            }
        }
        // FP64 polish of the converged N, D, G_inf fits (the pipeline's FP64
        // chi-square mode; fits the batch collected are polished after its
        // shared launch)
        if (!ctx->did_gls && chi2_64 && !collected) {
            if (n_acf > 0 && mask[0] == 0x13u)
                launch_lm_polish(n_acf, npts, fd, fw, fx, fo, fs, fc, st);
            if (n_ccf > 0 && mask[1] == 0x13u)
                launch_lm_polish(n_ccf, npts, fd + (size_t)n_acf * npts, fw + (size_t)n_acf * npts, fx,
                                 fo + (size_t)n_acf * 20, fs + n_acf, fc + n_acf, st);
        }
        for (int r = 0; r < 2; r++)
            ctx->chi2_norm[r] = ctx->did_gls
                                    // chi2 of GLSFit: sum of whitened residuals^2 / (chanum - nfree - 1) (IFCS:10733-10735)
                                    ? (double)(ctx->chanum - nfree[r] - 1)
                                    // chi2 normalization of GPU_ACF_Fit (IFCS:15234)
                                    : (double)((npts - 1) - nfree[r] - 1);
    }
    return 0;
}

static int run_config(imfcs_ctx *ctx, const uint16_t *d_stack, const imfcs_config *c, const int32_t *psums,
                      std::vector<LMGroup> *collect = nullptr) {
    if (!ctx || !d_stack || !c) return fail(IMFCS_ERR_ARG, "null argument");
    if (!c->do_acf && !c->do_ccf) return fail(IMFCS_ERR_ARG, "neither ACF nor CCF requested");
    HIP_TRY(hipSetDevice(ctx->device));
    int ow, oh, chanum;
    int rc = imfcs_plan(c, &ow, &oh, &chanum);
    if (rc) return rc;
    rc = check_geom(c, ow, oh);
    if (rc) return rc;
    EngineArgs ea{};
    SeriesGeom &g = ea.g;
    g.data = d_stack;
    g.src = kSrcU16;
    g.plane = (int64_t)c->width * c->height;
    g.stride = c->width;
    g.f0 = c->firstframe - 1;
    g.T = c->lastframe - c->firstframe + 1;
    g.bx = c->binning_x;
    g.by = c->binning_y;
    g.bckg_total = c->background * c->binning_x * c->binning_y;
    g.ow = ow;
    g.oh = oh;
    g.npix = ow * oh;
    g.ox0 = (int)std::ceil((double)c->roi_x / c->pixbin_x) * c->pixbin_x;
    g.oy0 = (int)std::ceil((double)c->roi_y / c->pixbin_y) * c->pixbin_y;
    g.pbx = c->pixbin_x;
    g.pby = c->pixbin_y;
    g.cfx = c->cf_x;
    g.cfy = c->cf_y;
    g.nseries = g.npix * (c->do_ccf ? 2 : 1);
    ea.order = c->poly_order;
    ea.fit_trace = 1;
    ea.nopit = std::min(g.T, 1000);   // IFCS:7323-7327
    ea.ave = g.T / ea.nopit;
    ea.frametime = c->frametime;
    ea.gls = c->gls ? 1 : 0;
    ea.t_off = 1.5;
    ea.p = c->correlator_p;
    ea.q = effective_q(c);
    ea.pd.npix = g.npix;
    ea.pd.acf_first = c->do_acf ? 1 : 0;
    ea.pd.has_ccf = c->do_ccf ? 1 : 0;
    ea.pd.npairs = g.npix * ((c->do_acf ? 1 : 0) + (c->do_ccf ? 1 : 0));
    ea.psums = psums;
    ctx->npix = g.npix;
    ctx->ow = ow;
    ctx->oh = oh;
    HostTables ht;
    CorrTables ct;
    hipStream_t st = ctx->stream;
    if (c->bleach_mode == IMFCS_BLEACH_LIN_SEGMENT) {
        if (c->poly_order != 0) return fail(IMFCS_ERR_ARG, "linear segments exclude the polynomial correction");
        ea.seg_swl = c->sliding_window_length;
    } else if (c->bleach_mode == IMFCS_BLEACH_SINGLE_EXP || c->bleach_mode == IMFCS_BLEACH_DOUBLE_EXP) {
        if (c->poly_order != 0) return fail(IMFCS_ERR_ARG, "exponential correction excludes the polynomial one");
        if (ea.nopit < 10) return fail(IMFCS_ERR_ARG, "too few frames for an exponential bleach fit");
        ea.exp_n = c->bleach_mode == IMFCS_BLEACH_SINGLE_EXP ? 1 : 2;
    } else if (c->bleach_mode != IMFCS_BLEACH_POLYNOMIAL && c->bleach_mode != IMFCS_BLEACH_SLIDING_WINDOW) {
        return fail(IMFCS_ERR_ARG, "unknown bleach correction mode");
    }
    if (c->bleach_mode == IMFCS_BLEACH_SLIDING_WINDOW) {
        // Sliding windows (IFCS:9146-9172): each window of L frames is
        // correlated on its own, without bleach correction (the intensities
        // of the window, getIntensity with no matching mode), and the ACF
        // and blocked variance are averaged over the windows, each term
        // divided by the window count; the blocking index reported is the
        // last window's.
        const int L = c->sliding_window_length;
        if (c->poly_order != 0 || c->gls)
            return fail(IMFCS_ERR_ARG, "sliding windows exclude the polynomial correction and GLS");
        if (L < 2 || L > g.T) return fail(IMFCS_ERR_ARG, "sliding window length outside the frame range");
        if (ea.q < 2) return fail(IMFCS_ERR_ARG, "sliding window too short for the correlator");
        const int numofsw = g.T / L;
        const int first = g.f0;
        ea.order = 0;
        ea.g.T = L;
        ea.nopit = std::min(L, 1000);
        ea.ave = L / ea.nopit;
        for (int i = 0; i < numofsw; i++) {
            ea.g.f0 = first + i * L;
            rc = run_engine(ctx, ea, ht, ct);
            if (rc) return rc;
            const int64_t n = (int64_t)ea.pd.npairs * ct.chanum;
            HIP_TRY(ctx->swacf.ensure(sizeof(double) * n));
            HIP_TRY(ctx->swvar.ensure(sizeof(double) * n));
            launch_accum_div(ctx->swacf.as<double>(), ctx->corav.as<double>(), n, (double)numofsw, i == 0, st);
            launch_accum_div(ctx->swvar.as<double>(), ctx->bvar.as<double>(), n, (double)numofsw, i == 0, st);
        }
        const int64_t n = (int64_t)ea.pd.npairs * ct.chanum;
        launch_copy(ctx->corav.as<double>(), ctx->swacf.as<double>(), sizeof(double) * n, st);
        launch_copy(ctx->bvar.as<double>(), ctx->swvar.as<double>(), sizeof(double) * n, st);
    } else {
        rc = run_engine(ctx, ea, ht, ct);
        if (rc) return rc;
        if (ctx->stage == 1) return 0;   // first part of a staged run
    }
    ctx->last_cfg = *c;
    ctx->last_lagtime = ht.lagtime;
    ctx->has_last = 1;
    ctx->did_fit = c->do_fit;
    if (c->do_fit) {
        rc = run_fit(ctx, c, collect);
        if (rc) return rc;
    }
    HIP_TRY(hipEventRecord(ctx->ev[5], st));
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int imfcs_run(imfcs_ctx *ctx, const uint16_t *d_stack, const imfcs_config *c) {
    if (ctx) ctx->last_noalloc = 0;
    return run_config(ctx, d_stack, c, nullptr);
}

static int noalloc_guard(imfcs_ctx *ctx) {
    if (ctx->last_noalloc)
        return fail(IMFCS_ERR_ARG, "the last run was imfcs_run_noalloc: its results are in the caller's buffers");
    return 0;
}

// Configurations that give the same correlations (every field but the fit's).
static bool same_correlation(const imfcs_config &a, const imfcs_config &b) {
    return a.width == b.width && a.height == b.height && a.frames == b.frames && a.firstframe == b.firstframe &&
           a.lastframe == b.lastframe && a.binning_x == b.binning_x && a.binning_y == b.binning_y &&
           a.pixbin_x == b.pixbin_x && a.pixbin_y == b.pixbin_y && a.cf_x == b.cf_x && a.cf_y == b.cf_y &&
           a.background == b.background && a.correlator_p == b.correlator_p &&
           a.correlator_q == b.correlator_q && a.frametime == b.frametime && a.poly_order == b.poly_order &&
           a.roi_x == b.roi_x && a.roi_y == b.roi_y && a.roi_w == b.roi_w && a.roi_h == b.roi_h &&
           a.do_acf == b.do_acf && a.do_ccf == b.do_ccf && a.gls == b.gls && a.bleach_mode == b.bleach_mode &&
           a.sliding_window_length == b.sliding_window_length;
}

// Fit stage only, on the correlations of the last imfcs_run of ctx: the
// PSF calibration (determinePSF, IFCS:8118-8160) repeats the whole
// diffusion-law pipeline for every trial PSF size although only the fit
// depends on it.
extern "C" int imfcs_refit(imfcs_ctx *ctx, const imfcs_config *c) {
    if (!ctx || !c) return fail(IMFCS_ERR_ARG, "null argument");
    if (!ctx->has_last) return fail(IMFCS_ERR_ARG, "no correlations to refit: run the pipeline first");
    if (noalloc_guard(ctx)) return IMFCS_ERR_ARG;
    if (!same_correlation(*c, ctx->last_cfg))
        return fail(IMFCS_ERR_ARG, "refit configuration differs from the last run outside the fit settings");
    HIP_TRY(hipSetDevice(ctx->device));
    // the host staging of the fit parameters is reused: wait for its last copies
    HIP_TRY(hipEventSynchronize(ctx->ev_h2d));
    HIP_TRY(hipEventRecord(ctx->ev[4], ctx->stream));
    ctx->did_fit = 1;
    const int rc = run_fit(ctx, c);
    if (rc) return rc;
    ctx->last_cfg = *c;
    HIP_TRY(hipEventRecord(ctx->ev[5], ctx->stream));
    HIP_TRY(hipGetLastError());
    return 0;
}

// Several configurations of one stack (binnings of a diffusion law, or the
// benchmark's three binnings), one context (stream) each.  When every
// configuration takes the polynomial bleach correction over the same frames
// and full frame size, one pass over the stack writes the per-pixel sums of
// each bleach-trace point (on the first context's stream) and every
// pipeline takes its trace from them; the other streams wait for that pass.
// Configurations whose polynomial bleach traces can all come from one pass
// over the same stack: the polynomial correction, one image size and frame
// range, one device, distinct contexts.
static bool can_share_point_sums(imfcs_ctx *const *ctxs, int n, const imfcs_config *cfgs) {
    bool share = true;
    for (int i = 0; i < n && share; i++) {
        const imfcs_config &c = cfgs[i];
        share = c.bleach_mode == IMFCS_BLEACH_POLYNOMIAL && c.poly_order > 0 && c.width == cfgs[0].width &&
                c.height == cfgs[0].height && c.firstframe == cfgs[0].firstframe &&
                c.lastframe == cfgs[0].lastframe && ctxs[i]->device == ctxs[0]->device;
        for (int j = 0; j < i && share; j++) share = ctxs[j] != ctxs[i];
    }
    return share;
}

// The per-pixel integer sums of every bleach point of a stack (nopit points
// of ave frames, as imfcs_run: IFCS:7323-7327), the input of every shared
// configuration's trace.
static int enqueue_point_sums(DevBuf &psums, const uint16_t *d_stack, const imfcs_config &k, hipStream_t st) {
    const int T = k.lastframe - k.firstframe + 1;
    if (T < 1 || k.firstframe < 1 || k.width < 1 || k.height < 1) return fail(IMFCS_ERR_ARG, "bad frame range");
    const int nopit = std::min(T, 1000), ave = T / nopit;
    const int64_t plane = (int64_t)k.width * k.height;
    HIP_TRY(psums.ensure(sizeof(int32_t) * (size_t)nopit * plane));
    launch_point_sums(d_stack, plane, k.firstframe - 1, nopit, ave, psums.as<int32_t>(), st);
    return 0;
}

// The binnings 1, 2 and 4 among configurations of one stack whose bleach
// passes one nested pass (k_apply_nested) can do: the same frames, panel,
// polynomial order and background, non-overlapping full grids from (0, 0),
// the default correlator kernels.  tri[b] = index of binning 2^b.
static bool nested_candidates(imfcs_ctx *const *ctxs, int n, const imfcs_config *cfgs, int (&tri)[3]) {
    tri[0] = tri[1] = tri[2] = -1;
    for (int i = 0; i < n; i++) {
        const imfcs_config &c = cfgs[i];
        const int b = c.binning_x;
        const int lb = b == 1 ? 0 : b == 2 ? 1 : b == 4 ? 2 : -1;
        if (lb < 0 || tri[lb] >= 0 || c.binning_y != b || c.pixbin_x != b || c.pixbin_y != b) continue;
        if (c.bleach_mode != IMFCS_BLEACH_POLYNOMIAL || c.poly_order < 1 || c.roi_x != 0 || c.roi_y != 0) continue;
        if (ctxs[i]->want_mfma) continue;
        tri[lb] = i;
    }
    if (tri[0] < 0 || tri[1] < 0 || tri[2] < 0) return false;
    const imfcs_config &a = cfgs[tri[0]];
    for (int lb = 1; lb < 3; lb++) {
        const imfcs_config &c = cfgs[tri[lb]];
        if (c.width != a.width || c.height != a.height || c.firstframe != a.firstframe ||
            c.lastframe != a.lastframe || c.background != a.background || c.poly_order != a.poly_order ||
            c.correlator_p != a.correlator_p || c.correlator_q != a.correlator_q || c.frametime != a.frametime)
            return false;
    }
    return a.width % 2 == 0 && ((int64_t)a.width * a.height) % 2 == 0;
}

// After the first parts of the three runs: their slot grids are the full
// nested grids (no padding, no second grid), the tables agree and every
// buffer resource of the pass spans its 16 rows.
static bool nested_ready(imfcs_ctx *const *ctxs, const int (&tri)[3], const uint16_t *d_stack) {
    int swp[3], sh[3];
    for (int lb = 0; lb < 3; lb++) {
        const StagedRun &S = ctxs[tri[lb]]->staged;
        const SeriesGeom &g = S.ea.g;
        const bool ccf = g.nseries > g.npix;
        const bool on_grid = ccf && (g.cfx % g.pbx == 0) && (g.cfy % g.pby == 0);
        if (ccf && !on_grid) return false;
        const int sw = g.ow + (ccf ? g.cfx / g.pbx : 0), shh = g.oh + (ccf ? g.cfy / g.pby : 0);
        if (sw != g.swp || (int64_t)sw * shh != g.nslots || g.ox0 != 0 || g.oy0 != 0) return false;
        if (S.bd.mode != 0 || S.bd.order != ctxs[tri[0]]->staged.bd.order || S.mfma) return false;
        if ((int64_t)g.nslots * 64 >= 0x7fffffff) return false;
        swp[lb] = g.swp;
        sh[lb] = shh;
        const CorrTables &c0 = ctxs[tri[0]]->staged.ct;
        if (S.ct.npos != c0.npos || std::memcmp(S.ct.pos, c0.pos, sizeof(int32_t) * c0.npos) != 0) return false;
    }
    const SeriesGeom &g1 = ctxs[tri[0]]->staged.ea.g;
    if (g1.plane * 2 * 16 >= 0x7fffffff || (uintptr_t)d_stack % 4 != 0) return false;
    return swp[0] == 2 * swp[1] && swp[1] == 2 * swp[2] && sh[0] == 2 * sh[1] && sh[1] == 2 * sh[2] &&
           swp[1] % 32 == 0 && sh[1] % 2 == 0;
}

// The runs of n configurations of one stack with a nested triple: the
// triple's first parts, its bleach pass on binning 1's stream once the other
// two have their polynomials, then every configuration in the given order
// (the triple's runs from the pass on, the others whole).  Without a nested
// layout each run does its own pass.
static int run_nested(imfcs_ctx *const *ctxs, int n, const int (&tri)[3], const uint16_t *d_stack,
                      const imfcs_config *cfgs, const int32_t *psums, std::vector<LMGroup> *collect) {
    for (int lb = 0; lb < 3; lb++) {
        imfcs_ctx *c = ctxs[tri[lb]];
        c->stage = 1;
        const int rc = run_config(c, d_stack, &cfgs[tri[lb]], psums, collect);
        c->stage = 0;
        if (rc) return rc;
    }
    const bool nest = nested_ready(ctxs, tri, d_stack);
    if (nest) {
        imfcs_ctx *c1 = ctxs[tri[0]];
        for (int lb = 1; lb < 3; lb++) {
            HIP_TRY(hipEventRecord(ctxs[tri[lb]]->ev_stage, ctxs[tri[lb]]->stream));
            HIP_TRY(hipStreamWaitEvent(c1->stream, ctxs[tri[lb]]->ev_stage, 0));
        }
        NestedApply A{};
        const SeriesGeom &g1 = c1->staged.ea.g;
        A.data = d_stack;
        A.plane = g1.plane;
        A.stride = g1.stride;
        A.f0 = g1.f0;
        A.T = g1.T;
        A.bg = cfgs[tri[0]].background;
        int plen = 4096;
        const int ns2 = ctxs[tri[1]]->staged.ea.g.nslots;
        while (plen > 64 && (int64_t)ns2 * ((g1.T + plen - 1) / plen) < ((int64_t)1 << 19)) plen >>= 1;
        for (int lb = 0; lb < 3; lb++) {
            imfcs_ctx *c = ctxs[tri[lb]];
            A.swp[lb] = c->staged.ea.g.swp;
            A.sh[lb] = c->staged.ea.g.nslots / c->staged.ea.g.swp;
            A.bd[lb] = c->staged.bd;
            A.ctot[lb] = c->ctot.as<double>();
            A.rec[lb] = c->rec.as<double>();
            A.X[lb] = c->X.as<float>();
        }
        launch_apply_nested(A, c1->ctab.as<CorrTables>(), c1->staged.bd.order, (g1.T + plen - 1) / plen, plen,
                            c1->stream);
        HIP_TRY(hipEventRecord(c1->ev_stage, c1->stream));
        for (int lb = 1; lb < 3; lb++) HIP_TRY(hipStreamWaitEvent(ctxs[tri[lb]]->stream, c1->ev_stage, 0));
    }
    for (int i = 0; i < n; i++) {
        imfcs_ctx *c = ctxs[i];
        const bool in_tri = i == tri[0] || i == tri[1] || i == tri[2];
        c->stage = in_tri ? (nest ? 2 : 3) : 0;
        c->run_nested = in_tri && nest;
        const int rc = run_config(c, d_stack, &cfgs[i], psums, collect);
        c->stage = 0;
        if (rc) return rc;
    }
    return 0;
}

extern "C" int imfcs_run_multi(imfcs_ctx *const *ctxs, int n, const uint16_t *d_stack, const imfcs_config *cfgs) {
    if (!ctxs || !cfgs || !d_stack || n < 1) return fail(IMFCS_ERR_ARG, "null argument");
    for (int i = 0; i < n; i++)
        if (!ctxs[i]) return fail(IMFCS_ERR_ARG, "null context");
    for (int i = 0; i < n; i++) ctxs[i]->last_noalloc = 0;
    const bool share = can_share_point_sums(ctxs, n, cfgs);
    if (!share) {
        for (int i = 0; i < n; i++) {
            const int rc = run_config(ctxs[i], d_stack, &cfgs[i], nullptr);
            if (rc) return rc;
        }
        return 0;
    }
    imfcs_ctx *c0 = ctxs[0];
    HIP_TRY(hipSetDevice(c0->device));
    // the previous runs of every context may still read the sums: wait for them
    for (int i = 1; i < n; i++) HIP_TRY(hipStreamWaitEvent(c0->stream, ctxs[i]->ev[5], 0));
    const int prc = enqueue_point_sums(c0->psums, d_stack, cfgs[0], c0->stream);
    if (prc) return prc;
    HIP_TRY(hipEventRecord(c0->ev_psums, c0->stream));
    for (int i = 1; i < n; i++) HIP_TRY(hipStreamWaitEvent(ctxs[i]->stream, c0->ev_psums, 0));
    // enqueued in the given order: a pipeline is ~25 launches (≈0.4 ms of
    // host time), so later configurations start later (enqueueing them from
    // parallel host threads measured no faster end to end); binnings 1, 2
    // and 4 of the stack share one bleach pass
    int tri[3];
    if (nested_candidates(ctxs, n, cfgs, tri))
        return run_nested(ctxs, n, tri, d_stack, cfgs, c0->psums.as<int32_t>(), nullptr);
    for (int i = 0; i < n; i++) {
        const int rc = run_config(ctxs[i], d_stack, &cfgs[i], c0->psums.as<int32_t>());
        if (rc) return rc;
    }
    return 0;
}

extern "C" int imfcs_set_fit_seeder(imfcs_ctx *ctx, imfcs_mlp *acf_net, imfcs_mlp *ccf_net) {
    if (!ctx) return fail(IMFCS_ERR_ARG, "null context");
    ctx->seeder[0] = mlp_of(acf_net);
    ctx->seeder[1] = mlp_of(ccf_net);
    return 0;
}

extern "C" int imfcs_get_stage_ms(imfcs_ctx *ctx, float *ms) {
    if (!ctx || !ms) return fail(IMFCS_ERR_ARG, "null argument");
    if (noalloc_guard(ctx)) return IMFCS_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < 5; i++) {
        float t = 0;
        HIP_TRY(hipEventElapsedTime(&t, ctx->ev[i], ctx->ev[i + 1]));
        ms[i] = t;
    }
    return 0;
}

extern "C" int imfcs_get_gls_covariance(imfcs_ctx *ctx, int pair, double *cov) {
    if (!ctx || !cov) return fail(IMFCS_ERR_ARG, "null argument");
    if (noalloc_guard(ctx)) return IMFCS_ERR_ARG;
    if (!ctx->did_gls) return fail(IMFCS_ERR_ARG, "the last run computed no GLS covariance");
    if (pair < 0 || pair >= ctx->npairs) return fail(IMFCS_ERR_ARG, "pair index out of range");
    const size_t m = ctx->chanum - 1;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipMemcpy(cov, ctx->glscov.as<double>() + (size_t)pair * m * m, sizeof(double) * m * m,
                      hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int imfcs_get_correlator_kernel_ms(imfcs_ctx *ctx, float *ms) {
    if (!ctx || !ms) return fail(IMFCS_ERR_ARG, "null argument");
    if (noalloc_guard(ctx)) return IMFCS_ERR_ARG;
    if (!ctx->evk_valid) return fail(IMFCS_ERR_ARG, "no correlation run yet");
    HIP_TRY(hipEventSynchronize(ctx->evk[1]));
    HIP_TRY(hipEventElapsedTime(ms, ctx->evk[0], ctx->evk[1]));
    return 0;
}

extern "C" int imfcs_get_results(imfcs_ctx *ctx, double *acf, double *var, int32_t *ind, int32_t *blocked,
                                 double *fitres, double *chi2, int32_t *states, int32_t *iters, double *coeffs) {
    if (!ctx) return fail(IMFCS_ERR_ARG, "null context");
    if (noalloc_guard(ctx)) return IMFCS_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const size_t np = ctx->npairs, nc = ctx->chanum;
    if (acf) HIP_TRY(hipMemcpy(acf, ctx->corav.p, sizeof(double) * np * nc, hipMemcpyDeviceToHost));
    if (var) HIP_TRY(hipMemcpy(var, ctx->bvar.p, sizeof(double) * np * nc, hipMemcpyDeviceToHost));
    if (ind) HIP_TRY(hipMemcpy(ind, ctx->ind.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost));
    if (blocked) HIP_TRY(hipMemcpy(blocked, ctx->bis.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost));
    if (coeffs && ctx->order > 0) {
        size_t n = ctx->order + 1;
        std::vector<double> cu(ctx->nslots * n);
        HIP_TRY(hipMemcpy(cu.data(), ctx->coef.p, sizeof(double) * cu.size(), hipMemcpyDeviceToHost));
        for (size_t s = 0; s < (size_t)ctx->nseries; s++) {
            const size_t k0 = (size_t)ctx->host_slots[s] * n;
            for (size_t k = 0; k < n; k++) coeffs[s * n + k] = cu[k0 + k] * ctx->coef_scale[k];
        }
    }
    if (ctx->did_fit) {
        std::vector<float> out(np * 20), c2(np);
        HIP_TRY(hipMemcpy(out.data(), ctx->fout.p, sizeof(float) * out.size(), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(c2.data(), ctx->fchi2.p, sizeof(float) * np, hipMemcpyDeviceToHost));
        if (fitres)
            for (size_t p = 0; p < np; p++)
                for (int k = 0; k < 11; k++) fitres[p * 11 + k] = out[p * 20 + k];
        if (chi2)
            for (size_t p = 0; p < np; p++) chi2[p] = c2[p] / ctx->chi2_norm[(int)p < ctx->fit_split ? 0 : 1];
        if (states) HIP_TRY(hipMemcpy(states, ctx->fstate.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost));
        if (iters) HIP_TRY(hipMemcpy(iters, ctx->fiter.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost));
    }
    return 0;
}

// Fit results of a run into the caller's device arrays: parameters 0..10
// (SI units) as double, chi-square normalised per correlation kind.
__global__ void k_export_fit(int np, const float *fout, const float *fchi2, int split, double norm0, double norm1,
                             double *fitres, double *chi2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    for (int k = 0; k < 11; k++) fitres[(size_t)p * 11 + k] = (double)fout[(size_t)p * 20 + k];
    chi2[p] = (double)fchi2[p] / (p < split ? norm0 : norm1);
}

// imfcs_run without device allocation: every buffer of the run is carved
// from the file-scope scratch array of the device (kScratchBytes), the
// results go to the caller's device arrays.  The context's own buffers are
// set aside for the call and restored after it (its enqueued work keeps the
// scratch addresses it was given).
extern "C" int imfcs_run_noalloc(imfcs_ctx *ctx, const uint16_t *d_stack, const imfcs_config *cfg, double *acf,
                                 double *var, int32_t *ind, int32_t *blocked, double *fitres, double *chi2,
                                 int32_t *states, int32_t *iters) {
    if (!ctx || !d_stack || !cfg || !acf || !var || !ind || !blocked) return fail(IMFCS_ERR_ARG, "null argument");
    if (cfg->do_fit && (!fitres || !chi2 || !states || !iters)) return fail(IMFCS_ERR_ARG, "null fit output");
    if (ctx->seeder[0] || ctx->seeder[1])
        return fail(IMFCS_ERR_ARG, "imfcs_run_noalloc does not seed fits (the networks hold their own buffers)");
    HIP_TRY(hipSetDevice(ctx->device));
    void *sp = scratch_ptr(ctx->device);
    if (!sp) return fail(IMFCS_ERR_DEVICE, "scratch address of this device unknown (imfcs_create could not look it up)");
    std::vector<DevBuf *> bufs = all_bufs(ctx);
    std::vector<DevBuf> saved;
    saved.reserve(bufs.size());
    for (DevBuf *b : bufs) {
        saved.push_back(*b);
        *b = DevBuf{};
    }
    invalidate_uploads(ctx);
    Arena ar{(unsigned char *)sp, kScratchBytes, 0, 0, false};
    g_arena = &ar;
    int rc = run_config(ctx, d_stack, cfg, nullptr);
    if (rc == 0) {
        hipStream_t st = ctx->stream;
        const size_t np = ctx->npairs, nc = ctx->chanum;
        hipError_t e = hipMemcpyAsync(acf, ctx->corav.p, sizeof(double) * np * nc, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(var, ctx->bvar.p, sizeof(double) * np * nc, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(ind, ctx->ind.p, sizeof(int32_t) * np, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(blocked, ctx->bis.p, sizeof(int32_t) * np, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess && cfg->do_fit) {
            hipLaunchKernelGGL(k_export_fit, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, (int)np,
                               ctx->fout.as<float>(), ctx->fchi2.as<float>(), ctx->fit_split, ctx->chi2_norm[0],
                               ctx->chi2_norm[1], fitres, chi2);
            e = hipGetLastError();
            if (e == hipSuccess)
                e = hipMemcpyAsync(states, ctx->fstate.p, sizeof(int32_t) * np, hipMemcpyDeviceToDevice, st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(iters, ctx->fiter.p, sizeof(int32_t) * np, hipMemcpyDeviceToDevice, st);
        }
        if (e != hipSuccess) rc = fail(IMFCS_ERR_DEVICE, std::string("imfcs_run_noalloc: ") + hipGetErrorString(e));
    }
    g_arena = nullptr;
    if (rc != 0 && ar.overflow)
        rc = fail(IMFCS_ERR_MEMORY, "imfcs_run_noalloc: this configuration needs more than " +
                                        std::to_string(kScratchBytes >> 20) + " MiB of scratch (at least " +
                                        std::to_string(ar.need) + " bytes)");
    for (size_t i = 0; i < bufs.size(); i++) *bufs[i] = saved[i];
    invalidate_uploads(ctx);
    ctx->has_last = 0;
    ctx->evk_valid = 0;
    ctx->last_noalloc = 1;
    return rc;
}


// ------------------------------------------------------------------
// Batches of stacks: one hipGraph of every stack's pipeline, one persistent
// fit launch over all of them, autotuned frame chunking
// ------------------------------------------------------------------
struct imfcs_batch {
    int device = 0, flags = 0;
    hipStream_t st = nullptr;
    std::vector<imfcs_ctx *> ctx;
    std::vector<hipEvent_t> ev_join;
    hipEvent_t ev_fork = nullptr, ev_fit = nullptr;
    DevBuf groups, prefix, counter;
    DevBuf psums;   // bleach-point sums of a stack every configuration shares
    CachedUpload up_groups, up_prefix;
    std::vector<LMGroup> host_groups;
    std::vector<int64_t> host_prefix;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    size_t graph_nodes = 0;
    std::vector<imfcs_config> last_cfgs;
    std::vector<const uint16_t *> last_stacks;
    std::vector<const void *> last_seeders;   // the seeding networks of each context at capture
    int lm_blocks = 0;
    double autotune_ms = 0.0;
    int tuned_shapes = 0;
};

namespace {

void batch_drop_graph(imfcs_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_nodes = 0;
}

// Every stack's pipeline on its own context stream, forked from and joined
// back into the batch stream, then the fits of all of them in one launch.
int batch_enqueue(imfcs_batch *b, const uint16_t *const *stacks, const imfcs_config *cfgs) {
    const int n = (int)b->ctx.size();
    // one stack under every configuration (the binnings of a diffusion law):
    // the bleach traces of all of them from one pass over it, as imfcs_run_multi
    bool one_stack = true;
    for (int i = 1; i < n && one_stack; i++) one_stack = stacks[i] == stacks[0];
    const bool shared = one_stack && can_share_point_sums(b->ctx.data(), n, cfgs);
    if (shared) {
        const int rc = enqueue_point_sums(b->psums, stacks[0], cfgs[0], b->st);
        if (rc) return rc;
    }
    HIP_TRY(hipEventRecord(b->ev_fork, b->st));
    b->host_groups.clear();
    const bool own_fits = (b->flags & IMFCS_BATCH_OWN_FITS) != 0;
    for (int i = 0; i < n; i++) {
        imfcs_ctx *c = b->ctx[i];
        HIP_TRY(hipStreamWaitEvent(c->stream, b->ev_fork, 0));
        c->last_noalloc = 0;
        const int rc = run_config(c, stacks[i], &cfgs[i], shared ? b->psums.as<int32_t>() : nullptr,
                                  own_fits ? nullptr : &b->host_groups);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(b->ev_join[i], c->stream));
        HIP_TRY(hipStreamWaitEvent(b->st, b->ev_join[i], 0));
    }
    const int ng = (int)b->host_groups.size();
    if (ng > 0) {
        b->host_prefix.assign(ng + 1, 0);
        for (int g = 0; g < ng; g++) b->host_prefix[g + 1] = b->host_prefix[g] + b->host_groups[g].nfits;
        HIP_TRY(b->up_groups.put(b->groups, b->host_groups.data(), sizeof(LMGroup) * ng, b->st));
        HIP_TRY(b->up_prefix.put(b->prefix, b->host_prefix.data(), sizeof(int64_t) * (ng + 1), b->st));
        HIP_TRY(b->counter.ensure(sizeof(unsigned long long)));
        launch_lm_queue(b->groups.as<LMGroup>(), b->prefix.as<int64_t>(), ng, b->host_prefix[ng],
                        b->counter.as<unsigned long long>(), b->lm_blocks, b->st);
        launch_lm_polish_groups(b->groups.as<LMGroup>(), b->prefix.as<int64_t>(), ng, b->host_prefix[ng], b->st);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

// Later work on each context (results, further runs) follows the batch.
int batch_release_contexts(imfcs_batch *b) {
    HIP_TRY(hipEventRecord(b->ev_fit, b->st));
    for (imfcs_ctx *c : b->ctx) HIP_TRY(hipStreamWaitEvent(c->stream, b->ev_fit, 0));
    return 0;
}

// Autotuner: for every stack whose series shape (slots, frames) has no
// entry yet, the bleach-correction/prefix pass and the blocking pass are
// timed (HIP events around each stage) with 2^17 .. 2^21 threads in flight,
// one pass after the other, and the fastest choices are recorded.
int batch_autotune(imfcs_batch *b, const uint16_t *const *stacks, const imfcs_config *cfgs) {
    const auto t0 = std::chrono::steady_clock::now();
    for (size_t i = 0; i < b->ctx.size(); i++) {
        imfcs_ctx *c = b->ctx[i];
        imfcs_config cf = cfgs[i];
        cf.do_fit = 0;
        c->force_tune = Tune{0, 0};
        int rc = run_config(c, stacks[i], &cf, nullptr);   // warm-up: allocations, and the slot count
        if (rc) return rc;
        HIP_TRY(hipStreamSynchronize(c->stream));
        const int T = cf.lastframe - cf.firstframe + 1;
        const std::pair<int, int> key{c->nslots, T};
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            if (g_tune.count(key)) continue;
        }
        Tune best{19, 19};
        float best_ms[2] = {1e30f, 1e30f};
        for (int pass = 0; pass < 2; pass++)
            for (int e = 17; e <= 21; e++) {
                Tune t = best;
                (pass == 0 ? t.apply_log2 : t.block_log2) = e;
                c->force_tune = t;
                rc = run_config(c, stacks[i], &cf, nullptr);
                if (rc) {
                    c->force_tune = Tune{0, 0};
                    return rc;
                }
                float ms[5];
                HIP_TRY(hipStreamSynchronize(c->stream));
                for (int k = 0; k < 5; k++) HIP_TRY(hipEventElapsedTime(&ms[k], c->ev[k], c->ev[k + 1]));
                const float m = pass == 0 ? ms[1] : ms[2];
                if (m < best_ms[pass]) {
                    best_ms[pass] = m;
                    best = t;
                }
            }
        c->force_tune = Tune{0, 0};
        tune_store(key.first, key.second, best);
        b->tuned_shapes++;
    }
    b->autotune_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

std::vector<const void *> batch_seeders(const imfcs_batch *b) {
    std::vector<const void *> v;
    for (const imfcs_ctx *c : b->ctx) {
        v.push_back(c->seeder[0]);
        v.push_back(c->seeder[1]);
    }
    return v;
}

bool same_batch(const imfcs_batch *b, const uint16_t *const *stacks, const imfcs_config *cfgs) {
    if (!b->exec || b->last_cfgs.size() != b->ctx.size()) return false;
    for (size_t i = 0; i < b->ctx.size(); i++)
        if (stacks[i] != b->last_stacks[i] || std::memcmp(&cfgs[i], &b->last_cfgs[i], sizeof(imfcs_config)) != 0)
            return false;
    return batch_seeders(b) == b->last_seeders;
}

}  // namespace

extern "C" imfcs_batch *imfcs_batch_create(int device, int n, int flags) {
    if (n < 1) {
        fail(IMFCS_ERR_ARG, "a batch needs at least one stack");
        return nullptr;
    }
    imfcs_batch *b = new imfcs_batch();
    b->device = device;
    b->flags = flags;
    for (int i = 0; i < n; i++) {
        imfcs_ctx *c = imfcs_create(device);
        if (!c) {
            imfcs_batch_destroy(b);
            return nullptr;
        }
        b->ctx.push_back(c);
        hipEvent_t e = nullptr;
        (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
        b->ev_join.push_back(e);
    }
    if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) {
        fail(IMFCS_ERR_DEVICE, "hipStreamCreate failed");
        imfcs_batch_destroy(b);
        return nullptr;
    }
    (void)hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming);
    (void)hipEventCreateWithFlags(&b->ev_fit, hipEventDisableTiming);
    b->lm_blocks = lm_queue_blocks();
    return b;
}

extern "C" void imfcs_batch_destroy(imfcs_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->st) (void)hipStreamSynchronize(b->st);
    batch_drop_graph(b);
    for (imfcs_ctx *c : b->ctx) imfcs_destroy(c);
    for (hipEvent_t e : b->ev_join)
        if (e) (void)hipEventDestroy(e);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_fit) (void)hipEventDestroy(b->ev_fit);
    b->groups.release();
    b->prefix.release();
    b->counter.release();
    if (b->st) (void)hipStreamDestroy(b->st);
    delete b;
}

extern "C" int imfcs_batch_run(imfcs_batch *b, const uint16_t *const *stacks, const imfcs_config *cfgs) {
    if (!b || !stacks || !cfgs) return fail(IMFCS_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(b->device));
    for (size_t i = 0; i < b->ctx.size(); i++) {
        if (!stacks[i]) return fail(IMFCS_ERR_ARG, "null stack " + std::to_string(i));
        if ((b->ctx[i]->seeder[0] || b->ctx[i]->seeder[1]) && !(b->flags & IMFCS_BATCH_OWN_FITS))
            return fail(IMFCS_ERR_ARG, "seeded pipelines need IMFCS_BATCH_OWN_FITS");
    }
    if (same_batch(b, stacks, cfgs)) {
        HIP_TRY(hipGraphLaunch(b->exec, b->st));
        return batch_release_contexts(b);
    }
    batch_drop_graph(b);
    if (b->flags & IMFCS_BATCH_AUTOTUNE) {
        const int rc = batch_autotune(b, stacks, cfgs);
        if (rc) return rc;
    }
    // first run of these stacks: eagerly (allocations, parameter uploads)
    int rc = batch_enqueue(b, stacks, cfgs);
    if (rc) return rc;
    rc = batch_release_contexts(b);
    if (rc) return rc;
    if (b->flags & IMFCS_BATCH_GRAPH) {
        // the same enqueue captured into a graph for the next runs: buffers
        // and uploads are in place, so only kernels, event records and
        // stream waits are recorded
        HIP_TRY(hipStreamSynchronize(b->st));
        for (imfcs_ctx *c : b->ctx) HIP_TRY(hipStreamSynchronize(c->stream));
        HIP_TRY(hipStreamBeginCapture(b->st, hipStreamCaptureModeRelaxed));
        rc = batch_enqueue(b, stacks, cfgs);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(b->st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess) return fail(IMFCS_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(ec));
        b->graph = g;
        HIP_TRY(hipGraphGetNodes(g, nullptr, &b->graph_nodes));
        HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    }
    b->last_cfgs.assign(cfgs, cfgs + b->ctx.size());
    b->last_stacks.assign(stacks, stacks + b->ctx.size());
    b->last_seeders = batch_seeders(b);
    return 0;
}

extern "C" int imfcs_batch_synchronize(imfcs_batch *b) {
    if (!b) return fail(IMFCS_ERR_ARG, "null batch");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->st));
    for (imfcs_ctx *c : b->ctx) HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" imfcs_ctx *imfcs_batch_context(imfcs_batch *b, int i) {
    if (!b || i < 0 || i >= (int)b->ctx.size()) {
        fail(IMFCS_ERR_ARG, "batch context index out of range");
        return nullptr;
    }
    return b->ctx[i];
}

extern "C" int imfcs_batch_info(imfcs_batch *b, int *graph_nodes, int *fit_groups, int *tuned_shapes,
                                double *autotune_ms) {
    if (!b) return fail(IMFCS_ERR_ARG, "null batch");
    if (graph_nodes) *graph_nodes = (int)b->graph_nodes;
    if (fit_groups) *fit_groups = (int)b->host_groups.size();
    if (tuned_shapes) *tuned_shapes = b->tuned_shapes;
    if (autotune_ms) *autotune_ms = b->autotune_ms;
    return 0;
}

extern "C" int imfcs_get_tuning(int nslots, int frames, int *apply_log2, int *block_log2) {
    const Tune t = tune_lookup(nslots, frames);
    if (apply_log2) *apply_log2 = t.apply_log2;
    if (block_log2) *block_log2 = t.block_log2;
    return 0;
}

extern "C" int imfcs_get_slot_count(imfcs_ctx *ctx, int *nslots) {
    if (!ctx || !nslots) return fail(IMFCS_ERR_ARG, "null argument");
    *nslots = ctx->nslots;
    return 0;
}

extern "C" int imfcs_get_run_flags(imfcs_ctx *ctx, int *flags) {
    if (!ctx || !flags) return fail(IMFCS_ERR_ARG, "null argument");
    *flags = ctx->run_nested ? IMFCS_RUN_NESTED : 0;
    return 0;
}

// ------------------------------------------------------------------
// Drop-in layer
// ------------------------------------------------------------------
namespace {

std::mutex g_dropin_mu;
imfcs_ctx *g_dropin_ctx = nullptr;

imfcs_ctx *dropin_ctx() {
    if (!g_dropin_ctx) g_dropin_ctx = imfcs_create(0);
    return g_dropin_ctx;
}

// free device memory fraction test used by the is*MemorySufficient calls
int mem_sufficient(double bytes) {
    size_t fr = 0, tot = 0;
    if (!device_present()) return fail(IMFCS_ERR_DEVICE, "no HIP device available");
    HIP_TRY(hipMemGetInfo(&fr, &tot));
    return (bytes > (double)fr * 0.9) ? 0 : 1;
}

}  // namespace

extern "C" int imfcs_reset_gpu(void) {
    std::lock_guard<std::mutex> lk(g_dropin_mu);
    if (g_dropin_ctx) {
        imfcs_destroy(g_dropin_ctx);
        g_dropin_ctx = nullptr;
    }
    return 0;
}

extern "C" int imfcs_reset_device(void) { return 0; }

extern "C" int imfcs_is_binning_memory_sufficient(const imfcs_acf_params *p) {
    if (!p) return fail(IMFCS_ERR_ARG, "null parameters");
    int T = p->lastframe - p->firstframe + 1;
    double bytes = ((double)p->win_star * p->hin_star + (double)p->w_temp * p->h_temp) * T * 4.0;
    return mem_sufficient(bytes);
}

extern "C" int imfcs_is_acf_memory_sufficient(const imfcs_acf_params *p) {
    if (!p) return fail(IMFCS_ERR_ARG, "null parameters");
    // float input, per-pair outputs and the per-chunk partial sums of the
    // streaming kernels (no T x pixels product buffer as in .cu:5031-5084)
    double T = p->lastframe - p->firstframe + 1;
    double pix = (double)p->width * p->height;
    double bytes = (double)p->w_temp * p->h_temp * T * 4.0 + pix * p->chanum * 8.0 * 6 + pix * 2 * 400 * 8.0 +
                   pix * 2 * 64.0 * (T / 4096 + 20);
    return mem_sufficient(bytes);
}

extern "C" int imfcs_calc_binning(const float *indata, float *outdata, const imfcs_acf_params *p) {
    if (!indata || !outdata || !p) return fail(IMFCS_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(g_dropin_mu);
    imfcs_ctx *ctx = dropin_ctx();
    if (!ctx) return IMFCS_ERR_DEVICE;
    int T = p->lastframe - p->firstframe + 1;
    size_t nin = (size_t)p->win_star * p->hin_star * T, nout = (size_t)p->w_temp * p->h_temp * T;
    HIP_TRY(ctx->part.ensure(nin * 4));
    HIP_TRY(ctx->fdata.ensure(nout * 4));
    HIP_TRY(hipMemcpyAsync(ctx->part.p, indata, nin * 4, hipMemcpyHostToDevice, ctx->stream));
    launch_binning_f32(ctx->part.as<float>(), ctx->fdata.as<float>(), p->win_star, p->hin_star, p->w_temp,
                       p->h_temp, T, p->binningX, p->binningY, ctx->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(outdata, ctx->fdata.p, nout * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

extern "C" int imfcs_calc_data_bleach_correction(const float *pixels, float *outdata, const imfcs_acf_params *p) {
    if (!pixels || !outdata || !p) return fail(IMFCS_ERR_ARG, "null argument");
    if (p->nopit <= 0 || p->ave <= 0) return fail(IMFCS_ERR_ARG, "nopit and ave must be positive");
    std::lock_guard<std::mutex> lk(g_dropin_mu);
    imfcs_ctx *ctx = dropin_ctx();
    if (!ctx) return IMFCS_ERR_DEVICE;
    int T = p->lastframe - p->firstframe + 1;
    size_t nin = (size_t)p->w_temp * p->h_temp * T, nout = (size_t)p->w_temp * p->h_temp * p->nopit;
    HIP_TRY(ctx->part.ensure(nin * 4));
    HIP_TRY(ctx->fdata.ensure(nout * 4));
    HIP_TRY(hipMemcpyAsync(ctx->part.p, pixels, nin * 4, hipMemcpyHostToDevice, ctx->stream));
    launch_bleach_average_f32(ctx->part.as<float>(), ctx->fdata.as<float>(), p->w_temp, p->h_temp, p->nopit, p->ave,
                              ctx->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(outdata, ctx->fdata.p, nout * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

extern "C" int imfcs_calc_acf(const float *pixels, double *pixels1, double *blockvararray, double *NBmean,
                              double *NBcov, double *blocked1D, const double *bleachcorr_params, const double *samp,
                              const int *lag, const imfcs_acf_params *p) {
    if (!pixels || !pixels1 || !blockvararray || !blocked1D || !p) return fail(IMFCS_ERR_ARG, "null argument");
    if (p->isNBcalculation && (!NBmean || !NBcov)) return fail(IMFCS_ERR_ARG, "null N&B output");
    // host-side argument checks, before any device work
    if (p->cfXDistance < 0 || p->cfYDistance < 0)
        return fail(IMFCS_ERR_ARG, "negative CF distances are not supported (as in IFCS:15296-15299)");
    if (p->width <= 0 || p->height <= 0 || p->pixbinX <= 0 || p->pixbinY <= 0)
        return fail(IMFCS_ERR_ARG, "empty output grid or non-positive pixel pitch");
    if (p->lastframe - p->firstframe + 1 < 2) return fail(IMFCS_ERR_ARG, "frame range holds fewer than 2 frames");
    if ((int64_t)(p->width - 1) * p->pixbinX + p->cfXDistance >= p->w_temp ||
        (int64_t)(p->height - 1) * p->pixbinY + p->cfYDistance >= p->h_temp)
        return fail(IMFCS_ERR_ARG, "output grid exceeds the binned input");
    if (p->bleachcorr_gpu && (!bleachcorr_params || p->bleachcorr_order < 1 || p->bleachcorr_order - 1 > kMaxOrder))
        return fail(IMFCS_ERR_ARG, "bleach correction needs 1..9 coefficients per pixel");
    if ((double)(int)p->correlatorp != p->correlatorp || (double)(int)p->correlatorq != p->correlatorq)
        return fail(IMFCS_ERR_ARG, "correlator p and q must be integers");
    {
        int rc0 = check_correlator_layout((int)p->correlatorp, (int)p->correlatorq);
        if (rc0) return rc0;
        HostTables ht0;
        rc0 = make_host_tables((int)p->correlatorp, (int)p->correlatorq, p->lastframe - p->firstframe + 1,
                               p->frametime, ht0);
        if (rc0) return rc0;
        // the output arrays are sized by the caller from p->chanum
        if (p->chanum != ht0.chanum)
            return fail(IMFCS_ERR_ARG, "ACFParameters.chanum = " + std::to_string(p->chanum) + " but p, q give " +
                                           std::to_string(ht0.chanum) + " channels");
        // the tables the plugin passes (setParameters, IFCS:7391-7424) must be
        // the ones the kernels derive from p and q
        for (int x = 0; x < ht0.chanum; x++) {
            if (lag && lag[x] != ht0.lag[x])
                return fail(IMFCS_ERR_ARG, "lag table differs from the p, q correlator at channel " + std::to_string(x));
            if (samp && samp[x] != ht0.samp[x])
                return fail(IMFCS_ERR_ARG, "samp table differs from the p, q correlator at channel " + std::to_string(x));
        }
    }
    std::lock_guard<std::mutex> lk(g_dropin_mu);
    imfcs_ctx *ctx = dropin_ctx();
    if (!ctx) return IMFCS_ERR_DEVICE;
    int T = p->lastframe - p->firstframe + 1;
    size_t nin = (size_t)p->w_temp * p->h_temp * T;
    DevBuf &input = ctx->dropin_in;
    HIP_TRY(input.ensure(nin * 4));
    HIP_TRY(hipMemcpyAsync(input.p, pixels, nin * 4, hipMemcpyHostToDevice, ctx->stream));
    EngineArgs ea{};
    SeriesGeom &g = ea.g;
    g.data = input.p;
    g.src = kSrcF32;
    g.plane = (int64_t)p->w_temp * p->h_temp;
    g.stride = p->w_temp;
    g.f0 = 0;
    g.T = T;
    g.bx = g.by = 1;
    g.bckg_total = 0;
    g.ow = p->width;
    g.oh = p->height;
    g.npix = p->width * p->height;
    g.ox0 = g.oy0 = 0;
    g.pbx = p->pixbinX;
    g.pby = p->pixbinY;
    g.cfx = p->cfXDistance;
    g.cfy = p->cfYDistance;
    bool ccf = (p->cfXDistance != 0 || p->cfYDistance != 0);
    g.nseries = g.npix * (ccf ? 2 : 1);
    int order = p->bleachcorr_gpu ? p->bleachcorr_order - 1 : 0;
    std::vector<double> hc;
    if (order > 0) {
        int n = order + 1;
        hc.resize((size_t)g.nseries * n);
        for (int s = 0; s < g.nseries; s++) {
            int k = s < g.npix ? s : s - g.npix;
            int ox = (k % g.ow) * g.pbx + (s >= g.npix ? g.cfx : 0);
            int oy = (k / g.ow) * g.pby + (s >= g.npix ? g.cfy : 0);
            const double *src = bleachcorr_params + ((size_t)oy * p->w_temp + ox) * p->bleachcorr_order;
            for (int i = 0; i < n; i++) hc[(size_t)s * n + i] = src[i];
        }
    }
    ea.order = order;
    ea.fit_trace = 0;
    ea.host_coef = hc.data();
    ea.frametime = p->frametime;
    ea.t_off = 0.5;
    ea.p = (int)p->correlatorp;
    ea.q = (int)p->correlatorq;
    ea.pd.npix = g.npix;
    ea.pd.acf_first = ccf ? 0 : 1;
    ea.pd.has_ccf = ccf ? 1 : 0;
    ea.pd.npairs = g.npix;
    ctx->npix = g.npix;
    HostTables ht;
    CorrTables ct;
    int rc = run_engine(ctx, ea, ht, ct);
    if (rc == 0 && p->isNBcalculation) {
        HIP_TRY(ctx->nbm.ensure(sizeof(double) * (size_t)g.npix));
        HIP_TRY(ctx->nbc.ensure(sizeof(double) * (size_t)g.npix));
        launch_nb_extract(ctx->ctab.as<CorrTables>(), ctx->last_g, ea.pd, ctx->F.as<double>(), ctx->corav.as<double>(),
                          ctx->nbm.as<double>(), ctx->nbc.as<double>(), ctx->stream);
        HIP_TRY(hipGetLastError());
    }
    if (rc == 0) {
        size_t np = g.npix, nc = ct.chanum;
        std::vector<double> cav(np * nc), bv(np * nc);
        std::vector<int32_t> iv(np), bi(np);
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess && p->isNBcalculation) {
            e = hipMemcpy(NBmean, ctx->nbm.p, sizeof(double) * np, hipMemcpyDeviceToHost);
            if (e == hipSuccess) e = hipMemcpy(NBcov, ctx->nbc.p, sizeof(double) * np, hipMemcpyDeviceToHost);
        }
        if (e == hipSuccess) e = hipMemcpy(cav.data(), ctx->corav.p, sizeof(double) * np * nc, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(bv.data(), ctx->bvar.p, sizeof(double) * np * nc, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(iv.data(), ctx->ind.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(bi.data(), ctx->bis.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            rc = fail(IMFCS_ERR_DEVICE, std::string("copy back: ") + hipGetErrorString(e));
        } else {
            // plane-major [z][y][x] outputs (agpufitjni.cu:5748-5774)
            for (size_t k = 0; k < np; k++)
                for (size_t z = 0; z < nc; z++) {
                    pixels1[z * np + k] = cav[k * nc + z];
                    blockvararray[z * np + k] = bv[k * nc + z];
                }
            for (size_t i = 0; i < np * nc; i++) blocked1D[i] = 0.0;
            for (size_t k = 0; k < np; k++) {
                blocked1D[k] = iv[k];
                blocked1D[np + k] = bi[k];
            }
        }
    }
    (void)hipStreamSynchronize(ctx->stream);
    return rc;
}

extern "C" int imfcs_fit(int n_fits, int n_points, const float *data, const float *weights, int model_id,
                         const float *initial_parameters, float tolerance, int max_n_iterations, int num_valid_coefs,
                         const int *parameters_to_fit, int estimator_id, size_t user_info_size, const void *user_info,
                         float *output_parameters, int *output_states, float *output_chi_squares,
                         int *output_n_iterations) {
    if (n_fits <= 0 || n_points <= 0) return fail(IMFCS_ERROR, "number of fits and points must be positive");
    if (!data || !initial_parameters || !parameters_to_fit || !output_parameters || !output_states ||
        !output_chi_squares || !output_n_iterations)
        return fail(IMFCS_ERROR, "null buffer");
    if (estimator_id != IMFCS_ESTIMATOR_LSE) return fail(IMFCS_ERROR, "only the LSE estimator is implemented");
    int npar;
    if (model_id == IMFCS_MODEL_ACF_1D) npar = 20;
    else if (model_id == IMFCS_MODEL_LINEAR_1D) npar = 11;
    else return fail(IMFCS_ERROR, "unsupported model id");
    if (!user_info || user_info_size < sizeof(float) * (size_t)n_points)
        return fail(IMFCS_ERROR, "user_info must hold one float abscissa per point");
    if (model_id == IMFCS_MODEL_LINEAR_1D && (num_valid_coefs < 1 || num_valid_coefs > 11))
        return fail(IMFCS_ERROR, "num_valid_coefs out of range");
    std::lock_guard<std::mutex> lk(g_dropin_mu);
    imfcs_ctx *ctx = dropin_ctx();
    if (!ctx) return IMFCS_ERROR;
    hipStream_t st = ctx->stream;
    size_t nd = (size_t)n_fits * n_points;
    HIP_TRY(ctx->fdata.ensure(nd * 4));
    HIP_TRY(ctx->fw.ensure(nd * 4));
    HIP_TRY(ctx->fx.ensure((size_t)n_points * 4));
    HIP_TRY(ctx->finit.ensure((size_t)n_fits * npar * 4));
    HIP_TRY(ctx->fpfit.ensure(npar * sizeof(int)));
    ctx->up_fx.invalidate();
    ctx->up_finit.invalidate();
    ctx->up_fpfit.invalidate();
    HIP_TRY(ctx->fout.ensure((size_t)n_fits * npar * 4));
    HIP_TRY(ctx->fstate.ensure((size_t)n_fits * 4));
    HIP_TRY(ctx->fchi2.ensure((size_t)n_fits * 4));
    HIP_TRY(ctx->fiter.ensure((size_t)n_fits * 4));
    HIP_TRY(hipMemcpyAsync(ctx->fdata.p, data, nd * 4, hipMemcpyHostToDevice, st));
    if (weights) HIP_TRY(hipMemcpyAsync(ctx->fw.p, weights, nd * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(ctx->fx.p, user_info, (size_t)n_points * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(ctx->finit.p, initial_parameters, (size_t)n_fits * npar * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(ctx->fpfit.p, parameters_to_fit, npar * sizeof(int), hipMemcpyHostToDevice, st));
    const float *w = weights ? ctx->fw.as<float>() : nullptr;
    unsigned mask = 0;
    for (int k = 0; k < npar && k < 32; k++) mask |= (parameters_to_fit[k] ? 1u : 0u) << k;
    if (model_id == IMFCS_MODEL_ACF_1D)
        launch_lm_acf(n_fits, n_points, ctx->fdata.as<float>(), w, ctx->fx.as<float>(), ctx->finit.as<float>(),
                      ctx->fpfit.as<int>(), mask & 0x7ffu, LMModelGeom{}, tolerance, max_n_iterations, 0,
                      ctx->fout.as<float>(), ctx->fstate.as<int>(), ctx->fchi2.as<float>(), ctx->fiter.as<int>(), st);
    else
        launch_lm_poly(n_fits, n_points, ctx->fdata.as<float>(), w, ctx->fx.as<float>(), ctx->finit.as<float>(),
                       ctx->fpfit.as<int>(), mask & ((1u << num_valid_coefs) - 1u), num_valid_coefs, tolerance,
                       max_n_iterations, 0, ctx->fout.as<float>(), ctx->fstate.as<int>(), ctx->fchi2.as<float>(),
                       ctx->fiter.as<int>(), st);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(output_parameters, ctx->fout.p, (size_t)n_fits * npar * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(output_states, ctx->fstate.p, (size_t)n_fits * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(output_chi_squares, ctx->fchi2.p, (size_t)n_fits * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(output_n_iterations, ctx->fiter.p, (size_t)n_fits * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}
