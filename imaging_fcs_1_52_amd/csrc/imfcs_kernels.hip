// HIP kernels of the imaging-FCS engine for CDNA4 (gfx950).
//
// Stage map (reference semantics followed, IFCS = Imaging_FCS_1_52.java):
//   k_trace_moments / k_polyfit : calcIntensityTrace (IFCS:9822-9855) and the
//        least-squares polynomial of PolynomFit (IFCS:10953-11004), solved
//        as FP64 normal equations in a normalized abscissa.
//   bleach correction (inline)  : getIntensity polynomial branch
//        (IFCS:9734-9750), applied on the fly to every loaded value.
//   k_prefix / k_prefix_combine : the window means ("direct/delayed
//        monitors") of every channel (IFCS:9487-9494, 9262-9270).
//   k_blocking / _finalize      : blockTransform (IFCS:9217-9422).
//   k_correlate                 : calculateCF (IFCS:9425-9521) -- G and the
//        blocked variance of every channel.
//   k_lm                        : batched Levenberg-Marquardt with the
//        step control of Gpufit (agpufitjni.cu:2044-2458), model FCS_3p
//        (IFCS:13131-13408) or a polynomial (agpufitjni.cu:362-420).
//
// All series arithmetic is FP64 (full rate on CDNA4 VALU); the LM model is
// evaluated in FP32 and its normal equations are accumulated and solved in
// FP64.
#include "imfcs_internal.h"
#include "imfcs_device.h"
#include "imfcs_rng.h"

#include <math.h>
#include <algorithm>
#include <type_traits>

namespace imfcs {

// ------------------------------------------------------------------
// Stage 1: intensity trace moments and polynomial fit
// ------------------------------------------------------------------
template <int NB>
__global__ void __launch_bounds__(64) k_trace_moments(SeriesGeom g, int nopit, int ave, double tscale,
                                                      double frametime, int order, int integer_trace,
                                                      int pts_per_chunk, double *part, float *trace_out) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;   // slot
    int chunk = blockIdx.y;
    if (s >= g.nslots) return;
    int ox, oy;
    slot_origin(g, s, ox, oy);
    int x0 = chunk * pts_per_chunk, x1 = min(nopit, x0 + pts_per_chunk);
    double acc[kMaxOrder + 1];
#pragma unroll
    for (int k = 0; k <= kMaxOrder; k++) acc[k] = 0.0;
    for (int x = x0; x < x1; x++) {
        double y;
        if (integer_trace && g.src == kSrcU16) {
            // Java int sum and int division by ave (IFCS:9841-9852)
            int isum = 0;
            const int fe = (x + 1) * ave;
            for (int f = x * ave; f < fe; f += 16) {
                int v[16];
                raw_ints16<NB>(g, ox, oy, f, fe - 1, v);
#pragma unroll
                for (int i = 0; i < 16; i++) isum += (f + i < fe) ? v[i] : 0;
            }
            y = (double)(isum / ave);
        } else {
            double sum = 0.0;
            for (int f = x * ave; f < (x + 1) * ave; f++) sum += raw_value(g, ox, oy, f);
            y = sum / ave;
        }
        if (trace_out) trace_out[(int64_t)s * nopit + x] = (float)y;
        double u = frametime * (x + 0.5) * ave / tscale, up = 1.0;
#pragma unroll
        for (int k = 0; k <= kMaxOrder; k++) {
            if (k <= order) acc[k] += y * up;
            up *= u;
        }
    }
    int64_t base = ((int64_t)chunk * g.nslots + s) * (order + 1);
#pragma unroll
    for (int k = 0; k <= kMaxOrder; k++)
        if (k <= order) part[base + k] = acc[k];
}

// Per-pixel integer sums of every bleach-trace point (frames x*ave ..
// (x+1)*ave - 1 after the first frame) over the whole uint16 frame, for
// several binnings of one stack: the trace of a window is the sum of its
// pixels' point sums (the Java integer division, IFCS:9852, comes after the
// sum), so one pass over the stack serves every binning.
// P[x][plane], int32; VEC consecutive pixels per lane (VEC = 4: one 8-byte
// load per frame).
template <int VEC>
__global__ void __launch_bounds__(256) k_point_sums(const uint16_t *__restrict__ data, int64_t plane, int f0,
                                                    int nopit, int ave, int pts_per_thread, int32_t *P) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;   // pixel group
    if (q * VEC >= plane) return;
    const int x0 = blockIdx.y * pts_per_thread, x1 = min(nopit, x0 + pts_per_thread);
    // the frames of the thread's points in batches of 16 rows (a point's
    // last batch is partial); the next batch's loads are issued before the
    // current one is summed
    using Word = typename std::conditional<VEC == 4, uint2, uint16_t>::type;
    const int nb_pt = (ave + 15) / 16, nbatch = (x1 - x0) * nb_pt;
    auto load = [&](int b, Word (&r)[16]) {
        const int x = x0 + b / nb_pt, f = (b % nb_pt) * 16;   // wave-uniform
        const int nv = min(16, ave - f);
        const uint16_t *base = data + ((int64_t)f0 + (int64_t)x * ave + f) * plane + q * VEC;
#pragma unroll
        for (int j = 0; j < 16; j++) r[j] = *(const Word *)(base + (int64_t)min(j, nv - 1) * plane);
    };
    Word cur[16];
    if (nbatch > 0) load(0, cur);
    int acc[VEC];
#pragma unroll
    for (int v = 0; v < VEC; v++) acc[v] = 0;
    for (int b = 0; b < nbatch; b++) {
        Word nxt[16];
        if (b + 1 < nbatch) load(b + 1, nxt);
        const int pb = b % nb_pt, nv = min(16, ave - pb * 16);
#pragma unroll
        for (int j = 0; j < 16; j++)
            if (j < nv) {
                if constexpr (VEC == 4) {
                    acc[0] += (int)(cur[j].x & 0xffffu);
                    acc[1] += (int)(cur[j].x >> 16);
                    acc[2] += (int)(cur[j].y & 0xffffu);
                    acc[3] += (int)(cur[j].y >> 16);
                } else {
                    acc[0] += cur[j];
                }
            }
        if (pb == nb_pt - 1) {
            int32_t *out = P + (int64_t)(x0 + b / nb_pt) * plane + q * VEC;
            if constexpr (VEC == 4)
                *(int4 *)out = make_int4(acc[0], acc[1], acc[2], acc[3]);
            else
                out[0] = acc[0];
#pragma unroll
            for (int v = 0; v < VEC; v++) acc[v] = 0;
        }
#pragma unroll
        for (int j = 0; j < 16; j++) cur[j] = nxt[j];
    }
}

// k_trace_moments from the point sums: the integer trace of each slot's
// window, y = (sum of P over the window - ave * background) / ave with
// Java's integer division, and its moments; the same points per chunk
// and the same summation order as k_trace_moments (identical results).
__global__ void __launch_bounds__(64) k_trace_points(SeriesGeom g, const int32_t *__restrict__ P, int nopit,
                                                     int ave, double tscale, double frametime, int order,
                                                     int pts_per_chunk, double *part, float *trace_out) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;   // slot
    int chunk = blockIdx.y;
    if (s >= g.nslots) return;
    int ox, oy;
    slot_origin(g, s, ox, oy);
    const int x0 = chunk * pts_per_chunk, x1 = min(nopit, x0 + pts_per_chunk);
    const int64_t lo = (int64_t)oy * g.stride + ox;
    double acc[kMaxOrder + 1];
#pragma unroll
    for (int k = 0; k <= kMaxOrder; k++) acc[k] = 0.0;
    for (int x = x0; x < x1; x++) {
        const int32_t *px = P + (int64_t)x * g.plane + lo;
        int isum = -ave * g.bckg_total;
        for (int k = 0; k < g.by; k++)
            for (int i = 0; i < g.bx; i++) isum += px[(int64_t)k * g.stride + i];
        const double y = (double)(isum / ave);
        if (trace_out) trace_out[(int64_t)s * nopit + x] = (float)y;
        double u = frametime * (x + 0.5) * ave / tscale, up = 1.0;
#pragma unroll
        for (int k = 0; k <= kMaxOrder; k++) {
            if (k <= order) acc[k] += y * up;
            up *= u;
        }
    }
    int64_t base = ((int64_t)chunk * g.nslots + s) * (order + 1);
#pragma unroll
    for (int k = 0; k <= kMaxOrder; k++)
        if (k <= order) part[base + k] = acc[k];
}

struct PolyL {
    double L[(kMaxOrder + 1) * (kMaxOrder + 1)];
};

// Solves M c = r with the shared Cholesky factor of the moment matrix.
// Workgroups of 64 series x kPolyQ chunk quarters: thread (s, j) sums the
// chunks j, j + kPolyQ, ... (the chunk totals are independent loads, so
// four short chains instead of one long one), the quarters meet in LDS and
// are added in the order j = 0, 1, 2, 3.
constexpr int kPolyQ = 4;
__global__ void __launch_bounds__(64 * kPolyQ) k_polyfit(const double *part, int nchunk, int nseries, int order,
                                                        PolyL Lf, double *coef) {
    __shared__ double red[kPolyQ - 1][kMaxOrder + 1][64];
    const int sl = threadIdx.x & 63, j = threadIdx.x >> 6;
    const int s = blockIdx.x * 64 + sl;
    const int n = order + 1;
    double r[kMaxOrder + 1], z[kMaxOrder + 1];
#pragma unroll
    for (int k = 0; k <= kMaxOrder; k++) r[k] = 0.0;
    if (s < nseries) {
#pragma unroll 8
        for (int c = j; c < nchunk; c += kPolyQ) {
            int64_t base = ((int64_t)c * nseries + s) * n;
#pragma unroll
            for (int k = 0; k <= kMaxOrder; k++)
                if (k < n) r[k] += part[base + k];
        }
    }
    if (j > 0)
#pragma unroll
        for (int k = 0; k <= kMaxOrder; k++) red[j - 1][k][sl] = r[k];
    __syncthreads();
    if (j > 0 || s >= nseries) return;
#pragma unroll
    for (int q = 0; q < kPolyQ - 1; q++)
#pragma unroll
        for (int k = 0; k <= kMaxOrder; k++) r[k] += red[q][k][sl];
#pragma unroll
    for (int i = 0; i <= kMaxOrder; i++) {
        if (i < n) {
            double v = r[i];
#pragma unroll
            for (int k = 0; k < i; k++) v -= Lf.L[i * n + k] * z[k];
            z[i] = v / Lf.L[i * n + i];
        }
    }
#pragma unroll
    for (int i = kMaxOrder; i >= 0; i--) {
        if (i < n) {
            double v = z[i];
#pragma unroll
            for (int k = kMaxOrder; k > i; k--)
                if (k < n) v -= Lf.L[k * n + i] * r[k];
            r[i] = v / Lf.L[i * n + i];
        }
    }
#pragma unroll
    for (int k = 0; k <= kMaxOrder; k++)
        if (k < n) coef[(int64_t)s * n + k] = r[k];
}

// ------------------------------------------------------------------
// Stage 2: bleach correction of every slot, written once as FP32
// deviations X = I' - mu, and the FP64 prefix sums of I' at the positions
// that form every channel's window means ("direct/delayed monitors",
// IFCS:9487-9494, 9262-9270).  One thread per (slot, frame chunk); frames
// are read in groups of 16 so the loads of a group are in flight together.
//
// The sums are taken per 64-frame sub-chunk (ctot[sub][slot], the sum of
// I' over the sub-chunk's frames in frame order) and rec[slot][k] is the sum
// from the start of position k's sub-chunk up to the position, so the prefix
// sums do not depend on the chunk length (any multiple of 64 frames): runs
// that split a stack differently give the same bits.
// ------------------------------------------------------------------
constexpr int kSubLog2 = 6;   // 64-frame sub-chunks of the prefix sums
//
// With l3 set, the pass also writes every slot's level-3 values (sums of 8
// stored frames, summed pairwise in FP32: ((x0+x1)+(x2+x3))+((x4+x5)+(x6+x7))),
// l3[y][slot] for y < numbin_3, and with cmax set the largest |X| of each
// (chunk, slot), cmax[chunk][slot] (NaN ignored).
template <class BL, int NB>
__global__ void __launch_bounds__(64) k_apply(SeriesGeom g, BleachDesc bd, const CorrTables *__restrict__ ctp,
                                              int chunk_len, double *ctot, double *rec, float *X, float *l3,
                                              float *cmax) {
    const CorrTables &ct = *ctp;
    const int k = blockIdx.x * 64 + threadIdx.x;
    const int chunk = blockIdx.y, nchunk = gridDim.y;
    if (k >= g.nslots) return;
    int ox, oy;
    slot_origin(g, k, ox, oy);
    BL bl;
    bl.load(bd, k);
    const double mu = slot_mu(bd.coef, bd.order, k);
    const int f0 = chunk * chunk_len, f1 = min(g.T, f0 + chunk_len);
    const int npos = ct.npos;
    int q = 0;
    while (q < npos && ct.pos[q] < f0) q++;
    int next = (q < npos) ? ct.pos[q] : 0x7fffffff;
    double run = 0.0;
    const uint32_t ld = (uint32_t)g.nslots;
    const bool bstore = (int64_t)ld * 64 < 0x7fffffff;
    // the next group's loads are issued before the current group is
    // corrected and stored (one group of loads always in flight)
    const int nb3 = ct.numbin[3];
    float mx = 0.f;
    uint32_t cur[16];
    if (f0 < f1) raw_words16<NB>(g, ox, oy, f0, f1 - 1, cur);
    for (int fb = f0; fb < f1; fb += 16) {
        double v[16];
        float xv[16];
        words_to_values16(g, cur, v);
        if (fb + 16 < f1) raw_words16<NB>(g, ox, oy, fb + 16, f1 - 1, cur);   // wave-uniform
        float *xw = X + (int64_t)fb * g.nslots;   // wave-uniform
        if (bstore && fb + 16 <= f1 && next >= fb + 16) {   // wave-uniform
            // a full group with no prefix record: straight-line code, buffer
            // stores with the frame's row as the scalar offset
            const __amdgpu_buffer_rsrc_t rs =
                __builtin_amdgcn_make_buffer_rsrc((void *)xw, (short)0, (int)(ld * 64), kBufferWord3);
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const double a = bl.apply(v[i], fb + i);
                run += a;
                xv[i] = (float)(a - mu);
                __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, xv[i]), rs, k * 4,
                                                      i * (int)ld * 4, 0);
            }
        } else {
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int f = fb + i;
                xv[i] = 0.f;
                if (f < f1) {
                    while (next == f) {
                        rec[(int64_t)k * npos + q] = run;
                        q++;
                        next = (q < npos) ? ct.pos[q] : 0x7fffffff;
                    }
                    const double a = bl.apply(v[i], f);
                    run += a;
                    xv[i] = (float)(a - mu);
                    xw[(uint32_t)i * ld + (uint32_t)k] = xv[i];
                }
            }
        }
        if (cmax) {
#pragma unroll
            for (int i = 0; i < 16; i++) mx = fmaxf(mx, fabsf(xv[i]));
        }
        if (l3) {
            // frames past the series end are 0 and their level-3 values unused
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const float *x = xv + 8 * h;
                const float s8 = ((x[0] + x[1]) + (x[2] + x[3])) + ((x[4] + x[5]) + (x[6] + x[7]));
                const int y3 = fb / 8 + h;
                if (y3 < nb3) l3[(int64_t)y3 * g.nslots + k] = s8;
            }
        }
        // end of a sub-chunk (chunks are whole sub-chunks up to the series
        // end): its sum, restarted unless the series ends here (the
        // positions at the end record the last sub-chunk's sum)
        const int fe = min(fb + 16, f1);
        if ((fe & ((1 << kSubLog2) - 1)) == 0 || fe == f1) {   // wave-uniform
            ctot[(int64_t)((fe - 1) >> kSubLog2) * g.nslots + k] = run;
            if (fe < g.T) run = 0.0;
        }
    }
    if (cmax) cmax[(int64_t)chunk * g.nslots + k] = mx;
    if (chunk == nchunk - 1)
        while (q < npos && ct.pos[q] >= f1) {
            rec[(int64_t)k * npos + q] = run;
            q++;
        }
}

// The bleach pass of binnings 1, 2 and 4 in one read of the stack.  A lane
// owns one binning-2 slot: its 2 x 2 raw pixels are the four binning-1 slots,
// and the four lanes of a quad are the 2 x 2 binning-2 slots of one
// binning-4 slot (window sums exchanged within the quad).  A wave covers 32
// x 2 binning-2 slots.  Every series is corrected, summed and stored exactly
// as k_apply does it (same integer window sums, same FP64 correction, same
// frame order of the sums per 64-frame sub-chunk), so the three binnings'
// X, ctot and rec hold the bits of their own passes.  The four lanes of a
// quad share the binning-4 series' corrections (two frames of each 8-frame
// group per lane) and every lane of the quad sums all of them in frame order;
// the quad's first lane writes its sums.
constexpr int kNG = 8;   // frames per group of the nested pass
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// Lane k's value (k < 4 wave-uniform) on every lane of its quad (DPP
// quad_perm broadcast of both halves).
template <int CTRL>
__device__ __forceinline__ double quad_bcast_t(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xf, 0xf, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double quad_bcast(double v, int k) {
    switch (k) {
        case 0: return quad_bcast_t<0x00>(v);
        case 1: return quad_bcast_t<0x55>(v);
        case 2: return quad_bcast_t<0xaa>(v);
        default: return quad_bcast_t<0xff>(v);
    }
}
template <int MAXO>
__global__ void __launch_bounds__(64) k_apply_nested(NestedApply A, const CorrTables *__restrict__ ctp,
                                                     int chunk_len) {
    const CorrTables &ct = *ctp;
    const int lane = threadIdx.x;
    const int wpr = A.swp[1] >> 5;   // waves per binning-2 row pair
    const int wy = (int)blockIdx.x / wpr, wx = (int)blockIdx.x - wy * wpr;
    const int qd = lane >> 2, li = lane & 1, lj = (lane >> 1) & 1;
    const int x2 = wx * 32 + 2 * qd + li, y2 = wy * 2 + lj;
    const int x4 = wx * 16 + qd, y4 = wy;
    const bool st4 = (lane & 3) == 0;
    int k1[4];
#pragma unroll
    for (int m = 0; m < 4; m++) k1[m] = (2 * y2 + (m >> 1)) * A.swp[0] + 2 * x2 + (m & 1);
    const int k2 = y2 * A.swp[1] + x2, k4 = y4 * A.swp[2] + x4;
    BleachP<MAXO> b1[4], b2, b4;
    double mu1[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        b1[m].load(A.bd[0], k1[m]);
        mu1[m] = slot_mu(A.bd[0].coef, A.bd[0].order, k1[m]);
    }
    b2.load(A.bd[1], k2);
    b4.load(A.bd[2], k4);
    const double mu2 = slot_mu(A.bd[1].coef, A.bd[1].order, k2);
    const double mu4 = slot_mu(A.bd[2].coef, A.bd[2].order, k4);
    const int chunk = blockIdx.y, nchunk = gridDim.y;
    const int f0 = chunk * chunk_len, f1 = min(A.T, f0 + chunk_len);
    const int npos = ct.npos;
    int q = 0;
    while (q < npos && ct.pos[q] < f0) q++;
    int next = (q < npos) ? ct.pos[q] : 0x7fffffff;
    double run1[4] = {0.0, 0.0, 0.0, 0.0}, run2 = 0.0, run4 = 0.0;
    const int64_t ns1 = (int64_t)A.swp[0] * A.sh[0], ns2 = (int64_t)A.swp[1] * A.sh[1],
                  ns4 = (int64_t)A.swp[2] * A.sh[2];
    // the two raw rows of the lane's 2 x 2 pixels, one 32-bit word each, by
    // buffer loads: the group's first frame in the resource, the frame as
    // the scalar offset, one lane offset per row (the host checks that 16
    // frames and 16 rows of every X fit one resource).  Groups of kNG
    // frames: the six series' coefficients and two groups of raw words
    // (the next one in flight) fill the registers of two waves per SIMD.
    const int lo0 = (2 * y2 * A.stride + 2 * x2) * 2, lo1 = lo0 + A.stride * 2;
    const int plb = (int)A.plane * 2;
    auto fetch = [&](int fb, uint32_t (&w)[2][kNG]) {
        const int nv = min(kNG, f1 - fb);
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(A.data + (int64_t)(A.f0 + fb) * A.plane), (short)0, plb * 16, kBufferWord3);
#pragma unroll
        for (int i = 0; i < kNG; i++) {
            const int so = min(i, nv - 1) * plb;
            w[0][i] = __builtin_amdgcn_raw_buffer_load_b32(rs, lo0, so, 0);
            w[1][i] = __builtin_amdgcn_raw_buffer_load_b32(rs, lo1, so, 0);
        }
    };
    auto pix = [](uint32_t w, int hi) { return (int)(hi ? (w >> 16) : (w & 0xffffu)); };
    auto sum4 = [&](const uint32_t (&w)[2][kNG], int i) {
        return (pix(w[0][i], 0) + pix(w[0][i], 1)) + (pix(w[1][i], 0) + pix(w[1][i], 1));
    };
    uint32_t cur[2][kNG], nxt[2][kNG];
    if (f0 < f1) fetch(f0, cur);
    for (int fb = f0; fb < f1; fb += kNG) {
        int s4[kNG];   // binning-4 window sums (the quad's four binning-2 sums)
#pragma unroll
        for (int i = 0; i < kNG; i++) {
            const int s2 = sum4(cur, i);
            const int t = s2 + __shfl_xor(s2, 1, 64);
            s4[i] = t + __shfl_xor(t, 2, 64);
        }
        if (fb + kNG < f1) fetch(fb + kNG, nxt);   // wave-uniform
        // this group's rows of the three X arrays (16 rows per resource)
        const __amdgpu_buffer_rsrc_t xr0 = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(A.X[0] + (int64_t)fb * ns1), (short)0, (int)(ns1 * 64), kBufferWord3);
        const __amdgpu_buffer_rsrc_t xr1 = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(A.X[1] + (int64_t)fb * ns2), (short)0, (int)(ns2 * 64), kBufferWord3);
        const __amdgpu_buffer_rsrc_t xr2 = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(A.X[2] + (int64_t)fb * ns4), (short)0, (int)(ns4 * 64), kBufferWord3);
        // binning 4: lane j of the quad corrects and stores frames j and j + 4
        // of the group; the values reach every lane of the quad for the sums
        // in frame order
        const int jq = lane & 3;
        double a4q[kNG / 4];
#pragma unroll
        for (int h = 0; h < kNG / 4; h++) {
            const int i = jq + 4 * h;
            // s4[i] without a dynamic register index
            const int *sv = s4 + 4 * h;
            const int s4i = jq == 0 ? sv[0] : jq == 1 ? sv[1] : jq == 2 ? sv[2] : sv[3];
            a4q[h] = b4.apply((double)(s4i - 16 * A.bg), fb + i);
            if (fb + i < f1)
                __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, (float)(a4q[h] - mu4)), xr2,
                                                      k4 * 4 + i * (int)ns4 * 4, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < kNG; i++) {
            const int f = fb + i;
            if (f < f1) {   // wave-uniform
                while (next == f) {
#pragma unroll
                    for (int m = 0; m < 4; m++) A.rec[0][(int64_t)k1[m] * npos + q] = run1[m];
                    A.rec[1][(int64_t)k2 * npos + q] = run2;
                    if (st4) A.rec[2][(int64_t)k4 * npos + q] = run4;
                    q++;
                    next = (q < npos) ? ct.pos[q] : 0x7fffffff;
                }
                // binning 1: the two pixels of each raw row as one 8-byte store
                // (slots k1[0], k1[0] + 1 and k1[2], k1[2] + 1)
                uint32_t xw[4];
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    const double a = b1[m].apply((double)(pix(cur[m >> 1][i], m & 1) - A.bg), f);
                    run1[m] += a;
                    xw[m] = __builtin_bit_cast(uint32_t, (float)(a - mu1[m]));
                }
#pragma unroll
                for (int r = 0; r < 2; r++) {
                    const u32x2 v = {xw[2 * r], xw[2 * r + 1]};
                    __builtin_amdgcn_raw_buffer_store_b64(v, xr0, k1[2 * r] * 4, i * (int)ns1 * 4, 0);
                }
                const double a2 = b2.apply((double)(sum4(cur, i) - 4 * A.bg), f);
                run2 += a2;
                __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, (float)(a2 - mu2)), xr1, k2 * 4,
                                                      i * (int)ns2 * 4, 0);
                run4 += quad_bcast(a4q[i >> 2], i & 3);
            }
            // one frame's six corrections at a time (they hold enough
            // independent work): bounds the registers of the unrolled group
            __builtin_amdgcn_sched_barrier(0);
        }
        const int fe = min(fb + kNG, f1);
        if ((fe & ((1 << kSubLog2) - 1)) == 0 || fe == f1) {   // wave-uniform
            const int64_t sub = (fe - 1) >> kSubLog2;
#pragma unroll
            for (int m = 0; m < 4; m++) A.ctot[0][sub * ns1 + k1[m]] = run1[m];
            A.ctot[1][sub * ns2 + k2] = run2;
            if (st4) A.ctot[2][sub * ns4 + k4] = run4;
            if (fe < A.T) {
#pragma unroll
                for (int m = 0; m < 4; m++) run1[m] = 0.0;
                run2 = run4 = 0.0;
            }
        }
#pragma unroll
        for (int i = 0; i < kNG; i++) {
            cur[0][i] = nxt[0][i];
            cur[1][i] = nxt[1][i];
        }
    }
    if (chunk == nchunk - 1)
        while (q < npos && ct.pos[q] >= f1) {
#pragma unroll
            for (int m = 0; m < 4; m++) A.rec[0][(int64_t)k1[m] * npos + q] = run1[m];
            A.rec[1][(int64_t)k2 * npos + q] = run2;
            if (st4) A.rec[2][(int64_t)k4 * npos + q] = run4;
            q++;
        }
}

// F[s][k] = S(ck) + rec[s][k], ck = the sub-chunk of position k (positions
// ascending; at most the last sub-chunk), S(m) the sum of the sub-chunk
// totals before m in a fixed order: runs of 16 totals, each summed in
// order; the runs before m's run in order; then m's own run up to m in
// order.  A workgroup takes 64 consecutive slots (lane = slot, so every load
// of ctot[sub][*] is one contiguous 512-byte row): its four waves sum the
// runs of kPcRuns-run segments into LDS, then each lane accumulates its
// slot's runs and the waves complete the positions of the segment (every
// kPcWaves-th position each, four at a time; the positions are the same for
// every slot).  Eight waves per workgroup: each wave's part is a few
// dependent round trips to memory.
//
// With xexp set, also the staging exponent of every slot for the matrix-core
// correlator: the smallest e >= 0 with max |X| * 2^-e < 2^14 (from the
// per-chunk maxima cmax of the nchunk bleach chunks), at most 60.
constexpr int kPcRuns = 64;    // runs of 16 sub-chunks per LDS segment
constexpr int kPcWaves = 8;    // waves per workgroup (64 slots)
__global__ void __launch_bounds__(64 * kPcWaves) k_prefix_combine(const CorrTables *__restrict__ ctp, int nslots,
                                                        const double *ctot, const double *rec, int nchunk,
                                                        double *F, const float *cmax, int32_t *xexp) {
    __shared__ double runsum[kPcRuns][64];
    __shared__ double runpre[kPcRuns][64];   // sum of the runs before run r (segment-local index)
    const CorrTables &ct = *ctp;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int s = blockIdx.x * 64 + lane;
    const bool live = s < nslots;
    const int sc = live ? s : nslots - 1;   // (dead lanes load a copy, write nothing)
    if (xexp && w == 0 && live) {
        float mx = 0.f;
        for (int c = 0; c < nchunk; c++) mx = fmaxf(mx, cmax[(int64_t)c * nslots + s]);
        int e = 0;
        if (mx > 0.f && mx <= 3.0e38f) {
            int ex;
            (void)frexpf(mx, &ex);   // mx < 2^ex
            e = min(max(ex - 14, 0), 60);
        }
        xexp[s] = e;
    }
    const int npos = ct.npos;
    const int nsub = (ct.T + (1 << kSubLog2) - 1) >> kSubLog2;
    const int nrun = (nsub + 15) >> 4;
    const double *recs = rec + (int64_t)sc * npos;
    double *Fs = F + (int64_t)sc * npos;
    double carry = 0.0;   // (wave 0) sum of the runs before this segment
    int k0 = 0;           // first position of this segment
    for (int r0 = 0; r0 < nrun; r0 += kPcRuns) {
        const int nr = min(kPcRuns, nrun - r0);
        // run sums of the segment, every kPcWaves-th run per wave
        for (int rr = w; rr < nr; rr += kPcWaves) {
            const int c0 = 16 * (r0 + rr);
            double v[16];
#pragma unroll
            for (int i = 0; i < 16; i++) v[i] = c0 + i < nsub ? ctot[(int64_t)(c0 + i) * nslots + sc] : 0.0;
            double t = 0.0;
#pragma unroll
            for (int i = 0; i < 16; i++) t += v[i];
            runsum[rr][lane] = t;
        }
        __syncthreads();
        if (w == 0) {
            for (int rr = 0; rr < nr; rr++) {
                runpre[rr][lane] = carry;
                carry += runsum[rr][lane];
            }
        }
        __syncthreads();
        // positions whose run lies in this segment (the last segment holds
        // the run of every remaining position)
        int k1 = k0;
        while (k1 < npos && (r0 + kPcRuns >= nrun || (min(ct.pos[k1] >> kSubLog2, nsub - 1) >> 4) < r0 + kPcRuns)) k1++;
        // four positions per wave at a time, their loads in flight together
        for (int kb = k0 + w; kb < k1; kb += 4 * kPcWaves) {
            int ck[4];
            double v[4][15], rv[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int k = kb + kPcWaves * u;
                ck[u] = k < k1 ? min(ct.pos[k] >> kSubLog2, nsub - 1) : 0;
                const int c0 = ck[u] & ~15;
#pragma unroll
                for (int i = 0; i < 15; i++)
                    v[u][i] = c0 + i < ck[u] ? ctot[(int64_t)(c0 + i) * nslots + sc] : 0.0;
                rv[u] = k < k1 ? recs[k] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int k = kb + kPcWaves * u;
                const int c0 = ck[u] & ~15;
                double part = 0.0;   // the run's totals before sub-chunk ck, in order
#pragma unroll
                for (int i = 0; i < 15; i++)
                    if (c0 + i < ck[u]) part += v[u][i];
                if (live && k < k1) Fs[k] = (runpre[(ck[u] >> 4) - r0][lane] + part) + rv[u];
            }
        }
        k0 = k1;
        __syncthreads();   // the next segment overwrites runsum and runpre
    }
}

// ------------------------------------------------------------------
// Stage 3: lag-1 blocking curve (blockTransform).  The Java copies the
// 1-based trace into a 0-based array, so the series is a'_y = a[y-1]
// (a'_0 = 0) against b'_{y+1} = b[y], y = 0..T-2 (IFCS:9259-9262).
//
// One thread per output pixel computes the curve of its ACF and of its CCF
// (both share the a series).  The products are cut into chunks of
// 2^kBlkJ = 256, each processed as 16 groups of 16 products held in
// registers.  Levels 0..3 are reduced inside a group, levels 4..7 by a
// cascade over the (wave-uniform) group index, and the level-8 block -- the
// chunk total -- goes to csum[chunk][pair] for the finalize kernel to
// cascade into the levels >= 8.  Every level keeps the pairwise tree of
// blockTransform, so each block value equals the Java block average times
// 2^level up to the FP32 rounding of the stored series.
// ------------------------------------------------------------------
constexpr int kBlkJ = 8;

// In a full group every block of levels 0..3 is complete, so the level sums
// s1[0..3] all grow by the group total h (the pairwise tree's root; the
// levels differ only in FP64 rounding order); a partial group keeps the
// complete blocks of each level.
template <bool FULL>
__device__ __forceinline__ void blk_group_stats(const double (&c)[16], int lim, double (&s1)[kBlkJ],
                                                double (&s2)[kBlkJ], double &h) {
    double d[8], e[4], f2[2];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        double v = (FULL || i < lim) ? c[i] : 0.0;
        if (!FULL) s1[0] += v;
        s2[0] = fma(v, v, s2[0]);
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
        d[j] = c[2 * j] + c[2 * j + 1];
        double v = (FULL || 2 * j + 2 <= lim) ? d[j] : 0.0;
        if (!FULL) s1[1] += v;
        s2[1] = fma(v, v, s2[1]);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        e[j] = d[2 * j] + d[2 * j + 1];
        double v = (FULL || 4 * j + 4 <= lim) ? e[j] : 0.0;
        if (!FULL) s1[2] += v;
        s2[2] = fma(v, v, s2[2]);
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
        f2[j] = e[2 * j] + e[2 * j + 1];
        double v = (FULL || 8 * j + 8 <= lim) ? f2[j] : 0.0;
        if (!FULL) s1[3] += v;
        s2[3] = fma(v, v, s2[3]);
    }
    h = f2[0] + f2[1];
    if (FULL) {
#pragma unroll
        for (int k = 0; k < 4; k++) s1[k] += h;
    }
}

// Levels 4..8 of one curve over the group index gi of a full group.
struct BlkCascade {
    double pend4, pend5, pend6, pend7;
    __device__ __forceinline__ void reset() { pend4 = pend5 = pend6 = pend7 = 0.0; }
    __device__ __forceinline__ void push(double carry, int gi, double (&s1)[kBlkJ], double (&s2)[kBlkJ],
                                         double *csum_out) {
        s1[4] += carry;
        s2[4] = fma(carry, carry, s2[4]);
        if (!(gi & 1)) { pend4 = carry; return; }
        carry += pend4;
        s1[5] += carry;
        s2[5] = fma(carry, carry, s2[5]);
        if (!(gi & 2)) { pend5 = carry; return; }
        carry += pend5;
        s1[6] += carry;
        s2[6] = fma(carry, carry, s2[6]);
        if (!(gi & 4)) { pend6 = carry; return; }
        carry += pend6;
        s1[7] += carry;
        s2[7] = fma(carry, carry, s2[7]);
        if (!(gi & 8)) { pend7 = carry; return; }
        *csum_out = carry + pend7;
    }
};

template <bool ACF, bool CCF>
__global__ void __launch_bounds__(64) k_blocking(SeriesGeom g, const CorrTables *__restrict__ ctp, PairDesc pd,
                                                 const double *F, const float *__restrict__ X, const double *coef,
                                                 int order, double *bpart, double *csum, int chunks_per_split) {
    const CorrTables &ct = *ctp;
    const int split = blockIdx.y;
    // a wave covers 64 consecutive slots of the pixel grid (a 256-byte
    // aligned row segment of X); lanes on padding slots have no pixel
    const int s = blockIdx.x * 64 + threadIdx.x, sy = s / g.swp, sx = s - sy * g.swp;
    if (sy >= g.oh || sx >= g.ow) return;
    const int p = sy * g.ow + sx;
    const int pa = p, pc = ACF ? pd.npix + p : p;   // pair indices of the ACF and the CCF
    const int ka = g.slot_of[p], kb = CCF ? g.slot_of[pd.npix + p] : ka;
    const int64_t ld = g.nslots;
    const int n0 = g.T - 1;
    const double mua = slot_mu(coef, order, ka), mub = slot_mu(coef, order, kb);
    // means of a'_0..a'_{n0-1} and b'_1..b'_{n0}, relative to the shifts
    const double alpha = F[(int64_t)ka * ct.npos + ct.pos_blk_a] / n0 - mua;
    const double beta_a = F[(int64_t)ka * ct.npos + ct.pos_blk_b] / n0 - mua;
    const double beta_b = F[(int64_t)kb * ct.npos + ct.pos_blk_b] / n0 - mub;
    const int nchunk = (n0 + (1 << kBlkJ) - 1) >> kBlkJ;
    const int t0 = split * chunks_per_split, t1 = min(nchunk, t0 + chunks_per_split);
    const bool bload = ld * 64 < 0x7fffffff;   // 16 rows addressable by one buffer resource
    double s1a[kBlkJ], s2a[kBlkJ], s1c[kBlkJ], s2c[kBlkJ];
#pragma unroll
    for (int k = 0; k < kBlkJ; k++) s1a[k] = s2a[k] = s1c[k] = s2c[k] = 0.0;
    BlkCascade cas_a, cas_c;
    for (int t = t0; t < t1; t++) {
        const int y0 = t << kBlkJ;
        // a'_{y-1} - abar for the first product of the chunk (a'_{-1} := 0)
        double aprev = (y0 == 0) ? -mua - alpha : (double)X[(int64_t)(y0 - 1) * ld + ka] - alpha;
        cas_a.reset();
        cas_c.reset();
        const int ngroups = min(16, (n0 - y0 + 15) >> 4);
        for (int gi = 0; gi < ngroups; gi++) {
            const int yb = y0 + 16 * gi;
            const int lim = min(16, n0 - yb);   // valid products, wave-uniform
            float xa[16], xb[16];
            const float *xq = X + (int64_t)yb * ld;   // wave-uniform
            const int nv = min(16, n0 - yb);
            if (bload) {
                // buffer loads: one lane offset per series, the row as the
                // scalar offset (no address arithmetic per load)
                const __amdgpu_buffer_rsrc_t rs =
                    __builtin_amdgcn_make_buffer_rsrc((void *)xq, (short)0, (int)(ld * 64), kBufferWord3);
#pragma unroll
                for (int i = 0; i < 16; i++) {
                    const int so = min(i, nv - 1) * (int)ld * 4;
                    xa[i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, ka * 4, so, 0));
                    if (CCF) xb[i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, kb * 4, so, 0));
                }
            } else {
#pragma unroll
                for (int i = 0; i < 16; i++) {
                    const uint32_t o = (uint32_t)min(i, nv - 1) * (uint32_t)ld;
                    xa[i] = xq[o + (uint32_t)ka];
                    if (CCF) xb[i] = xq[o + (uint32_t)kb];
                }
            }
            double am[16];
#pragma unroll
            for (int i = 0; i < 16; i++) am[i] = i ? (double)xa[i - 1] - alpha : aprev;
            aprev = (double)xa[15] - alpha;
            double h;
            if (ACF) {
                double c[16];
#pragma unroll
                for (int i = 0; i < 16; i++) c[i] = am[i] * ((double)xa[i] - beta_a);
                if (lim == 16) {
                    blk_group_stats<true>(c, 16, s1a, s2a, h);
                    cas_a.push(h, gi, s1a, s2a, &csum[(int64_t)t * pd.npairs + pa]);
                } else {
                    blk_group_stats<false>(c, lim, s1a, s2a, h);
                }
            }
            if (CCF) {
                double c[16];
#pragma unroll
                for (int i = 0; i < 16; i++) c[i] = am[i] * ((double)xb[i] - beta_b);
                if (lim == 16) {
                    blk_group_stats<true>(c, 16, s1c, s2c, h);
                    cas_c.push(h, gi, s1c, s2c, &csum[(int64_t)t * pd.npairs + pc]);
                } else {
                    blk_group_stats<false>(c, lim, s1c, s2c, h);
                }
            }
        }
    }
    if (ACF) {
        const int64_t base = ((int64_t)split * pd.npairs + pa) * kBlkJ * 2;
#pragma unroll
        for (int k = 0; k < kBlkJ; k++) {
            bpart[base + 2 * k] = s1a[k];
            bpart[base + 2 * k + 1] = s2a[k];
        }
    }
    if (CCF) {
        const int64_t base = ((int64_t)split * pd.npairs + pc) * kBlkJ * 2;
#pragma unroll
        for (int k = 0; k < kBlkJ; k++) {
            bpart[base + 2 * k] = s1c[k];
            bpart[base + 2 * k + 1] = s2c[k];
        }
    }
}

// bpart holds [nsplit][pair][kBlkJ][2] partial sums of the levels < kBlkJ;
// the levels >= kBlkJ are cascaded from the 256-product chunk totals
// csum[chunk][pair].
// Four lanes per pair (a wave holds 16 pairs): quarter j = lane & 3 sums
// the split partials j, j + 4, ... and cascades its own segment of the chunk
// totals (segments of a whole number of the longest blocks, so that every
// block lies in one segment); the quarters' sums meet by two lane swaps
// ((q0 + q1) + (q2 + q3) on every lane of the quad) and every lane applies
// the criteria, quarter 0 writing the result.
constexpr int kBfinQ = 4;
__global__ void __launch_bounds__(64) k_blocking_finalize(const CorrTables *__restrict__ ctp, SeriesGeom g,
                                                          PairDesc pd, const double *F, const double *bpart,
                                                          int nsplit, const double *csum, int32_t *ind_out,
                                                          int32_t *bis_out) {
    const CorrTables &ct = *ctp;
    const int j = threadIdx.x & (kBfinQ - 1);
    const int p0 = blockIdx.x * (64 / kBfinQ) + (threadIdx.x / kBfinQ);
    const bool live = p0 < pd.npairs;
    const int p = live ? p0 : pd.npairs - 1;   // (dead quads compute a copy, write nothing)
    const int n0 = g.T - 1;
    const int K = ct.blocknum;
    double S1[kMaxBlock], S2[kMaxBlock], pend[kMaxBlock];
#pragma unroll
    for (int k = 0; k < kMaxBlock; k++) S1[k] = S2[k] = pend[k] = 0.0;
    // split partials (split-major, so a split's 2 kBlkJ values are contiguous);
    // unrolled so that several splits' loads are in flight together
#pragma unroll 4
    for (int c = j; c < nsplit; c += kBfinQ) {
        const double *bp = bpart + ((int64_t)c * pd.npairs + p) * kBlkJ * 2;
#pragma unroll
        for (int k = 0; k < kBlkJ; k++) {
            S1[k] += bp[2 * k];
            S2[k] += bp[2 * k + 1];
        }
    }
#pragma unroll
    for (int k = 0; k < kBlkJ; k++)
        if (k >= K) S1[k] = S2[k] = 0.0;
    if (K > kBlkJ) {
        const int nfull = n0 >> kBlkJ;
        // segment length: a multiple of the longest block, 2^(K - 1 - kBlkJ) totals
        const int blk = 1 << (K - 1 - kBlkJ);
        const int seg = ((nfull + kBfinQ - 1) / kBfinQ + blk - 1) / blk * blk;
        const int tb = min(nfull, j * seg), te = min(nfull, tb + seg);
#pragma unroll 4
        for (int t = tb; t < te; t++) {
            double carry = csum[(int64_t)t * pd.npairs + p];
            bool open = true;
#pragma unroll
            for (int k = kBlkJ; k < kMaxBlock; k++) {
                if (open && k < K) {
                    S1[k] += carry;
                    S2[k] = fma(carry, carry, S2[k]);
                    if (((t >> (k - kBlkJ)) & 1) == 0) {
                        pend[k] = carry;
                        open = false;
                    } else {
                        carry += pend[k];
                    }
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kMaxBlock; k++) {
        S1[k] += __shfl_xor(S1[k], 1, 64);
        S2[k] += __shfl_xor(S2[k], 1, 64);
        S1[k] += __shfl_xor(S1[k], 2, 64);
        S2[k] += __shfl_xor(S2[k], 2, 64);
    }
    int sa, sb;
    pair_series(pd, p, sa, sb);
    const int ka = g.slot_of[sa], kb = g.slot_of[sb];
    const double am = F[(int64_t)ka * ct.npos + ct.pos_blk_a] / n0;
    const double bm = F[(int64_t)kb * ct.npos + ct.pos_blk_b] / n0;
    double sd[kMaxBlock], upper[kMaxBlock], lower[kMaxBlock];
    int crt[kMaxBlock], cr12[kMaxBlock], cr3[kMaxBlock], diffpos[kMaxBlock];
    double dd = am * bm;
    double nb = (double)n0;
#pragma unroll
    for (int k = 0; k < kMaxBlock; k++) {
        sd[k] = upper[k] = lower[k] = 0.0;
        crt[k] = cr12[k] = cr3[k] = diffpos[k] = 0;
        if (k < K) {
            double s1 = S1[k], s2 = S2[k];
            if (k > 0) nb = floor(nb / 2);
            double scale = ldexp(1.0, -k);
            double sp = s1 * scale, sp2 = s2 * scale * scale;
            double v = (sp2 / nb - (sp / nb) * (sp / nb)) / (nb * (dd * dd));
            sd[k] = sqrt(v);
            double err = sd[k] / sqrt(2 * (nb - 1));
            upper[k] = sd[k] + err;
            lower[k] = sd[k] - err;
        }
    }
    // criteria of blockTransform (IFCS:9364-9406)
#pragma unroll
    for (int x = 0; x < kMaxBlock - 1; x++)
        if (x < K - 1 && upper[x] > lower[x + 1] && upper[x + 1] > lower[x]) crt[x] = 1;
#pragma unroll
    for (int x = 0; x < kMaxBlock - 2; x++)
        if (x < K - 2 && crt[x] * crt[x + 1] == 1) cr12[x] = 1;
#pragma unroll
    for (int x = 0; x < kMaxBlock - 1; x++)
        if (x < K - 1 && sd[x + 1] - sd[x] > 0) diffpos[x] = 1;
#pragma unroll
    for (int x = 0; x < kMaxBlock - 2; x++)
        if (x < K - 2 && diffpos[x] * diffpos[x + 1] == 1) cr3[x] = 1;
    int last0 = 0, ind = 0, bis;
#pragma unroll
    for (int x = 0; x < kMaxBlock - 2; x++)
        if (x < K - 2 && cr3[x] == 1 && cr12[x] == 0) last0 = x;
#pragma unroll
    for (int x = 0; x < kMaxBlock; x++)
        if (x <= last0 || x == K - 3 || x == K - 4) cr12[x] = 0;
#pragma unroll
    for (int x = kMaxBlock - 1; x > 0; x--)
        if (x <= K - 5 && cr12[x] == 1) ind = x + 1;
    if (ind == 0) {
        bis = 0;
        ind = (K - 3 > 0) ? K - 3 : K - 1;
    } else {
        bis = 1;
    }
    if (ind < ct.q - 1) ind = ct.q - 1;
    if (live && j == 0) {
        ind_out[p] = ind;
        bis_out[p] = bis;
    }
}

// ------------------------------------------------------------------
// Stage 4: multi-tau correlation with blocked variance (calculateCF).
//
// One thread per (pair, level).  Level-c values are sums of 2^c frames
// (the pairwise-sum rebinning of IFCS:9477-9478).  Products are grouped by
// their a-index in groups of G values; block boundaries of the blocked
// variance (2^(ind-c) products) are multiples of G.  Values are shifted by
// the series mean so FP64 block sums carry no large offsets; every channel's
// exact window means (alpha, beta relative to the shift) are folded in
// analytically:
//   sum_y (a_y - abar)(b_{y+d} - bbar) = sum_y a~ b~ - n alpha beta
//   block sum S = Sab - beta Sa - alpha Sb + B alpha beta
// ------------------------------------------------------------------
struct LevelCtx {
    const float *X;
    const double *lv;   // level c >= 1: [numbin][ld] FP64 pairwise sums (k_level_sums)
    int64_t ld;
    int ka, kb, c, numbin;
    // level-c value: sum of 2^c stored deviations, i.e. I'-sum minus 2^c mu.
    // y is wave-uniform: the row's base goes into a scalar buffer resource
    // and the lane offset is the slot's; values past the series end read 0.
    // CHECK false: the caller has shown y < numbin; F32: level 0 (X)
    template <bool CHECK, bool F32>
    __device__ __forceinline__ double get(int k, int y) const {
        if (CHECK && y >= numbin) return 0.0;
        if (F32) {
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                (void *)(X + (int64_t)y * ld), (short)0, (int)(ld * 4), kBufferWord3);
            return (double)__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, k * 4, 0, 0));
        }
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(lv + (int64_t)y * ld), (short)0, (int)(ld * 8), kBufferWord3);
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, k * 8, 0, 0));
    }
};

// One run of nl <= NL channels of a level with the consecutive delays
// d0 .. d0 + nl - 1 over the values [ybeg, yend): a register window of b~
// values [y0 + d0, y0 + d0 + WN) slides by G.  Values past the series end
// read as 0, so products with y + d >= numbin add nothing.  Per block only
// Sab (per channel), Sa and the channel-d0 sum Sb0 are accumulated; the
// other channels' Sb follow from Sb0 at the block's end,
//   Sb_i = Sb0 + sum_{k<i} (b~[ye + d0 + k] - b~[ys + d0 + k]),
// with the first values past the block in the window and those at its start
// read again.  Every block j < pnum holds all of its B products.  The run's
// sums {cross, a1, a2} of channel i live at op[(3 i + m) ps] (m = 0, 1, 2)
// and are added to at block ends; alpha_i, beta_i are at ab[(2 i + m) ps].
template <int G, int NL, bool F32>
__device__ __forceinline__ void correlate_level(const LevelCtx &L, int d0, int nl, int B, int pnum, int ybeg, int yend,
                                                const double *ab, double *op, int64_t ps, double *sl) {
    constexpr int WN = G + NL - 1;
    double w[WN];   // b~ values [y0 + d0, y0 + d0 + WN)
    double ga[G];   // a~ values [y0, y0 + G)
    double sab[NL], sa = 0.0, sb0 = 0.0;
#pragma unroll
    for (int i = 0; i < NL; i++) {
        sab[i] = 0.0;
        if (i < nl) op[(3 * i) * ps] = op[(3 * i + 1) * ps] = op[(3 * i + 2) * ps] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < WN; k++) w[k] = L.template get<true, F32>(L.kb, ybeg + d0 + k);
#pragma unroll
    for (int k = 0; k < G; k++) ga[k] = L.template get<true, F32>(L.ka, ybeg + k);
    // one group of G values; the next group's a~ values and window values
    // are loaded first, in flight while this group's products run.  CHECK:
    // those loads reach past the series end.
    auto step = [&](int y0, auto check) {
        constexpr bool CHECK = decltype(check)::value;
        double na[G], nb[G];
#pragma unroll
        for (int k = 0; k < G; k++) {
            na[k] = L.template get<CHECK, F32>(L.ka, y0 + G + k);
            nb[k] = L.template get<CHECK, F32>(L.kb, y0 + d0 + WN + k);
        }
        double sg = 0.0, tg = 0.0;
#pragma unroll
        for (int k = 0; k < G; k++) {
            sg += ga[k];
            tg += w[k];
        }
        sa += sg;
        sb0 += tg;
#pragma unroll
        for (int i = 0; i < NL; i++) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < G; k++) s = fma(ga[k], w[k + i], s);
            sab[i] += s;
        }
        const int yb = y0 + G;
        if ((yb & (B - 1)) == 0) {
            // block end: the channels one at a time (not unrolled) from
            // the lane's LDS copy of Sab, so the loads and updates of the
            // block sums do not all claim registers at once
#pragma unroll
            for (int i = 0; i < NL; i++) {
                sl[i * 64 + threadIdx.x] = sab[i];
                sab[i] = 0.0;
            }
            const bool inblk = yb / B - 1 < pnum;
            double hd = 0.0;   // Sb_i - Sb0
#pragma unroll 1
            for (int i = 0; i < nl; i++) {
                const double sv = sl[i * 64 + threadIdx.x];
                double *o = op + (3 * i) * ps;
                o[0] += sv;
                if (inblk) {
                    const double al = ab[(2 * i) * ps], be = ab[(2 * i + 1) * ps];
                    const double S = sv - be * sa - al * (sb0 + hd) + (double)B * al * be;
                    o[ps] += S;
                    o[2 * ps] += S * S;
                    hd += L.template get<true, F32>(L.kb, yb + d0 + i) -
                          L.template get<true, F32>(L.kb, yb - B + d0 + i);
                }
            }
            sa = sb0 = 0.0;
        }
        // slide the window by G
#pragma unroll
        for (int k = 0; k < NL - 1; k++) w[k] = w[k + G];
#pragma unroll
        for (int k = 0; k < G; k++) {
            w[NL - 1 + k] = nb[k];
            ga[k] = na[k];
        }
    };
    for (int y0 = ybeg; y0 < yend; y0 += G) {
        if (y0 + 2 * G + d0 + NL - 1 <= L.numbin && y0 + 2 * G <= L.numbin)
            step(y0, std::false_type{});
        else
            step(y0, std::true_type{});
    }
#pragma unroll
    for (int i = 0; i < NL; i++)
        if (i < nl) op[(3 * i) * ps] += sab[i];
}

template <int NL, bool F32>
__device__ __forceinline__ void correlate_run_t(const LevelCtx &L, int G, int d0, int nl, int B, int pnum, int ybeg,
                                                int yend, const double *ab, double *op, int64_t ps, double *sl) {
    switch (G) {
        case 8: correlate_level<8, NL, F32>(L, d0, nl, B, pnum, ybeg, yend, ab, op, ps, sl); break;
        case 4: correlate_level<4, NL, F32>(L, d0, nl, B, pnum, ybeg, yend, ab, op, ps, sl); break;
        case 2: correlate_level<2, NL, F32>(L, d0, nl, B, pnum, ybeg, yend, ab, op, ps, sl); break;
        default: correlate_level<1, NL, F32>(L, d0, nl, B, pnum, ybeg, yend, ab, op, ps, sl); break;
    }
}

template <int NL>
__device__ __forceinline__ void correlate_run(const LevelCtx &L, int G, int d0, int nl, int B, int pnum, int ybeg,
                                              int yend, const double *ab, double *op, int64_t ps, double *sl) {
    if (L.c == 0)
        correlate_run_t<NL, true>(L, G, d0, nl, B, pnum, ybeg, yend, ab, op, ps, sl);
    else
        correlate_run_t<NL, false>(L, G, d0, nl, B, pnum, ybeg, yend, ab, op, ps, sl);
}

// Levels s+1 .. s+3 of every slot for the table-driven correlator in one
// pass over level s (s = 0: the stored FP32 deviations X; else the FP64
// level array src): each value is the FP64 sum of the two values below it,
// dj[y][k] = d(j-1)[2y][k] + d(j-1)[2y+1][k].  One thread per (group of 4
// level-(s+1) values, slot); a null dj (past the last level) is skipped and
// n(j) = floor(n(j-1) / 2) bounds each level.
__global__ void __launch_bounds__(256) k_level_sums(const float *X, const double *src, double *d1, double *d2,
                                                    double *d3, int n1, int n2, int n3, int nslots) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t ngroup = (n1 + 3) / 4;
    if (i >= ngroup * nslots) return;
    const int64_t gi = i / nslots, k = i - gi * nslots;
    double v1[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int64_t y = 4 * gi + j;
        if (y < n1) {
            const int64_t a = (2 * y) * nslots + k, b = a + nslots;
            v1[j] = X ? (double)X[a] + (double)X[b] : src[a] + src[b];
            d1[y * nslots + k] = v1[j];
        }
    }
    if (!d2) return;
    double v2[2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int64_t y = 2 * gi + j;
        v2[j] = v1[2 * j] + v1[2 * j + 1];
        if (y < n2) d2[y * nslots + k] = v2[j];
    }
    if (d3 && gi < n3) d3[gi * nslots + k] = v2[0] + v2[1];
}

// Window means of every (pair, channel): am, bmn of the channel's products
// and alpha, beta relative to the level's shift; ab[(2 x + m) npairs + p].
__device__ __forceinline__ void channel_means(const SeriesGeom &g, const BleachDesc &bd, const CorrTables &ct,
                                              const PairDesc &pd, const double *F, int p, int x, double &am,
                                              double &bmn, double &alpha, double &beta) {
    int sa, sb;
    pair_series(pd, p, sa, sb);
    const int ka = g.slot_of[sa], kb = g.slot_of[sb];
    const int c = ct.level[x];
    const double *Fa = F + (int64_t)ka * ct.npos, *Fb = F + (int64_t)kb * ct.npos;
    const double n = (double)(ct.numbin[c] - ct.del[x]);
    am = Fa[ct.pos_a_end[x]] / n;
    bmn = (Fb[ct.pos_b_end[x]] - Fb[ct.pos_b_beg[x]]) / n;
    alpha = am - slot_mu(bd.coef, bd.order, ka) * (double)(1 << c);
    beta = bmn - slot_mu(bd.coef, bd.order, kb) * (double)(1 << c);
}

__global__ void __launch_bounds__(64) k_correlate_means(SeriesGeom g, BleachDesc bd,
                                                        const CorrTables *__restrict__ ctp, PairDesc pd,
                                                        const double *F, double *ab) {
    const CorrTables &ct = *ctp;
    const int p = blockIdx.x * 64 + threadIdx.x, x = blockIdx.y;
    if (p >= pd.npairs) return;
    double am, bmn, alpha, beta;
    channel_means(g, bd, ct, pd, F, p, x, am, bmn, alpha, beta);
    ab[(2 * (int64_t)x) * pd.npairs + p] = alpha;
    ab[(2 * (int64_t)x + 1) * pd.npairs + p] = beta;
}

// One thread per (pair, level, time chunk) for every channel layout of
// setParameters: the level's channels are taken in runs of at most 17
// consecutive delays (p = 16: delays 0..16 on level 0, 9..16 on the others;
// a repeated delay starts a new run), each run one pass over the chunk's
// values (and the delays past its end).  Chunks of clen frames hold whole
// blocks of the blocked variance at every level; the chunk's sums go to
// part[((chunk chanum + x) 3 + m) npairs + p] (m: cross, a1, a2) and
// k_correlate_finalize adds the chunks.
__global__ void __launch_bounds__(64) k_correlate(SeriesGeom g, const CorrTables *__restrict__ ctp, PairDesc pd,
                                                  const float *X, const int32_t *ind, const double *lvg,
                                                  LevelOffsets lo, const double *ab, int clen, int nrun, double *part) {
    const CorrTables &ct = *ctp;
    // blockIdx.x = ((pair block / 8) nrun + run) 8 + pair block % 8: the
    // runs of one pair block read the same values and sit on one XCD
    // (workgroups go to the XCDs round-robin), back to back
    const int xcd = blockIdx.x & 7, t = blockIdx.x >> 3;
    const int run = t % nrun, pblk = (t / nrun) * 8 + xcd;
    const int p = pblk * 64 + threadIdx.x;
    const int c = blockIdx.y, chunk = blockIdx.z;
    if (p >= pd.npairs || c >= ct.nlevels) return;
    const int len = clen >> c;   // values of level c per chunk
    const int ybeg = chunk * len;
    if (ybeg >= ct.numbin[c]) return;
    const int yend = min(ct.numbin[c], ybeg + len);
    // the run's channels: runs of at most 8 consecutive delays
    const int chan0 = ct.first_chan[c], nch = ct.nchan[c];
    int r0 = 0, nl = 0;
    for (int r = 0; r <= run; r++) {
        r0 += nl;
        if (r0 >= nch) return;
        nl = 1;
        while (r0 + nl < nch && nl < 8 && ct.del[chan0 + r0 + nl] == ct.del[chan0 + r0] + nl) nl++;
    }
    int sa, sb;
    pair_series(pd, p, sa, sb);
    LevelCtx L;
    L.X = X;
    L.lv = c > 0 ? lvg + lo.off[c] : nullptr;
    L.ld = g.nslots;
    L.ka = g.slot_of[sa];
    L.kb = g.slot_of[sb];
    L.c = c;
    L.numbin = ct.numbin[c];
    const int indp = ind[p];
    const int B = 1 << (indp - c);
    const int pnum = (int)floor(ct.mtab_last / pow(2.0, fmax(indp - ct.log2_samp_last, 0.0)));
    // products per group: 8, or fewer on levels whose blocks are shorter
    // (B >= 2^(q-1-c) since ind >= q - 1)
    int G = 8;
    const int bmin = 1 << max(ct.q - 1 - c, 0);
    if (bmin < G) G = bmin;
    const int64_t ps = pd.npairs;
    __shared__ double sabl[8][64];   // each lane's Sab at block ends
    const int x0 = chan0 + r0;
    const double *abr = ab + 2 * x0 * ps + p;
    double *op = part + ((int64_t)chunk * ct.chanum + x0) * 3 * ps + p;
    correlate_run<8>(L, G, ct.del[x0], nl, B, pnum, ybeg, yend, abr, op, ps, &sabl[0][0]);
}

// G and the blocked variance of every (pair, channel) from the chunks' sums,
// added in chunk order.
__global__ void __launch_bounds__(64) k_correlate_finalize(SeriesGeom g, BleachDesc bd,
                                                           const CorrTables *__restrict__ ctp, PairDesc pd,
                                                           const double *F, const int32_t *ind, const double *part,
                                                           int clen, double *corav, double *bvar) {
    const CorrTables &ct = *ctp;
    const int p = blockIdx.x * 64 + threadIdx.x, x = blockIdx.y;
    if (p >= pd.npairs) return;
    const int c = ct.level[x];
    double am, bmn, alpha, beta;
    channel_means(g, bd, ct, pd, F, p, x, am, bmn, alpha, beta);
    const double n = (double)(ct.numbin[c] - ct.del[x]);
    const int64_t ps = pd.npairs;
    double cross = 0.0, a1 = 0.0, a2 = 0.0;
    // chunks that hold level-c values
    const int nchunk = (ct.numbin[c] + (clen >> c) - 1) / (clen >> c);
    for (int k = 0; k < nchunk; k++) {
        const double *o = part + ((int64_t)k * ct.chanum + x) * 3 * ps + p;
        cross += o[0];
        a1 += o[ps];
        a2 += o[2 * ps];
    }
    const int indp = ind[p];
    const int B = 1 << (indp - c);
    const int pnum = (int)floor(ct.mtab_last / pow(2.0, fmax(indp - ct.log2_samp_last, 0.0)));
    const double tot = cross - n * alpha * beta;
    const double abm = am * bmn;
    const int64_t i = (int64_t)p * ct.chanum + x;
    corav[i] = tot / (n * am * bmn);
    const double sp = a1 / B, sp2 = a2 / ((double)B * B);
    const double pn = (double)pnum;
    bvar[i] = (sp2 / pn - (sp / pn) * (sp / pn)) / ((pn - 1) * (abm * abm));
}

// ------------------------------------------------------------------
// Stage 5: batched Levenberg-Marquardt, one wavefront per fit.
// ------------------------------------------------------------------
constexpr int kLMMaxFree = 11;
// fit states (GpufitImFCS.java:817-866)
constexpr int kStateConverged = 0, kStateMaxIter = 1, kStateSingular = 2;

// FCS_3p value and gradient (IFCS:13147-13408) in micrometre/second units.
struct Comp {
    float plat, dDplat, pspim, dDpspim, dvx, dvy, acf;
};

// exp and 1/x from the hardware exp2 / reciprocal (1 ulp each); the model
// is evaluated ~100x per point and fit, so the full-precision library
// sequences (~10 instructions each) are the bulk of an LM iteration.
__device__ __forceinline__ float fexp(float v) { return __builtin_amdgcn_exp2f(v * 1.4426950408889634f); }
__device__ __forceinline__ float frcp(float v) { return __builtin_amdgcn_rcpf(v); }

// One lateral dimension of the pixel-integrated Gaussian (IFCS:13163-13190
// for x, the same with a_y, r_y, v_y for y): l = p0/sqrt(pi) pexp + perf,
// dDpexp (the exponential part of d l / dD, exponents combined) and
// ev = erf(p2) + 2 erf(p3) - erf(p1) (the flow derivative factor).
struct DimTerms {
    float e1, e2, e3, pexp, l, dDpexp;
};

// erf(u) from exp(-u^2), which the model needs anyway.  |u| >= 0.5:
// Abramowitz & Stegun, Handbook of Mathematical Functions (1964), formula
// 7.1.26, erf(u) = 1 - (a1 t + ... + a5 t^5) exp(-u^2), t = 1 / (1 + p u),
// absolute error below 1.5e-7 (relative below 1e-6 there); |u| < 0.5: the
// Maclaurin series 2/sqrt(pi) sum (-1)^n u^(2n+1) / (n! (2n+1)) to n = 5
// (truncation below 3e-8 relative).  Both are evaluated and one selected.
__device__ __forceinline__ float erf_as(float u, float expmu2) {
    const float au = fabsf(u);
    const float t = frcp(fmaf(0.3275911f, au, 1.f));
    float poly = fmaf(t, 1.061405429f, -1.453152027f);
    poly = fmaf(t, poly, 1.421413741f);
    poly = fmaf(t, poly, -0.284496736f);
    poly = fmaf(t, poly, 0.254829592f);
    const float e_as = fmaf(-poly * t, expmu2, 1.f);
    const float s = au * au;
    float ser = fmaf(s, -1.f / 1320.f, 1.f / 216.f);
    ser = fmaf(s, ser, -1.f / 42.f);
    ser = fmaf(s, ser, 1.f / 10.f);
    ser = fmaf(s, ser, -1.f / 3.f);
    ser = fmaf(s, ser, 1.f);
    const float e_ser = 1.1283791670955126f * au * ser;
    return copysignf(au < 0.5f ? e_ser : e_as, u);
}

__device__ __forceinline__ void fcs_dim(float a, float r, float v, float x, float p0t, float ip0, float ip02,
                                        bool grad, DimTerms &o) {
    const float isqrpi = 0.56418958354775628f;
    if (r == 0.f && v == 0.f) {
        // centred (ACF without flow): p1 = p2 = a, p3 = 0, so erf(p1) = erf(p2),
        // erf(p3) = 0 and exp(-p3^2 / p0^2) = exp((p7 - p4) / p0^2) = 1 exactly
        float u = a * ip0;
        const float E = fexp(-u * u);
        float e = erf_as(u, E);
        o.e1 = o.e2 = e;
        o.e3 = 0.f;
        o.pexp = 2.f * E - 2.f;
        o.l = p0t * isqrpi * o.pexp + 2.f * (a * e);
        if (grad) o.dDpexp = 2.f * (2.f * fexp(-(a * a) * ip02) - 2.f);
        return;
    }
    float p1t = a + r - v * x, p2t = a - r + v * x, p3t = r - v * x;
    const float u1 = p1t * ip0, u2 = p2t * ip0, u3 = p3t * ip0;
    const float E1 = fexp(-u1 * u1), E2 = fexp(-u2 * u2), E3 = fexp(-u3 * u3);
    o.e1 = erf_as(u1, E1);
    o.e2 = erf_as(u2, E2);
    o.e3 = erf_as(u3, E3);
    o.pexp = E1 + E2 - 2.f * E3;
    float perf = p1t * o.e1 + p2t * o.e2 - 2.f * p3t * o.e3;
    o.l = p0t * isqrpi * o.pexp + perf;
    if (grad) {
        // dDpexp = 2 exp(-p4/p0^2)(exp(p5/p0^2) + exp(p6/p0^2) - 2 exp(p7/p0^2)), exponents combined
        float p4 = 2.f * a * a + 3.f * r * r - 6.f * x * r * v + 3.f * (x * v) * (x * v);
        float p5 = p3t * p3t + p1t * p1t, p6 = p3t * p3t + p2t * p2t;
        float p7 = 2.f * (a * a + r * r - 2.f * x * r * v + (x * v) * (x * v));
        o.dDpexp = 2.f * (fexp((p5 - p4) * ip02) + fexp((p6 - p4) * ip02) - 2.f * fexp((p7 - p4) * ip02));
    }
}

__device__ __forceinline__ void fcs_component(const LMModelGeom &m, float x, float D, float vx, float vy, bool grad,
                                              Comp &o) {
    const float isqrpi = 0.56418958354775628f;
    // p0 = sqrt(4 D x + s^2) and 1/p0 from one hardware reciprocal square
    // root (1 ulp), the model's precision elsewhere
    const float p0sq = 4.f * D * x + m.s * m.s;
    const float ip0 = __builtin_amdgcn_rsqf(p0sq), ip02 = ip0 * ip0;
    const float p0t = p0sq * ip0;
    DimTerms tx, ty;
    fcs_dim(m.ax, m.rx, vx, x, p0t, ip0, ip02, grad, tx);
    if (m.ay == m.ax && m.ry == m.rx && vy == vx)
        ty = tx;   // square pixel, symmetric shift and flow: identical y factor
    else
        fcs_dim(m.ay, m.ry, vy, x, p0t, ip0, ip02, grad, ty);
    float inorm = m.obsvol * frcp(4.f * (m.ax * m.ay) * (m.ax * m.ay));
    o.plat = tx.l * ty.l * inorm;
    float szi2 = frcp(m.sz * m.sz);
    float q = 1.f + 4.f * D * x * szi2;
    o.pspim = __builtin_amdgcn_rsqf(q);
    o.acf = o.plat * o.pspim;
    if (grad) {
        o.dDplat = (ip0 * isqrpi) * (ty.dDpexp * x * tx.l + tx.dDpexp * x * ty.l) * inorm;
        // -4 x / (2 sz^2 q^(3/2)), q^(-3/2) = pspim^3
        o.dDpspim = -2.f * x * szi2 * (o.pspim * o.pspim * o.pspim);
        o.dvx = ty.l * ((tx.e2 + 2.f * tx.e3 - tx.e1) * x);
        o.dvy = tx.l * ((ty.e2 + 2.f * ty.e3 - ty.e1) * x);
    }
}

// pq: N, D, vx, vy, G, F2, D2, F3, D3, Ftrip, Ttrip (um, s units)
__device__ float fcs3p(const LMModelGeom &m, float x, const float *pq, bool grad, bool c2on, bool c3on,
                       float *gr) {
    Comp c1, c2 = {0, 0, 0, 0, 0, 0, 0}, c3 = {0, 0, 0, 0, 0, 0, 0};
    fcs_component(m, x, pq[1], pq[2], pq[3], grad, c1);
    if (c2on) fcs_component(m, x, pq[6], pq[2], pq[3], grad, c2);
    if (c3on) fcs_component(m, x, pq[8], pq[2], pq[3], grad, c3);
    float q2 = m.q2, q3 = m.q3, F2 = pq[5], F3 = pq[7];
    // reciprocals (1 ulp) instead of full-precision divisions
    const float rT = (pq[10] != 0.f) ? frcp(pq[10]) : 0.f;
    float et = (pq[10] != 0.f) ? fexp(-x * rT) : 0.f;
    const float r1F = frcp(1.f - pq[9]);
    float triplet = 1.f + pq[9] * r1F * et;
    float den = 1.f - F2 - F3 + q2 * F2 + q3 * F3;
    const float rden = frcp(den);
    float invN = frcp(pq[0]);
    float mix = (1.f - F2 - F3) * c1.acf + q2 * q2 * F2 * c2.acf + q3 * q3 * F3 * c3.acf;
    float pacf = invN * mix * (rden * rden) * triplet + pq[4];
    if (grad) {
        float pf1 = (1.f - F2 - F3) * rden, pf2 = q2 * q2 * F2 * rden, pf3 = q3 * q3 * F3 * rden;
        const float rdfnom = rden * rden * rden;
        float df21 = 1.f - F2 - F3 + q2 * F2 - q3 * F3 + 2.f * q2 * F3 - 2.f * q2;
        float df22 = q2 * q2 * (1.f + F2 - F3 - q2 * F2 + q3 * F3);
        float df23 = 2.f * F3 * q3 * q3 * (1.f - q2);
        float df31 = 1.f - F2 - F3 - q2 * F2 + 2.f * q3 * F2 - 2.f * q3 + q3 * F3;
        float df32 = 2.f * F2 * q2 * q2 * (1.f - q3);
        float df33 = q3 * q3 * (1.f - F2 + F3 + q2 * F2 - q3 * F3);
        const float inorm = m.obsvol * frcp(4.f * (m.ax * m.ay) * (m.ax * m.ay));
        float dtF = et * (r1F + pq[9] * (r1F * r1F));
        float dtT = (pq[10] != 0.f) ? et * (pq[9] * x) * r1F * (rT * rT) : 0.f;
        gr[0] = -invN * invN * (pf1 * c1.acf + pf2 * c2.acf + pf3 * c3.acf) * triplet;
        gr[1] = invN * pf1 * (c1.plat * c1.dDpspim + c1.pspim * c1.dDplat);   // no triplet, IFCS:13310
        gr[2] = invN * (pf1 * c1.dvx * c1.pspim + pf2 * c2.dvx * c2.pspim + pf3 * c3.dvx * c3.pspim) * inorm * triplet;
        gr[3] = invN * (pf1 * c1.dvy * c1.pspim + pf2 * c2.dvy * c2.pspim + pf3 * c3.dvy * c3.pspim) * inorm * triplet;
        gr[4] = 1.f;
        gr[5] = invN * rdfnom * (df21 * c1.acf + df22 * c2.acf + df23 * c3.acf) * triplet;
        gr[6] = invN * pf2 * (c2.plat * c2.dDpspim + c2.pspim * c2.dDplat) * triplet;
        gr[7] = invN * rdfnom * (df31 * c1.acf + df32 * c2.acf + df33 * c3.acf) * triplet;
        gr[8] = invN * pf3 * (c3.plat * c3.dDpspim + c3.pspim * c3.dDplat) * triplet;
        gr[9] = dtF * pacf;
        gr[10] = dtT * pacf;
    }
    return pacf;
}

// One DPP move of a double (both halves); lanes the row mask excludes, or
// whose source lane is out of the row, read 0.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wave (all lanes active), returned to every
// lane.  DPP steps: pairs and quads (quad_perm), 8 and 16 lanes (half-row
// and row mirror), then rows 0+1 and 2+3 (row_bcast:15 into rows 1 and 3)
// and the two halves (row_bcast:31 into rows 2 and 3); lane 63 holds the
// total, which is read back as a wave-uniform value.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<0xB1, 0xf>(v);    // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E, 0xf>(v);    // quad_perm [2,3,0,1]
    v += dpp_f64<0x141, 0xf>(v);   // row_half_mirror
    v += dpp_f64<0x140, 0xf>(v);   // row_mirror
    v += dpp_f64<0x142, 0xa>(v);   // row_bcast:15
    v += dpp_f64<0x143, 0xc>(v);   // row_bcast:31
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_f32(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xf, false));
}

// FP32 version of wave_sum (same DPP steps).
__device__ __forceinline__ float wave_sum_f32(float v) {
    v += dpp_f32<0xB1, 0xf>(v);
    v += dpp_f32<0x4E, 0xf>(v);
    v += dpp_f32<0x141, 0xf>(v);
    v += dpp_f32<0x140, 0xf>(v);
    v += dpp_f32<0x142, 0xa>(v);
    v += dpp_f32<0x143, 0xc>(v);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// FP32 chi-square with the reduction order of Gpufit: points are grouped in
// pieces of 512, each piece is summed as a binary tree in FP32 (strides
// 256, 128, ..., 1, agpufitjni.cu:1326-1345, 1546-1603, with 4 x the next
// power of two threads per fit, 2722-2731) and the pieces are added in FP64
// (1385-1405).  Lane l holds the piece's elements l + 64 k in e[k].
__device__ __forceinline__ float piece_tree(float (&e)[8]) {
    e[0] += e[4];
    e[1] += e[5];
    e[2] += e[6];
    e[3] += e[7];
    e[0] += e[2];
    e[1] += e[3];
    e[0] += e[1];
    float v = e[0];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return __shfl(v, 0, 64);
}

// In-place Cholesky solve of the n x n SPD system A x = b (FP64). Returns
// false if A is not positive definite.
__device__ bool chol_solve(double *A, double *b, int n) {
    for (int j = 0; j < n; j++) {
        double s = A[j * kLMMaxFree + j];
        for (int k = 0; k < j; k++) s -= A[j * kLMMaxFree + k] * A[j * kLMMaxFree + k];
        if (!(s > 0.0)) return false;
        double d = sqrt(s);
        A[j * kLMMaxFree + j] = d;
        for (int i = j + 1; i < n; i++) {
            double t = A[i * kLMMaxFree + j];
            for (int k = 0; k < j; k++) t -= A[i * kLMMaxFree + k] * A[j * kLMMaxFree + k];
            A[i * kLMMaxFree + j] = t / d;
        }
    }
    for (int i = 0; i < n; i++) {
        double t = b[i];
        for (int k = 0; k < i; k++) t -= A[i * kLMMaxFree + k] * b[k];
        b[i] = t / A[i * kLMMaxFree + i];
    }
    for (int i = n - 1; i >= 0; i--) {
        double t = b[i];
        for (int k = i + 1; k < n; k++) t -= A[k * kLMMaxFree + i] * b[k];
        b[i] = t / A[i * kLMMaxFree + i];
    }
    return true;
}

// kModelExp1/2: the bleach models SingleExp A e^{-t/t1} + C and DoubleExp
// A e^{-t/t1} + B e^{-t/t2} + C (IFCS:13028-13086), fitted by k_lm_exp64.
enum ModelKind { kModelFCS = 0, kModelPoly = 1, kModelExp1 = 2, kModelExp2 = 3 };

struct LMArgs {
    int nfits, npts, nparams, kind, ncoef;
    const float *data, *weights, *x, *init;
    const int *pfit;
    LMModelGeom geom;
    float tol;
    int maxit;
    int chi2_fp64;   // 0: Gpufit's FP32 chi-square tree; 1: FP64 sum of FP64 residual squares
    float *out_params;
    int *states;
    float *chi2;
    int *iters;
    int init_two_rows;   // 0: init holds a row per fit; 1: two rows, split at init_split
    int init_split;
    // 1: fit again only the fits whose state is not CONVERGED (from init),
    // adding the iterations of the earlier attempt
    int only_failed;
};

// Whether this fit is skipped by an only_failed launch, and the iterations
// an earlier attempt spent on it.
__device__ __forceinline__ bool lm_skip(const LMArgs &a, int fit, int &prev_iters) {
    prev_iters = 0;
    if (!a.only_failed) return false;
    // converged, or data without a finite chi-square (NaN correlations: the
    // fit fails from any start)
    const float c2 = a.chi2[fit];
    if (a.states[fit] == kStateConverged || !(c2 == c2)) return true;
    prev_iters = a.iters[fit];
    return false;
}

__device__ __forceinline__ const float *init_row(const LMArgs &a, int fit) {
    if (a.init_two_rows) return a.init + (fit < a.init_split ? 0 : a.nparams);
    return a.init + (int64_t)fit * a.nparams;
}

// Evaluates chi^2 and (optionally) gradient / hessian sums for the fit, with
// every lane handling points lane, lane+64, ...  Results are wave totals.
__device__ void lm_eval(const LMArgs &a, int fit, const float *par, const int *fidx, int nfree, bool need_jac,
                        double &chi2, double *grad, double *hess) {
    const int lane = threadIdx.x & 63;
    const float *y = a.data + (int64_t)fit * a.npts;
    const float *w = a.weights ? a.weights + (int64_t)fit * a.npts : nullptr;
    double c2 = 0.0, cx = 0.0, gl[kLMMaxFree], hl[kLMMaxFree * (kLMMaxFree + 1) / 2];
    int nh = nfree * (nfree + 1) / 2;
    for (int i = 0; i < nfree; i++) gl[i] = 0.0;
    for (int i = 0; i < nh; i++) hl[i] = 0.0;
    bool c2on = (par[5] != 0.f) || a.pfit[5] || a.pfit[6];
    bool c3on = (par[7] != 0.f) || a.pfit[7] || a.pfit[8];
    for (int j0 = 0; j0 < a.npts; j0 += 512) {
      float e[8];
#pragma unroll
      for (int k = 0; k < 8; k++) e[k] = 0.f;
#pragma unroll 1
      for (int k = 0; k < 8; k++) {
        const int i = j0 + 64 * k + lane;
        if (i >= a.npts) continue;
        float xi = a.x[i];
        float jf[kLMMaxFree];
        float f;
        if (a.kind == kModelFCS) {
            float gfull[11];
            f = fcs3p(a.geom, xi, par, need_jac, c2on, c3on, gfull);
            if (need_jac)
                for (int k = 0; k < nfree; k++) jf[k] = gfull[fidx[k]];
        } else {
            // LINEAR_1D: value = sum p_i x^i with FP64 accumulation (agpufitjni.cu:402-418)
            double tx = 1.0, sv = (double)par[0];
            float gfull[kLMMaxFree];
            gfull[0] = 1.f;
            for (int k = 1; k < a.ncoef; k++) {
                tx *= xi;
                gfull[k] = (float)tx;
                sv += (double)par[k] * tx;
            }
            for (int k = a.ncoef; k < kLMMaxFree; k++) gfull[k] = 0.f;
            f = (float)sv;
            if (need_jac)
                for (int k = 0; k < nfree; k++) jf[k] = gfull[fidx[k]];
        }
        const float dev = f - y[i];
        const float ek = w ? dev * dev * w[i] : dev * dev;
#pragma unroll
        for (int q = 0; q < 8; q++) e[q] = (q == k) ? ek : e[q];
        double wi = w ? (double)w[i] : 1.0;
        double r = (double)y[i] - (double)f;
        cx += wi * r * r;
        if (need_jac) {
            int h = 0;
            for (int q = 0; q < nfree; q++) {
                double jq = (double)jf[q];
                gl[q] += wi * r * jq;
                for (int l = 0; l <= q; l++) hl[h++] += wi * jq * (double)jf[l];
            }
        }
      }
      c2 += (double)piece_tree(e);
    }
    chi2 = a.chi2_fp64 ? wave_sum(cx) : (double)(float)c2;
    if (need_jac) {
        for (int k = 0; k < nfree; k++) grad[k] = wave_sum(gl[k]);
        for (int k = 0; k < nh; k++) hess[k] = wave_sum(hl[k]);
    }
}

__global__ void __launch_bounds__(64) k_lm(LMArgs a) {
    const int fit = blockIdx.x;
    if (fit >= a.nfits) return;
    int prev_iters;
    if (lm_skip(a, fit, prev_iters)) return;
    const int lane = threadIdx.x & 63;
    // parameters in model units
    float par[20], prev[20];
    int fidx[kLMMaxFree], nfree = 0;
    for (int k = 0; k < a.nparams; k++) {
        par[k] = init_row(a, fit)[k];
        if (a.pfit[k] && nfree < kLMMaxFree && (a.kind == kModelPoly ? k < a.ncoef : k < 11)) fidx[nfree++] = k;
    }
    for (int k = a.nparams; k < 20; k++) par[k] = 0.f;
    LMArgs ac = a;
    if (a.kind == kModelFCS) {
        // SI -> micrometre/second units; geometry from slots 11..19
        // (the ACF_1D parameter layout of GPU_ACF_Fit, IFCS:15090-15110)
        par[1] *= 1e12f;
        par[6] *= 1e12f;
        par[8] *= 1e12f;
        par[2] *= 1e6f;
        par[3] *= 1e6f;
        ac.geom.ax = par[11] * 1e6f;
        ac.geom.ay = par[12] * 1e6f;
        ac.geom.s = par[13] * 1e6f;
        ac.geom.sz = par[14] * 1e6f;
        ac.geom.rx = par[15] * 1e6f;
        ac.geom.ry = par[16] * 1e6f;
        ac.geom.obsvol = par[17] * 1e12f;
        ac.geom.q2 = par[18];
        ac.geom.q3 = par[19];
    }
    double chi2;
    double grad[kLMMaxFree], hess[kLMMaxFree * (kLMMaxFree + 1) / 2];
    double H[kLMMaxFree * kLMMaxFree], scale[kLMMaxFree], delta[kLMMaxFree];
    lm_eval(ac, fit, par, fidx, nfree, true, chi2, grad, hess);
    int state = kStateConverged;
    int iters = 0;
    if (nfree > 0) {
        double lambda = 1e-3;
        double prev_chi2 = chi2;
        bool failed = false;
        for (int k = 0; k < nfree; k++) scale[k] = 0.0;
        // hessian as a full matrix; the diagonal keeps its lambda shift
        // between failed iterations, as Gpufit does
        for (int k = 0, h = 0; k < nfree; k++)
            for (int l = 0; l <= k; l++, h++) H[k * kLMMaxFree + l] = H[l * kLMMaxFree + k] = hess[h];
        bool finished = false;
        for (int it = 0; it < a.maxit && !finished; it++) {
            // cuda_modify_step_widths (agpufitjni.cu:2044-2076)
            for (int k = 0; k < nfree; k++) {
                double &hd = H[k * kLMMaxFree + k];
                if (failed) hd -= scale[k] * lambda / 10.0;
                scale[k] = fmax(scale[k], hd);
                hd += scale[k] * lambda;
            }
            double A[kLMMaxFree * kLMMaxFree];
            for (int k = 0; k < nfree * kLMMaxFree; k++) A[k] = H[k];
            for (int k = 0; k < nfree; k++) delta[k] = grad[k];
            bool ok = chol_solve(A, delta, nfree);
            iters = it + 1;
            if (!ok) {
                state = kStateSingular;
                break;
            }
            for (int k = 0; k < 20; k++) prev[k] = par[k];
            for (int k = 0; k < nfree; k++) par[fidx[k]] = (float)((double)par[fidx[k]] + delta[k]);
            double nchi2;
            double ngrad[kLMMaxFree], nhess[kLMMaxFree * (kLMMaxFree + 1) / 2];
            lm_eval(ac, fit, par, fidx, nfree, false, nchi2, ngrad, nhess);
            const bool decreased = nchi2 < prev_chi2;
            // cuda_check_fit_improvement (agpufitjni.cu:1453-1476): a step
            // from chi^2 = 0 never counts as failed
            const bool step_failed = (prev_chi2 != 0.0) && !decreased;
            if (!step_failed) lm_eval(ac, fit, par, fidx, nfree, true, nchi2, ngrad, nhess);
            // cuda_check_for_convergence (agpufitjni.cu:2289-2318)
            if (fabs(nchi2 - prev_chi2) < (double)a.tol * fmax(1.0, nchi2)) finished = true;
            else if (it == a.maxit - 1) state = kStateMaxIter;
            if (!step_failed) {
                for (int k = 0; k < nfree; k++) grad[k] = ngrad[k];
                for (int k = 0, h = 0; k < nfree; k++)
                    for (int l = 0; l <= k; l++, h++) H[k * kLMMaxFree + l] = H[l * kLMMaxFree + k] = nhess[h];
            }
            failed = step_failed;
            // cuda_prepare_next_iteration (agpufitjni.cu:2458-2490)
            if (decreased) {
                lambda *= 0.1;
                prev_chi2 = nchi2;
            } else {
                lambda *= 10.0;
                for (int k = 0; k < 20; k++) par[k] = prev[k];
            }
        }
        chi2 = prev_chi2;
    }
    if (a.kind == kModelFCS) {
        par[1] *= 1e-12f;
        par[6] *= 1e-12f;
        par[8] *= 1e-12f;
        par[2] *= 1e-6f;
        par[3] *= 1e-6f;
    }
    if (lane == 0) {
        for (int k = 0; k < a.nparams; k++) a.out_params[(int64_t)fit * a.nparams + k] = par[k];
        a.states[fit] = state;
        a.chi2[fit] = (float)chi2;
        a.iters[fit] = iters + prev_iters;
    }
}

// ------------------------------------------------------------------
// Specialised LM for a compile-time set of free parameters: the plugin's
// default fit of N, D, G_inf (FCS, slots 0, 1, 4) and the contiguous
// polynomial coefficients of the bleach fit.  All arrays are statically
// indexed, so the state stays in registers.
// ------------------------------------------------------------------
template <int KIND, int NF>
struct FixedSlots {
    static __device__ __forceinline__ constexpr int slot(int k) {
        return KIND == kModelFCS ? (k == 0 ? 0 : (k == 1 ? 1 : 4)) : k;
    }
};

template <int KIND, int NF>
__device__ void lm_eval_fixed(const LMArgs &a, int fit, const float *par, bool need_jac, double &chi2, double *grad,
                              double *hess) {
    constexpr int NH = NF * (NF + 1) / 2;
    const int lane = threadIdx.x & 63;
    const float *y = a.data + (int64_t)fit * a.npts;
    const float *w = a.weights ? a.weights + (int64_t)fit * a.npts : nullptr;
    double c2 = 0.0, cx = 0.0, gl[NF], hl[NH];
#pragma unroll
    for (int i = 0; i < NF; i++) gl[i] = 0.0;
#pragma unroll
    for (int i = 0; i < NH; i++) hl[i] = 0.0;
    bool c2on = false, c3on = false;
    if (KIND == kModelFCS) {
        c2on = (par[5] != 0.f) || a.pfit[6];
        c3on = (par[7] != 0.f) || a.pfit[8];
    }
    for (int j0 = 0; j0 < a.npts; j0 += 512) {
      float e[8];
#pragma unroll
      for (int k = 0; k < 8; k++) e[k] = 0.f;
      const int kend = min(8, (a.npts - j0 + 63) >> 6);   // wave-uniform
#pragma unroll 2
      for (int k = 0; k < kend; k++) {
        const int i = j0 + 64 * k + lane;
        if (i >= a.npts) continue;
        const float xi = a.x[i];
        float jf[NF];
        float f;
        if (KIND == kModelFCS) {
            float gfull[11];
            f = fcs3p(a.geom, xi, par, need_jac, c2on, c3on, gfull);
#pragma unroll
            for (int k = 0; k < NF; k++) jf[k] = gfull[FixedSlots<KIND, NF>::slot(k)];
        } else {
            double tx = 1.0, sv = (double)par[0];
            jf[0] = 1.f;
#pragma unroll
            for (int k = 1; k < NF; k++) {
                tx *= xi;
                jf[k] = (float)tx;
                sv += (double)par[k] * tx;
            }
            f = (float)sv;
        }
        const float dev = f - y[i];
        const float ek = w ? dev * dev * w[i] : dev * dev;
#pragma unroll
        for (int q = 0; q < 8; q++) e[q] = (q == k) ? ek : e[q];
        const double wi = w ? (double)w[i] : 1.0;
        const double r = (double)y[i] - (double)f;
        cx += wi * r * r;
        if (need_jac) {
            int h = 0;
#pragma unroll
            for (int q = 0; q < NF; q++) {
                const double jq = (double)jf[q];
                gl[q] += wi * r * jq;
#pragma unroll
                for (int l = 0; l <= q; l++) hl[h++] += wi * jq * (double)jf[l];
            }
        }
      }
      if (!a.chi2_fp64) c2 += (double)piece_tree(e);
    }
    chi2 = a.chi2_fp64 ? wave_sum(cx) : (double)(float)c2;
    if (need_jac) {
#pragma unroll
        for (int k = 0; k < NF; k++) grad[k] = wave_sum(gl[k]);
#pragma unroll
        for (int k = 0; k < NH; k++) hess[k] = wave_sum(hl[k]);
    }
}

// Cholesky solve with one reciprocal per column (the solve latency is on
// the critical path of every LM iteration of a fit).
template <int NF>
__device__ __forceinline__ bool chol_solve_fixed(double (&A)[NF][NF], double (&b)[NF]) {
    double inv[NF];
#pragma unroll
    for (int j = 0; j < NF; j++) {
        double s = A[j][j];
#pragma unroll
        for (int k = 0; k < j; k++) s -= A[j][k] * A[j][k];
        if (!(s > 0.0)) return false;
        // hardware FP64 reciprocal square root and one Newton step
        const double y = __builtin_amdgcn_rsq(s);
        inv[j] = fma(0.5 * y, fma(-(s * y), y, 1.0), y);
#pragma unroll
        for (int i = j + 1; i < NF; i++) {
            double t = A[i][j];
#pragma unroll
            for (int k = 0; k < j; k++) t -= A[i][k] * A[j][k];
            A[i][j] = t * inv[j];
        }
    }
#pragma unroll
    for (int i = 0; i < NF; i++) {
        double t = b[i];
#pragma unroll
        for (int k = 0; k < i; k++) t -= A[i][k] * b[k];
        b[i] = t * inv[i];
    }
#pragma unroll
    for (int i = NF - 1; i >= 0; i--) {
        double t = b[i];
#pragma unroll
        for (int k = i + 1; k < NF; k++) t -= A[k][i] * b[k];
        b[i] = t * inv[i];
    }
    return true;
}

// FCS parameters: SI at the boundary (D m^2/s, v m/s, geometry m), micro-
// metres and seconds inside the model; the fixed geometry travels in
// parameter slots 11..19 (the ACF_1D layout of GPU_ACF_Fit, IFCS:15090-15110).
__device__ __forceinline__ void fcs_params_in(float *par, LMModelGeom &geom) {
    par[1] *= 1e12f;
    par[6] *= 1e12f;
    par[8] *= 1e12f;
    par[2] *= 1e6f;
    par[3] *= 1e6f;
    geom.ax = par[11] * 1e6f;
    geom.ay = par[12] * 1e6f;
    geom.s = par[13] * 1e6f;
    geom.sz = par[14] * 1e6f;
    geom.rx = par[15] * 1e6f;
    geom.ry = par[16] * 1e6f;
    geom.obsvol = par[17] * 1e12f;
    geom.q2 = par[18];
    geom.q3 = par[19];
}
__device__ __forceinline__ void fcs_params_out(float *par) {
    par[1] *= 1e-12f;
    par[6] *= 1e-12f;
    par[8] *= 1e-12f;
    par[2] *= 1e-6f;
    par[3] *= 1e-6f;
}

// Gpufit's LM iteration over the fixed free slots.  eval(par, chi2, grad,
// hess) returns the chi-square, gradient and packed normal matrix of the
// current parameters.
template <int KIND, int NF, class Eval, class PT>
__device__ void lm_fixed_iterate(const LMArgs &a, PT *par, Eval &eval, double &chi2, int &state, int &iters) {
    constexpr int NH = NF * (NF + 1) / 2;
    double grad[NF], hess[NH], H[NF][NF], scale[NF];
    eval(par, chi2, grad, hess);
    state = kStateConverged;
    iters = 0;
    double lambda = 1e-3;
    double prev_chi2 = chi2;
    bool failed = false;
#pragma unroll
    for (int k = 0; k < NF; k++) scale[k] = 0.0;
    {
        int h = 0;
#pragma unroll
        for (int k = 0; k < NF; k++)
#pragma unroll
            for (int l = 0; l <= k; l++, h++) H[k][l] = H[l][k] = hess[h];
    }
    bool finished = false;
    for (int it = 0; it < a.maxit && !finished; it++) {
        // cuda_modify_step_widths (agpufitjni.cu:2044-2076)
#pragma unroll
        for (int k = 0; k < NF; k++) {
            if (failed) H[k][k] -= scale[k] * (lambda * 0.1);   // the previous lambda
            scale[k] = fmax(scale[k], H[k][k]);
            H[k][k] += scale[k] * lambda;
        }
        double A[NF][NF], delta[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) {
            delta[k] = grad[k];
#pragma unroll
            for (int l = 0; l < NF; l++) A[k][l] = H[k][l];
        }
        iters = it + 1;
        if (!chol_solve_fixed<NF>(A, delta)) {
            state = kStateSingular;
            break;
        }
        PT prev[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) {
            const int sl = FixedSlots<KIND, NF>::slot(k);
            prev[k] = par[sl];
            par[sl] = (PT)((double)par[sl] + delta[k]);
        }
        // values and derivatives at the trial point in one pass, as
        // calc_curve_values does (agpufitjni.cu:4073-4076); the sums are
        // dropped when the step fails
        double nchi2;
        double ngrad[NF], nhess[NH];
        eval(par, nchi2, ngrad, nhess);
        const bool decreased = nchi2 < prev_chi2;
        // cuda_check_fit_improvement (agpufitjni.cu:1453-1476)
        const bool step_failed = (prev_chi2 != 0.0) && !decreased;
        // cuda_check_for_convergence (agpufitjni.cu:2289-2318)
        if (fabs(nchi2 - prev_chi2) < (double)a.tol * fmax(1.0, nchi2)) finished = true;
        else if (it == a.maxit - 1) state = kStateMaxIter;
        if (!step_failed) {
#pragma unroll
            for (int k = 0; k < NF; k++) grad[k] = ngrad[k];
            int h = 0;
#pragma unroll
            for (int k = 0; k < NF; k++)
#pragma unroll
                for (int l = 0; l <= k; l++, h++) H[k][l] = H[l][k] = nhess[h];
        }
        failed = step_failed;
        // cuda_prepare_next_iteration (agpufitjni.cu:2458-2490)
        if (decreased) {
            lambda *= 0.1;
            prev_chi2 = nchi2;
        } else {
            lambda *= 10.0;
#pragma unroll
            for (int k = 0; k < NF; k++) par[FixedSlots<KIND, NF>::slot(k)] = prev[k];
        }
    }
    chi2 = prev_chi2;
}

template <int KIND, int NF>
__global__ void __launch_bounds__(64) k_lm_fixed(LMArgs a) {
    constexpr int NP = (KIND == kModelFCS) ? 20 : 11;
    const int fit = blockIdx.x;
    if (fit >= a.nfits) return;
    int prev_iters;
    if (lm_skip(a, fit, prev_iters)) return;
    const int lane = threadIdx.x & 63;
    float par[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) par[k] = init_row(a, fit)[k];
    LMArgs ac = a;
    if (KIND == kModelFCS) fcs_params_in(par, ac.geom);
    auto eval = [&](const float *p, double &c2, double *g, double *h) {
        lm_eval_fixed<KIND, NF>(ac, fit, p, true, c2, g, h);
    };
    double chi2;
    int state, iters;
    lm_fixed_iterate<KIND, NF>(a, par, eval, chi2, state, iters);
    if (KIND == kModelFCS) fcs_params_out(par);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NP; k++) a.out_params[(int64_t)fit * NP + k] = par[k];
        a.states[fit] = state;
        a.chi2[fit] = (float)chi2;
        a.iters[fit] = iters + prev_iters;
    }
}

// Exponential bleach fits with FP64 parameters, model and sums (the Java
// CPU fit is FP64; FP32 parameters cannot follow the shallow A-C-t valley of
// a weakly decaying trace).  Gpufit's step control as above; one wave per
// trace, points lane, lane + 64, ...; the result goes to out64[fit][5].
template <int NEXP>
__global__ void __launch_bounds__(64) k_lm_exp64(LMArgs a, double *out64) {
    constexpr int NF = NEXP == 1 ? 3 : 5, NH = NF * (NF + 1) / 2;
    const int fit = blockIdx.x;
    if (fit >= a.nfits) return;
    const int lane = threadIdx.x & 63;
    double par[11];
#pragma unroll
    for (int k = 0; k < 11; k++) par[k] = (double)init_row(a, fit)[k];
    const float *y = a.data + (int64_t)fit * a.npts;
    auto eval = [&](const double *p, double &c2, double *g, double *h) {
        double cx = 0.0, gl[NF], hl[NH];
#pragma unroll
        for (int q = 0; q < NF; q++) gl[q] = 0.0;
#pragma unroll
        for (int q = 0; q < NH; q++) hl[q] = 0.0;
        for (int i = lane; i < a.npts; i += 64) {
            const double t = (double)a.x[i];
            const double e1 = exp(-t / p[1]);
            double jf[NF], v = p[0] * e1;
            jf[0] = e1;
            jf[1] = p[0] * t * e1 / (p[1] * p[1]);
            if constexpr (NEXP == 2) {
                const double e2 = exp(-t / p[3]);
                v += p[2] * e2 + p[4];
                jf[2] = e2;
                jf[3] = p[2] * t * e2 / (p[3] * p[3]);
                jf[4] = 1.0;
            } else {
                v += p[2];
                jf[2] = 1.0;
            }
            const double r = (double)y[i] - v;
            cx = fma(r, r, cx);
            int hh = 0;
#pragma unroll
            for (int q = 0; q < NF; q++) {
                gl[q] = fma(r, jf[q], gl[q]);
#pragma unroll
                for (int l = 0; l <= q; l++, hh++) hl[hh] = fma(jf[q], jf[l], hl[hh]);
            }
        }
        c2 = wave_sum(cx);
#pragma unroll
        for (int q = 0; q < NF; q++) g[q] = wave_sum(gl[q]);
#pragma unroll
        for (int q = 0; q < NH; q++) h[q] = wave_sum(hl[q]);
    };
    double chi2;
    int state, iters;
    lm_fixed_iterate<kModelPoly, NF>(a, par, eval, chi2, state, iters);
    // Undamped Gauss-Newton steps from the LM end point, kept while chi^2
    // falls: the damped steps crawl along the shallow A-C-t valley, and the
    // reference's MINPACK-style LM reaches its bottom.
    for (int it = 0; it < 16; it++) {
        double c2, g[NF], hp[NH], H[NF][NF];
        eval(par, c2, g, hp);
        int h = 0;
#pragma unroll
        for (int k = 0; k < NF; k++)
#pragma unroll
            for (int l = 0; l <= k; l++, h++) H[k][l] = H[l][k] = hp[h];
        if (!chol_solve_fixed<NF>(H, g)) break;
        double trial[11];
#pragma unroll
        for (int k = 0; k < 11; k++) trial[k] = k < NF ? par[k] + g[k] : par[k];
        double c2t, gt[NF], ht[NH];
        eval(trial, c2t, gt, ht);
        if (!(c2t < c2)) break;
#pragma unroll
        for (int k = 0; k < NF; k++) par[k] = trial[k];
        chi2 = c2t;
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 5; k++) out64[(int64_t)fit * 5 + k] = k < NF ? par[k] : 0.0;
        a.states[fit] = state;
        a.chi2[fit] = (float)chi2;
        a.iters[fit] = iters;
    }
}

// One dimension of the one-component model in FP64 (the device library's
// exp and erf): the lateral factor l and the D derivative of its
// exponential terms, the same expressions as fcs_dim (agpufitjni.cu:440-535
// evaluates the model in FP64 as well).
__device__ void fcs_dim64(double a, double r, double v, double x, double p0, double &l, double &dDpexp) {
    const double isqrpi = 0.56418958354775628;
    if (r == 0.0 && v == 0.0) {
        const double u = a / p0, E = exp(-u * u);
        l = p0 * isqrpi * (2.0 * E - 2.0) + 2.0 * a * erf(u);
        dDpexp = 2.0 * (2.0 * E - 2.0);
        return;
    }
    const double p1 = a + r - v * x, p2 = a - r + v * x, p3 = r - v * x;
    const double u1 = p1 / p0, u2 = p2 / p0, u3 = p3 / p0;
    const double pexp = exp(-u1 * u1) + exp(-u2 * u2) - 2.0 * exp(-u3 * u3);
    l = p0 * isqrpi * pexp + (p1 * erf(u1) + p2 * erf(u2) - 2.0 * p3 * erf(u3));
    const double ip02 = 1.0 / (p0 * p0);
    const double p4 = 2.0 * a * a + 3.0 * r * r - 6.0 * x * r * v + 3.0 * (x * v) * (x * v);
    const double p5 = p3 * p3 + p1 * p1, p6 = p3 * p3 + p2 * p2;
    const double p7 = 2.0 * (a * a + r * r - 2.0 * x * r * v + (x * v) * (x * v));
    dDpexp = 2.0 * (exp((p5 - p4) * ip02) + exp((p6 - p4) * ip02) - 2.0 * exp((p7 - p4) * ip02));
}

// The one-component model (no second or third component, no triplet; flow
// fixed) in FP64: value and derivatives in N, D, G_inf (fcs3p's slots 0, 1,
// 4), parameters in the fit's internal units.
__device__ double fcs1_f64(const LMModelGeom &m, double x, double N, double D, double vx, double vy, double G,
                           double (&j)[3], double &acf_out) {
    const double isqrpi = 0.56418958354775628;
    const double ax = m.ax, ay = m.ay, s = m.s, sz = m.sz;
    const double p0 = sqrt(4.0 * D * x + s * s);
    double lx, dx, ly, dy;
    fcs_dim64(ax, m.rx, vx, x, p0, lx, dx);
    if (m.ay == m.ax && m.ry == m.rx && vy == vx) {
        ly = lx;
        dy = dx;
    } else {
        fcs_dim64(ay, m.ry, vy, x, p0, ly, dy);
    }
    const double inorm = (double)m.obsvol / (4.0 * (ax * ay) * (ax * ay));
    const double plat = lx * ly * inorm;
    const double szi2 = 1.0 / (sz * sz);
    const double pspim = 1.0 / sqrt(1.0 + 4.0 * D * x * szi2);
    const double acf = plat * pspim;
    const double dDplat = (isqrpi / p0) * (dy * x * lx + dx * x * ly) * inorm;
    const double dDpspim = -2.0 * x * szi2 * (pspim * pspim * pspim);
    const double iN = 1.0 / N;
    j[0] = -acf * iN * iN;
    j[1] = iN * (plat * dDpspim + pspim * dDplat);
    j[2] = 1.0;
    acf_out = acf;
    return acf * iN + G;
}

// The same LM fit of the FCS model with NF fixed free slots for curves of
// more than 64 points and an FP64 chi-square: two wavefronts per fit, one
// point per lane (up to 128 points, the p = 16 layouts up to q = 15; the
// lanes loop over longer curves), so one model evaluation per iteration
// instead of two passes of one wave (an iteration of a lone long-running
// fit is latency-bound).  The
// per-wave sums meet in LDS (double-buffered, one barrier per evaluation);
// both waves then take identical LM steps from identical sums.
template <int NF, bool LONG = false>
__device__ void lm_fcs_2w_fit(const LMArgs &a, int fit) {
    constexpr int NP = 20, NH = NF * (NF + 1) / 2, NS = NF + NH;
    // per buffer and wave: the FP64 chi-square and the FP32 gradient and
    // normal-matrix sums (FP32 as in Gpufit; only the chi-square, which
    // decides step acceptance and convergence, needs FP64)
    __shared__ double red_c2[2][2];
    __shared__ float red[2][2][NS];
    int prev_iters;
    if (lm_skip(a, fit, prev_iters)) return;
    const int tid = threadIdx.x, wv = tid >> 6;
    float par[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) par[k] = init_row(a, fit)[k];
    LMArgs ac = a;
    fcs_params_in(par, ac.geom);
    const bool c2on = (par[5] != 0.f) || a.pfit[6];
    const bool c3on = (par[7] != 0.f) || a.pfit[8];
    const bool act = tid < a.npts;
    const float xi = act ? a.x[tid] : 0.f;
    const float yi = act ? a.data[(int64_t)fit * a.npts + tid] : 0.f;
    const double wi = !act ? 0.0 : (a.weights ? (double)a.weights[(int64_t)fit * a.npts + tid] : 1.0);
    int buf = 0;
    const float wf = (float)wi;
    auto eval = [&](const float *p, double &c2, double *g, double *h) {
        double cx = 0.0;
        float sv[NS];
#pragma unroll
        for (int k = 0; k < NS; k++) sv[k] = 0.f;
        // a point's terms: set (the lane's first point) or added (the
        // points past 128 of a LONG curve)
        auto point = [&](float xx, float yy, double ww, float wq, auto add) {
            constexpr bool ADD = decltype(add)::value;
            float gfull[11];
            const float f = fcs3p(ac.geom, xx, p, true, c2on, c3on, gfull);
            const double r = (double)yy - (double)f;   // exact
            if constexpr (ADD) cx += ww * r * r;
            else cx = ww * r * r;
            const float wr = wq * (float)r;
            int hh = NF;
#pragma unroll
            for (int q = 0; q < NF; q++) {
                const float jq = gfull[FixedSlots<kModelFCS, NF>::slot(q)];
                if constexpr (ADD) sv[q] += wr * jq;
                else sv[q] = wr * jq;
#pragma unroll
                for (int l = 0; l <= q; l++, hh++) {
                    const float e = wq * jq * gfull[FixedSlots<kModelFCS, NF>::slot(l)];
                    if constexpr (ADD) sv[hh] += e;
                    else sv[hh] = e;
                }
            }
        };
        if (act) point(xi, yi, wi, wf, std::false_type{});
        if constexpr (LONG)
            for (int i = tid + 128; i < a.npts; i += 128) {
                const double ww = a.weights ? (double)a.weights[(int64_t)fit * a.npts + i] : 1.0;
                point(a.x[i], a.data[(int64_t)fit * a.npts + i], ww, (float)ww, std::true_type{});
            }
        cx = wave_sum(cx);
#pragma unroll
        for (int k = 0; k < NS; k++) sv[k] = wave_sum_f32(sv[k]);
        if ((tid & 63) == 0) {
            red_c2[buf][wv] = cx;
#pragma unroll
            for (int k = 0; k < NS; k++) red[buf][wv][k] = sv[k];
        }
        __syncthreads();
        c2 = red_c2[buf][0] + red_c2[buf][1];
#pragma unroll
        for (int q = 0; q < NF; q++) g[q] = (double)(red[buf][0][q] + red[buf][1][q]);
#pragma unroll
        for (int q = 0; q < NH; q++) h[q] = (double)(red[buf][0][NF + q] + red[buf][1][NF + q]);
        buf ^= 1;
    };
    double chi2;
    int state, iters;
    lm_fixed_iterate<kModelFCS, NF>(a, par, eval, chi2, state, iters);
    fcs_params_out(par);
    if (tid == 0) {
#pragma unroll
        for (int k = 0; k < NP; k++) a.out_params[(int64_t)fit * NP + k] = par[k];
        a.states[fit] = state;
        a.chi2[fit] = (float)chi2;
        a.iters[fit] = iters + prev_iters;
    }
}

// FP64 polish of a converged N, D, G_inf fit of the one-component model
// (no second or third component, no triplet): Newton steps on chi^2 with
// the FP64 model and FP64 sums from the LM's end point, kept while chi^2
// falls.  The FP32 model's rounding (about 1e-6 relative) moves the LM's
// end point along a shallow chi^2 valley by up to ~1e-3 in D on short
// series; the steps end at the FP64 optimum, where the reference's FP64 CPU
// fit ends.  Along that valley the residuals are not small, so Gauss-Newton
// steps (J^T W J alone) overshoot and zig-zag; the Hessian here adds the
// second-order term -sum w r d2f, whose N-N and N-D entries are analytic
// and whose D-D entry comes from a forward difference of the analytic D
// derivative (step 1e-5 D), taken once at the start point.  Two waves, one
// point per lane (points past 128, layouts other than p = 16, q = 8, loop
// over the lanes again); a separate launch, so that the FP64 model's
// registers do not lower the LM kernel's occupancy.  Reads and rewrites the
// fit's output row and chi-square.
template <bool LONG>
__device__ void lm_polish_fit(const LMArgs &a, int fit) {
    __shared__ double red64[2][2][13];
    const int tid = threadIdx.x, wv = tid >> 6;
    if (a.states[fit] != kStateConverged) return;   // uniform per workgroup
    float *row = a.out_params + (int64_t)fit * 20;
    float par[20];
#pragma unroll
    for (int k = 0; k < 20; k++) par[k] = row[k];
    if (par[5] != 0.f || par[7] != 0.f || par[9] != 0.f || !(par[0] != 0.f) || !(par[1] > 0.f)) return;
    LMModelGeom geom;
    fcs_params_in(par, geom);
    const bool act = tid < a.npts;
    const double x64 = act ? (double)a.x[tid] : 0.0;
    const double y64 = act ? (double)a.data[(int64_t)fit * a.npts + tid] : 0.0;
    const double wi = !act ? 0.0 : (a.weights ? (double)a.weights[(int64_t)fit * a.npts + tid] : 1.0);
    const double vx = (double)par[2], vy = (double)par[3];
    int buf = 0;
    // chi^2, gradient g = J^T W r, packed J^T W J and (second) the
    // second-order sums sum w r d2f/dN2, d2f/dNdD, d2f/dD2
    auto eval64 = [&](const double (&p)[3], bool second, double &c2, double (&g)[3], double (&h)[6],
                      double (&sm)[3]) {
        double v[13];
#pragma unroll
        for (int k = 0; k < 13; k++) v[k] = 0.0;
        // a point's terms: set (the lane's first point) or added (the
        // points past 128 of a LONG curve)
        auto point = [&](double xx, double yy, double ww, auto add) {
            constexpr bool ADD = decltype(add)::value;
            auto put = [&](int k, double e) {
                if constexpr (ADD) v[k] += e;
                else v[k] = e;
            };
            double jq[3], acf;
            const double r = yy - fcs1_f64(geom, xx, p[0], p[1], vx, vy, p[2], jq, acf);
            put(0, ww * r * r);
            int hh = 4;
#pragma unroll
            for (int q = 0; q < 3; q++) {
                put(1 + q, ww * r * jq[q]);
#pragma unroll
                for (int l = 0; l <= q; l++, hh++) put(hh, ww * jq[q] * jq[l]);
            }
            if (second) {
                const double hD = 1e-5 * p[1], iN = 1.0 / p[0];
                double jh[3], acfh;
                (void)fcs1_f64(geom, xx, 1.0, p[1] + hD, vx, vy, 0.0, jh, acfh);
                const double aD = jq[1] * p[0];   // d acf / dD
                put(10, ww * r * (2.0 * acf * iN * iN * iN));
                put(11, ww * r * (-aD * iN * iN));
                put(12, ww * r * ((jh[1] - aD) / hD * iN));
            }
        };
        if (act) point(x64, y64, wi, std::false_type{});
        if constexpr (LONG)
            for (int i = tid + 128; i < a.npts; i += 128)
                point((double)a.x[i], (double)a.data[(int64_t)fit * a.npts + i],
                      a.weights ? (double)a.weights[(int64_t)fit * a.npts + i] : 1.0, std::true_type{});
        const int nv = second ? 13 : 10;   // uniform
        for (int k = 0; k < nv; k++) v[k] = wave_sum(v[k]);
        if ((tid & 63) == 0)
            for (int k = 0; k < nv; k++) red64[buf][wv][k] = v[k];
        __syncthreads();
        c2 = red64[buf][0][0] + red64[buf][1][0];
#pragma unroll
        for (int q = 0; q < 3; q++) g[q] = red64[buf][0][1 + q] + red64[buf][1][1 + q];
#pragma unroll
        for (int q = 0; q < 6; q++) h[q] = red64[buf][0][4 + q] + red64[buf][1][4 + q];
        if (second)
#pragma unroll
            for (int q = 0; q < 3; q++) sm[q] = red64[buf][0][10 + q] + red64[buf][1][10 + q];
        buf ^= 1;
    };
    double p[3] = {(double)par[0], (double)par[1], (double)par[4]};
    double c2, g[3], h[6], sm[3];
    eval64(p, true, c2, g, h, sm);
    for (int it = 0; it < 3; it++) {
        double d[3];
        bool solved = false;
        for (int full = 1; full >= 0 && !solved; full--) {
            // Jacobi-scaled Hessian (the columns differ by orders of
            // magnitude); without the second-order term if that one is not
            // positive definite
            double sc[3], H[3][3];
            bool ok = true;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double hk = h[k * (k + 1) / 2 + k];
                ok = ok && hk > 0.0;
                sc[k] = hk > 0.0 ? 1.0 / sqrt(hk) : 0.0;
            }
            if (!ok) break;
            int hh = 0;
#pragma unroll
            for (int k = 0; k < 3; k++)
#pragma unroll
                for (int l = 0; l <= k; l++, hh++) {
                    double e = h[hh];
                    if (full) e -= (k == 0 && l == 0) ? sm[0] : (k == 1 && l == 0) ? sm[1] : (k == 1 && l == 1) ? sm[2] : 0.0;
                    H[k][l] = H[l][k] = e * sc[k] * sc[l];
                }
#pragma unroll
            for (int k = 0; k < 3; k++) d[k] = g[k] * sc[k];
            solved = chol_solve_fixed<3>(H, d);
#pragma unroll
            for (int k = 0; k < 3; k++) d[k] *= sc[k];
        }
        if (!solved) break;
        double t[3];
        bool moves = false;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            t[k] = p[k] + d[k];
            moves = moves || fabs(d[k]) > 1e-12 * fabs(p[k]);
        }
        if (!moves) break;
        double c2t, gt[3], ht[6], unused[3];
        eval64(t, false, c2t, gt, ht, unused);
        if (!(c2t < c2)) break;
        c2 = c2t;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            p[k] = t[k];
            g[k] = gt[k];
        }
#pragma unroll
        for (int k = 0; k < 6; k++) h[k] = ht[k];
    }
    if (tid == 0) {
        // back to SI units, as fcs_params_out
        row[0] = (float)p[0];
        row[1] = (float)p[1] * 1e-12f;
        row[4] = (float)p[2];
        a.chi2[fit] = (float)c2;
    }
}

template <bool LONG>
__global__ void __launch_bounds__(128) k_lm_polish(LMArgs a) {
    const int fit = blockIdx.x;
    if (fit < a.nfits) lm_polish_fit<LONG>(a, fit);
}

// The polish over the fits of many groups (a batch's persistent fit
// launch): one workgroup per fit, its group found in the prefix table.
__global__ void __launch_bounds__(128) k_lm_polish_groups(const LMGroup *groups, const int64_t *prefix, int ngroups,
                                                          int64_t total) {
    const int64_t idx = blockIdx.x;
    if (idx >= total) return;
    int lo = 0, hi = ngroups - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= idx) lo = mid;
        else hi = mid - 1;
    }
    const LMGroup &G = groups[lo];
    LMArgs a{G.nfits, G.npts, 20, kModelFCS, 0, G.data, G.weights, G.x, G.init, G.pfit, LMModelGeom{}, G.tol,
             G.maxit, 1, G.out_params, G.states, G.chi2, G.iters, G.init_two_rows, G.init_split, 0};
    if (a.npts > 128)   // uniform per workgroup
        lm_polish_fit<true>(a, (int)(idx - prefix[lo]));
    else
        lm_polish_fit<false>(a, (int)(idx - prefix[lo]));
}

template <int NF, bool LONG>
__global__ void __launch_bounds__(128) k_lm_fcs_2w(LMArgs a) {
    // last fit first: the CCF fits (the second half), whose few longest
    // fits set the kernel's duration, are dispatched before the ACF fits
    const int fit = a.nfits - 1 - (int)blockIdx.x;
    if (fit < 0) return;
    lm_fcs_2w_fit<NF, LONG>(a, fit);
}

// One persistent launch over the fits of many pipelines (a batch of
// stacks): workgroups take fits from a global counter until every fit of
// every group is taken, so the long fits of one stack overlap the others'
// work instead of each stack's launch waiting for its slowest fit.  Within a
// group the last fit is dispatched first, as in k_lm_fcs_2w.
__global__ void __launch_bounds__(128) k_lm_queue(const LMGroup *groups, const int64_t *prefix, int ngroups,
                                                  int64_t total, unsigned long long *counter) {
    __shared__ long long s_idx;
    for (;;) {
        __syncthreads();   // the previous fit's reads of shared memory are done
        if (threadIdx.x == 0) s_idx = (long long)atomicAdd(counter, 1ull);
        __syncthreads();
        const int64_t idx = s_idx;
        if (idx >= total) break;
        int lo = 0, hi = ngroups - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (prefix[mid] <= idx) lo = mid;
            else hi = mid - 1;
        }
        const LMGroup &G = groups[lo];
        LMArgs a{G.nfits, G.npts, 20, kModelFCS, 0, G.data, G.weights, G.x, G.init, G.pfit, LMModelGeom{}, G.tol,
                 G.maxit, 1, G.out_params, G.states, G.chi2, G.iters, G.init_two_rows, G.init_split, 0};
        lm_fcs_2w_fit<3>(a, G.nfits - 1 - (int)(idx - prefix[lo]));
    }
}

__global__ void k_zero_u64(unsigned long long *p) { *p = 0ull; }

bool lm_queue_eligible(int npts, unsigned free_mask, int chi2_fp64) {
    return free_mask == 0x13u && chi2_fp64 && npts > 64 && npts <= 128;
}

int lm_queue_blocks() {
    int dev = 0, ncu = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void *)k_lm_queue, 128, 0);
    return std::max(1, ncu) * std::max(1, per);
}

void launch_lm_queue(const LMGroup *groups, const int64_t *prefix, int ngroups, int64_t total,
                     unsigned long long *counter, int blocks, hipStream_t st) {
    if (total <= 0 || ngroups <= 0) return;
    // the counter is cleared by a kernel (no memset node in captured graphs)
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, counter);
    const int64_t nb = std::min<int64_t>(total, std::max(1, blocks));
    hipLaunchKernelGGL(k_lm_queue, dim3((unsigned)nb), dim3(128), 0, st, groups, prefix, ngroups, total, counter);
}

// GLS fit of the FCS model with N, D, G_inf free (GLSFit, IFCS:10622-10789):
// the model values and Jacobian at the m lag times are whitened by W = L^-1
// (dataTransform / GLS_fitFunction, IFCS:13569-13729) and fitted to the
// whitened data yw = W mcov with unit weights; chi-square in FP64.  W lives
// in LDS (m x m doubles) for the fit's lifetime.
__global__ void __launch_bounds__(64) k_lm_gls(LMArgs a, const double *W, const double *yw, const int32_t *okv,
                                               int m) {
    constexpr int NF = 3, NH = 6, NP = 20;
    extern __shared__ double gls_lds[];
    double *Wl = gls_lds;              // [m][m]
    double *fv = gls_lds + m * m;      // [m]
    double *jv = fv + m;               // [m][NF]
    const int fit = blockIdx.x;
    if (fit >= a.nfits) return;
    const int lane = threadIdx.x & 63;
    float par[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) par[k] = init_row(a, fit)[k];
    if (!okv[fit]) {
        // the Cholesky factorization failed: Java's fit throws (no result)
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < NP; k++) a.out_params[(int64_t)fit * NP + k] = par[k];
            a.states[fit] = kStateSingular;
            a.chi2[fit] = __builtin_nanf("");
            a.iters[fit] = 0;
        }
        return;
    }
    const double *Wg = W + (int64_t)fit * m * m;
    for (int i = lane; i < m * m; i += 64) Wl[i] = Wg[i];
    const double *ywf = yw + (int64_t)fit * m;
    LMArgs ac = a;
    fcs_params_in(par, ac.geom);
    const bool c2on = (par[5] != 0.f) || a.pfit[6];
    const bool c3on = (par[7] != 0.f) || a.pfit[8];
    __syncthreads();
    auto eval = [&](const float *p, double &c2, double *g, double *h) {
        for (int i = lane; i < m; i += 64) {
            float gfull[11];
            const float f = fcs3p(ac.geom, a.x[i], p, true, c2on, c3on, gfull);
            fv[i] = (double)f;
#pragma unroll
            for (int q = 0; q < NF; q++) jv[i * NF + q] = (double)gfull[FixedSlots<kModelFCS, NF>::slot(q)];
        }
        __syncthreads();
        double cx = 0.0, gl[NF], hl[NH];
#pragma unroll
        for (int q = 0; q < NF; q++) gl[q] = 0.0;
#pragma unroll
        for (int q = 0; q < NH; q++) hl[q] = 0.0;
        for (int r = lane; r < m; r += 64) {
            double fw = 0.0, jw[NF] = {0.0, 0.0, 0.0};
            for (int k = 0; k <= r; k++) {
                const double w = Wl[r * m + k];
                fw = fma(w, fv[k], fw);
#pragma unroll
                for (int q = 0; q < NF; q++) jw[q] = fma(w, jv[k * NF + q], jw[q]);
            }
            const double res = ywf[r] - fw;
            cx = fma(res, res, cx);
            int hh = 0;
#pragma unroll
            for (int q = 0; q < NF; q++) {
                gl[q] = fma(res, jw[q], gl[q]);
#pragma unroll
                for (int l = 0; l <= q; l++, hh++) hl[hh] = fma(jw[q], jw[l], hl[hh]);
            }
        }
        __syncthreads();
        c2 = wave_sum(cx);
#pragma unroll
        for (int q = 0; q < NF; q++) g[q] = wave_sum(gl[q]);
#pragma unroll
        for (int q = 0; q < NH; q++) h[q] = wave_sum(hl[q]);
    };
    double chi2;
    int state, iters;
    lm_fixed_iterate<kModelFCS, NF>(a, par, eval, chi2, state, iters);
    fcs_params_out(par);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NP; k++) a.out_params[(int64_t)fit * NP + k] = par[k];
        a.states[fit] = state;
        a.chi2[fit] = (float)chi2;
        a.iters[fit] = iters;
    }
}

// ------------------------------------------------------------------
// Drop-in helpers: float binning and bleach-trace averaging
// ------------------------------------------------------------------
__global__ void k_binning_f32(const float *in, float *out, int win, int hin, int wt, int ht, int T, int bx, int by) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int yz = blockIdx.y;
    int y = yz % ht, t = yz / ht;
    if (x >= wt || t >= T) return;
    const float *b = in + ((int64_t)t * hin + y) * win + x;
    float s = 0.f;
    for (int i = 0; i < bx; i++)
        for (int j = 0; j < by; j++) s += b[(int64_t)j * win + i];
    out[((int64_t)t * ht + y) * wt + x] = s;
}

__global__ void k_bleach_average_f32(const float *in, float *out, int wt, int ht, int nopit, int ave) {
    int pix = blockIdx.x * blockDim.x + threadIdx.x;
    int z = blockIdx.y;
    if (pix >= wt * ht || z >= nopit) return;
    double s = 0.0;
    for (int f = z * ave; f < (z + 1) * ave; f++) s += (float)in[(int64_t)f * wt * ht + pix];
    out[(int64_t)pix * nopit + z] = (float)(s / ave);
}

__global__ void k_pack_fit_inputs(int chanum, int npairs, const double *corav, const double *bvar, float *data,
                                  float *weights) {
    const int p = blockIdx.x;
    if (p >= npairs) return;
    for (int z = threadIdx.x; z < chanum - 1; z += blockDim.x) {
        const int64_t src = (int64_t)p * chanum + z + 1, dst = (int64_t)p * (chanum - 1) + z;
        data[dst] = (float)corav[src];
        weights[dst] = 1.0f / (float)bvar[src];
    }
}

// N&B of the drop-in calcACF (isNBcalculation, agpufitjni.cu:5167-5179):
// the mean <n(t)> over frames 0..T-2 and the lag-1 covariance
// <n(t) n(t+1)> - <n(t)><n(t+1)>, i.e. channel 1's direct monitor and its
// correlation numerator: nbcov = G(lag 1) * abar * bbar.
__global__ void k_nb_extract(const CorrTables *__restrict__ ctp, SeriesGeom g, PairDesc pd, const double *F,
                             const double *corav, double *nbmean, double *nbcov) {
    const CorrTables &ct = *ctp;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pd.npairs) return;
    int sa, sb;
    pair_series(pd, p, sa, sb);
    const double *Fa = F + (int64_t)g.slot_of[sa] * ct.npos, *Fb = F + (int64_t)g.slot_of[sb] * ct.npos;
    const double n = (double)(ct.numbin[0] - ct.del[1]);
    const double am = Fa[ct.pos_a_end[1]] / n;
    const double bm = (Fb[ct.pos_b_end[1]] - Fb[ct.pos_b_beg[1]]) / n;
    nbmean[p] = am;
    nbcov[p] = corav[(int64_t)p * ct.chanum + 1] * am * bm;
}

// Plain device copy (16-byte vectors, grid-stride, byte tail).
__global__ void k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16, const uint8_t *src_tail,
                       uint8_t *dst_tail, int ntail) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
    if (blockIdx.x == 0 && (int)threadIdx.x < ntail) dst_tail[threadIdx.x] = src_tail[threadIdx.x];
}

// The same copy in 4-byte words by buffer loads and stores, one word per
// lane and 64 rows of 64 lanes per step, with the row as the scalar offset:
// the access pattern of the pipeline's FP32 series passes (the known-size
// reference of the traffic counters for such kernels).  Words past n32 are
// not touched.
__global__ void k_copy_words(const uint32_t *src, uint32_t *dst, int64_t n32) {
    const int64_t base = (int64_t)blockIdx.x * 4096;   // 64 rows of 64 words
    const int lane = threadIdx.x;
    const int64_t left = n32 - base;
    if (left <= 0) return;
    const int nbytes = (int)(std::min<int64_t>(left, 4096) * 4);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void *)(src + base), (short)0, nbytes,
                                                                        kBufferWord3);
    const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc((void *)(dst + base), (short)0, nbytes,
                                                                        kBufferWord3);
    uint32_t v[64];
#pragma unroll
    for (int r = 0; r < 64; r++)
        v[r] = (64 * r + lane) * 4 < nbytes ? __builtin_amdgcn_raw_buffer_load_b32(rs, lane * 4, r * 256, 0) : 0u;
#pragma unroll
    for (int r = 0; r < 64; r++)
        if ((64 * r + lane) * 4 < nbytes) __builtin_amdgcn_raw_buffer_store_b32(v[r], rd, lane * 4, r * 256, 0);
}

// ------------------------------------------------------------------
// Launchers
// ------------------------------------------------------------------
static inline int cdiv(int64_t a, int b) { return (int)((a + b - 1) / b); }

void launch_nb_extract(const CorrTables *dct, const SeriesGeom &g, const PairDesc &pd, const double *F,
                       const double *corav, double *nbmean, double *nbcov, hipStream_t st) {
    hipLaunchKernelGGL(k_nb_extract, dim3(cdiv(pd.npairs, 64)), dim3(64), 0, st, dct, g, pd, F, corav, nbmean, nbcov);
}

__global__ void k_accum_div(double *dst, const double *src, int64_t n, double div, int first) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (first ? 0.0 : dst[i]) + src[i] / div;
}

void launch_accum_div(double *dst, const double *src, int64_t n, double div, int first, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_accum_div, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dst, src, n, div, first);
}

void launch_copy(void *dst, const void *src, size_t bytes, hipStream_t st) {
    const size_t n16 = bytes / 16;
    const int ntail = (int)(bytes - n16 * 16);
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>((n16 + 255) / 256, 256 * 64));
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, st, (const uint4 *)src, (uint4 *)dst, n16,
                       (const uint8_t *)src + n16 * 16, (uint8_t *)dst + n16 * 16, ntail);
}

void launch_copy_words(void *dst, const void *src, size_t bytes, hipStream_t st) {
    const int64_t n32 = (int64_t)(bytes / 4);
    if (n32 <= 0) return;
    hipLaunchKernelGGL(k_copy_words, dim3((unsigned)((n32 + 4095) / 4096)), dim3(64), 0, st, (const uint32_t *)src,
                       (uint32_t *)dst, n32);
}

void launch_trace_moments(const SeriesGeom &g, int nopit, int ave, double tscale, double frametime, int order,
                          int integer_trace, double *part, int nchunk, hipStream_t st, float *trace_out) {
    int ppc = cdiv(nopit, nchunk);
    dim3 grid(cdiv(g.nslots, 64), nchunk);
    switch (raw_batch_rows(g)) {
        case 2:
            hipLaunchKernelGGL(k_trace_moments<2>, grid, dim3(64), 0, st, g, nopit, ave, tscale, frametime, order,
                               integer_trace, ppc, part, trace_out);
            break;
        default:
            hipLaunchKernelGGL(k_trace_moments<0>, grid, dim3(64), 0, st, g, nopit, ave, tscale, frametime, order,
                               integer_trace, ppc, part, trace_out);
            break;
    }
}

// Initial guesses of SingleExpFit / DoubleExpFit (IFCS:10862-10864,
// 10913-10917) from each slot's trace: [tr[1], t[n/2], 0] or
// [tr[1]/2, t[n/10], tr[1]/2, t[n/2], 0], in the 11-slot parameter layout.
__global__ void k_exp_init(const float *trace, const float *itime, int nslots, int nopit, int nexp, float *init) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const float t1 = trace[(int64_t)s * nopit + 1];
    float *q = init + (int64_t)s * 11;
#pragma unroll
    for (int k = 0; k < 11; k++) q[k] = 0.f;
    if (nexp == 1) {
        q[0] = t1;
        q[1] = itime[nopit / 2];
    } else {
        q[0] = t1 / 2;
        q[1] = itime[nopit / 10];
        q[2] = t1 / 2;
        q[3] = itime[nopit / 2];
    }
}

// Fitted exponentials -> coef[slot] = {f(0), params}: f(0) = A + C or
// A + B + C; 0 when the reference would skip the correction (A t1 == 0,
// or A t1 B t2 == 0; IFCS:9689, 9713).
__global__ void k_exp_pack(const double *out64, int nslots, int nexp, double *coef) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const int np = nexp == 1 ? 3 : 5;
    const double *p = out64 + (int64_t)s * 5;
    double *c = coef + (int64_t)s * (np + 1);
    for (int k = 0; k < np; k++) c[1 + k] = p[k];
    const bool on = nexp == 1 ? (p[0] * p[1] != 0.0) : (p[0] * p[1] * p[2] * p[3] != 0.0);
    c[0] = !on ? 0.0 : (nexp == 1 ? p[0] + p[2] : p[0] + p[2] + p[4]);
}

void launch_exp_init(const float *trace, const float *itime, int nslots, int nopit, int nexp, float *init,
                     hipStream_t st) {
    hipLaunchKernelGGL(k_exp_init, dim3(cdiv(nslots, 64)), dim3(64), 0, st, trace, itime, nslots, nopit, nexp,
                       init);
}

void launch_exp_pack(const double *out64, int nslots, int nexp, double *coef, hipStream_t st) {
    hipLaunchKernelGGL(k_exp_pack, dim3(cdiv(nslots, 64)), dim3(64), 0, st, out64, nslots, nexp, coef);
}

void launch_point_sums(const uint16_t *data, int64_t plane, int f0, int nopit, int ave, int32_t *P,
                       hipStream_t st) {
    const int ppt = 8;
    const bool vec4 = plane % 4 == 0 && (uintptr_t)data % 8 == 0;
    const int64_t groups = vec4 ? plane / 4 : plane;
    dim3 grid((unsigned)((groups + 255) / 256), (unsigned)cdiv(nopit, ppt));
    if (vec4)
        hipLaunchKernelGGL(k_point_sums<4>, grid, dim3(256), 0, st, data, plane, f0, nopit, ave, ppt, P);
    else
        hipLaunchKernelGGL(k_point_sums<1>, grid, dim3(256), 0, st, data, plane, f0, nopit, ave, ppt, P);
}

void launch_trace_points(const SeriesGeom &g, const int32_t *P, int nopit, int ave, double tscale,
                         double frametime, int order, double *part, int nchunk, hipStream_t st, float *trace_out) {
    int ppc = cdiv(nopit, nchunk);
    dim3 grid(cdiv(g.nslots, 64), nchunk);
    hipLaunchKernelGGL(k_trace_points, grid, dim3(64), 0, st, g, P, nopit, ave, tscale, frametime, order, ppc, part,
                       trace_out);
}

void launch_polyfit(const double *part, int nchunk, int nseries, int nopit, int ave, double tscale,
                    double frametime, int order, double *coef, hipStream_t st) {
    // moment matrix of the normalized abscissa and its Cholesky factor
    int n = order + 1;
    double M[(kMaxOrder + 1) * (kMaxOrder + 1)] = {0};
    for (int x = 0; x < nopit; x++) {
        double u = frametime * (x + 0.5) * ave / tscale, up[2 * kMaxOrder + 1];
        up[0] = 1.0;
        for (int k = 1; k < 2 * n - 1; k++) up[k] = up[k - 1] * u;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) M[i * n + j] += up[i + j];
    }
    PolyL Lf;
    for (int i = 0; i < (kMaxOrder + 1) * (kMaxOrder + 1); i++) Lf.L[i] = 0.0;
    for (int j = 0; j < n; j++) {
        double s = M[j * n + j];
        for (int k = 0; k < j; k++) s -= Lf.L[j * n + k] * Lf.L[j * n + k];
        Lf.L[j * n + j] = sqrt(s);
        for (int i = j + 1; i < n; i++) {
            double t = M[i * n + j];
            for (int k = 0; k < j; k++) t -= Lf.L[i * n + k] * Lf.L[j * n + k];
            Lf.L[i * n + j] = t / Lf.L[j * n + j];
        }
    }
    hipLaunchKernelGGL(k_polyfit, dim3(cdiv(nseries, 64)), dim3(64 * kPolyQ), 0, st, part, nchunk, nseries, order,
                       Lf, coef);
}

template <int NB>
static void launch_apply_t(const SeriesGeom &g, const BleachDesc &bd, const CorrTables *dct, double *ctot,
                           double *rec, float *X, float *l3, float *cmax, int nchunk, int chunk_len, hipStream_t st) {
    dim3 grid(cdiv(g.nslots, 64), nchunk);
    if (bd.mode == 1) {
        hipLaunchKernelGGL((k_apply<BleachSeg, NB>), grid, dim3(64), 0, st, g, bd, dct, chunk_len, ctot, rec, X, l3, cmax);
        return;
    }
    if (bd.mode == 2) {
        hipLaunchKernelGGL((k_apply<BleachExp<1>, NB>), grid, dim3(64), 0, st, g, bd, dct, chunk_len, ctot, rec,
                           X, l3, cmax);
        return;
    }
    if (bd.mode == 3) {
        hipLaunchKernelGGL((k_apply<BleachExp<2>, NB>), grid, dim3(64), 0, st, g, bd, dct, chunk_len, ctot, rec,
                           X, l3, cmax);
        return;
    }
    switch (bleach_maxo(bd.order)) {
        case 0:
            hipLaunchKernelGGL((k_apply<BleachP<0>, NB>), grid, dim3(64), 0, st, g, bd, dct, chunk_len, ctot, rec,
                               X, l3, cmax);
            break;
        case 4:
            hipLaunchKernelGGL((k_apply<BleachP<4>, NB>), grid, dim3(64), 0, st, g, bd, dct, chunk_len, ctot, rec,
                               X, l3, cmax);
            break;
        default:
            hipLaunchKernelGGL((k_apply<BleachP<kMaxOrder>, NB>), grid, dim3(64), 0, st, g, bd, dct, chunk_len,
                               ctot, rec, X, l3, cmax);
            break;
    }
}

void launch_apply(const SeriesGeom &g, const BleachDesc &bd, const CorrTables *dct, double *ctot, double *rec,
                  float *X, float *l3, float *cmax, int nchunk, int chunk_len, hipStream_t st) {
    switch (raw_batch_rows(g)) {
        case 2: launch_apply_t<2>(g, bd, dct, ctot, rec, X, l3, cmax, nchunk, chunk_len, st); break;
        default: launch_apply_t<0>(g, bd, dct, ctot, rec, X, l3, cmax, nchunk, chunk_len, st); break;
    }
}

// Segment means of getIntensity's "Lin Segment" branch (IFCS:9779-9787):
// segment k sums the 1-based frames 1 + k len .. (k+1) len - 1 (len - 1
// frames) and divides by len; one thread per (slot, segment), exact integer
// window values summed in FP64.
__global__ void k_seg_means(SeriesGeom g, int seg_len, int nseg, double *coef) {
    const int s = blockIdx.x * 64 + threadIdx.x, k = blockIdx.y;
    if (s >= g.nslots) return;
    int ox, oy;
    slot_origin(g, s, ox, oy);
    const int fa = k * seg_len, fe = (k + 1) * seg_len - 1;   // 0-based [fa, fe)
    double sum = 0.0;
    for (int f = fa; f < fe; f += 16) {
        int v[16];
        raw_ints16<0>(g, ox, oy, f, fe - 1, v);
#pragma unroll
        for (int i = 0; i < 16; i++)
            if (f + i < fe) sum += (double)v[i];
    }
    coef[(int64_t)s * (nseg + 1) + 1 + k] = sum / seg_len;
}

// bcInt0 = bcTrace[0] + (bcTrace[0] - bcTrace[1]) / 2 (IFCS:9789)
__global__ void k_seg_int0(int nslots, int nseg, double *coef) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    double *c = coef + (int64_t)s * (nseg + 1);
    c[0] = c[1] + (c[1] - c[2]) / 2;
}

void launch_seg_means(const SeriesGeom &g, int seg_len, int nseg, double *coef, hipStream_t st) {
    hipLaunchKernelGGL(k_seg_means, dim3(cdiv(g.nslots, 64), nseg), dim3(64), 0, st, g, seg_len, nseg, coef);
    hipLaunchKernelGGL(k_seg_int0, dim3(cdiv(g.nslots, 64)), dim3(64), 0, st, g.nslots, nseg, coef);
}

void launch_prefix_combine(const CorrTables *dct, int nslots, const double *ctot, const double *rec, int nchunk,
                           int chunk_len, double *F, const float *cmax, int32_t *xexp, hipStream_t st) {
    (void)chunk_len;   // the prefix sums are taken per 64-frame sub-chunk whatever the chunking
    hipLaunchKernelGGL(k_prefix_combine, dim3(cdiv(nslots, 64)), dim3(64 * kPcWaves), 0, st, dct, nslots, ctot, rec,
                       nchunk, F, cmax, xexp);
}

int prefix_subchunks(int T) { return (T + (1 << kSubLog2) - 1) >> kSubLog2; }

void launch_apply_nested(const NestedApply &A, const CorrTables *dct, int order, int nchunk, int chunk_len,
                         hipStream_t st) {
    dim3 grid((A.swp[1] >> 5) * (A.sh[1] >> 1), nchunk);
    if (bleach_maxo(order) <= 4)
        hipLaunchKernelGGL(k_apply_nested<4>, grid, dim3(64), 0, st, A, dct, chunk_len);
    else
        hipLaunchKernelGGL(k_apply_nested<kMaxOrder>, grid, dim3(64), 0, st, A, dct, chunk_len);
}

BlockingPlan plan_blocking(const CorrTables &ct, int npix, int threads_log2) {
    BlockingPlan bp;
    const int n0 = ct.T - 1;
    bp.nchunk_prod = (n0 + (1 << kBlkJ) - 1) >> kBlkJ;
    // about 2^threads_log2 threads in flight, whole chunks per split
    int want = std::max(1, (1 << threads_log2) / std::max(1, npix));
    bp.nsplit = std::min(bp.nchunk_prod, want);
    bp.chunks_per_split = (bp.nchunk_prod + bp.nsplit - 1) / bp.nsplit;
    bp.nsplit = (bp.nchunk_prod + bp.chunks_per_split - 1) / bp.chunks_per_split;
    return bp;
}

void launch_blocking(const SeriesGeom &g, const BleachDesc &bd, const CorrTables *dct, const PairDesc &pd,
                     const double *F, const float *X, const BlockingPlan &bp, double *bpart, double *csum,
                     hipStream_t st) {
    dim3 grid(cdiv((int64_t)g.oh * g.swp, 64), bp.nsplit);
    if (pd.acf_first && pd.has_ccf)
        hipLaunchKernelGGL((k_blocking<true, true>), grid, dim3(64), 0, st, g, dct, pd, F, X, bd.coef, bd.order,
                           bpart, csum, bp.chunks_per_split);
    else if (pd.acf_first)
        hipLaunchKernelGGL((k_blocking<true, false>), grid, dim3(64), 0, st, g, dct, pd, F, X, bd.coef, bd.order,
                           bpart, csum, bp.chunks_per_split);
    else
        hipLaunchKernelGGL((k_blocking<false, true>), grid, dim3(64), 0, st, g, dct, pd, F, X, bd.coef, bd.order,
                           bpart, csum, bp.chunks_per_split);
}

void launch_blocking_finalize(const CorrTables &ct, const CorrTables *dct, const SeriesGeom &g, const PairDesc &pd,
                              const double *F, const BlockingPlan &bp, const double *bpart, const double *csum,
                              int32_t *ind, int32_t *bis, hipStream_t st) {
    (void)ct;
    hipLaunchKernelGGL(k_blocking_finalize, dim3(cdiv(pd.npairs, 64 / kBfinQ)), dim3(64), 0, st, dct, g, pd, F,
                       bpart, bp.nsplit, csum, ind, bis);
}

size_t generic_level_values(const CorrTables &ct, int nslots, LevelOffsets &lo) {
    size_t tot = 0;
    for (int c = 0; c < kMaxLevel; c++) lo.off[c] = 0;
    for (int c = 1; c < ct.nlevels; c++) {
        lo.off[c] = (int64_t)tot;
        tot += (size_t)ct.numbin[c] * nslots;
    }
    return tot;
}

int generic_chunk_len(const CorrTables &ct);

void launch_correlate(const SeriesGeom &g, const BleachDesc &bd, const CorrTables &ct, const CorrTables *dct,
                      const PairDesc &pd, const double *F, const float *X, const int32_t *ind, double *corav,
                      double *bvar, double *lvg, const LevelOffsets &lo, double *part, hipStream_t st) {
    const int ns = g.nslots;
    for (int c = 1; c < ct.nlevels; c += 3) {
        const int64_t n = (int64_t)((ct.numbin[c] + 3) / 4) * ns;
        double *d2 = c + 1 < ct.nlevels ? lvg + lo.off[c + 1] : nullptr;
        double *d3 = c + 2 < ct.nlevels ? lvg + lo.off[c + 2] : nullptr;
        hipLaunchKernelGGL(k_level_sums, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, st, c == 1 ? X : nullptr,
                           c == 1 ? nullptr : lvg + lo.off[c - 1], lvg + lo.off[c], d2, d3, ct.numbin[c],
                           d2 ? ct.numbin[c + 1] : 0, d3 ? ct.numbin[c + 2] : 0, ns);
    }
    const int clen = generic_chunk_len(ct);
    const int nchunk = (ct.T + clen - 1) / clen;
    double *ab = part + (size_t)nchunk * pd.npairs * ct.chanum * 3;
    hipLaunchKernelGGL(k_correlate_means, dim3(cdiv(pd.npairs, 64), ct.chanum), dim3(64), 0, st, g, bd, dct, pd, F,
                       ab);
    // most runs of at most 8 consecutive delays on one level
    int nrun = 1;
    for (int c = 0; c < ct.nlevels; c++) {
        int n = 0;
        for (int r0 = 0, nl; r0 < ct.nchan[c]; r0 += nl, n++) {
            const int x = ct.first_chan[c] + r0;
            nl = 1;
            while (r0 + nl < ct.nchan[c] && nl < 8 && ct.del[x + nl] == ct.del[x] + nl) nl++;
        }
        nrun = std::max(nrun, n);
    }
    dim3 grid((unsigned)(cdiv(cdiv(pd.npairs, 64), 8) * 8 * nrun), ct.nlevels, nchunk);
    hipLaunchKernelGGL(k_correlate, grid, dim3(64), 0, st, g, dct, pd, X, ind, lvg, lo, (const double *)ab, clen,
                       nrun, part);
    hipLaunchKernelGGL(k_correlate_finalize, dim3(cdiv(pd.npairs, 64), ct.chanum), dim3(64), 0, st, g, bd, dct, pd, F,
                       ind, (const double *)part, clen, corav, bvar);
}

int generic_chunk_len(const CorrTables &ct) {
    // a multiple of every block of the blocked variance on every level:
    // blocks of 2^ind frames, ind <= max(blocknum - 1, q - 1)
    return 1 << std::max(ct.blocknum - 1, ct.q - 1);
}

static void launch_lm(const LMArgs &a, unsigned free_mask, hipStream_t st) {
    if (a.nfits <= 0) return;
    if (a.kind == kModelFCS && free_mask == 0x13u) {   // N, D, G_inf
        if (a.chi2_fp64 && a.npts > 128)
            hipLaunchKernelGGL((k_lm_fcs_2w<3, true>), dim3(a.nfits), dim3(128), 0, st, a);
        else if (a.chi2_fp64 && a.npts > 64)
            hipLaunchKernelGGL((k_lm_fcs_2w<3, false>), dim3(a.nfits), dim3(128), 0, st, a);
        else
            hipLaunchKernelGGL((k_lm_fixed<kModelFCS, 3>), dim3(a.nfits), dim3(64), 0, st, a);
        return;
    }
    if (a.kind == kModelPoly) {
        unsigned contiguous = (1u << a.ncoef) - 1u;
        if (free_mask == contiguous) {
            switch (a.ncoef) {
                case 1: hipLaunchKernelGGL((k_lm_fixed<kModelPoly, 1>), dim3(a.nfits), dim3(64), 0, st, a); return;
                case 2: hipLaunchKernelGGL((k_lm_fixed<kModelPoly, 2>), dim3(a.nfits), dim3(64), 0, st, a); return;
                case 3: hipLaunchKernelGGL((k_lm_fixed<kModelPoly, 3>), dim3(a.nfits), dim3(64), 0, st, a); return;
                case 4: hipLaunchKernelGGL((k_lm_fixed<kModelPoly, 4>), dim3(a.nfits), dim3(64), 0, st, a); return;
                case 5: hipLaunchKernelGGL((k_lm_fixed<kModelPoly, 5>), dim3(a.nfits), dim3(64), 0, st, a); return;
                case 6: hipLaunchKernelGGL((k_lm_fixed<kModelPoly, 6>), dim3(a.nfits), dim3(64), 0, st, a); return;
                case 7: hipLaunchKernelGGL((k_lm_fixed<kModelPoly, 7>), dim3(a.nfits), dim3(64), 0, st, a); return;
                default: break;
            }
        }
    }
    hipLaunchKernelGGL(k_lm, dim3(a.nfits), dim3(64), 0, st, a);
}

void launch_lm_gls(int nfits, int m, const float *x, const float *init, const int *pfit, float tol, int maxit,
                   const double *W, const double *yw, const int32_t *ok, float *out_params, int *states, float *chi2,
                   int *iters, hipStream_t st, int init_split) {
    LMArgs a{nfits, m, 20, kModelFCS, 0, nullptr, nullptr, x, init, pfit, LMModelGeom{}, tol, maxit, 1, out_params,
             states, chi2, iters, init_split >= 0 ? 1 : 0, init_split};
    const size_t lds = sizeof(double) * ((size_t)m * m + (size_t)m * 4);
    if (lds > 65536) (void)hipFuncSetAttribute((const void *)k_lm_gls, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)lds);
    hipLaunchKernelGGL(k_lm_gls, dim3(nfits), dim3(64), lds, st, a, W, yw, ok, m);
}

void launch_lm_acf(int nfits, int npts, const float *data, const float *weights, const float *x, const float *init,
                   const int *pfit, unsigned free_mask, const LMModelGeom &geom, float tol, int maxit,
                   int chi2_fp64, float *out_params, int *states, float *chi2, int *iters, hipStream_t st,
                   int init_split, int only_failed) {
    LMArgs a{nfits, npts, 20, kModelFCS, 0, data, weights, x, init, pfit, geom, tol, maxit, chi2_fp64, out_params,
             states, chi2, iters, init_split >= 0 ? 1 : 0, init_split, only_failed};
    launch_lm(a, free_mask, st);
}

void launch_lm_polish(int nfits, int npts, const float *data, const float *weights, const float *x,
                      float *out_params, int *states, float *chi2, hipStream_t st) {
    if (nfits <= 0 || npts <= 0) return;
    LMArgs a{nfits, npts, 20, kModelFCS, 0, data, weights, x, nullptr, nullptr, LMModelGeom{}, 0.f, 0, 1, out_params,
             states, chi2, nullptr, 0, 0, 0};
    if (npts > 128)
        hipLaunchKernelGGL(k_lm_polish<true>, dim3(nfits), dim3(128), 0, st, a);
    else
        hipLaunchKernelGGL(k_lm_polish<false>, dim3(nfits), dim3(128), 0, st, a);
}

void launch_lm_polish_groups(const LMGroup *groups, const int64_t *prefix, int ngroups, int64_t total,
                             hipStream_t st) {
    if (total <= 0 || ngroups <= 0) return;
    hipLaunchKernelGGL(k_lm_polish_groups, dim3((unsigned)total), dim3(128), 0, st, groups, prefix, ngroups, total);
}

void launch_lm_exp(int nfits, int npts, const float *data, const float *x, const float *init, const int *pfit,
                   int nexp, float tol, int maxit, double *out64, int *states, float *chi2, int *iters,
                   hipStream_t st) {
    LMModelGeom g{};
    LMArgs a{nfits, npts, 11, nexp == 1 ? kModelExp1 : kModelExp2, 0, data, nullptr, x, init, pfit, g, tol, maxit, 1,
             nullptr, states, chi2, iters, 0, 0};
    if (nfits <= 0) return;
    if (nexp == 1)
        hipLaunchKernelGGL(k_lm_exp64<1>, dim3(nfits), dim3(64), 0, st, a, out64);
    else
        hipLaunchKernelGGL(k_lm_exp64<2>, dim3(nfits), dim3(64), 0, st, a, out64);
}

void launch_lm_poly(int nfits, int npts, const float *data, const float *weights, const float *x, const float *init,
                    const int *pfit, unsigned free_mask, int ncoef, float tol, int maxit, int chi2_fp64,
                    float *out_params, int *states, float *chi2, int *iters, hipStream_t st) {
    LMModelGeom g{};
    LMArgs a{nfits, npts, 11, kModelPoly, ncoef, data, weights, x, init, pfit, g, tol, maxit, chi2_fp64, out_params,
             states, chi2, iters, 0, 0};
    launch_lm(a, free_mask, st);
}

void launch_binning_f32(const float *in, float *out, int win, int hin, int wt, int ht, int T, int bx, int by,
                        hipStream_t st) {
    dim3 grid(cdiv(wt, 64), (unsigned)((int64_t)ht * T));
    hipLaunchKernelGGL(k_binning_f32, grid, dim3(64), 0, st, in, out, win, hin, wt, ht, T, bx, by);
}

void launch_bleach_average_f32(const float *in, float *out, int wt, int ht, int nopit, int ave, hipStream_t st) {
    dim3 grid(cdiv((int64_t)wt * ht, 64), nopit);
    hipLaunchKernelGGL(k_bleach_average_f32, grid, dim3(64), 0, st, in, out, wt, ht, nopit, ave);
}

void launch_pack_fit_inputs(int chanum, int npairs, const double *corav, const double *bvar, float *data,
                            float *weights, hipStream_t st) {
    if (npairs <= 0) return;
    hipLaunchKernelGGL(k_pack_fit_inputs, dim3(npairs), dim3(128), 0, st, chanum, npairs, corav, bvar, data, weights);
}

// ------------------------------------------------------------------
// Training curves for the learned initial guesses: one thread per curve
// draws its parameters (log-uniform N and D, uniform F2 and G_inf N),
// evaluates FCS_3p at the lag times, adds Gaussian noise whose scale relative
// to 1/N is log-uniform and grows with the lag as sqrt(mtab_1 / mtab_i) (the
// number of products behind a channel), normalises by the largest magnitude
// and writes bf16 inputs and standardised targets.
// ------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_mlp_curves(CurveGenArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint64_t key = rkey(a.seed, (uint64_t)(a.base + i), 0, 11);
    const float lnN = a.lnN[0] + (a.lnN[1] - a.lnN[0]) * u01(key, 0);
    const float lnD = a.lnD[0] + (a.lnD[1] - a.lnD[0]) * u01(key, 1);
    const float N = expf(lnN), D = expf(lnD);
    float lnD2 = lnD, F2 = 0.f;
    if (a.ncomp == 2) {
        lnD2 = a.lnD2[0] + (a.lnD2[1] - a.lnD2[0]) * u01(key, 2);
        F2 = u01(key, 3);
    }
    const float G = a.ginf * (2.f * u01(key, 4) - 1.f) / N;
    const float noise = exp10f(a.lg_noise[0] + (a.lg_noise[1] - a.lg_noise[0]) * u01(key, 5)) / N;
    float pq[11] = {N, D, 0.f, 0.f, G, F2, expf(lnD2), 0.f, 0.f, 0.f, 0.f};
    float g[128];
    float s = 0.f;
    const int np = a.npts < 128 ? a.npts : 128;
    for (int k = 0; k < np; k++) {
        const float v = fcs3p(a.geom, a.x[k], pq, false, a.ncomp == 2, false, nullptr) +
                        noise * a.nw[k] * nrm(key, 8 + 2 * k);
        g[k] = v;
        s = fmaxf(s, fabsf(v));
    }
    const float is = 1.f / s;
    uint16_t *xr = a.X + (int64_t)i * a.ldx;
    for (int k = 0; k < a.ldx; k++) {
        const float v = k < np ? g[k] * is : 0.f;
        uint32_t u = __float_as_uint(v);
        u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
        xr[k] = (uint16_t)(u >> 16);
    }
    const float t[5] = {logf(N * s), lnD, lnD2, F2, G * is};
    for (int k = 0; k < 5; k++) a.Tg[(int64_t)i * 5 + k] = (t[k] - a.tmu[k]) / a.tsd[k];
}

void launch_mlp_curves(const CurveGenArgs &a, hipStream_t st) {
    if (a.n <= 0) return;
    hipLaunchKernelGGL(k_mlp_curves, dim3(cdiv(a.n, 64)), dim3(64), 0, st, a);
}

}  // namespace imfcs
