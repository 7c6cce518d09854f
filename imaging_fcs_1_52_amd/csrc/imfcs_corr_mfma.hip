// Levels 0..2 of the multi-tau correlator with blocked variance
// (calculateCF, IFCS:9425-9521) on the matrix cores.
//
// A correlation pair (series a, b) has one product form that a matrix unit
// can evaluate: the Hankel product of its 16-frame groups,
//   C_k[i][j] = sum_kappa a(16 kappa + i) * b(16 kappa + 16 k + j),
// i, j in [0, 16), whose entry (i, j) holds the lag L = 16 k + j - i
// products of the frames t = 16 kappa + i.  One v_mfma_f32_16x16x32_f16
// sums 32 groups (a 512-frame span) of one 16 x 16 tile; five tiles
// (k = 0..4) cover every lag up to 79 frames.  All three levels come out of
// these level-0 products, since a level-c value is the sum of its 2^c
// frames: the level-1 channel d1 (lag 2 d1 frames) is the sum of the
// entries with (L + (i & 1)) >> 1 == d1, L in [17, 33], and the level-2
// channel d2 the sum of the entries with (L + (i & 3)) >> 2 == d2,
// L >= 33 (each entry belongs to at most one channel).
//
// Precision: every staged value x (the bleach-corrected FP32 deviation,
// scaled by a per-slot power of two so that |x| < 2^14) is split into
// FP16 hi = fp16(x) and lo = fp16(x - hi) (|x - hi - lo| <= 2^-22 |x|), and
// a product is hi_a hi_b + hi_a lo_b + lo_a hi_b, accumulated in FP32 by the
// matrix unit (three products per tile).
//
// Workgroup: 8 slots of one row of the slot grid (8 output pixels) and
// their CCF partners at slot offset delta (0..4: the partner to the right),
// over one time chunk.  Sixteen waves, one correlation pair each (the ACF
// and the CCF of a pixel), keeping the pair's tile accumulators across
// spans.  The staged series are double-buffered in LDS as FP16 rows
// [series][i][kappa] (36 groups: the span and the 64 frames its lags
// reach); nine waves also load a 64-frame share of the next span while the
// current one is multiplied.
//
// Speed: on MI355X this kernel is slower than the VALU kernel k_corr_main
// (binning 1 of a 128 x 128 x 50000 stack: about 5.1 ms against 2.0 ms).
// The products take little time; the block ends (33 channel sums from 1280
// tile entries through LDS, the head and tail sums) and the per-span
// staging, prefix sums and barriers of a workgroup that holds only 8
// pixels cost the rest.  It is selected with IMFCS_CORRELATOR=mfma.
//
// Blocked variance: a pair's accumulators cover exactly one block of the
// blocked variance (2^ind frames; a block shorter than a span masks the
// A operand's groups to the block).  At a block end the tiles are reduced
// to the 33 channel sums through a per-wave LDS table (each channel has
// exactly 16 partial sums after in-lane pre-reduction), and lane c forms the
// block's mean-corrected sum
//   S = Sab - beta Sa - alpha (Sb + tail - head) + B alpha beta
// (Sa, Sb: block sums of a and b; head / tail: the sums of b over the lag's
// first frames after the block start / end).
#include "imfcs_device.h"
#include "imfcs_internal.h"

namespace imfcs {

typedef _Float16 mf_half8 __attribute__((ext_vector_type(8)));
typedef float mf_float4 __attribute__((ext_vector_type(4)));
typedef uint32_t mf_u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t mf_u32x2 __attribute__((ext_vector_type(2)));

constexpr int kMfSpanG = 32;              // 16-frame groups per span (K of one product)
constexpr int kMfSpanF = 16 * kMfSpanG;   // frames per span
constexpr int kMfStageG = kMfSpanG + 4;   // staged groups: the span and 64 frames of lags
constexpr int kMfRowH = 40;               // halves per staged row (80 bytes: aligned 16-byte reads)
constexpr int kMfTiles = 5;               // product tiles: lags up to 79 frames
constexpr int kMfNP = 8;                  // slots (pixels) per workgroup
constexpr int kMfNW = 2 * kMfNP;          // waves per workgroup: one correlation pair each
constexpr int kMfUnitG = 4;               // 16-frame groups per loader unit
constexpr int kMfUnits = kMfStageG / kMfUnitG;   // loader units per span (9)
constexpr int kMfSerH = 16 * kMfRowH + 8; // halves per staged series (padded: stores spread over the banks)
constexpr int kMfMaxDelta = 4;            // partner slot offset handled
constexpr int kMfChan = 33;               // channels of levels 0..2 (p = 16)
constexpr int kMfCs = 20;                 // channel-table row stride in floats (80 bytes: rows spread over the banks)

struct MfArgs {
    SeriesGeom g;
    BleachDesc bd;
    const CorrTables *ctp;
    PairDesc pd;
    const double *F;
    const float *X;
    const int32_t *ind;
    const int32_t *xexp;   // per slot: values are staged as x * 2^-xexp
    int chunk_len, nchunk;
    int has_acf, has_ccf;
    int delta;   // partner slot = pixel slot + delta (0 without partners)
    int nxb;     // workgroups per slot-grid row
    int ns;      // staged series per workgroup (kMfNP + delta)
    float *part;   // [chunk][pair][33][3]: cross, A1, A2
    float *gls;    // [pair][block][channel - 1] or null
    int gls_nblk;
};

// LDS layout of one workgroup (ns staged series).
struct MfLds {
    uint16_t *hi, *lo;   // [2][ns][kMfSerH]: rows [16][kMfRowH]
    float *gs;           // [ns][kMfSpanG]: 16-frame sums of the span being staged
    float *ps;           // [ns][kMfSpanG + 1]: exclusive prefix sums of the current span's
    float *scr;          // [kMfNW][33 x kMfCs + 64]: channel partial sums, a trash word per lane
    uint16_t *xmap;      // [64][24]: per lane, the channel-table word of each tile entry (20 used)
    int ns;
    __device__ MfLds(char *base, int ns_) : ns(ns_) {
        const size_t rows = (size_t)2 * ns * kMfSerH;
        hi = (uint16_t *)base;
        lo = hi + rows;
        gs = (float *)(lo + rows);
        ps = gs + (size_t)ns * kMfSpanG;
        scr = ps + (size_t)ns * (kMfSpanG + 1);
        xmap = (uint16_t *)(scr + (size_t)kMfNW * (kMfChan * kMfCs + 64));
    }
    __device__ uint16_t *hrow(int buf, int s, int j) const { return hi + ((size_t)buf * ns + s) * kMfSerH + j * kMfRowH; }
    __device__ uint16_t *lrow(int buf, int s, int j) const { return lo + ((size_t)buf * ns + s) * kMfSerH + j * kMfRowH; }
};

constexpr int corr_mfma_lds_bytes(int ns) {
    return 2 * 2 * ns * kMfSerH * 2 + ns * kMfSpanG * 4 + ns * (kMfSpanG + 1) * 4 + kMfNW * (kMfChan * kMfCs + 64) * 4 +
           64 * 24 * 2;
}
// static LDS for the largest partner offset (104 KB)
constexpr int kMfLdsMax = corr_mfma_lds_bytes(kMfNP + kMfMaxDelta);

__device__ __forceinline__ float pow2f(int e) { return __int_as_float((127 + e) << 23); }   // |e| < 127

// Quad (4-lane) sums with DPP: lanes 4s + r, r = 0..3, all receive the sum.
__device__ __forceinline__ float quad_sum(float x) {
    x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, false));   // [1,0,3,2]
    x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x4E, 0xF, 0xF, false));   // [2,3,0,1]
    return x;
}

// Inclusive prefix sum over the 64 lanes of a wave with DPP (all lanes
// active): row_shr 1, 2, 4, 8 inside each 16-lane row, then row_bcast:15
// (each row's total into rows 1 and 3) and row_bcast:31 (lanes 0..31's total
// into rows 2 and 3).  Lanes without a source lane add 0.
template <int CTRL, int ROWS>
__device__ __forceinline__ float mf_dpp(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xf, false));
}
__device__ __forceinline__ float wave_incl_scan(float v) {
    v += mf_dpp<0x111, 0xf>(v);   // row_shr:1
    v += mf_dpp<0x112, 0xf>(v);   // row_shr:2
    v += mf_dpp<0x114, 0xf>(v);   // row_shr:4
    v += mf_dpp<0x118, 0xf>(v);   // row_shr:8
    v += mf_dpp<0x142, 0xa>(v);   // row_bcast:15
    v += mf_dpp<0x143, 0xc>(v);   // row_bcast:31
    return v;
}

// Workgroup geometry: row y of the slot grid, slots s0 .. s0 + ns - 1.
// Workgroups go to the 8 XCDs round-robin by id, so id w runs on XCD w % 8:
// XCD x takes the rows y = 8 i + x, and its consecutive workgroups take
// consecutive blocks of one row, which share the 128-byte lines of each
// frame row in that XCD's L2.  Ids of rows past the grid exit.
struct MfGeo {
    int y, xb, s0, chunk, cstart, cend;
    __device__ MfGeo(const MfArgs &A) {
        const int w = (int)blockIdx.x, k = w >> 3;
        const int ri = k / A.nxb;
        y = 8 * ri + (w & 7);
        xb = k - ri * A.nxb;
        s0 = y * A.g.swp + xb * kMfNP;
        chunk = (int)blockIdx.y;
        cstart = chunk * A.chunk_len;
        cend = min(A.g.T, cstart + A.chunk_len);
    }
    // staged series si is read by some pair of this workgroup
    __device__ bool needed(const MfArgs &A, int si) const {
        const int x = xb * kMfNP + si;
        const bool a = si < kMfNP && x < A.g.ow;
        const bool b = A.has_ccf && si >= A.delta && si - A.delta < kMfNP && x - A.delta < A.g.ow;
        return a || b;
    }
};

// One loader unit: staged series lane >> 2, frames F0 + 64 u + 4 k + (lane & 3),
// k < 16 (groups 4u .. 4u + 3 of the staged span).  Buffer loads: the
// resource ends at the series' last frame and unused series get an offset
// past it, so those read 0 with no branch.
__device__ __forceinline__ void unit_load(const MfArgs &A, const MfGeo &G, int F0, int u, int lane, float (&v)[16]) {
    const int si = lane >> 2, r = lane & 3;
    const bool use = si < A.ns && G.needed(A, si);
    const int ld4 = A.g.nslots * 4;
    const int nrow = max(0, min(A.g.T - F0, kMfStageG * 16));   // frames of the resource
    // (the descriptor is wave-uniform; readfirstlane keeps it in scalar
    // registers, a per-lane descriptor would cost a loop per load)
    const uint64_t base = (uint64_t)(A.X + (int64_t)F0 * A.g.nslots);
    const uint64_t ub = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(base >> 32)) << 32) |
                        (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)base);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)ub, (short)0, __builtin_amdgcn_readfirstlane(nrow * ld4), kBufferWord3);
    // the whole offset in the lane offset, which the range check covers
    // (at most 0x7ffffff0 + 60 * 932k * 4 < 2^32: no wrap)
    const uint32_t voff = use ? (uint32_t)((64 * u + r) * ld4 + (G.s0 + si) * 4) : 0x7ffffff0u;
#pragma unroll
    for (int k = 0; k < 16; k++)
        v[k] = __builtin_bit_cast(float,
                                  __builtin_amdgcn_raw_buffer_load_b32(rs, (int)(voff + (uint32_t)(4 * k * ld4)), 0, 0));
}

// Stores a loaded unit into LDS buffer buf as FP16 hi/lo rows, and (for the
// span's own groups) its 16-frame sums.
__device__ __forceinline__ void unit_store(const MfArgs &A, const MfGeo &G, const MfLds &L, int buf, int u, int lane,
                                           const float (&v)[16]) {
    typedef _Float16 half4 __attribute__((ext_vector_type(4)));
    const int si = lane >> 2, r = lane & 3;
    const bool inb = si < A.ns;
    const float sc = pow2f(-((inb && G.needed(A, si)) ? A.xexp[G.s0 + si] : 0));
    // lane (si, r) holds the frames 16 kappa + 4 jq + r of the unit's groups
#pragma unroll
    for (int jq = 0; jq < 4; jq++) {
        const int j = 4 * jq + r;
        half4 hv, lv;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float x = v[4 * q + jq] * sc;
            hv[q] = (_Float16)x;
            lv[q] = (_Float16)(x - (float)hv[q]);
        }
        if (inb) {
            *(half4 *)(L.hrow(buf, si, j) + kMfUnitG * u) = hv;
            *(half4 *)(L.lrow(buf, si, j) + kMfUnitG * u) = lv;
        }
    }
    if (u < kMfSpanG / kMfUnitG) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float gsum = quad_sum((v[4 * q] + v[4 * q + 1]) + (v[4 * q + 2] + v[4 * q + 3]));
            if (inb && r == 0) L.gs[si * kMfSpanG + kMfUnitG * u + q] = gsum;
        }
    }
}

// A 12-element window (groups 8 g .. 8 g + 11 of row j, as 6 dwords of two
// halves) of one staged series' hi or lo plane.
struct MfWin {
    uint32_t w[6];
    __device__ void read(const uint16_t *row, int g) {
        const mf_u32x4 a = *(const mf_u32x4 *)(row + 8 * g);
        const mf_u32x2 b = *(const mf_u32x2 *)(row + 8 * g + 8);
        w[0] = a[0];
        w[1] = a[1];
        w[2] = a[2];
        w[3] = a[3];
        w[4] = b[0];
        w[5] = b[1];
    }
    // elements k .. k + 7 as an operand
    template <int K>
    __device__ __forceinline__ mf_half8 op() const {
        mf_u32x4 o;
#pragma unroll
        for (int d = 0; d < 4; d++) {
            if (K % 2 == 0)
                o[d] = w[K / 2 + d];
            else
                o[d] = __builtin_amdgcn_alignbit(w[(K + 1) / 2 + d], w[(K - 1) / 2 + d], 16);
        }
        return __builtin_bit_cast(mf_half8, o);
    }
};

// Where each of a lane's 20 tile entries (tile k, row 4 (lane >> 4) + r,
// column lane & 15) goes in the wave's channel table (33 channels x 16
// partial sums in rows of kMfCs floats, then one trash word per lane): level-2 rows are summed in
// the lane first (four rows of one channel), level-1 row pairs likewise,
// level-0 entries go alone.  Written once per workgroup into LDS as
// xmap[lane][4 k + r] (words 20..22: the level flags of the five tiles).
__device__ void mf_xmap_init(uint16_t *xm, int lane) {
    const int j = lane & 15, rg = lane >> 4, trash = kMfChan * kMfCs + lane;
    uint32_t sel = 0;
#pragma unroll
    for (int k = 0; k < kMfTiles; k++) {
        const int base = 16 * k + j - 4 * rg;
        const bool l2 = base >= 36 && base <= 67, none = base >= 36 && !l2;
        const int b0 = base, b1 = base - 2;
        const bool l1p0 = !l2 && !none && b0 >= 18 && b0 <= 33;
        const bool l1p1 = !l2 && !none && b1 >= 18 && b1 <= 33;
        sel |= (l2 ? 1u : 0u) << (3 * k) | (l1p0 ? 1u : 0u) << (3 * k + 1) | (l1p1 ? 1u : 0u) << (3 * k + 2);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int Lg = base - r;   // lag of row r
            int o = (!l2 && !none && Lg >= 0 && Lg <= 16) ? Lg * kMfCs + 4 * rg + r : trash;
            if (r < 2 && l1p0) o = r == 0 ? (8 + (b0 >> 1)) * kMfCs + 4 * rg + (b0 & 1) : trash;
            if (r >= 2 && l1p1) o = r == 2 ? (8 + (b1 >> 1)) * kMfCs + 4 * rg + 2 + (b1 & 1) : trash;
            if (l2) o = r == 0 ? (16 + (base >> 2)) * kMfCs + 4 * rg + (base & 3) : trash;
            if (none) o = trash;
            xm[lane * 24 + 4 * k + r] = (uint16_t)o;
        }
    }
    xm[lane * 24 + 20] = (uint16_t)(sel & 0xffff);
    xm[lane * 24 + 21] = (uint16_t)(sel >> 16);
}
constexpr int kMfScr = kMfChan * kMfCs + 64;   // floats of a wave's channel table

// Channel sums (lane c < 33) of a pair's tiles.
__device__ __forceinline__ float mf_extract(const mf_float4 (&acc)[kMfTiles], const uint16_t *xm, float *scr,
                                            int lane) {
    // the whole map first: one wait for the reads, none between the writes
    const mf_u32x4 w0 = *(const mf_u32x4 *)(xm + lane * 24), w1 = *(const mf_u32x4 *)(xm + lane * 24 + 8);
    const mf_u32x4 w2 = *(const mf_u32x4 *)(xm + lane * 24 + 16);
    const uint32_t wd[12] = {w0[0], w0[1], w0[2], w0[3], w1[0], w1[1], w1[2], w1[3], w2[0], w2[1], w2[2], w2[3]};
    const uint32_t sel = wd[10];
#pragma unroll
    for (int k = 0; k < kMfTiles; k++) {
        const mf_float4 a = acc[k];
        const float p0 = a[0] + a[1], p1 = a[2] + a[3];
        const bool l2 = (sel >> (3 * k)) & 1, l1a = (sel >> (3 * k + 1)) & 1, l1b = (sel >> (3 * k + 2)) & 1;
        const uint32_t oa = wd[2 * k], ob = wd[2 * k + 1];
        scr[oa & 0xffff] = l2 ? p0 + p1 : (l1a ? p0 : a[0]);
        scr[oa >> 16] = a[1];
        scr[ob & 0xffff] = l1b ? p1 : a[2];
        scr[ob >> 16] = a[3];
    }
    const int c = lane < kMfChan ? lane : 0;
    const mf_float4 *t = (const mf_float4 *)(scr + c * kMfCs);
    const mf_float4 x0 = t[0], x1 = t[1], x2 = t[2], x3 = t[3];
    const float s0 = (x0[0] + x0[1]) + (x0[2] + x0[3]);
    const float s1 = (x1[0] + x1[1]) + (x1[2] + x1[3]);
    const float s2 = (x2[0] + x2[1]) + (x2[2] + x2[3]);
    const float s3 = (x3[0] + x3[1]) + (x3[2] + x3[3]);
    return (s0 + s1) + (s2 + s3);
}

// Lag in frames and level of channel c (levels 0..2, p = 16).
__device__ __forceinline__ int mf_lag(int c) { return c <= 16 ? c : c <= 24 ? 2 * (c - 8) : 4 * (c - 16); }
__device__ __forceinline__ int mf_level(int c) { return c <= 16 ? 0 : c <= 24 ? 1 : 2; }

// Sum of series si's first m frames from frame offset o of the staged span,
// m = the lag of lane c's channel.
__device__ __forceinline__ float mf_head(const MfLds &L, int buf, int si, int e, int o, int lane) {
    const int t = o + lane;   // < 512 + 64: staged
    const int j = t & 15, kap = t >> 4;
    const _Float16 h = __builtin_bit_cast(_Float16, L.hrow(buf, si, j)[kap]);
    const _Float16 l = __builtin_bit_cast(_Float16, L.lrow(buf, si, j)[kap]);
    const float incl = wave_incl_scan(((float)h + (float)l) * pow2f(e));
    const int m = lane < kMfChan ? mf_lag(lane) : 0;
    const float hm = __shfl(incl, m > 0 ? m - 1 : 0, 64);
    return m > 0 ? hm : 0.f;
}

struct MfPair {
    mf_float4 acc[kMfTiles];                   // tiles of the open block
    float alpha, beta, a1, a2, cross, head;    // lane c < 33: channel c
    float Sa, Sb;                              // block sums of a and b
    int pair, lb, pnum, ea, eb;
    bool live, pending;
};

// The products of one segment [kap0, kap1) of the span added to P's tiles.
__device__ __forceinline__ void mf_products(MfPair &P, const MfWin &ah, const MfWin &al, const MfWin &bh,
                                            const MfWin &bl, int kap0, int kap1, int lane) {
    mf_half8 Ah = ah.op<0>(), Al = al.op<0>();
    if (kap0 > 0 || kap1 < kMfSpanG) {   // wave-uniform: mask the A groups outside the segment
        const int g = lane >> 4;
        mf_u32x4 mh = __builtin_bit_cast(mf_u32x4, Ah), ml = __builtin_bit_cast(mf_u32x4, Al);
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const int k0 = 8 * g + 2 * d;
            const uint32_t m = ((k0 >= kap0 && k0 < kap1) ? 0x0000FFFFu : 0u) |
                               ((k0 + 1 >= kap0 && k0 + 1 < kap1) ? 0xFFFF0000u : 0u);
            mh[d] &= m;
            ml[d] &= m;
        }
        Ah = __builtin_bit_cast(mf_half8, mh);
        Al = __builtin_bit_cast(mf_half8, ml);
    }
#define MF_TILE(K)                                                                            \
    {                                                                                         \
        const mf_half8 Bh = bh.op<K>(), Bl = bl.op<K>();                                      \
        P.acc[K] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah, Bh, P.acc[K], 0, 0, 0);         \
        P.acc[K] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al, Bh, P.acc[K], 0, 0, 0);         \
        P.acc[K] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah, Bl, P.acc[K], 0, 0, 0);         \
        __builtin_amdgcn_sched_barrier(0);                                                    \
    }
    MF_TILE(0)
    MF_TILE(1)
    MF_TILE(2)
    MF_TILE(3)
    MF_TILE(4)
#undef MF_TILE
    P.pending = true;
}

// Block end at frame offset o of the span (absolute frame vabs), or the
// chunk's last partial block (close = false): channel sums into cross, and
// for a block its S into A1, A2 (and the GLS table).
__device__ __forceinline__ void mf_close(const MfArgs &A, const MfLds &L, MfPair &P, float *scr, int ib, int buf,
                                         int o, int vabs, bool close, int lane) {
    const float Sab = mf_extract(P.acc, L.xmap, scr, lane) * pow2f(P.ea + P.eb);
    P.cross += Sab;
    if (close) {
        const float tail = mf_head(L, buf, ib, P.eb, o, lane);
        if (lane < kMfChan) {
            const int c = lane, lv = mf_level(c);
            const float B = (float)(1 << (P.lb - lv));
            const float S = Sab - P.beta * P.Sa - P.alpha * (P.Sb + (tail - P.head)) + B * P.alpha * P.beta;
            const int blk = (vabs >> P.lb) - 1;
            if (blk < P.pnum) {
                P.a1 += S;
                P.a2 = fmaf(S, S, P.a2);
                if (A.gls && c > 0) {
                    const int ld = A.ctp->chanum - 1;
                    A.gls[((int64_t)P.pair * A.gls_nblk + blk) * ld + c - 1] = S;
                }
            }
        }
        P.head = tail;
        P.Sa = P.Sb = 0.f;
    }
#pragma unroll
    for (int k = 0; k < kMfTiles; k++) P.acc[k] = mf_float4{0.f, 0.f, 0.f, 0.f};
    P.pending = false;
}

// Level-1 / level-2 products whose b frame lies past the last complete
// level-1 (2 numbin_1) or level-2 (4 numbin_2) value: the level-0 products
// count them, the level's own series has no such value.  Removed from the
// channel sum of the chunk that holds their a frame (they lie in the last
// 72 frames, far past the blocks of the blocked variance).
__device__ __forceinline__ float mf_tail_excess(const MfArgs &A, int slot_a, int slot_b, int cstart, int cend,
                                                int lane) {
    const CorrTables &ct = *A.ctp;
    const int T = A.g.T;
    if (lane < 17 || lane >= kMfChan) return 0.f;
    const int c = lane, lv = mf_level(c);
    const int d = lv == 1 ? c - 8 : c - 16;
    const int tlim = lv == 1 ? 2 * ct.numbin[1] : 4 * ct.numbin[2];
    const int64_t ld = A.g.nslots;
    float ex = 0.f;
    for (int tb = tlim; tb < T; tb++) {
        const float xb = A.X[(int64_t)tb * ld + slot_b];
        for (int Lg = (d << lv) - (1 << lv) + 1; Lg <= (d << lv) + (1 << lv) - 1; Lg++) {
            const int ta = tb - Lg;
            if (ta < cstart || ta >= cend) continue;
            if (((Lg + (ta & ((1 << lv) - 1))) >> lv) != d) continue;
            ex = fmaf(A.X[(int64_t)ta * ld + slot_a], xb, ex);
        }
    }
    return ex;
}

__global__ void __launch_bounds__(kMfNW * 64) k_corr_mfma(MfArgs A) {
    __shared__ __attribute__((aligned(16))) char mf_smem[kMfLdsMax];
    const MfLds L(mf_smem, A.ns);
    const MfGeo G(A);
    if (G.y >= A.g.oh) return;   // workgroup-uniform
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int lane = (int)threadIdx.x & 63;
    const CorrTables &ct = *A.ctp;
    float *scr = L.scr + wave * kMfScr;
    if (wave == 0) mf_xmap_init(L.xmap, lane);

    // this wave's pair: pixel ia = wave / 2 of the workgroup, its ACF (even
    // waves) or its CCF (odd waves: b series staged delta further)
    const int ia = wave >> 1, kd = wave & 1;
    const int ib = ia + (kd ? A.delta : 0);
    const int x = G.xb * kMfNP + ia;
    MfPair Q;
    {
        const int p = G.y * A.g.ow + x;
        Q.live = x < A.g.ow && (kd == 0 ? A.has_acf : A.has_ccf);
        Q.pair = kd == 0 ? p : (A.has_acf ? A.g.npix : 0) + p;
        const int slot_a = G.s0 + ia, slot_b = G.s0 + ib;
        Q.ea = Q.live ? A.xexp[slot_a] : 0;
        Q.eb = Q.live ? A.xexp[slot_b] : 0;
        const int indp = Q.live ? A.ind[Q.pair] : ct.q - 1;
        Q.lb = indp;
        Q.pnum = (int)floor(ct.mtab_last / pow(2.0, fmax(indp - ct.log2_samp_last, 0.0)));
        Q.alpha = Q.beta = 0.f;
        if (Q.live && lane < kMfChan) {
            const int c = lane, lv = ct.level[c];
            const double *Fa = A.F + (int64_t)slot_a * ct.npos, *Fb = A.F + (int64_t)slot_b * ct.npos;
            const double mua = slot_mu(A.bd.coef, A.bd.order, slot_a) * (double)(1 << lv);
            const double mub = slot_mu(A.bd.coef, A.bd.order, slot_b) * (double)(1 << lv);
            const double n = (double)(ct.numbin[lv] - ct.del[c]);
            Q.alpha = (float)(Fa[ct.pos_a_end[c]] / n - mua);
            Q.beta = (float)((Fb[ct.pos_b_end[c]] - Fb[ct.pos_b_beg[c]]) / n - mub);
        }
        Q.a1 = Q.a2 = Q.cross = Q.head = 0.f;
        Q.Sa = Q.Sb = 0.f;
#pragma unroll
        for (int k = 0; k < kMfTiles; k++) Q.acc[k] = mf_float4{0.f, 0.f, 0.f, 0.f};
        Q.pending = false;
    }

    // loader: wave u < kMfUnits stages groups 4u .. 4u + 3 of every span
    const bool loader = wave < kMfUnits;
    const int nspan = (G.cend - G.cstart + kMfSpanF - 1) / kMfSpanF;
    float v[16];
    if (loader) {
        unit_load(A, G, G.cstart, wave, lane, v);
        unit_store(A, G, L, 0, wave, lane, v);
    }
    __syncthreads();
    if (Q.live) Q.head = mf_head(L, 0, ib, Q.eb, 0, lane);   // sums of b from the chunk start
    for (int s = 0; s < nspan; s++) {
        const int cur = s & 1;
        const int F0 = G.cstart + s * kMfSpanF;
        // exclusive prefix sums of the span's group sums (wave si: series si)
        if (wave < A.ns) {
            const int si = wave;
            const float incl = wave_incl_scan(lane < kMfSpanG ? L.gs[si * kMfSpanG + lane] : 0.f);
            if (lane < kMfSpanG) L.ps[si * (kMfSpanG + 1) + lane + 1] = incl;
            if (lane == 0) L.ps[si * (kMfSpanG + 1)] = 0.f;
        }
        const bool more = s + 1 < nspan;
        if (more && loader) unit_load(A, G, F0 + kMfSpanF, wave, lane, v);
        __syncthreads();
        if (Q.live) {   // wave-uniform
            const int vend = min(F0 + kMfSpanF, G.cend);
            const int j = lane & 15, g = lane >> 4;
            int u = F0;
            while (u < vend) {
                const int vv = min(vend, ((u >> Q.lb) + 1) << Q.lb);
                const int k0 = (u - F0) >> 4, k1 = (vv - F0 + 15) >> 4;
                {
                    // operand windows, read per segment so that they are dead
                    // across the block ends
                    MfWin ah, al, bh, bl;
                    ah.read(L.hrow(cur, ia, j), g);
                    al.read(L.lrow(cur, ia, j), g);
                    bh.read(L.hrow(cur, ib, j), g);
                    bl.read(L.lrow(cur, ib, j), g);
                    mf_products(Q, ah, al, bh, bl, k0, k1, lane);
                }
                Q.Sa += L.ps[ia * (kMfSpanG + 1) + k1] - L.ps[ia * (kMfSpanG + 1) + k0];
                Q.Sb += L.ps[ib * (kMfSpanG + 1) + k1] - L.ps[ib * (kMfSpanG + 1) + k0];
                if ((vv & ((1 << Q.lb) - 1)) == 0) mf_close(A, L, Q, scr, ib, cur, vv - F0, vv, true, lane);
                u = vv;
            }
        }
        if (more && loader) unit_store(A, G, L, cur ^ 1, wave, lane, v);
        __syncthreads();
    }
    // the last chunk's partial block; products past the level-1/2 series
    // ends; per-chunk sums out
    if (Q.live) {
        if (Q.pending) mf_close(A, L, Q, scr, ib, 0, 0, 0, false, lane);
        if (G.cend + 80 > A.g.T) Q.cross -= mf_tail_excess(A, G.s0 + ia, G.s0 + ib, G.cstart, G.cend, lane);
        if (lane < kMfChan) {
            float *out = A.part + (((int64_t)G.chunk * A.pd.npairs + Q.pair) * kMfChan + lane) * 3;
            out[0] = Q.cross;
            out[1] = Q.a1;
            out[2] = Q.a2;
        }
    }
}

bool corr_mfma_eligible(const SeriesGeom &g, const PairDesc &pd) {
    if (!pd.has_ccf) return true;
    const bool on_grid = (g.cfx % g.pbx == 0) && (g.cfy % g.pby == 0);
    if (!on_grid || g.cfy != 0) return false;
    const int dx = g.cfx / g.pbx;
    return dx >= 0 && dx <= kMfMaxDelta;
}

void launch_corr_mfma(const SeriesGeom &g, const BleachDesc &bd, const CorrTables *dct, const PairDesc &pd,
                      const double *F, const float *X, const int32_t *ind, const int32_t *xexp, int chunk_len,
                      int nchunk, float *part, float *gls, int gls_nblk, hipStream_t st) {
    MfArgs A;
    A.g = g;
    A.bd = bd;
    A.ctp = dct;
    A.pd = pd;
    A.F = F;
    A.X = X;
    A.ind = ind;
    A.xexp = xexp;
    A.chunk_len = chunk_len;
    A.nchunk = nchunk;
    A.has_acf = pd.acf_first;
    A.has_ccf = pd.has_ccf;
    A.delta = pd.has_ccf ? g.cfx / g.pbx : 0;
    A.nxb = (g.ow + kMfNP - 1) / kMfNP;
    A.ns = kMfNP + A.delta;
    A.part = part;
    A.gls = gls;
    A.gls_nblk = gls_nblk;
    hipLaunchKernelGGL(k_corr_mfma, dim3(8 * ((g.oh + 7) / 8) * A.nxb, nchunk), dim3(kMfNW * 64), 0, st, A);
}

}  // namespace imfcs
