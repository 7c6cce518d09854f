// Device helpers shared by the HIP translation units: series access and the
// on-the-fly polynomial bleach correction.
#pragma once
#include "imfcs_internal.h"

namespace imfcs {

// Third word of a raw (unswizzled, 32-bit) buffer resource on gfx9 parts:
// the buffer loads and stores below take one lane offset and one scalar
// (wave-uniform) offset, instead of a 64-bit address per lane.
constexpr int kBufferWord3 = 0x00020000;

// ------------------------------------------------------------------
// Series access
// ------------------------------------------------------------------
__device__ __forceinline__ void series_origin(const SeriesGeom &g, int s, int &ox, int &oy) {
    int k = s < g.npix ? s : s - g.npix;
    int x = k % g.ow, y = k / g.ow;
    ox = g.ox0 + x * g.pbx;
    oy = g.oy0 + y * g.pby;
    if (s >= g.npix) {
        ox += g.cfx;
        oy += g.cfy;
    }
}

__device__ __forceinline__ void slot_origin(const SeriesGeom &g, int slot, int &ox, int &oy) {
    ox = g.slot_org[2 * slot];
    oy = g.slot_org[2 * slot + 1];
}

// Exact integer window sum minus background (IFCS:9668-9678).
// Rows of a window that raw_ints16 reads in one batch, else 0: two-row
// windows (four-row windows measured no faster batched: their passes run
// near the memory bandwidth either way).
__host__ __device__ inline int raw_batch_rows(const SeriesGeom &g) {
    const bool vec = g.src == kSrcU16 && ((g.vec == 4 && g.bx == 4) || (g.vec == 2 && g.bx == 2));
    return (vec && g.by == 2) ? 2 : 0;
}

__device__ __forceinline__ int raw_u16(const SeriesGeom &g, int ox, int oy, int f) {
    const uint16_t *base = (const uint16_t *)g.data + (int64_t)(g.f0 + f) * g.plane +
                           (int64_t)oy * g.stride + ox;
    if (g.bx == 1 && g.by == 1) return (int)base[0] - g.bckg_total;
    int acc = 0;
    for (int k = 0; k < g.by; k++)
        for (int i = 0; i < g.bx; i++) acc += base[(int64_t)k * g.stride + i];
    return acc - g.bckg_total;
}

// Window sums (uint16 source, minus background) of 16 consecutive frames
// f0 .. f0+15; frames past fmax repeat frame fmax.  The frame base address
// is wave-uniform (f0 is), so each load needs one 32-bit lane offset.  With
// aligned rows of 2 or 4 pixels and windows of NB = 2 or 4 rows, the loads
// of all rows are issued before any is summed (one round trip per group,
// at the cost of 32..128 registers: separate instantiations, NB =
// raw_batch_rows(g)); with NB = 0 the window loops are outermost so the 16
// loads of a window element are in flight together.
template <int NB>
__device__ __forceinline__ void raw_ints16(const SeriesGeom &g, int ox, int oy, int f0, int fmax, int (&acc)[16]) {
    const int nv = min(16, fmax - f0 + 1);
    const uint32_t pl = (uint32_t)g.plane;
    const uint32_t lo = (uint32_t)(oy * g.stride + ox);
    const uint16_t *q = (const uint16_t *)g.data + (int64_t)(g.f0 + f0) * g.plane;
#pragma unroll
    for (int j = 0; j < 16; j++) acc[j] = -g.bckg_total;
    if (NB > 0 && g.vec == 4) {
        const uint2 *q4 = (const uint2 *)q;
        uint2 r[NB > 0 ? NB : 1][16];
#pragma unroll
        for (int k = 0; k < NB; k++) {
            const uint32_t d = (lo + (uint32_t)(k * g.stride)) >> 2;
#pragma unroll
            for (int j = 0; j < 16; j++) r[k][j] = q4[d + (uint32_t)min(j, nv - 1) * (pl >> 2)];
        }
#pragma unroll
        for (int k = 0; k < NB; k++)
#pragma unroll
            for (int j = 0; j < 16; j++)
                acc[j] += (int)((r[k][j].x & 0xffffu) + (r[k][j].x >> 16) + (r[k][j].y & 0xffffu) + (r[k][j].y >> 16));
        return;
    }
    if (NB > 0) {   // g.vec == 2
        const uint32_t *q2 = (const uint32_t *)q;
        uint32_t r[NB > 0 ? NB : 1][16];
#pragma unroll
        for (int k = 0; k < NB; k++) {
            const uint32_t d = (lo + (uint32_t)(k * g.stride)) >> 1;
#pragma unroll
            for (int j = 0; j < 16; j++) r[k][j] = q2[d + (uint32_t)min(j, nv - 1) * (pl >> 1)];
        }
#pragma unroll
        for (int k = 0; k < NB; k++)
#pragma unroll
            for (int j = 0; j < 16; j++) acc[j] += (int)((r[k][j] & 0xffffu) + (r[k][j] >> 16));
        return;
    }
    if (g.vec == 4 && g.bx == 4) {
        // aligned rows of 4 pixels: one 8-byte load per row and frame
        const uint2 *q4 = (const uint2 *)q;
        for (int k = 0; k < g.by; k++) {
            const uint32_t d = (lo + (uint32_t)(k * g.stride)) >> 2;
            uint2 r[16];
#pragma unroll
            for (int j = 0; j < 16; j++) r[j] = q4[d + (uint32_t)min(j, nv - 1) * (pl >> 2)];
#pragma unroll
            for (int j = 0; j < 16; j++)
                acc[j] += (int)((r[j].x & 0xffffu) + (r[j].x >> 16) + (r[j].y & 0xffffu) + (r[j].y >> 16));
        }
        return;
    }
    if (g.vec == 2 && g.bx == 2) {
        // aligned pixel pairs: one 4-byte load per row and frame
        const uint32_t *q2 = (const uint32_t *)q;
        for (int k = 0; k < g.by; k++) {
            const uint32_t d = (lo + (uint32_t)(k * g.stride)) >> 1;
            uint32_t r[16];
#pragma unroll
            for (int j = 0; j < 16; j++) r[j] = q2[d + (uint32_t)min(j, nv - 1) * (pl >> 1)];
#pragma unroll
            for (int j = 0; j < 16; j++) acc[j] += (int)((r[j] & 0xffffu) + (r[j] >> 16));
        }
        return;
    }
    if ((int64_t)pl * 32 < 0x7fffffff) {   // wave-uniform
        // buffer loads: the 16 frames differ by a scalar offset, so a load
        // needs no address arithmetic of its own
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc((void *)q, (short)0, (int)(pl * 32), kBufferWord3);
        for (int k = 0; k < g.by; k++)
            for (int i = 0; i < g.bx; i++) {
                const int d2 = (int)(lo + (uint32_t)(k * g.stride + i)) * 2;
                uint16_t r[16];
#pragma unroll
                for (int j = 0; j < 16; j++)
                    r[j] = (uint16_t)__builtin_amdgcn_raw_buffer_load_b16(rs, d2, min(j, nv - 1) * (int)pl * 2, 0);
#pragma unroll
                for (int j = 0; j < 16; j++) acc[j] += r[j];
            }
        return;
    }
    for (int k = 0; k < g.by; k++)
        for (int i = 0; i < g.bx; i++) {
            const uint32_t d = lo + (uint32_t)(k * g.stride + i);
            uint16_t r[16];
#pragma unroll
            for (int j = 0; j < 16; j++) r[j] = q[d + (uint32_t)min(j, nv - 1) * pl];
#pragma unroll
            for (int j = 0; j < 16; j++) acc[j] += r[j];
        }
}

// The 16 values of raw_ints16 (or of a float source) in two steps: as
// 32-bit words (integer window sums, or FP32 values), so that a caller can
// hold the next group's loads in flight while it works on the current
// group, and their conversion to FP64.
template <int NB>
__device__ __forceinline__ void raw_words16(const SeriesGeom &g, int ox, int oy, int f0, int fmax, uint32_t (&r)[16]) {
    if (g.src == kSrcU16) {
        int acc[16];
        raw_ints16<NB>(g, ox, oy, f0, fmax, acc);
#pragma unroll
        for (int j = 0; j < 16; j++) r[j] = (uint32_t)acc[j];
    } else {
        const int nv = min(16, fmax - f0 + 1);
        const uint32_t pl = (uint32_t)g.plane;
        const uint32_t lo = (uint32_t)(oy * g.stride + ox);
        const float *q = (const float *)g.data + (int64_t)(g.f0 + f0) * g.plane;
#pragma unroll
        for (int j = 0; j < 16; j++) r[j] = __builtin_bit_cast(uint32_t, q[lo + (uint32_t)min(j, nv - 1) * pl]);
    }
}

__device__ __forceinline__ void words_to_values16(const SeriesGeom &g, const uint32_t (&r)[16], double (&v)[16]) {
    if (g.src == kSrcU16) {
#pragma unroll
        for (int j = 0; j < 16; j++) v[j] = (double)(int)r[j];
    } else {
#pragma unroll
        for (int j = 0; j < 16; j++) v[j] = (double)__builtin_bit_cast(float, r[j]);
    }
}

__device__ __forceinline__ double raw_value(const SeriesGeom &g, int ox, int oy, int f) {
    if (g.src == kSrcU16) return (double)raw_u16(g, ox, oy, f);
    const float *base = (const float *)g.data + (int64_t)(g.f0 + f) * g.plane + (int64_t)oy * g.stride + ox;
    return (double)base[0];
}

// Per-slot bleach polynomial held in registers, normalized by its value at
// t = 0 so that g(u) = f(u) / c0.  MAXO is a compile-time degree >= the run's
// order; the coefficients above the order are zero, which leaves Horner's
// result unchanged (0 * u + 0 = 0 until the leading coefficient).  MAXO = 0
// is the uncorrected series; MAXO > 0 is instantiated only for order >= 1
// (bleach_maxo), so apply() needs no per-value test of the order.
template <int MAXO>
struct BleachP {
    double c[MAXO + 1];
    double c0, scale, off;
    __device__ void load(const BleachDesc &bd, int slot) {
        const int order = bd.order;
        c0 = MAXO > 0 ? bd.coef[(int64_t)slot * (order + 1)] : 1.0;
#pragma unroll
        for (int k = 0; k <= MAXO; k++)
            c[k] = (MAXO > 0 && k <= order) ? bd.coef[(int64_t)slot * (order + 1) + k] / c0 : 0.0;
        scale = bd.frametime / bd.tscale;
        off = bd.t_off;
    }
    // I' = I / sqrt(f/c0) + c0 (1 - sqrt(f/c0))   (IFCS:9744-9750), with
    // g = f/c0 and r = 1/sqrt(g) written as r (I - c0 g) + c0 (sqrt(g) =
    // g r).  r is the hardware FP64 reciprocal square root (about 2^-23
    // relative) refined by one Newton step (relative error ~1e-14).  g < 0
    // yields NaN as in Java; g = 0 yields NaN where Java has +-inf (the
    // pixel's correlation is NaN either way).
    __device__ __forceinline__ double apply(double v, int f) const {
        if (MAXO == 0) return v;
        const double u = ((double)f + off) * scale;
        double g = 0.0;
#pragma unroll
        for (int k = MAXO; k >= 0; k--) g = fma(g, u, c[k]);
        const double y = __builtin_amdgcn_rsq(g);
        const double e = fma(-(g * y), y, 1.0);   // 1 - g y^2
        const double r = fma(0.5 * y, e, y);
        return fma(r, v - c0 * g, c0);
    }
};

// "Lin Segment" bleach correction (IFCS:9767-9805) of one slot: bcTrace[k]
// are the segment means, c0 = bcInt0 their extrapolation to the start; the
// reference function g is linear between segment centres, extrapolated
// before the first centre and, with the reference's index nf = x - (n-1)
// len + len/2 + len, after the last.  x is Java's 1-based frame index.
struct BleachSeg {
    const double *t;
    double c0;
    int n, len, h;
    __device__ void load(const BleachDesc &bd, int slot) {
        n = bd.order;
        len = bd.seg_len;
        h = len / 2;
        const double *c = bd.coef + (int64_t)slot * (n + 1);
        c0 = c[0];
        t = c + 1;
    }
    __device__ __forceinline__ double apply(double v, int f) const {
        const int x = f + 1;
        double g;
        if (x < h) {
            g = t[0] + (t[0] - t[1]) / len * (double)(h - x);
        } else if (x < (n - 1) * len + h) {
            const int xs = (x - h) / len + 1, y = (x - h) - (xs - 1) * len;
            g = t[xs - 1] + (t[xs] - t[xs - 1]) * y / len;
        } else {
            const int nf = x - (n - 1) * len + h + len;
            g = t[n - 2] + (t[n - 1] - t[n - 2]) * nf / len;
        }
        const double r = sqrt(g / c0);
        return v / r + c0 * (1.0 - r);
    }
};

// "Single Exp" / "Double Exp" bleach correction (IFCS:9682-9738) of one
// slot: f(t) = A e^{-t/t1} + C (+ B e^{-t/t2}), f0 = f(0), t = frametime
// (x + 0.5) with Java's 1-based frame index x; I' = I / sqrt(f/f0) +
// f0 (1 - sqrt(f/f0)).  f0 = 0 marks a fit the reference would not apply.
// FP64 throughout, as the Java path.
template <int NE>
struct BleachExp {
    double a, t1, b, t2, c, f0, scale, off;
    __device__ void load(const BleachDesc &bd, int slot) {
        const int np = NE == 1 ? 3 : 5;
        const double *q = bd.coef + (int64_t)slot * (np + 1);
        f0 = q[0];
        a = q[1];
        t1 = q[2];
        if (NE == 1) {
            b = 0.0;
            t2 = 1.0;
            c = q[3];
        } else {
            b = q[3];
            t2 = q[4];
            c = q[5];
        }
        scale = bd.frametime;
        off = bd.t_off;
    }
    __device__ __forceinline__ double curve(double t) const {
        double v = a * exp(-t / t1) + c;
        if (NE == 2) v += b * exp(-t / t2);
        return v;
    }
    __device__ __forceinline__ double apply(double v, int f) const {
        if (f0 == 0.0) return v;
        const double r = sqrt(curve(((double)f + off) * scale) / f0);
        return v / r + f0 * (1.0 - r);
    }
};

// Compile-time degree used for a run's polynomial order.
inline int bleach_maxo(int order) { return order <= 0 ? 0 : (order <= 4 ? 4 : kMaxOrder); }

}  // namespace imfcs
