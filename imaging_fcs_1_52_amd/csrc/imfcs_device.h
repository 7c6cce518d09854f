// Device helpers shared by the HIP translation units: series access and the
// on-the-fly polynomial bleach correction.
#pragma once
#include "imfcs_internal.h"

namespace imfcs {

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

// Exact integer window sum minus background (IFCS:9668-9678).
__device__ __forceinline__ int raw_u16(const SeriesGeom &g, int ox, int oy, int f) {
    const uint16_t *base = (const uint16_t *)g.data + (int64_t)(g.f0 + f) * g.plane +
                           (int64_t)oy * g.stride + ox;
    if (g.bx == 1 && g.by == 1) return (int)base[0] - g.bckg_total;
    int acc = 0;
    for (int k = 0; k < g.by; k++)
        for (int i = 0; i < g.bx; i++) acc += base[(int64_t)k * g.stride + i];
    return acc - g.bckg_total;
}

__device__ __forceinline__ double raw_value(const SeriesGeom &g, int ox, int oy, int f) {
    if (g.src == kSrcU16) return (double)raw_u16(g, ox, oy, f);
    const float *base = (const float *)g.data + (int64_t)(g.f0 + f) * g.plane + (int64_t)oy * g.stride + ox;
    return (double)base[0];
}

// Per-series bleach polynomial held in registers, normalized by its value
// at t = 0 so that g(u) = f(u) / c0.
struct Bleach {
    double c[kMaxOrder + 1];
    int order;
    double c0, scale;   // u = (f + off) * scale
    double off;
    __device__ void load(const BleachDesc &bd, int s) {
        order = bd.order;
        c0 = (order > 0) ? bd.coef[(int64_t)s * (order + 1)] : 1.0;
#pragma unroll
        for (int k = 0; k <= kMaxOrder; k++)
            c[k] = (k <= order && order > 0) ? bd.coef[(int64_t)s * (order + 1) + k] / c0 : 0.0;
        scale = bd.frametime / bd.tscale;
        off = bd.t_off;
    }
    // I' = I / sqrt(f/c0) + c0 (1 - sqrt(f/c0))   (IFCS:9744-9750), with
    // 1/sqrt(g) from the FP32 reciprocal square root refined by one FP64
    // Newton step (relative error ~1e-14).  g <= 0 yields NaN as in Java.
    __device__ __forceinline__ double apply(double v, int f) const {
        if (order == 0) return v;
        double u = ((double)f + off) * scale;
        double g = 0.0;
#pragma unroll
        for (int k = kMaxOrder; k >= 0; k--)
            if (k <= order) g = fma(g, u, c[k]);
        double r = (double)rsqrtf((float)g);
        r = r * fma(-0.5 * g * r, r, 1.5);
        double sq = g * r;
        return fma(v, r, c0 * (1.0 - sq));
    }
};

}  // namespace imfcs
