// Seeded 2D particle simulator on the GPU (the physics of the plugin's
// simulateACF2D, Imaging_FCS_1_52.java:13963-14393): emitters random-walk
// with Gaussian steps of variance 2 D dt per axis inside periodic boxes
// ("domains"), bleach irreversibly with probability 1 - exp(-dt/tau) per
// frame, and emit photons through a Gaussian PSF onto the detector pixels
// above a camera offset.
//
// Photon detection is drawn per pixel: the photons of one emitter in one
// frame are Poisson(lambda) with lambda the emitter's brightness, and by
// Poisson splitting the counts of the detector pixels are independent
// Poisson draws with means lambda * P(pixel), P the PSF mass over the pixel
// (separable: a difference of two erf per axis).  The 7 x 7 pixels around the
// emitter hold the mass (sigma <= 0.95 pixel keeps > 99.9 % inside).
//
// One wavefront per emitter: every lane repeats the emitter's walk (the
// same counter-based random numbers, so the values are wave-uniform) and
// lane l < 49 draws the counts of pixel (l % 7 - 3, l / 7 - 3) around it.
// Random numbers are a hash of (seed, emitter, frame, purpose, lane)
// (imfcs_rng.h), so a stack depends only on its parameters and seed.  Frames are simulated in
// chunks into an int32 accumulation buffer, then written as uint16 with the
// offset (saturating at 65535); the emitter state carries across chunks.
#include "../../include/imfcs.h"
#include "imfcs_internal.h"
#include "imfcs_rng.h"

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

namespace imfcs {

namespace {

// Poisson(mu) draw from the uniforms of key.  mu < 12: multiplication of
// uniforms until the product falls below e^-mu (Knuth, The Art of Computer
// Programming vol. 2, 3.4.1 F); mu >= 12: the transformed rejection method
// PTRS of W. Hoermann, "The transformed rejection method for generating
// Poisson random variables", Insurance: Mathematics and Economics 12 (1993)
// 39-45.
__device__ int poisson(float mu, uint64_t key) {
    if (!(mu > 0.f)) return 0;
    uint32_t k = 0;
    if (mu < 12.f) {
        const float L = expf(-mu);
        float p = 1.f;
        int n = -1;
        do {
            p *= u01(key, k++);
            n++;
        } while (p > L && n < 200);
        return n;
    }
    const float slam = sqrtf(mu), loglam = logf(mu);
    const float b = 0.931f + 2.53f * slam, a = -0.059f + 0.02483f * b;
    const float invalpha = 1.1239f + 1.1328f / (b - 3.4f), vr = 0.9277f - 3.6224f / (b - 2.f);
    for (int it = 0; it < 64; it++) {
        const float U = u01(key, k++) - 0.5f, V = u01(key, k++);
        const float us = 0.5f - fabsf(U);
        const float kf = floorf((2.f * a / us + b) * U + mu + 0.43f);
        if (us >= 0.07f && V <= vr) return (int)kf;
        if (kf < 0.f || (us < 0.013f && V > us)) continue;
        if (logf(V) + logf(invalpha) - logf(a / (us * us) + b) <= -mu + kf * loglam - lgammaf(kf + 1.f))
            return (int)kf;
    }
    return (int)(mu + 0.5f);
}

struct SimDomainDev {
    double x0, y0, w, h;       // periodic box (um)
    double step[4];            // sqrt(2 D dt) per quadrant of the box, component 1
    double step2;              // component 2
    float frac2;
    float lam_lo, log_ratio;   // brightness lam_lo * exp(log_ratio * (x - x0) / w)
    int32_t first;             // first emitter index of the domain
};

struct SimArgs {
    const SimDomainDev *dom;
    int ndom;
    int nemit;
    int W, H;
    double pix;
    float sig_px;              // PSF sigma in pixels
    double pbleach;
    uint64_t seed;
    double *state;             // [nemit][3]: x, y, alive
    int32_t *acc;              // [nfr][H][W]
    int t0, nfr;
    int init;
};

__global__ void __launch_bounds__(256) k_sim_emitters(SimArgs a) {
    const int e = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (e >= a.nemit) return;
    int d = 0;
    while (d + 1 < a.ndom && a.dom[d + 1].first <= e) d++;
    const SimDomainDev D = a.dom[d];
    const bool comp2 = u01(rkey(a.seed, e, 0xFFFFFFFFull, 7), 0) <= D.frac2;
    double x, y, alive;
    if (a.init) {
        const uint64_t k0 = rkey(a.seed, e, 0xFFFFFFFFull, 1);
        x = D.x0 + u01d(k0, 0) * D.w;
        y = D.y0 + u01d(k0, 1) * D.h;
        alive = 1.0;
    } else {
        x = a.state[(size_t)e * 3];
        y = a.state[(size_t)e * 3 + 1];
        alive = a.state[(size_t)e * 3 + 2];
    }
    const double cx = D.x0 + 0.5 * D.w, cy = D.y0 + 0.5 * D.h;
    const int dx = lane % 7 - 3, dy = lane / 7 - 3;
    const float inv_s2 = 0.70710678118654752f / a.sig_px;
    const size_t plane = (size_t)a.W * a.H;
    for (int f = 0; f < a.nfr; f++) {
        const uint64_t t = (uint64_t)(a.t0 + f);
        // walk: one Box-Muller pair per frame
        const uint64_t kw = rkey(a.seed, e, t, 2);
        const double r = sqrt(-2.0 * log(u01d(kw, 0))), th = 6.283185307179586 * u01d(kw, 1);
        const int qd = (x >= cx ? 1 : 0) + (y >= cy ? 2 : 0);
        const double st = comp2 ? D.step2 : D.step[qd];
        x += st * r * cos(th);
        y += st * r * sin(th);
        x -= D.w * floor((x - D.x0) / D.w);
        y -= D.h * floor((y - D.y0) / D.h);
        if (alive != 0.0 && a.pbleach > 0.0 && u01d(kw, 2) < a.pbleach) alive = 0.0;
        if (alive == 0.0 || lane >= 49) continue;
        // detector pixel of this lane and its PSF mass
        const double xp = x / a.pix, yp = y / a.pix;
        const int ix = (int)floor(xp) + dx, iy = (int)floor(yp) + dy;
        if (ix < 0 || ix >= a.W || iy < 0 || iy >= a.H) continue;
        const float fx = (float)(xp - ix), fy = (float)(yp - iy);   // emitter relative to the pixel corner
        const float mx = 0.5f * (erff((1.f - fx) * inv_s2) - erff(-fx * inv_s2));
        const float my = 0.5f * (erff((1.f - fy) * inv_s2) - erff(-fy * inv_s2));
        const float lam = D.lam_lo * expf(D.log_ratio * (float)((x - D.x0) / D.w));
        const int n = poisson(lam * mx * my, rkey(a.seed, e, t, 16 + (uint64_t)lane));
        if (n > 0) atomicAdd(&a.acc[(size_t)f * plane + (size_t)iy * a.W + ix], n);
    }
    if (lane == 0) {
        a.state[(size_t)e * 3] = x;
        a.state[(size_t)e * 3 + 1] = y;
        a.state[(size_t)e * 3 + 2] = alive;
    }
}

// out = min(offset + acc, 65535); acc cleared for the next chunk
__global__ void k_sim_write(int32_t *acc, uint16_t *out, size_t n, int offset) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int v = offset + acc[i];
        out[i] = (uint16_t)(v > 65535 ? 65535 : v);
        acc[i] = 0;
    }
}

__global__ void k_fill_i32(int32_t *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

}  // namespace

}  // namespace imfcs

using namespace imfcs;

// Declared in include/imfcs.h.
extern "C" int imfcs_simulate_2d_device(uint16_t *d_out, int W, int H, int T, const imfcs_sim_domain *domains,
                                        int ndomains, double pixel_um, double psf_sigma_um, double frametime_s,
                                        int offset, double tau_bleach_s, uint64_t seed, void *stream) {
    if (!d_out || W <= 0 || H <= 0 || T <= 0 || !domains || ndomains <= 0 || pixel_um <= 0 || frametime_s <= 0 ||
        offset < 0 || offset > 65535 || psf_sigma_um <= 0)
        return set_error(IMFCS_ERR_ARG, "simulate: bad stack dimensions, domains or optics");
    const float sig_px = (float)(psf_sigma_um / pixel_um);
    if (sig_px > 0.95f)   // the 7 x 7 pixel footprint needs sigma <= 0.95 pixel
        return set_error(IMFCS_ERR_ARG, "simulate: PSF sigma above 0.95 pixel");
    std::vector<SimDomainDev> dd(ndomains);
    int64_t nemit = 0;
    for (int i = 0; i < ndomains; i++) {
        const imfcs_sim_domain &s = domains[i];
        if (s.w_um <= 0 || s.h_um <= 0 || s.n_emitters < 0 || s.counts_lo < 0 || s.counts_hi < 0 ||
            (s.counts_lo == 0) != (s.counts_hi == 0) || s.frac2 < 0 || s.frac2 > 1)
            return set_error(IMFCS_ERR_ARG, "simulate: bad domain " + std::to_string(i));
        SimDomainDev &d = dd[i];
        d.x0 = s.x0_um;
        d.y0 = s.y0_um;
        d.w = s.w_um;
        d.h = s.h_um;
        for (int q = 0; q < 4; q++) d.step[q] = std::sqrt(2.0 * s.D_um2s[q] * frametime_s);
        d.step2 = std::sqrt(2.0 * s.D2_um2s * frametime_s);
        d.frac2 = (float)s.frac2;
        d.lam_lo = (float)s.counts_lo;
        d.log_ratio = s.counts_lo > 0 ? (float)std::log(s.counts_hi / s.counts_lo) : 0.f;
        d.first = (int32_t)nemit;
        nemit += s.n_emitters;
    }
    if (nemit > (1 << 30)) return set_error(IMFCS_ERR_ARG, "simulate: too many emitters");
    hipStream_t st = (hipStream_t)stream;
    const size_t plane = (size_t)W * H;
    // frames per chunk: an int32 accumulation buffer of at most 2^28 values (1 GiB)
    const int cf = (int)std::max<size_t>(1, std::min<size_t>((size_t)T, ((size_t)1 << 28) / plane));
    SimDomainDev *d_dom = nullptr;
    double *d_state = nullptr;
    int32_t *d_acc = nullptr;
    auto cleanup = [&]() {
        if (d_dom) (void)hipFree(d_dom);
        if (d_state) (void)hipFree(d_state);
        if (d_acc) (void)hipFree(d_acc);
    };
    if (hipMalloc(&d_dom, sizeof(SimDomainDev) * ndomains) != hipSuccess ||
        hipMalloc(&d_state, sizeof(double) * 3 * std::max<int64_t>(nemit, 1)) != hipSuccess ||
        hipMalloc(&d_acc, sizeof(int32_t) * plane * cf) != hipSuccess) {
        cleanup();
        return set_error(IMFCS_ERR_MEMORY, "simulate: device allocation failed");
    }
    hipError_t e = hipMemcpyAsync(d_dom, dd.data(), sizeof(SimDomainDev) * ndomains, hipMemcpyHostToDevice, st);
    const size_t nacc = plane * cf;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)std::min<size_t>((nacc + 255) / 256, 65536)), dim3(256), 0, st,
                       d_acc, nacc);
    SimArgs a{};
    a.dom = d_dom;
    a.ndom = ndomains;
    a.nemit = (int)nemit;
    a.W = W;
    a.H = H;
    a.pix = pixel_um;
    a.sig_px = sig_px;
    a.pbleach = tau_bleach_s > 0 ? 1.0 - std::exp(-frametime_s / tau_bleach_s) : 0.0;
    a.seed = seed;
    a.state = d_state;
    a.acc = d_acc;
    for (int t0 = 0; t0 < T && e == hipSuccess; t0 += cf) {
        a.t0 = t0;
        a.nfr = std::min(cf, T - t0);
        a.init = t0 == 0;
        if (nemit > 0) hipLaunchKernelGGL(k_sim_emitters, dim3((unsigned)((nemit + 3) / 4)), dim3(256), 0, st, a);
        const size_t n = plane * a.nfr;
        hipLaunchKernelGGL(k_sim_write, dim3((unsigned)std::min<size_t>((n + 255) / 256, 65536)), dim3(256), 0, st,
                           d_acc, d_out + (size_t)t0 * plane, n, offset);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    cleanup();
    return e == hipSuccess ? 0 : set_error(IMFCS_ERR_DEVICE, std::string("simulate: ") + hipGetErrorString(e));
}
