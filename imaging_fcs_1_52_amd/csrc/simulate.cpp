// Seeded 2D particle simulator producing uint16 imaging-FCS stacks.
//
// The physics follows the plugin's 2D simulator (Imaging_FCS_1_52.java
// simulateACF2D, IFCS:13963-14393): emitters random-walk with Gaussian steps
// of variance 2*D*dt per axis, bleach irreversibly with probability
// 1 - exp(-dt/tau) per frame, emit Poisson(cps*dt) photons per frame, and
// each photon lands at the emitter position plus a Gaussian PSF offset
// (sigma), binned into detector pixels over a camera offset.  Boundaries are
// periodic over a field that extends `margin` pixels beyond the detector.
//
// Random numbers come from a SplitMix64 stream per emitter, so a stack
// depends only on (parameters, seed), not on the thread count.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

struct SplitMix {
    uint64_t s;
    inline uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    inline double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    // uniform in (0, 1]
    inline double uniform_pos() { return ((double)(next() >> 11) + 1.0) * (1.0 / 9007199254740992.0); }
    inline float uniformf_pos() { return ((float)(next() >> 40) + 1.0f) * (1.0f / 16777216.0f); }
    // Box-Muller pair in double
    inline void normal2(double &a, double &b) {
        double r = std::sqrt(-2.0 * std::log(uniform_pos()));
        double t = 6.283185307179586 * uniform();
        a = r * std::cos(t);
        b = r * std::sin(t);
    }
    inline void normal2f(float &a, float &b) {
        float r = std::sqrt(-2.0f * std::log(uniformf_pos()));
        float t = 6.2831853f * (float)(next() >> 40) * (1.0f / 16777216.0f);
        a = r * std::cos(t);
        b = r * std::sin(t);
    }
};

inline uint64_t mix_seed(uint64_t seed, uint64_t k) {
    SplitMix m{seed * 0x2545F4914F6CDD1Dull + k * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull};
    m.next();
    return m.next();
}

}  // namespace

extern "C" {

// Returns 0 on success, -1 on bad arguments.
// out: uint16 [T][H][W], frame-major.  D_um2s: four diffusion coefficients,
// one per field quadrant (x<cx,y<cy), (x>=cx,y<cy), (x<cx,y>=cy), (x>=cx,y>=cy);
// pass four equal values for a homogeneous sample.  tau_bleach_s <= 0
// disables bleaching.
int imfcs_simulate_2d(uint16_t *out, int W, int H, int T, int margin, int n_emitters,
                      double pixel_um, double psf_sigma_um, double frametime_s,
                      double counts_per_frame, int offset, double tau_bleach_s,
                      const double *D_um2s, uint64_t seed, int nthreads) {
    if (!out || W <= 0 || H <= 0 || T <= 0 || margin < 0 || n_emitters < 0 || pixel_um <= 0 ||
        frametime_s <= 0 || offset < 0 || offset > 65535 || !D_um2s)
        return -1;
    const size_t plane = (size_t)W * H;
    const size_t total = plane * (size_t)T;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long long i = 0; i < (long long)total; i++) out[i] = (uint16_t)offset;

    const double fw = (W + 2.0 * margin) * pixel_um, fh = (H + 2.0 * margin) * pixel_um;
    const double cx = fw / 2, cy = fh / 2;
    const double lam = counts_per_frame;
    // Poisson inverse-CDF table for the constant per-frame photon rate
    std::vector<double> cdf;
    {
        double pk = std::exp(-lam), c = pk;
        cdf.push_back(c);
        for (int k = 1; k < 100000 && c < 1.0 - 1e-15; k++) {
            pk *= lam / k;
            c += pk;
            cdf.push_back(c);
        }
    }
    const int kmax = (int)cdf.size();
    const double pbleach = tau_bleach_s > 0 ? 1.0 - std::exp(-frametime_s / tau_bleach_s) : 0.0;
    double step[4];
    for (int q = 0; q < 4; q++) step[q] = std::sqrt(2.0 * D_um2s[q] * frametime_s);
    const float sig = (float)psf_sigma_um;
    const float inv_pix = (float)(1.0 / pixel_um);

#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
    for (int e = 0; e < n_emitters; e++) {
        SplitMix rng{mix_seed(seed, (uint64_t)e)};
        double x = rng.uniform() * fw, y = rng.uniform() * fh;
        bool alive = true;
        for (int t = 0; t < T; t++) {
            int qd = (x >= cx ? 1 : 0) + (y >= cy ? 2 : 0);
            double dx, dy;
            rng.normal2(dx, dy);
            x += step[qd] * dx;
            y += step[qd] * dy;
            x -= fw * std::floor(x / fw);
            y -= fh * std::floor(y / fh);
            if (alive && pbleach > 0 && rng.uniform() < pbleach) alive = false;
            if (!alive) continue;
            double u = rng.uniform();
            int n = 0;
            while (n < kmax - 1 && u > cdf[n]) n++;
            uint16_t *fr = out + (size_t)t * plane;
            const float xm = (float)(x / pixel_um) - margin, ym = (float)(y / pixel_um) - margin;
            for (int p = 0; p < n; p += 2) {
                float gx0, gy0, gx1, gy1;
                rng.normal2f(gx0, gy0);
                rng.normal2f(gx1, gy1);
                for (int h = 0; h < 2 && p + h < n; h++) {
                    float px = xm + (h ? gx1 : gx0) * sig * inv_pix;
                    float py = ym + (h ? gy1 : gy0) * sig * inv_pix;
                    int ix = (int)std::floor(px), iy = (int)std::floor(py);
                    if (ix >= 0 && ix < W && iy >= 0 && iy < H) {
                        uint16_t *cell = fr + (size_t)iy * W + ix;
                        // saturating add: never wraps past 65535
                        uint16_t old = __atomic_load_n(cell, __ATOMIC_RELAXED);
                        while (old < 65535 &&
                               !__atomic_compare_exchange_n(cell, &old, (uint16_t)(old + 1), true,
                                                            __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
                        }
                    }
                }
            }
        }
    }
    return 0;
}

}  // extern "C"
