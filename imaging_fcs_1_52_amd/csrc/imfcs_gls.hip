// Generalized least squares: the regularized channel covariance of every
// correlation pair and the GLS fit on it.  The reference computes both on
// the CPU only: calculateCF's GLS branch (IFCS:9524-9639), dataTransform
// (IFCS:13705-13729), GLS_fitFunction (IFCS:13569-13666), GLSFit
// (IFCS:10622-10789).
//
// k_gls_cov (one workgroup per pair): the correlator has stored every block
// sum S of the blocked variance; P = S / (B abar bbar) is Java's normalized
// blocked product prod[x][z] / (directm delayedm).  Two passes over the
// pnum x m block table (m = chanum - 1 fitted channels): the means mcov, then
// per channel pair (x, y <= x) C = sum d_x d_y and Q = sum (d_x d_y)^2 of the
// deviations d.  The Schaefer-Strimmer weights follow from C and Q in closed
// form:  sum_z (d_x^2 - Cxx)^2 = Q_xx - (pnum - 2) Cxx^2 and
// sum_z (c_x c_y - R_xy)^2 = Q_xy / (Cxx Cyy) - (pnum - 2) R_xy^2, with the
// normalized C = sum / (pnum - 1), c = d / sqrt(Cxx), R = Cxy / sqrt(Cxx Cyy).
// Java rescales its running covariance sum inside the y loop
// (tmpnum *= pnum / (pnum-1)^3 after every y, IFCS:9593-9597), so channel
// pair (x, k) enters the numerator with weight f (1 - f^(x-k)) / (1 - f),
// f = pnum / (pnum - 1)^3.  The shrunk matrix / pnum is factorized in FP64
// with the checks of commons-math3's CholeskyDecomposition (pivot <= 1e-10
// fails, as Java throws there); W = L^-1 and y = W mcov are stored for the
// fit.
//
// k_lm_gls (one wave per pair): the LM of k_lm_fixed on the whitened
// residuals W (mcov - f), with W held in LDS.
#include "imfcs_device.h"
#include "imfcs_internal.h"

#include <math.h>

namespace imfcs {

constexpr int kGlsThreads = 256;
constexpr int kGlsZT = 32;             // block rows staged per pass-2 tile
constexpr int kGlsEnt = (kGlsMaxM * (kGlsMaxM + 1) / 2 + kGlsThreads - 1) / kGlsThreads;   // entries per thread


__global__ void __launch_bounds__(kGlsThreads) k_gls_cov(GlsArgs A) {
    const CorrTables &ct = *A.ctp;
    const int pair = blockIdx.x;
    const int tid = threadIdx.x;
    const int m = ct.chanum - 1;
    __shared__ double norm[kGlsMaxM], mean[kGlsMaxM], cdiag[kGlsMaxM];
    __shared__ double tile[kGlsZT][kGlsMaxM];
    __shared__ double S[kGlsMaxM][kGlsMaxM + 1];
    __shared__ double red[2][kGlsThreads];
    __shared__ double sh_median, sh_lamvar, sh_lamcov;
    __shared__ int sh_ok;
    int sa, sb;
    pair_series(A.pd, pair, sa, sb);
    const int ka = A.g.slot_of[sa], kb = A.g.slot_of[sb];
    const int indp = A.ind[pair];
    const int pnum = (int)floor(ct.mtab_last / pow(2.0, fmax(indp - ct.log2_samp_last, 0.0)));
    const float *blk = A.blk + (int64_t)pair * A.nblk * m;
    // channel normalization 1 / (B abar bbar) and the means
    for (int j = tid; j < m; j += kGlsThreads) {
        const int x = j + 1, c = ct.level[x];
        const double n = (double)(ct.numbin[c] - ct.del[x]);
        const double *Fa = A.F + (int64_t)ka * ct.npos, *Fb = A.F + (int64_t)kb * ct.npos;
        const double am = Fa[ct.pos_a_end[x]] / n;
        const double bm = (Fb[ct.pos_b_end[x]] - Fb[ct.pos_b_beg[x]]) / n;
        const double B = (double)(1 << (indp - c));
        norm[j] = 1.0 / (B * am * bm);
        double s = 0.0;
        for (int z = 0; z < pnum; z++) s += (double)blk[(int64_t)z * m + j] * norm[j];
        mean[j] = s / pnum;
        if (A.mcov) A.mcov[(int64_t)pair * ct.chanum + x] = mean[j];
    }
    if (A.mcov && tid == 0) A.mcov[(int64_t)pair * ct.chanum] = 0.0;
    // this thread's lower-triangle entries e = tid + k * threads
    const int ne = m * (m + 1) / 2;
    int ex[kGlsEnt], ey[kGlsEnt];
    double C[kGlsEnt], Q[kGlsEnt];
#pragma unroll
    for (int k = 0; k < kGlsEnt; k++) {
        int e = tid + k * kGlsThreads;
        int x = 0;
        if (e < ne) {
            x = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
            while ((x + 1) * (x + 2) / 2 <= e) x++;
            while (x * (x + 1) / 2 > e) x--;
        }
        ex[k] = x;
        ey[k] = e < ne ? e - x * (x + 1) / 2 : 0;
        C[k] = Q[k] = 0.0;
    }
    __syncthreads();
    for (int z0 = 0; z0 < pnum; z0 += kGlsZT) {
        const int nz = min(kGlsZT, pnum - z0);
        for (int i = tid; i < nz * m; i += kGlsThreads) {
            const int zz = i / m, j = i - zz * m;
            tile[zz][j] = (double)blk[(int64_t)(z0 + zz) * m + j] * norm[j] - mean[j];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kGlsEnt; k++) {
            if (tid + k * kGlsThreads < ne) {
                double c = C[k], q = Q[k];
                for (int zz = 0; zz < nz; zz++) {
                    const double p = tile[zz][ex[k]] * tile[zz][ey[k]];
                    c += p;
                    q = fma(p, p, q);
                }
                C[k] = c;
                Q[k] = q;
            }
        }
        __syncthreads();
    }
    const double pm1 = (double)(pnum - 1);
    // normalized covariance into S (both triangles), its diagonal into cdiag
#pragma unroll
    for (int k = 0; k < kGlsEnt; k++) {
        if (tid + k * kGlsThreads < ne) {
            const double v = C[k] / pm1;
            S[ex[k]][ey[k]] = v;
            S[ey[k]][ex[k]] = v;
            if (ex[k] == ey[k]) cdiag[ex[k]] = v;
        }
    }
    __syncthreads();
    // median of the diagonal together with Java's unused slot 0 (value 0):
    // m + 1 values, element m / 2 of the sorted list when m + 1 is odd
    if (tid < 64) {
        const int cnt = m + 1;
        double med = 0.0;
        // rank selection: the value whose rank brackets the middle position(s)
        const int lo = (cnt - 1) / 2, hi = cnt / 2;
        double vlo = 0.0, vhi = 0.0;
        for (int i = tid; i < cnt; i += 64) {
            const double v = i < m ? cdiag[i] : 0.0;
            int less = 0;
            for (int j = 0; j < cnt; j++) {
                const double u = j < m ? cdiag[j] : 0.0;
                less += (u < v) || (u == v && j < i);
            }
            if (less == lo) vlo = v;
            if (less == hi) vhi = v;
        }
        // exactly one lane found each rank
        for (int off = 32; off > 0; off >>= 1) {
            vlo += __shfl_xor(vlo, off, 64);
            vhi += __shfl_xor(vhi, off, 64);
        }
        med = (lo == hi) ? vlo : 0.5 * (vlo + vhi);
        if (tid == 0) sh_median = med;
    }
    __syncthreads();
    const double median = sh_median;
    const double f = (double)pnum / (pm1 * pm1 * pm1);
    // variance of the variance (lamvar) and of the covariance (lamcov)
    double nv = 0.0, dv = 0.0, nc = 0.0, dc = 0.0;
#pragma unroll
    for (int k = 0; k < kGlsEnt; k++) {
        if (tid + k * kGlsThreads < ne) {
            const int x = ex[k], y = ey[k];
            const double cxx = cdiag[x], cyy = cdiag[y];
            if (x == y) {
                nv += (Q[k] - (double)(pnum - 2) * cxx * cxx) * f;
                dv += (cxx - median) * (cxx - median);
            } else {
                const double r = S[x][y] / sqrt(cxx * cyy);
                const double sxy = Q[k] / (cxx * cyy) - (double)(pnum - 2) * r * r;
                // channel indices are 1-based in Java: x + 1, y + 1
                const double w = f * (1.0 - pow(f, (double)(x - y))) / (1.0 - f);
                nc += sxy * w;
                dc += r * r;
            }
        }
    }
    for (int pass = 0; pass < 2; pass++) {
        red[0][tid] = pass ? nc : nv;
        red[1][tid] = pass ? dc : dv;
        __syncthreads();
        for (int s = kGlsThreads / 2; s > 0; s >>= 1) {
            if (tid < s) {
                red[0][tid] += red[0][tid + s];
                red[1][tid] += red[1][tid + s];
            }
            __syncthreads();
        }
        if (tid == 0) {
            const double lam = fmax(fmin(1.0, red[0][0] / red[1][0]), 0.0);
            if (pass) sh_lamcov = lam;
            else sh_lamvar = lam;
        }
        __syncthreads();
    }
    const double lamvar = sh_lamvar, lamcov = sh_lamcov;
    // regularized covariance / pnum in place (IFCS:9601-9613): each element
    // depends on itself and the saved diagonal only
    for (int i = tid; i < m * m; i += kGlsThreads) {
        const int x = i / m, y = i - x * m;
        const double cmx = lamvar * median + (1.0 - lamvar) * cdiag[x];
        double v;
        if (x == y) {
            v = cmx / pnum;
        } else {
            const double cmy = lamvar * median + (1.0 - lamvar) * cdiag[y];
            const double r = S[x][y] / sqrt(cdiag[x] * cdiag[y]);
            v = (1.0 - lamcov) * r * sqrt(cmx * cmy) / pnum;
        }
        S[x][y] = v;
        if (A.cov) A.cov[((int64_t)pair * m + x) * m + y] = v;
    }
    if (tid == 0) sh_ok = 1;
    __syncthreads();
    // Cholesky in the column order of commons-math3 CholeskyDecomposition:
    // row i of lT is finalized at step i; pivot <= 1e-10 fails
    for (int i = 0; i < m; i++) {
        if (tid == 0) {
            if (!(S[i][i] > 1e-10)) sh_ok = 0;
        }
        __syncthreads();
        if (!sh_ok) break;
        const double piv = sqrt(S[i][i]);
        const double inv = 1.0 / piv;
        __syncthreads();
        // lT[i][q] *= inv for q > i; lT[q][p] -= lT[i][q] lT[i][p] for p >= q
        for (int qq = i + 1 + tid; qq < m; qq += kGlsThreads) S[i][qq] *= inv;
        __syncthreads();
        for (int idx = tid; idx < (m - i - 1) * (m - i - 1); idx += kGlsThreads) {
            const int qq = i + 1 + idx / (m - i - 1), pp = i + 1 + idx % (m - i - 1);
            if (pp >= qq) S[qq][pp] -= S[i][qq] * S[i][pp];
        }
        __syncthreads();
        if (tid == 0) S[i][i] = piv;
        __syncthreads();
    }
    const int ok = sh_ok;
    if (tid == 0) A.ok[pair] = ok;
    if (!ok) return;
    // L = lT^T: L[r][c] = S[c][r] for c <= r.  W = L^-1 by columns (one
    // column per thread, forward substitution), then y = W mcov.
    double *Wp = A.W + (int64_t)pair * m * m;
    for (int col = tid; col < m; col += kGlsThreads) {
        for (int r = 0; r < m; r++) {
            double v;
            if (r < col) {
                v = 0.0;
            } else {
                double s = (r == col) ? 1.0 : 0.0;
                for (int k2 = col; k2 < r; k2++) s -= S[k2][r] * Wp[(int64_t)k2 * m + col];
                v = s / S[r][r];
            }
            Wp[(int64_t)r * m + col] = v;
        }
    }
    __syncthreads();
    for (int r = tid; r < m; r += kGlsThreads) {
        double s = 0.0;
        for (int k2 = 0; k2 <= r; k2++) s += Wp[(int64_t)r * m + k2] * mean[k2];
        A.yw[(int64_t)pair * m + r] = s;
    }
}

void launch_gls_cov(const GlsArgs &A, int npairs, hipStream_t st) {
    hipLaunchKernelGGL(k_gls_cov, dim3(npairs), dim3(kGlsThreads), 0, st, A);
}

}  // namespace imfcs
