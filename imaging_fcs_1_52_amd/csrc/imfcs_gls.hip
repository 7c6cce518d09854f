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
// deviations d -- two SYRKs of the table, D^T D and (D o D)^T (D o D), on
// the matrix cores: the lower-triangle 16 x 16 tiles are spread over the
// workgroup's 4 wavefronts, each accumulating its C and Q tiles with
// v_mfma_f64_16x16x4_f64 from operands read from the staged table (one
// double per lane and block column; squared in registers for Q).  The Schaefer-Strimmer weights follow from C and Q in closed
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
constexpr int kGlsZT = 32;             // block rows staged per tile
constexpr int kGlsNB = (kGlsMaxM + 15) / 16;             // 16-channel blocks
constexpr int kGlsLd = kGlsNB * 16;                       // staged row length (zero padded)
constexpr int kGlsTiles = kGlsNB * (kGlsNB + 1) / 2;     // lower-triangle tiles
constexpr int kGlsTilesPerWave = (kGlsTiles + 3) / 4;

typedef double f64x4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(kGlsThreads) k_gls_cov(GlsArgs A) {
    const CorrTables &ct = *A.ctp;
    const int pair = blockIdx.x;
    const int tid = threadIdx.x;
    const int m = ct.chanum - 1;
    __shared__ double norm[kGlsMaxM], mean[kGlsMaxM], cdiag[kGlsMaxM];
    __shared__ double tile[kGlsZT][kGlsLd];
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
    // channel normalization 1 / (B abar bbar)
    for (int j = tid; j < m; j += kGlsThreads) {
        const int x = j + 1, c = ct.level[x];
        const double n = (double)(ct.numbin[c] - ct.del[x]);
        const double *Fa = A.F + (int64_t)ka * ct.npos, *Fb = A.F + (int64_t)kb * ct.npos;
        const double am = Fa[ct.pos_a_end[x]] / n;
        const double bm = (Fb[ct.pos_b_end[x]] - Fb[ct.pos_b_beg[x]]) / n;
        const double B = (double)(1 << (indp - c));
        norm[j] = 1.0 / (B * am * bm);
    }
    // the means: thread (g, j) sums rows z = g, g + G, ... of column j
    // (G = threads / m groups, each row read coalesced by a group), with
    // four partial sums in flight; the groups' sums are then added in order
    {
        const int G = kGlsThreads / m, g = tid / m, j = tid - g * m;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        if (g < G) {
            int z = g;
            for (; z + 3 * G < pnum; z += 4 * G) {
                s0 += (double)blk[(int64_t)z * m + j];
                s1 += (double)blk[(int64_t)(z + G) * m + j];
                s2 += (double)blk[(int64_t)(z + 2 * G) * m + j];
                s3 += (double)blk[(int64_t)(z + 3 * G) * m + j];
            }
            for (; z < pnum; z += G) s0 += (double)blk[(int64_t)z * m + j];
        }
        red[0][tid] = (s0 + s1) + (s2 + s3);
        __syncthreads();
        for (int jj = tid; jj < m; jj += kGlsThreads) {
            double s = 0.0;
            for (int gg = 0; gg < G; gg++) s += red[0][gg * m + jj];
            mean[jj] = s * norm[jj] / pnum;
            if (A.mcov) A.mcov[(int64_t)pair * ct.chanum + jj + 1] = mean[jj];
        }
    }
    if (A.mcov && tid == 0) A.mcov[(int64_t)pair * ct.chanum] = 0.0;
    // this wave's tiles: t = wave + 4 u of the lower triangle (bi >= bj)
    const int wv = tid >> 6, lane = tid & 63;
    const int nb = (m + 15) / 16, ntiles = nb * (nb + 1) / 2;
    int tbi[kGlsTilesPerWave], tbj[kGlsTilesPerWave];
    f64x4 accC[kGlsTilesPerWave], accQ[kGlsTilesPerWave];
#pragma unroll
    for (int u = 0; u < kGlsTilesPerWave; u++) {
        const int t = wv + 4 * u;
        int bi = 0, bj = 0;
        if (t < ntiles) {
            int r = t;
            while (r > bi) {
                r -= bi + 1;
                bi++;
            }
            bj = r;
        }
        tbi[u] = bi;
        tbj[u] = bj;
        accC[u] = accQ[u] = f64x4{0.0, 0.0, 0.0, 0.0};
    }
    const int arow = lane >> 4, acol = lane & 15;   // MFMA operand: k = lane / 16, row / column lane % 16
    __syncthreads();
    for (int z0 = 0; z0 < pnum; z0 += kGlsZT) {
        const int nz = min(kGlsZT, pnum - z0), nz4 = (nz + 3) & ~3;
        for (int i = tid; i < nz4 * kGlsLd; i += kGlsThreads) {
            const int zz = i / kGlsLd, j = i - zz * kGlsLd;
            tile[zz][j] = (zz < nz && j < m) ? (double)blk[(int64_t)(z0 + zz) * m + j] * norm[j] - mean[j] : 0.0;
        }
        __syncthreads();
        for (int k0 = 0; k0 < nz4; k0 += 4) {
            double fa[kGlsNB], fq[kGlsNB];
#pragma unroll
            for (int bb = 0; bb < kGlsNB; bb++) {
                fa[bb] = bb < nb ? tile[k0 + arow][16 * bb + acol] : 0.0;
                fq[bb] = fa[bb] * fa[bb];
            }
#pragma unroll
            for (int u = 0; u < kGlsTilesPerWave; u++) {
                if (wv + 4 * u < ntiles) {
                    double ai = 0.0, aj = 0.0, qi = 0.0, qj = 0.0;
#pragma unroll
                    for (int bb = 0; bb < kGlsNB; bb++) {
                        if (bb == tbi[u]) {
                            ai = fa[bb];
                            qi = fq[bb];
                        }
                        if (bb == tbj[u]) {
                            aj = fa[bb];
                            qj = fq[bb];
                        }
                    }
                    accC[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, aj, accC[u], 0, 0, 0);
                    accQ[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(qi, qj, accQ[u], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    const double pm1 = (double)(pnum - 1);
    // normalized covariance into S (both triangles), its diagonal into
    // cdiag; result element r of a v_mfma_f64_16x16x4_f64 tile: row
    // lane / 16 + 4 r, column lane % 16
#pragma unroll
    for (int u = 0; u < kGlsTilesPerWave; u++) {
        if (wv + 4 * u < ntiles) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int x = 16 * tbi[u] + arow + 4 * r, y = 16 * tbj[u] + acol;
                if (x < m && y <= x) {
                    const double v = accC[u][r] / pm1;
                    S[x][y] = v;
                    S[y][x] = v;
                    if (x == y) cdiag[x] = v;
                }
            }
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
    for (int u = 0; u < kGlsTilesPerWave; u++) {
        if (wv + 4 * u < ntiles) {
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                const int x = 16 * tbi[u] + arow + 4 * rr, y = 16 * tbj[u] + acol;
                if (!(x < m && y <= x)) continue;
                const double q = accQ[u][rr];
                const double cxx = cdiag[x], cyy = cdiag[y];
                if (x == y) {
                    nv += (q - (double)(pnum - 2) * cxx * cxx) * f;
                    dv += (cxx - median) * (cxx - median);
                } else {
                    const double r = S[x][y] / sqrt(cxx * cyy);
                    const double sxy = q / (cxx * cyy) - (double)(pnum - 2) * r * r;
                    // channel indices are 1-based in Java: x + 1, y + 1
                    const double w = f * (1.0 - pow(f, (double)(x - y))) / (1.0 - f);
                    nc += sxy * w;
                    dc += r * r;
                }
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
    // L = lT^T: L[r][c] = S[c][r] for c <= r.  W = L^-1 (lower) by rows of
    // forward substitution, all columns of a row at once, in LDS: W[r][c]
    // (r > c) overwrites the no longer needed lower triangle S[r][c], the
    // diagonal 1 / L[r][r] goes to cdiag.  Then W and y = W mcov are stored.
    for (int r = tid; r < m; r += kGlsThreads) cdiag[r] = 1.0 / S[r][r];
    __syncthreads();
    for (int r = 1; r < m; r++) {
        for (int col = tid; col < r; col += kGlsThreads) {
            double acc = S[col][r] * cdiag[col];   // k2 = col: L[r][col] W[col][col]
            for (int k2 = col + 1; k2 < r; k2++) acc += S[k2][r] * S[k2][col];
            S[r][col] = -acc * cdiag[r];
        }
        __syncthreads();
    }
    double *Wp = A.W + (int64_t)pair * m * m;
    for (int i = tid; i < m * m; i += kGlsThreads) {
        const int r = i / m, c = i - r * m;
        Wp[i] = r < c ? 0.0 : (r == c ? cdiag[r] : S[r][c]);
    }
    for (int r = tid; r < m; r += kGlsThreads) {
        double s = cdiag[r] * mean[r];
        for (int k2 = 0; k2 < r; k2++) s += S[r][k2] * mean[k2];
        A.yw[(int64_t)pair * m + r] = s;
    }
}

void launch_gls_cov(const GlsArgs &A, int npairs, hipStream_t st) {
    hipLaunchKernelGGL(k_gls_cov, dim3(npairs), dim3(kGlsThreads), 0, st, A);
}

}  // namespace imfcs
