// Learned initial guesses for the LM fit: a small BF16 multilayer
// perceptron on the matrix cores.
//
// The network maps a correlation curve (channels 1..72, divided by its
// largest magnitude s) to standardised LM starting values: ln(N s), ln D,
// ln D2, F2 and G_inf / s.  Layers: 72 (padded to 80) -> 256 GELU -> 256 GELU
// -> 5 (padded to 16).  One network serves one fit geometry (pixel size,
// PSF, CCF shift), as the LM's own model does: it is trained on the GPU, in
// a few seconds, on FCS_3p curves of that geometry with random parameters
// and noise, by the kernels below.
//
// Every matrix product is one kernel, k_gemm_bf16: 64 x 64 output tiles per
// workgroup of 4 wavefronts, each wavefront a 32 x 32 block of 2 x 2
// v_mfma_f32_16x16x16_bf16 tiles, K in steps of 16 staged through LDS, FP32
// accumulation, and an epilogue that adds the bias and applies GELU (storing
// the FP32 pre-activation for the backward pass), multiplies by GELU'(Z) in
// the backward pass, or adds split-K partial weight gradients atomically.
// Optimizer: AdamW on FP32 master weights, BF16 copies for the products.
#include "../../include/imfcs.h"
#include "imfcs_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <type_traits>
#include <string>
#include <vector>

namespace imfcs {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 80;        // 72 channels padded to a multiple of 16
constexpr int kHid = 256;
constexpr int kOut = 16;       // 5 outputs padded
constexpr int kNOut = 5;

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

enum { kEpiStore = 0, kEpiGelu = 1, kEpiGeluGrad = 2, kEpiAtomic = 3 };

// C[M][N] = op(A)[M][K] op(B)[K][N]; A(m, k) at A[m sam + k sak], B(k, n) at
// B[k sbk + n sbn] (strides express transposed operands).
struct GemmArgs {
    const uint16_t *A;
    int64_t sam, sak;
    const uint16_t *B;
    int64_t sbk, sbn;
    int M, N, K;
    int kslice;          // K range of one grid.z slice (multiple of 16)
    int epi;
    const float *bias;   // [N] (kEpiStore, kEpiGelu) or null
    float *Cf;           // FP32 output (pre-activation / value / gradient)
    int64_t ldcf;
    uint16_t *Cb;        // BF16 output (activation / value) or null
    int64_t ldcb;
    const float *Z;      // pre-activation for kEpiGeluGrad
    int64_t ldz;
};

__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs g) {
    // tiles staged as [row][k] and [col][k]: a lane's 4 consecutive k of one
    // row or column are one 8-byte LDS read (row pitch 20 elements = 40 B)
    __shared__ __attribute__((aligned(16))) uint16_t As[64][20];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[64][20];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int kbeg = blockIdx.z * g.kslice;
    const int kend = min(g.K, kbeg + g.kslice);
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int lr = tid >> 2, lk = (tid & 3) * 4;   // this thread's staging row and k offset
    for (int k0 = kbeg; k0 < kend; k0 += 16) {
        const int m = m0 + lr, n = n0 + lr;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = k0 + lk + j;
            As[lr][lk + j] = (m < g.M && k < kend) ? g.A[(int64_t)m * g.sam + (int64_t)k * g.sak] : (uint16_t)0;
            Bs[lr][lk + j] = (n < g.N && k < kend) ? g.B[(int64_t)k * g.sbk + (int64_t)n * g.sbn] : (uint16_t)0;
        }
        __syncthreads();
        // v_mfma_f32_16x16x16_bf16 operands: lane l holds row / column l % 16,
        // k = 4 (l / 16) .. 4 (l / 16) + 3
        bf16x4 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++) a[i] = *(const bf16x4 *)&As[wm + i * 16 + (lane & 15)][4 * (lane >> 4)];
#pragma unroll
        for (int j = 0; j < 2; j++) b[j] = *(const bf16x4 *)&Bs[wn + j * 16 + (lane & 15)][4 * (lane >> 4)];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    // result element r of tile (i, j): row 4 (l / 16) + r, column l % 16
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + i * 16 + 4 * (lane >> 4) + r;
                const int n = n0 + wn + j * 16 + (lane & 15);
                if (m >= g.M || n >= g.N) continue;
                float v = acc[i][j][r];
                if (g.epi == kEpiAtomic) {
                    atomicAdd(&g.Cf[(int64_t)m * g.ldcf + n], v);
                    continue;
                }
                if (g.epi == kEpiGeluGrad) {
                    v *= gelu_grad(g.Z[(int64_t)m * g.ldz + n]);
                    if (g.Cf) g.Cf[(int64_t)m * g.ldcf + n] = v;
                    if (g.Cb) g.Cb[(int64_t)m * g.ldcb + n] = f2bf(v);
                    continue;
                }
                if (g.bias) v += g.bias[n];
                if (g.Cf) g.Cf[(int64_t)m * g.ldcf + n] = v;
                if (g.Cb) g.Cb[(int64_t)m * g.ldcb + n] = f2bf(g.epi == kEpiGelu ? gelu(v) : v);
            }
}

// Loss sum_k w_k (y_k - t_k)^2 / B over the 5 outputs; dY (BF16, zero in the
// padding columns) and the loss (atomically added into *loss).
__global__ void k_mlp_loss(const float *Y, const float *T, int B, const float *w, uint16_t *dY, float *loss) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < B) {
        for (int k = 0; k < kOut; k++) {
            float d = 0.f;
            if (k < kNOut) {
                const float e = Y[(int64_t)i * kOut + k] - T[(int64_t)i * kNOut + k];
                l += w[k] * e * e;
                d = 2.f * w[k] * e / (float)B;
            }
            dY[(int64_t)i * kOut + k] = f2bf(d);
        }
    }
    // wave sum, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, l / (float)B);
}

// db[n] += sum_b dZ[b][n] (BF16 rows of length N): blockIdx.y takes rows
// [128 y, 128 y + 128), one thread per column, one atomic add per thread
__global__ void k_colsum(const uint16_t *dZ, int B, int N, float *db) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int b0 = blockIdx.y * 128, b1 = min(B, b0 + 128);
    float s = 0.f;
    for (int b = b0; b < b1; b++) s += bf2f(dZ[(int64_t)b * N + n]);
    atomicAdd(&db[n], s);
}

__global__ void k_fill_f32(float *p, int64_t n, float v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// AdamW (Loshchilov & Hutter, "Decoupled weight decay regularization",
// ICLR 2019) on the FP32 master parameters, and the BF16 copies of the
// weight matrices: [K][N] for the training products (copy_of: index of a
// parameter's copy, -1 for biases) and [N][K] for the inference kernel
// (copyT_of).
__global__ void k_adamw(float *p, const float *gr, float *m, float *v, int64_t n, float lr, float b1, float b2,
                        float eps, float wd, float c1, float c2, const int32_t *copy_of, uint16_t *pb,
                        const int32_t *copyT_of, uint16_t *pbT) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = gr[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const int32_t c = copy_of[i];
        float pi = p[i];
        pi -= lr * ((mi * c1) / (sqrtf(vi * c2) + eps) + (c >= 0 ? wd * pi : 0.f));
        p[i] = pi;
        if (c >= 0) {
            pb[c] = f2bf(pi);
            pbT[copyT_of[i]] = pb[c];
        }
    }
}


// Standardised network outputs -> physical values: N, D (um^2/s), D2, F2, G_inf.
__device__ __forceinline__ void decode(const float *y, const float *mu, const float *sd, float s, float *o) {
    float t[kNOut];
    for (int k = 0; k < kNOut; k++) t[k] = y[k] * sd[k] + mu[k];
    o[0] = expf(t[0]) / s;
    o[1] = expf(t[1]);
    o[2] = expf(t[2]);
    o[3] = fminf(fmaxf(t[3], 0.f), 1.f);
    o[4] = t[4] * s;
}



// The whole forward pass for kFusedRows curves per workgroup (4 wavefronts), for
// seeding fits and predictions: inputs normalised in registers, the three
// layers on v_mfma_f32_16x16x16_bf16 with the activations kept in LDS
// (BF16, as the training path stores them) and the weights read as [N][K]
// BF16 copies straight into the MFMA B operand (one 8-byte load per lane
// and step); the outputs decoded into LM starting rows (init, from the
// template tmpl) and/or physical values (out5).
constexpr int kFusedRows = 32;   // curves per workgroup (LDS ~42 KB: 3 workgroups per CU)

struct FusedArgs {
    const float *data;
    int n, npts;
    const uint16_t *W1T, *W2T, *W3T;
    const float *b1, *b2, *b3, *mu, *sd;
    int ncomp;
    const float *tmpl;
    float *init;   // [n][20] or null
    float *out5;   // [n][5] or null
};

__global__ void __launch_bounds__(256) k_mlp_fused(FusedArgs a) {
    __shared__ __attribute__((aligned(16))) uint16_t Xs[kFusedRows][kIn + 8];
    __shared__ __attribute__((aligned(16))) uint16_t H1s[kFusedRows][kHid + 8];
    __shared__ __attribute__((aligned(16))) uint16_t H2s[kFusedRows][kHid + 8];
    __shared__ float Ys[kFusedRows][kOut + 1];
    __shared__ float Ss[kFusedRows];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int r0 = blockIdx.x * kFusedRows;
    // inputs: wave w normalises rows w, w + 4, ...; lane l holds channels l
    // and l + 64; the loads of all its rows are issued first
    constexpr int RPW = kFusedRows / 4;
    float in0[RPW], in1[RPW];
#pragma unroll
    for (int q = 0; q < RPW; q++) {
        const int row = r0 + w + 4 * q;
        in0[q] = in1[q] = 0.f;
        if (row < a.n) {
            const float *d = a.data + (int64_t)row * a.npts;
            if (lane < a.npts) in0[q] = d[lane];
            if (lane + 64 < a.npts) in1[q] = d[lane + 64];
        }
    }
#pragma unroll
    for (int q = 0; q < RPW; q++) {
        const int rr = w + 4 * q, row = r0 + rr;
        const float v0 = in0[q], v1 = in1[q];
        float mx = fmaxf(fabsf(v0), fabsf(v1));
        int bad = !(isfinite(v0) && isfinite(v1));
        for (int off = 32; off > 0; off >>= 1) {
            mx = fmaxf(mx, __shfl_xor(mx, off, 64));
            bad |= __shfl_xor(bad, off, 64);
        }
        const bool ok = row < a.n && !bad && mx > 0.f;
        const float is = ok ? 1.f / mx : 0.f;
        Xs[rr][lane] = f2bf(v0 * is);
        if (lane < kIn - 64) Xs[rr][64 + lane] = f2bf(v1 * is);
        if (lane == 0) Ss[rr] = ok ? mx : 0.f;
    }
    __syncthreads();
    const int lr = lane & 15, lq = 4 * (lane >> 4);
    // hidden layers: wave w computes columns [64 w, 64 w + 64) of all rows;
    // K is a compile-time constant so the unrolled steps issue their weight
    // loads ahead of the MFMAs instead of one L2 round trip per step
    auto hidden = [&](auto kconst, const uint16_t(*In)[kHid + 8], const uint16_t *WT, const float *bias,
                      uint16_t(*Out)[kHid + 8], bool from_x) {
        constexpr int K = decltype(kconst)::value;
        constexpr int RT = kFusedRows / 16;   // row tiles
        f32x4 acc[RT][4];
#pragma unroll
        for (int i = 0; i < RT; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
        // every weight operand of the wave's columns is loaded first (K / 16
        // steps x 4 column tiles, 8 bytes each per lane): one wait for L2
        // instead of one per step
        bf16x4 bv[K / 16][4];
#pragma unroll
        for (int ks = 0; ks < K / 16; ks++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                bv[ks][j] = *(const bf16x4 *)&WT[(int64_t)(64 * w + j * 16 + lr) * K + 16 * ks + lq];
        __builtin_amdgcn_sched_barrier(0);   // keep the loads ahead of the products
#pragma unroll
        for (int ks = 0; ks < K / 16; ks++) {
            const int k0 = 16 * ks;
            bf16x4 av[RT];
#pragma unroll
            for (int i = 0; i < RT; i++)
                av[i] = from_x ? *(const bf16x4 *)&Xs[i * 16 + lr][k0 + lq] : *(const bf16x4 *)&In[i * 16 + lr][k0 + lq];
#pragma unroll
            for (int i = 0; i < RT; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[i], bv[ks][j], acc[i][j], 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < RT; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int col = 64 * w + j * 16 + lr;
                    Out[i * 16 + lq + r][col] = f2bf(gelu(acc[i][j][r] + bias[col]));
                }
    };
    hidden(std::integral_constant<int, kIn>{}, H2s, a.W1T, a.b1, H1s, true);
    __syncthreads();
    hidden(std::integral_constant<int, kHid>{}, H1s, a.W2T, a.b2, H2s, false);
    __syncthreads();
    // output layer: wave w < kFusedRows / 16 computes rows [16 w, 16 w + 16) x 16 outputs
    if (w < kFusedRows / 16) {
        f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k0 = 0; k0 < kHid; k0 += 16) {
            const bf16x4 av = *(const bf16x4 *)&H2s[16 * w + lr][k0 + lq];
            const bf16x4 bv = *(const bf16x4 *)&a.W3T[(int64_t)lr * kHid + k0 + lq];
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; r++) Ys[16 * w + lq + r][lr] = acc[r] + a.b3[lr];
    }
    __syncthreads();
    if (tid < kFusedRows) {
        const int row = r0 + tid;
        if (row >= a.n) return;
        const float s = Ss[tid];
        float o[kNOut];
        bool ok = s > 0.f;
        if (ok) {
            decode(&Ys[tid][0], a.mu, a.sd, s, o);
            ok = isfinite(o[0]) && o[0] > 0.f && isfinite(o[1]) && o[1] > 0.f && isfinite(o[4]) &&
                 (a.ncomp == 1 || (isfinite(o[2]) && o[2] > 0.f));
        }
        if (a.out5)
            for (int k = 0; k < kNOut; k++) a.out5[(int64_t)row * kNOut + k] = s > 0.f ? o[k] : __builtin_nanf("");
        if (a.init) {
            float rowv[20];
            for (int k = 0; k < 20; k++) rowv[k] = a.tmpl[k];
            if (ok) {
                rowv[0] = o[0];
                rowv[1] = o[1] * 1e-12f;
                rowv[4] = o[4];
                if (a.ncomp == 2) {
                    rowv[6] = o[2] * 1e-12f;
                    rowv[5] = o[3];
                }
            }
            for (int k = 0; k < 20; k++) a.init[(int64_t)row * 20 + k] = rowv[k];
        }
    }
}

inline unsigned cdivu(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

// Parameter block layout (FP32 master, and BF16 copies of the weights):
// W1 [kIn][kHid], b1 [kHid], W2 [kHid][kHid], b2 [kHid], W3 [kHid][kOut], b3 [kOut]
struct MlpNet {
    int device = 0;
    hipStream_t st = nullptr;
    int ncomp = 1, npts = 72;
    LMModelGeom geom{};
    std::vector<float> lag, nw;
    float tmu[kNOut], tsd[kNOut];
    uint64_t seed = 1;
    int64_t drawn = 0;        // curves generated so far (random stream position)
    int64_t adam_t = 0;
    static constexpr int64_t oW1 = 0, ob1 = oW1 + kIn * kHid, oW2 = ob1 + kHid, ob2 = oW2 + kHid * kHid,
                             oW3 = ob2 + kHid, ob3 = oW3 + kHid * kOut, nP = ob3 + kOut;
    static constexpr int64_t bW1 = 0, bW2 = bW1 + kIn * kHid, bW3 = bW2 + kHid * kHid, nB = bW3 + kHid * kOut;
    float *P = nullptr, *G = nullptr, *Mo = nullptr, *Vo = nullptr;
    uint16_t *Pb = nullptr, *PbT = nullptr;
    int32_t *copy_of = nullptr, *copyT_of = nullptr;
    float *consts = nullptr;   // lag[128], nw[128], tmu[5], tsd[5], lossw[5]
    float *loss = nullptr;     // [max steps]
    int loss_cap = 0;
    // activations for up to `cap` rows
    int cap = 0;
    uint16_t *X = nullptr, *H1 = nullptr, *H2 = nullptr, *dY = nullptr, *dZ2 = nullptr, *dZ1 = nullptr;
    float *Z1 = nullptr, *Z2 = nullptr, *Y = nullptr, *Tg = nullptr, *scale = nullptr;

    const float *d_lag() const { return consts; }
    const float *d_nw() const { return consts + 128; }
    const float *d_mu() const { return consts + 256; }
    const float *d_sd() const { return consts + 261; }
    const float *d_lw() const { return consts + 266; }

    hipError_t ensure_rows(int rows) {
        if (rows <= cap) return hipSuccess;
        free_rows();
        hipError_t e = hipSuccess;
        auto al = [&](void **p, size_t b) {
            if (e == hipSuccess) e = hipMalloc(p, b);
        };
        al((void **)&X, sizeof(uint16_t) * rows * kIn);
        al((void **)&H1, sizeof(uint16_t) * rows * kHid);
        al((void **)&H2, sizeof(uint16_t) * rows * kHid);
        al((void **)&dY, sizeof(uint16_t) * rows * kOut);
        al((void **)&dZ2, sizeof(uint16_t) * rows * kHid);
        al((void **)&dZ1, sizeof(uint16_t) * rows * kHid);
        al((void **)&Z1, sizeof(float) * rows * kHid);
        al((void **)&Z2, sizeof(float) * rows * kHid);
        al((void **)&Y, sizeof(float) * rows * kOut);
        al((void **)&Tg, sizeof(float) * rows * kNOut);
        al((void **)&scale, sizeof(float) * rows);
        if (e == hipSuccess) cap = rows;
        return e;
    }
    void free_rows() {
        void *ps[] = {X, H1, H2, dY, dZ2, dZ1, Z1, Z2, Y, Tg, scale};
        for (void *p : ps)
            if (p) (void)hipFree(p);
        X = H1 = H2 = dY = dZ2 = dZ1 = nullptr;
        Z1 = Z2 = Y = Tg = scale = nullptr;
        cap = 0;
    }
    ~MlpNet() {
        free_rows();
        void *ps[] = {P, G, Mo, Vo, Pb, PbT, copy_of, copyT_of, consts, loss};
        for (void *p : ps)
            if (p) (void)hipFree(p);
    }

    void gemm(const uint16_t *A, int64_t sam, int64_t sak, const uint16_t *B, int64_t sbk, int64_t sbn, int M,
              int N, int K, int epi, const float *bias, float *Cf, int64_t ldcf, uint16_t *Cb, int64_t ldcb,
              const float *Z = nullptr, int64_t ldz = 0, int ksplit = 1) const {
        GemmArgs g{A, sam, sak, B, sbk, sbn, M, N, K, 0, epi, bias, Cf, ldcf, Cb, ldcb, Z, ldz};
        g.kslice = ((K + ksplit - 1) / ksplit + 15) / 16 * 16;
        const unsigned kz = cdivu(K, g.kslice);
        hipLaunchKernelGGL(k_gemm_bf16, dim3(cdivu(N, 64), cdivu(M, 64), kz), dim3(256), 0, st, g);
    }

    // forward pass over rows of X: Z1/H1, Z2/H2, Y
    void forward(int rows) const {
        gemm(X, kIn, 1, Pb + bW1, kHid, 1, rows, kHid, kIn, kEpiGelu, P + ob1, Z1, kHid, H1, kHid);
        gemm(H1, kHid, 1, Pb + bW2, kHid, 1, rows, kHid, kHid, kEpiGelu, P + ob2, Z2, kHid, H2, kHid);
        gemm(H2, kHid, 1, Pb + bW3, kOut, 1, rows, kOut, kHid, kEpiStore, P + ob3, Y, kOut, nullptr, 0);
    }

    // loss and parameter gradients (G zeroed here) of rows of X against Tg
    void backward(int rows, float *loss_slot) const {
        hipLaunchKernelGGL(k_fill_f32, dim3(256), dim3(256), 0, st, G, (int64_t)nP, 0.f);
        hipLaunchKernelGGL(k_mlp_loss, dim3(cdivu(rows, 256)), dim3(256), 0, st, Y, Tg, rows, d_lw(), dY, loss_slot);
        const int ks = std::max(1, std::min(32, rows / 256));
        // layer 3: dW3 = H2^T dY, db3, dH2 = dY W3^T -> dZ2 = dH2 * GELU'(Z2)
        gemm(H2, 1, kHid, dY, kOut, 1, kHid, kOut, rows, kEpiAtomic, nullptr, G + oW3, kOut, nullptr, 0, nullptr,
             0, ks);
        hipLaunchKernelGGL(k_colsum, dim3(1, cdivu(rows, 128)), dim3(64), 0, st, dY, rows, kOut, G + ob3);
        gemm(dY, kOut, 1, Pb + bW3, 1, kOut, rows, kHid, kOut, kEpiGeluGrad, nullptr, nullptr, 0, dZ2, kHid, Z2,
             kHid);
        // layer 2
        gemm(H1, 1, kHid, dZ2, kHid, 1, kHid, kHid, rows, kEpiAtomic, nullptr, G + oW2, kHid, nullptr, 0, nullptr,
             0, ks);
        hipLaunchKernelGGL(k_colsum, dim3(kHid / 64, cdivu(rows, 128)), dim3(64), 0, st, dZ2, rows, kHid, G + ob2);
        gemm(dZ2, kHid, 1, Pb + bW2, 1, kHid, rows, kHid, kHid, kEpiGeluGrad, nullptr, nullptr, 0, dZ1, kHid, Z1,
             kHid);
        // layer 1
        gemm(X, 1, kIn, dZ1, kHid, 1, kIn, kHid, rows, kEpiAtomic, nullptr, G + oW1, kHid, nullptr, 0, nullptr, 0,
             ks);
        hipLaunchKernelGGL(k_colsum, dim3(kHid / 64, cdivu(rows, 128)), dim3(64), 0, st, dZ1, rows, kHid, G + ob1);
    }

    void curves(int rows) {
        CurveGenArgs a{};
        a.geom = geom;
        a.x = d_lag();
        a.nw = d_nw();
        a.npts = npts;
        a.ncomp = ncomp;
        if (ncomp == 1) {
            a.lnN[0] = std::log(0.05f);
            a.lnN[1] = std::log(500.f);
            a.lnD[0] = std::log(0.03f);
            a.lnD[1] = std::log(60.f);
        } else {
            a.lnN[0] = std::log(0.5f);
            a.lnN[1] = std::log(100.f);
            a.lnD[0] = std::log(0.1f);
            a.lnD[1] = std::log(2.f);
            a.lnD2[0] = std::log(1.f);
            a.lnD2[1] = std::log(40.f);
        }
        a.ginf = 0.05f;
        a.lg_noise[0] = -3.f;
        a.lg_noise[1] = -0.3f;
        a.tmu = d_mu();
        a.tsd = d_sd();
        a.seed = seed;
        a.base = drawn;
        a.X = X;
        a.ldx = kIn;
        a.Tg = Tg;
        a.n = rows;
        launch_mlp_curves(a, st);
        drawn += rows;
    }

    void fused(const float *data, int n, const float *tmpl, float *init, float *out5, hipStream_t s) const;

    void adamw(float lr) {
        adam_t++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f / (1.f - std::pow(b1, (float)adam_t)), c2 = 1.f / (1.f - std::pow(b2, (float)adam_t));
        hipLaunchKernelGGL(k_adamw, dim3(cdivu(nP, 256)), dim3(256), 0, st, P, G, Mo, Vo, (int64_t)nP, lr, b1, b2,
                           1e-8f, 1e-4f, c1, c2, copy_of, Pb, copyT_of, PbT);
    }
};

int mlp_points(const MlpNet *net) { return net->npts; }

void MlpNet::fused(const float *data, int n, const float *tmpl, float *init, float *out5, hipStream_t s) const {
    if (n <= 0) return;
    FusedArgs f{data, n, npts, PbT + bW1, PbT + bW2, PbT + bW3, P + ob1, P + ob2, P + ob3, d_mu(), d_sd(), ncomp,
                tmpl, init, out5};
    hipLaunchKernelGGL(k_mlp_fused, dim3(cdivu(n, kFusedRows)), dim3(256), 0, s, f);
}

int mlp_seed_init(MlpNet *net, const float *data, int n, const float *tmpl, float *init, hipStream_t st) {
    if (!net || n <= 0) return 0;
    net->fused(data, n, tmpl, init, nullptr, st);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : set_error(IMFCS_ERR_DEVICE, std::string("mlp seed: ") + hipGetErrorString(e));
}

}  // namespace imfcs

using namespace imfcs;

struct imfcs_mlp {
    MlpNet net;
};

MlpNet *imfcs::mlp_of(imfcs_mlp *m) { return m ? &m->net : nullptr; }


extern "C" {

imfcs_mlp *imfcs_mlp_create(imfcs_ctx *ctx, int ncomp, const imfcs_fit_model *geom, const double *lagtime,
                            int npts, const double *mtab, uint64_t seed) {
    if (!ctx || !geom || !lagtime || !mtab || (ncomp != 1 && ncomp != 2) || npts < 8 || npts > 72) {
        set_error(IMFCS_ERR_ARG, "mlp: bad arguments (1 or 2 components, 8..72 points)");
        return nullptr;
    }
    if (hipSetDevice(ctx_device(ctx)) != hipSuccess) {
        set_error(IMFCS_ERR_DEVICE, "mlp: hipSetDevice failed");
        return nullptr;
    }
    imfcs_mlp *m = new imfcs_mlp();
    MlpNet &n = m->net;
    n.device = ctx_device(ctx);
    n.st = ctx_stream(ctx);
    n.ncomp = ncomp;
    n.npts = npts;
    n.seed = seed;
    n.geom = LMModelGeom{(float)(geom->ax * 1e6), (float)(geom->ay * 1e6), (float)(geom->s * 1e6),
                         (float)(geom->sz * 1e6), (float)(geom->rx * 1e6), (float)(geom->ry * 1e6),
                         (float)(geom->fitobsvol * 1e12), (float)geom->q2, (float)geom->q3};
    // target standardisation from the sampling ranges
    const float mu[kNOut] = {-0.5f, ncomp == 1 ? 0.3f : -0.8f, ncomp == 1 ? 0.3f : 1.8f, ncomp == 1 ? 0.f : 0.5f, 0.f};
    const float sd[kNOut] = {0.5f, ncomp == 1 ? 2.2f : 0.9f, ncomp == 1 ? 2.2f : 1.1f, ncomp == 1 ? 1.f : 0.29f,
                             0.05f};
    std::vector<float> consts(512, 0.f);
    for (int k = 0; k < npts; k++) {
        consts[k] = (float)lagtime[k];
        consts[128 + k] = (float)std::sqrt(mtab[0] / std::max(1.0, mtab[k]));
    }
    for (int k = 0; k < kNOut; k++) {
        consts[256 + k] = n.tmu[k] = mu[k];
        consts[261 + k] = n.tsd[k] = sd[k];
        consts[266 + k] = (ncomp == 1 && (k == 2 || k == 3)) ? 0.f : 1.f;   // one component: D2, F2 untrained
    }
    // He-normal weights, zero biases (host RNG, seeded)
    std::vector<float> P(MlpNet::nP, 0.f);
    std::vector<int32_t> copy_of(MlpNet::nP, -1), copyT_of(MlpNet::nP, -1);
    std::mt19937_64 rng(seed * 0x9E3779B97F4A7C15ull + 7);
    auto fill = [&](int64_t off, int64_t boff, int rows, int cols, int used_rows, int used_cols) {
        std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (float)used_rows));
        for (int r = 0; r < rows; r++)
            for (int c = 0; c < cols; c++) {
                const int64_t i = off + (int64_t)r * cols + c;
                P[i] = (r < used_rows && c < used_cols) ? nd(rng) : 0.f;
                copy_of[i] = (int32_t)(boff + (int64_t)r * cols + c);
                copyT_of[i] = (int32_t)(boff + (int64_t)c * rows + r);
            }
    };
    fill(MlpNet::oW1, MlpNet::bW1, kIn, kHid, npts, kHid);
    fill(MlpNet::oW2, MlpNet::bW2, kHid, kHid, kHid, kHid);
    fill(MlpNet::oW3, MlpNet::bW3, kHid, kOut, kHid, kNOut);
    for (int64_t i = MlpNet::oW3; i < MlpNet::ob3; i++) P[i] *= 0.5f;
    std::vector<uint16_t> Pb(MlpNet::nB, 0), PbT(MlpNet::nB, 0);
    for (int64_t i = 0; i < MlpNet::nP; i++)
        if (copy_of[i] >= 0) {
            uint32_t u;
            std::memcpy(&u, &P[i], 4);
            u += 0x7fffu + ((u >> 16) & 1u);
            Pb[copy_of[i]] = PbT[copyT_of[i]] = (uint16_t)(u >> 16);
        }
    hipError_t e = hipSuccess;
    auto al = [&](void **p, size_t b) {
        if (e == hipSuccess) e = hipMalloc(p, b);
    };
    al((void **)&n.P, sizeof(float) * MlpNet::nP);
    al((void **)&n.G, sizeof(float) * MlpNet::nP);
    al((void **)&n.Mo, sizeof(float) * MlpNet::nP);
    al((void **)&n.Vo, sizeof(float) * MlpNet::nP);
    al((void **)&n.Pb, sizeof(uint16_t) * MlpNet::nB);
    al((void **)&n.copy_of, sizeof(int32_t) * MlpNet::nP);
    al((void **)&n.PbT, sizeof(uint16_t) * MlpNet::nB);
    al((void **)&n.copyT_of, sizeof(int32_t) * MlpNet::nP);
    al((void **)&n.consts, sizeof(float) * consts.size());
    if (e == hipSuccess) e = hipMemcpy(n.P, P.data(), sizeof(float) * P.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(n.Pb, Pb.data(), sizeof(uint16_t) * Pb.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(n.copy_of, copy_of.data(), sizeof(int32_t) * copy_of.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(n.PbT, PbT.data(), sizeof(uint16_t) * PbT.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(n.copyT_of, copyT_of.data(), sizeof(int32_t) * copyT_of.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(n.consts, consts.data(), sizeof(float) * consts.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(n.Mo, 0, sizeof(float) * MlpNet::nP);
    if (e == hipSuccess) e = hipMemset(n.Vo, 0, sizeof(float) * MlpNet::nP);
    if (e != hipSuccess) {
        set_error(IMFCS_ERR_MEMORY, std::string("mlp: ") + hipGetErrorString(e));
        delete m;
        return nullptr;
    }
    return m;
}

void imfcs_mlp_destroy(imfcs_mlp *m) {
    if (!m) return;
    (void)hipSetDevice(m->net.device);
    (void)hipStreamSynchronize(m->net.st);
    delete m;
}

int imfcs_mlp_train(imfcs_mlp *m, int steps, int batch, float lr, float *loss_out) {
    if (!m || steps < 1 || batch < 64 || batch > (1 << 16)) return set_error(IMFCS_ERR_ARG, "mlp train: bad arguments");
    MlpNet &n = m->net;
    if (hipSetDevice(n.device) != hipSuccess) return set_error(IMFCS_ERR_DEVICE, "mlp: hipSetDevice failed");
    hipError_t e = n.ensure_rows(batch);
    if (e == hipSuccess && n.loss_cap < steps) {
        if (n.loss) (void)hipFree(n.loss);
        n.loss = nullptr;
        e = hipMalloc(&n.loss, sizeof(float) * steps);
        n.loss_cap = e == hipSuccess ? steps : 0;
    }
    if (e != hipSuccess) return set_error(IMFCS_ERR_MEMORY, std::string("mlp train: ") + hipGetErrorString(e));
    hipLaunchKernelGGL(k_fill_f32, dim3(cdivu(steps, 256)), dim3(256), 0, n.st, n.loss, (int64_t)steps, 0.f);
    for (int s = 0; s < steps; s++) {
        // cosine decay of the learning rate to 5 % over the run
        const float f = 0.05f + 0.95f * 0.5f * (1.f + std::cos(3.14159265f * (float)s / (float)steps));
        n.curves(batch);
        n.forward(batch);
        n.backward(batch, n.loss + s);
        n.adamw(lr * f);
    }
    e = hipGetLastError();
    if (e == hipSuccess && loss_out) e = hipMemcpyAsync(loss_out, n.loss, sizeof(float) * steps, hipMemcpyDeviceToHost, n.st);
    if (e == hipSuccess) e = hipStreamSynchronize(n.st);
    return e == hipSuccess ? 0 : set_error(IMFCS_ERR_DEVICE, std::string("mlp train: ") + hipGetErrorString(e));
}

int imfcs_mlp_param_count(void) { return (int)MlpNet::nP; }

int imfcs_mlp_target_scale(imfcs_mlp *m, float *mu, float *sd) {
    if (!m || !mu || !sd) return set_error(IMFCS_ERR_ARG, "mlp: null argument");
    for (int k = 0; k < kNOut; k++) {
        mu[k] = m->net.tmu[k];
        sd[k] = m->net.tsd[k];
    }
    return 0;
}

int imfcs_mlp_get_params(imfcs_mlp *m, float *params) {
    if (!m || !params) return set_error(IMFCS_ERR_ARG, "mlp: null argument");
    MlpNet &n = m->net;
    hipError_t e = hipStreamSynchronize(n.st);
    if (e == hipSuccess) e = hipMemcpy(params, n.P, sizeof(float) * MlpNet::nP, hipMemcpyDeviceToHost);
    return e == hipSuccess ? 0 : set_error(IMFCS_ERR_DEVICE, hipGetErrorString(e));
}

int imfcs_mlp_gradient(imfcs_mlp *m, const float *inputs, const float *targets, int rows, float *grad,
                       float *outputs, float *loss) {
    if (!m || !inputs || !targets || rows < 1 || rows > (1 << 16)) return set_error(IMFCS_ERR_ARG, "mlp gradient: bad arguments");
    MlpNet &n = m->net;
    if (hipSetDevice(n.device) != hipSuccess) return set_error(IMFCS_ERR_DEVICE, "mlp: hipSetDevice failed");
    hipError_t e = n.ensure_rows(rows);
    if (e == hipSuccess && n.loss_cap < 1) {
        e = hipMalloc(&n.loss, sizeof(float));
        n.loss_cap = e == hipSuccess ? 1 : 0;
    }
    if (e != hipSuccess) return set_error(IMFCS_ERR_MEMORY, hipGetErrorString(e));
    std::vector<uint16_t> xb((size_t)rows * kIn);
    for (size_t i = 0; i < xb.size(); i++) {
        uint32_t u;
        std::memcpy(&u, &inputs[i], 4);
        u += 0x7fffu + ((u >> 16) & 1u);
        xb[i] = (uint16_t)(u >> 16);
    }
    e = hipMemcpyAsync(n.X, xb.data(), sizeof(uint16_t) * xb.size(), hipMemcpyHostToDevice, n.st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(n.Tg, targets, sizeof(float) * rows * kNOut, hipMemcpyHostToDevice, n.st);
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, n.st, n.loss, (int64_t)1, 0.f);
    n.forward(rows);
    n.backward(rows, n.loss);
    if (e == hipSuccess) e = hipGetLastError();
    std::vector<float> y((size_t)rows * kOut);
    if (e == hipSuccess && grad) e = hipMemcpyAsync(grad, n.G, sizeof(float) * MlpNet::nP, hipMemcpyDeviceToHost, n.st);
    if (e == hipSuccess && outputs) e = hipMemcpyAsync(y.data(), n.Y, sizeof(float) * y.size(), hipMemcpyDeviceToHost, n.st);
    if (e == hipSuccess && loss) e = hipMemcpyAsync(loss, n.loss, sizeof(float), hipMemcpyDeviceToHost, n.st);
    if (e == hipSuccess) e = hipStreamSynchronize(n.st);
    if (e != hipSuccess) return set_error(IMFCS_ERR_DEVICE, std::string("mlp gradient: ") + hipGetErrorString(e));
    if (outputs)
        for (int i = 0; i < rows; i++)
            for (int k = 0; k < kNOut; k++) outputs[(size_t)i * kNOut + k] = y[(size_t)i * kOut + k];
    return 0;
}

int imfcs_mlp_curves(imfcs_mlp *m, int rows, float *inputs, float *targets) {
    if (!m || !inputs || !targets || rows < 1 || rows > (1 << 16)) return set_error(IMFCS_ERR_ARG, "mlp curves: bad arguments");
    MlpNet &n = m->net;
    if (hipSetDevice(n.device) != hipSuccess) return set_error(IMFCS_ERR_DEVICE, "mlp: hipSetDevice failed");
    hipError_t e = n.ensure_rows(rows);
    if (e != hipSuccess) return set_error(IMFCS_ERR_MEMORY, hipGetErrorString(e));
    n.curves(rows);
    std::vector<uint16_t> xb((size_t)rows * kIn);
    e = hipMemcpyAsync(xb.data(), n.X, sizeof(uint16_t) * xb.size(), hipMemcpyDeviceToHost, n.st);
    if (e == hipSuccess) e = hipMemcpyAsync(targets, n.Tg, sizeof(float) * rows * kNOut, hipMemcpyDeviceToHost, n.st);
    if (e == hipSuccess) e = hipStreamSynchronize(n.st);
    if (e != hipSuccess) return set_error(IMFCS_ERR_DEVICE, hipGetErrorString(e));
    for (size_t i = 0; i < xb.size(); i++) {
        const uint32_t u = (uint32_t)xb[i] << 16;
        std::memcpy(&inputs[i], &u, 4);
    }
    return 0;
}

int imfcs_mlp_predict(imfcs_mlp *m, const float *d_curves, int rows, float *d_out) {
    if (!m || !d_curves || !d_out || rows < 1) return set_error(IMFCS_ERR_ARG, "mlp predict: bad arguments");
    MlpNet &n = m->net;
    if (hipSetDevice(n.device) != hipSuccess) return set_error(IMFCS_ERR_DEVICE, "mlp: hipSetDevice failed");
    n.fused(d_curves, rows, nullptr, nullptr, d_out, n.st);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(n.st);
    return e == hipSuccess ? 0 : set_error(IMFCS_ERR_DEVICE, std::string("mlp predict: ") + hipGetErrorString(e));
}

}  // extern "C"
