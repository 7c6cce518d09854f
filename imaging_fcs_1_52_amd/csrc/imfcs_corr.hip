// Multi-tau correlator with blocked variance (calculateCF, IFCS:9425-9521),
// LDS-staged and wave-specialised for CDNA4.
//
// k_corr_main: one workgroup per (64 output pixels, time chunk).  Loader
//   waves (one per staged half: a series, partner b series) read the
//   bleach-corrected FP32 deviations X and stage 64 frames in LDS, together
//   with the pair-summed level-1 and level-2 values and 16-frame group sums;
//   the level-3 (8-frame) sums go to global memory for k_corr_high.  The
//   LDS tile is double-buffered: the loaders fetch tile t+2 and store tile
//   t+1 while the role waves work on tile t (one barrier per tile).  Five
//   role waves per correlation type consume the tile:
//     level 0, delays 0-4 / 5-9 / 10-16
//     level 1, delays 9-16
//     level 2, delays 9-16
//   Each lane keeps a 32-value register window of its level and processes
//   products in groups of 16 a-indices; block sums of the blocked variance
//   (2^(ind-level) products, multiples of 16) are closed at group ends.
// k_corr_high: levels 3..q-1 from the level-3 series, one thread per pair.
// k_corr_finalize: combines the per-chunk sums into G and the variance.
//
// With the series shifted by its slot's shift mu, a~ = a - mu, every channel's exact
// window means enter through alpha = abar - mu_c, beta = bbar - mu_c:
//   sum_y (a_y - abar)(b_{y+d} - bbar) = sum a~ b~ - n alpha beta
//   block sum S_j = Sab - beta Sa - alpha Sb + B alpha beta
#include "imfcs_device.h"
#include "imfcs_internal.h"

namespace imfcs {

constexpr int kTileF = 64;     // frames per LDS tile
constexpr int kMainChan = 33;  // channels of levels 0..2 (p = 16)
constexpr int kNS = 128;       // staged series per tile (a and partner b)

// The level >= 3 arrays: level 3 + k at (k == 0 ? l3 : lvd + off[k]), numbin[k] rows
struct LevelPads {
    int nlev, nslots;
    int numbin[kMaxLevel];
    int64_t off[kMaxLevel];
};

struct MainArgs {
    SeriesGeom g;
    BleachDesc bd;
    const CorrTables *ctp;
    PairDesc pd;
    const double *F;
    const float *X;
    const int32_t *ind;
    int chunk_len, nchunk;
    int has_acf, has_ccf;
    float *part;    // [chunk][pair][kMainChan][3] : cross, A1, A2
    float *l3;      // level-3 values [numbin_3][nslots]
    float *lvd;     // levels >= 4, level c at lvd + loff[c]: [numbin_c][nslots]
    int64_t loff[kMaxLevel];
    // GLS: every block sum of the blocked variance, [pair][block][channel-1]
    // (channels 1..chanum-1); null when no covariance is requested
    float *gls;
    int gls_nblk;
};

// Level rows stored per series ([series][value], padded to 4 mod 32 words
// so that 8 lanes' 16-byte accesses cover 32 distinct banks): a lane reads
// a group's 16 values with four 16-byte loads and the loader writes them
// the same way.
constexpr int kL0Ld = kTileF + 4, kL1Ld = kTileF / 2 + 4, kL2Ld = kTileF / 4 + 4;
struct Tile {
    float l0[kNS][kL0Ld];
    float l1[kNS][kL1Ld];
    float l2[kNS][kL2Ld];
    // sums of each 16-frame block's staged values on levels 0, 1 and 2
    // (16, 8 and 4 values): the group sums of the role waves
    float gs0[kTileF / 16][kNS];
    float gs1[kTileF / 16][kNS];
    float gs2[kTileF / 16][kNS];
};

// Shift of series s at level c (2^c frames per value).
__device__ __forceinline__ double level_shift(const MainArgs &A, int s, int c) {
    return slot_mu(A.bd.coef, A.bd.order, A.g.slot_of[s]) * (double)(1 << c);
}

// Loader wave: one staged series per lane, all 64 frames of a tile.  The
// next tile is fetched into registers while the role waves work on the
// current one, then written to the other LDS buffer.
struct TileLoader {
    bool own;   // this series writes its slot's level-3 values
    int sl, slot;
    int nb3;    // level-3 values (held in a register: the level-3 stores
                // would otherwise force a reload and a wait on every store)
    float xr[kTileF];
    __device__ void init(const MainArgs &A, int half, int lane, int p0) {
        sl = half * 64 + lane;
        int p = p0 + lane;
        if (p >= A.g.npix) p = A.g.npix - 1;   // padding lanes duplicate the last pixel
        const int s = half ? A.g.npix + p : p;
        slot = A.g.slot_of[s];
        own = A.g.slot_own[s] != 0;
        nb3 = A.ctp->numbin[3];
    }
    // frames f0 .. f0+63 (rows past the end of the series read the last
    // row; store() zeroes them).  Nothing consumes the values before the
    // next tile's store, so all 64 loads are in flight together.  A loader
    // wave issues every instruction of its tile, so the interior tiles use
    // buffer loads: one lane offset for all rows and one scalar offset per
    // row, instead of a 64-bit address per load.
    __device__ void fetch(const MainArgs &A, int f0) {
        const int T = A.g.T;
        const int64_t ld = A.g.nslots;
        if (f0 + kTileF <= T) {   // wave-uniform
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                (void *)(A.X + (int64_t)f0 * ld), (short)0, (int)(kTileF * ld * 4), kBufferWord3);
            const int lo = slot * 4, rowb = (int)ld * 4;
#pragma unroll
            for (int k = 0; k < kTileF; k++)   // (the builtin returns the 32 bits as an integer)
                xr[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, lo, k * rowb, 0));
            return;
        }
#pragma unroll
        for (int k = 0; k < kTileF; k++) {
            const float *row = A.X + (int64_t)min(f0 + k, T - 1) * ld;   // wave-uniform
            xr[k] = row[(uint32_t)slot];
        }
    }
    // level 0..2 rows and group sums into LDS, level-3 sums of the chunk's
    // own frames to global memory; values past the end of the series are
    // zero (and incomplete pairs on levels 1, 2), checked only in the tiles
    // that reach the end
    __device__ void store(const MainArgs &A, Tile &tl, int f0, int chunk_end) const {
        if (f0 + kTileF <= A.g.T)   // wave-uniform
            store_t<false>(A, tl, f0, chunk_end);
        else
            store_t<true>(A, tl, f0, chunk_end);
    }
    // Pair sums in FP32 (each level-1 value is the correctly rounded sum
    // of its two frames; levels 2 and 3 are sums of rounded pairs, within
    // 2 ulp of the exact sums): the loader wave issues every instruction
    // of its tile, so it avoids FP64 conversions.  Interior 16-frame blocks
    // have one group sum on all three levels; in the last tile each level
    // sums its own (masked) values.
    template <bool TAIL>
    __device__ __forceinline__ void store_t(const MainArgs &A, Tile &tl, int f0, int chunk_end) const {
        const int T = A.g.T;
#pragma unroll
        for (int fb = 0; fb < kTileF / 16; fb++) {
            const int fbase = f0 + 16 * fb;
            float x[16], u1[8], u2[4];
#pragma unroll
            for (int k = 0; k < 16; k++) {
                x[k] = (!TAIL || fbase + k < T) ? xr[16 * fb + k] : 0.f;
                tl.l0[sl][16 * fb + k] = x[k];
            }
#pragma unroll
            for (int k = 0; k < 8; k++) {
                u1[k] = x[2 * k] + x[2 * k + 1];
                tl.l1[sl][8 * fb + k] = (!TAIL || fbase + 2 * k + 1 < T) ? u1[k] : 0.f;
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                u2[k] = u1[2 * k] + u1[2 * k + 1];
                tl.l2[sl][4 * fb + k] = (!TAIL || fbase + 4 * k + 3 < T) ? u2[k] : 0.f;
            }
            const float s8a = u2[0] + u2[1], s8b = u2[2] + u2[3];
            if (own && fbase < chunk_end) {
                const int y3 = fbase / 8;   // complete 8-frame sums only (y3 < T / 8)
                if (y3 < nb3) A.l3[(int64_t)y3 * A.g.nslots + slot] = s8a;
                if (y3 + 1 < nb3) A.l3[(int64_t)(y3 + 1) * A.g.nslots + slot] = s8b;
            }
            if (!TAIL) {
                const float g = s8a + s8b;
                tl.gs0[fb][sl] = g;
                tl.gs1[fb][sl] = g;
                tl.gs2[fb][sl] = g;
            } else {
                float g0 = 0.f, g1 = 0.f, g2 = 0.f;
#pragma unroll
                for (int k = 0; k < 16; k++) g0 += x[k];
#pragma unroll
                for (int k = 0; k < 8; k++) g1 += (fbase + 2 * k + 1 < T) ? u1[k] : 0.f;
#pragma unroll
                for (int k = 0; k < 4; k++) g2 += (fbase + 4 * k + 3 < T) ? u2[k] : 0.f;
                tl.gs0[fb][sl] = g0;
                tl.gs1[fb][sl] = g1;
                tl.gs2[fb][sl] = g2;
            }
        }
    }
};

// Tiles of one workgroup's chunk: its own frames plus two tiles beyond, so
// that the lag windows of the last groups are complete.
__device__ __forceinline__ int chunk_tiles(const MainArgs &A, int chunk, int &cstart, int &cend) {
    cstart = chunk * A.chunk_len;
    cend = min(A.g.T, cstart + A.chunk_len);
    return (cend - cstart + kTileF - 1) / kTileF + 2;
}

// Loader waves: tile t+1 goes to LDS buffer (t+1)&1 while the roles read
// buffer t&1; one barrier per tile (the same count as the role waves).
__device__ void run_loader(const MainArgs &A, Tile *tl2, int half, int lane, int p0, int chunk) {
    TileLoader L;
    L.init(A, half, lane, p0);
    int cstart, cend;
    const int ntiles = chunk_tiles(A, chunk, cstart, cend);
    L.fetch(A, cstart);
    L.store(A, tl2[0], cstart, cend);
    L.fetch(A, cstart + kTileF);   // ntiles >= 3
    __syncthreads();
    for (int t = 0; t < ntiles; t++) {
        if (t + 1 < ntiles) {
            L.store(A, tl2[(t + 1) & 1], cstart + (t + 1) * kTileF, cend);
            if (t + 2 < ntiles) L.fetch(A, cstart + (t + 2) * kTileF);
        }
        __syncthreads();
    }
}

// Per-lag sums of one chunk.  A chunk spans at most 2^(blocknum-1) frames,
// i.e. at most 32 blocks of the blocked variance, so FP32 sums stay within
// ~1e-6 relative; chunks are combined in FP64 by k_corr_finalize.
template <int NL>
struct LagAcc {
    float sab[NL], head[NL], alpha[NL], beta[NL];
    float a1[NL], a2[NL], cross[NL];
};

// Sum of the 16 values of group q of a tile on level LEVEL, series s.
template <int LEVEL>
__device__ __forceinline__ float group_sum(const Tile &tl, int q, int s) {
    if (LEVEL == 0) return tl.gs0[q][s];
    if (LEVEL == 1) return tl.gs1[2 * q][s] + tl.gs1[2 * q + 1][s];
    return (tl.gs2[0][s] + tl.gs2[1][s]) + (tl.gs2[2][s] + tl.gs2[3][s]);
}

// One role: NL delays starting at DMIN on level LEVEL (group size 16).
template <int NL, int DMIN, int LEVEL, bool CCF>
__device__ void run_role(const MainArgs &A, const Tile *tl2, int lane, int p0, int chunk) {
    constexpr int type = CCF ? 1 : 0;
    constexpr int GPT = kTileF / (16 << LEVEL);   // groups per tile
    const CorrTables &ct = *A.ctp;
    const int p = p0 + lane;
    const bool live = (p < A.g.npix) && (type == 0 ? A.has_acf : A.has_ccf);
    constexpr bool ccf = CCF;
    const int pair = type == 0 ? p : (A.has_acf ? A.g.npix + p : p);
    const int sa = p < A.g.npix ? p : A.g.npix - 1;
    const int sb = ccf ? A.g.npix + sa : sa;
    const int asl = lane, bsl = ccf ? 64 + lane : lane;
    LagAcc<NL> acc;
    // block sums of a and b over the current blocking block; the b sum at
    // delay d is sb_blk + sum_{k<d} b[ye+k] - sum_{k<d} b[ys+k] for the block
    // [ys, ye), the head term kept per delay from the previous block end
    float sa_blk = 0.f, sb_blk = 0.f;
    int B = 1, lb = 0, pnum = 0;   // block length B = 2^lb values
    {
        const double *Fa = A.F + (int64_t)A.g.slot_of[sa] * ct.npos, *Fb = A.F + (int64_t)A.g.slot_of[sb] * ct.npos;
        double mua = level_shift(A, sa, LEVEL);
        double mub = level_shift(A, sb, LEVEL);
        const int x0 = ct.first_chan[LEVEL] + (LEVEL == 0 ? DMIN : DMIN - 9);
#pragma unroll
        for (int i = 0; i < NL; i++) {
            int x = x0 + i;
            double n = (double)(ct.numbin[LEVEL] - ct.del[x]);
            acc.alpha[i] = (float)(Fa[ct.pos_a_end[x]] / n - mua);
            acc.beta[i] = (float)((Fb[ct.pos_b_end[x]] - Fb[ct.pos_b_beg[x]]) / n - mub);
            acc.sab[i] = acc.head[i] = 0.f;
            acc.a1[i] = acc.a2[i] = acc.cross[i] = 0.f;
        }
        int indp = live ? A.ind[pair] : ct.q - 1;
        lb = indp - LEVEL;
        B = 1 << lb;
        pnum = (int)floor(ct.mtab_last / pow(2.0, fmax(indp - ct.log2_samp_last, 0.0)));
    }
    int cstart, cend;
    const int ntiles = chunk_tiles(A, chunk, cstart, cend);
    const bool last_chunk = (chunk == A.nchunk - 1);
    const int ystart = cstart >> LEVEL;
    const int yend = last_chunk ? ct.numbin[LEVEL] : (cend >> LEVEL);
    float w[32], ga_cur[CCF ? 16 : 1], ga_nxt[CCF ? 16 : 1];
    float gsa_cur = 0.f, gsa_nxt = 0.f, gsb_cur = 0.f, gsb_nxt = 0.f;
#pragma unroll
    for (int k = 0; k < 32; k++) w[k] = 0.f;
    if (CCF) {
#pragma unroll
        for (int k = 0; k < 16; k++) ga_cur[k] = ga_nxt[k] = 0.f;
    }
    // GLS block sums of this role's channels: row [pair][block], column
    // channel - 1 (channel 0 has none); hoisted so that the stores inside
    // the loop force no reloads of the tables
    const int xch0 = ct.first_chan[LEVEL] + (LEVEL == 0 ? DMIN : DMIN - 9);
    const int gls_ld = ct.chanum - 1;
    float *const gls_row = A.gls ? A.gls + (int64_t)pair * A.gls_nblk * gls_ld : nullptr;
    int ynext = ystart;   // level index of the next value to load
    for (int t = 0; t < ntiles; t++) {
        const Tile &tl = tl2[t & 1];
        // unrolled by two: the window shift becomes register renaming, with
        // half the code of a full unroll (the role loops of one workgroup
        // share the CU's instruction cache)
#pragma unroll 2
        for (int q = 0; q < GPT; q++) {
            // shift in 16 new level values
#pragma unroll
            for (int k = 0; k < 16; k++) {
                w[k] = w[k + 16];
                if (CCF) ga_cur[k] = ga_nxt[k];
            }
            gsb_cur = gsb_nxt;
            if (CCF) gsa_cur = gsa_nxt;
            const float *rb = (LEVEL == 0) ? &tl.l0[bsl][16 * q] : (LEVEL == 1) ? &tl.l1[bsl][16 * q] : &tl.l2[bsl][16 * q];
            const float *ra = (LEVEL == 0) ? &tl.l0[asl][16 * q] : (LEVEL == 1) ? &tl.l1[asl][16 * q] : &tl.l2[asl][16 * q];
#pragma unroll
            for (int k = 0; k < 16; k += 4) {
                const float4 vb = *(const float4 *)(rb + k);
                w[16 + k] = vb.x;
                w[17 + k] = vb.y;
                w[18 + k] = vb.z;
                w[19 + k] = vb.w;
                if (CCF) {
                    const float4 va = *(const float4 *)(ra + k);
                    ga_nxt[k] = va.x;
                    ga_nxt[k + 1] = va.y;
                    ga_nxt[k + 2] = va.z;
                    ga_nxt[k + 3] = va.w;
                }
            }
            gsb_nxt = group_sum<LEVEL>(tl, q, bsl);
            if (CCF) gsa_nxt = group_sum<LEVEL>(tl, q, asl);
            ynext += 16;
            const int y0 = ynext - 32;   // group whose window is complete
            if (live && y0 >= ystart && y0 < yend) {
                const float *ga = CCF ? ga_cur : w;
                if (y0 == ystart) {
                    // the chunk starts a block: head terms sum_{k<d} b[ystart+k]
                    float h = 0.f;
#pragma unroll
                    for (int k = 0; k < DMIN + NL; k++) {
                        if (k >= DMIN) acc.head[k - DMIN] = h;
                        if (k + 1 < DMIN + NL) h += w[k];
                    }
                }
                sa_blk += CCF ? gsa_cur : gsb_cur;
                if (CCF) sb_blk += gsb_cur;
#pragma unroll
                for (int i = 0; i < NL; i++) {
                    const int d = DMIN + i;
                    // the group's 16 products summed on their own, then added
                    // to the block sum once (one rounding at the block sum's
                    // magnitude per group)
                    float s0 = 0.f;
#pragma unroll
                    for (int k = 0; k < 16; k++) s0 = fmaf(ga[k], w[k + d], s0);
                    acc.sab[i] += s0;
                }
                const int ye = y0 + 16;
                if ((ye & (B - 1)) == 0) {
                    const int blk = (ye >> lb) - 1;   // ye >= 16 > 0
                    const bool inblk = blk < pnum;
                    // tail terms sum_{k<d} b[ye+k]: the next group's values
                    float tail[NL];
                    {
                        float run = 0.f;
#pragma unroll
                        for (int k = 0; k < DMIN + NL; k++) {
                            if (k >= DMIN) tail[k - DMIN] = run;
                            if (k + 1 < DMIN + NL) run += w[16 + k];
                        }
                    }
                    const float sbb = CCF ? sb_blk : sa_blk;
#pragma unroll
                    for (int i = 0; i < NL; i++) {
                        const float sb = sbb + (tail[i] - acc.head[i]);
                        acc.head[i] = tail[i];
                        float S = acc.sab[i] - acc.beta[i] * sa_blk - acc.alpha[i] * sb +
                                  (float)B * acc.alpha[i] * acc.beta[i];
                        if (inblk) {
                            acc.a1[i] += S;
                            acc.a2[i] = fmaf(S, S, acc.a2[i]);
                            if (gls_row && xch0 + i > 0) gls_row[(int64_t)blk * gls_ld + xch0 + i - 1] = S;
                        }
                        acc.cross[i] += acc.sab[i];
                        acc.sab[i] = 0.f;
                    }
                    sa_blk = 0.f;
                    sb_blk = 0.f;
                }
            }
        }
        __syncthreads();
    }
    if (live) {
        const int x0 = ct.first_chan[LEVEL] + (LEVEL == 0 ? DMIN : DMIN - 9);
        float *out = A.part + ((int64_t)chunk * A.pd.npairs + pair) * kMainChan * 3;
#pragma unroll
        for (int i = 0; i < NL; i++) {
            out[(x0 + i) * 3 + 0] = acc.cross[i] + acc.sab[i];
            out[(x0 + i) * 3 + 1] = acc.a1[i];
            out[(x0 + i) * 3 + 2] = acc.a2[i];
        }
    }
}

// 12 waves with partner series (10 roles + 2 loaders), 6 without (5 + 1).
// Waves go to the four SIMDs round-robin, so wave w shares its SIMD with
// w +- 4 and w +- 8; the role order below gives every SIMD about the same
// work per tile: two 5-delay level-0 roles and a level-2 role on two SIMDs;
// the 7-delay level-0 role, a level-1 role and a loader (few arithmetic
// instructions) on the other two.  Each role has one call site, so that
// every role is inlined.
__global__ void __launch_bounds__(768) k_corr_main(MainArgs A) {
    __shared__ __attribute__((aligned(16))) Tile tl2[2];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int p0 = blockIdx.x * 64;
    const int chunk = blockIdx.y;
    const bool pairs2 = (blockDim.x >> 6) == 12;   // partner series staged
    const int nrole = pairs2 ? 10 : 5;
    if (wave >= nrole) {
        run_loader(A, tl2, wave - nrole, lane, p0, chunk);
        return;
    }
    // without partners: waves 0..4 take roles 0, 4, 2, 6, 8
    const int role = pairs2 ? wave : (wave == 1 ? 4 : wave == 3 ? 6 : wave == 4 ? 8 : wave);
    __syncthreads();
    switch (role) {
        case 0: run_role<5, 0, 0, false>(A, tl2, lane, p0, chunk); break;
        case 1: run_role<5, 0, 0, true>(A, tl2, lane, p0, chunk); break;
        case 2: run_role<7, 10, 0, false>(A, tl2, lane, p0, chunk); break;
        case 3: run_role<7, 10, 0, true>(A, tl2, lane, p0, chunk); break;
        case 4: run_role<5, 5, 0, false>(A, tl2, lane, p0, chunk); break;
        case 5: run_role<5, 5, 0, true>(A, tl2, lane, p0, chunk); break;
        case 6: run_role<8, 9, 1, false>(A, tl2, lane, p0, chunk); break;
        case 7: run_role<8, 9, 1, true>(A, tl2, lane, p0, chunk); break;
        case 8: run_role<8, 9, 2, false>(A, tl2, lane, p0, chunk); break;
        default: run_role<8, 9, 2, true>(A, tl2, lane, p0, chunk); break;
    }
}

// ------------------------------------------------------------------
// Levels 3..q-1 from the level-3 series.
// ------------------------------------------------------------------
// Values of one level >= 3 (stored as FP32 sums of 2^c frames, computed
// from there on in FP64).  Branch-free (clamped index, zero past the end)
// so that a group's loads issue back to back.
struct LevelCtx {
    const float *lv;   // this level's values [numbin + kLevelPad][nslots], zero past numbin
    int nslots, sa, sb, c, numbin;   // sa, sb: slots of the pair's series
    // y < numbin + kLevelPad: no clamp and no select, so that nothing uses a
    // loaded value before the group that needs it (the loads stay in flight)
    __device__ __forceinline__ float val(int s, int y) const { return lv[(int64_t)y * nslots + s]; }
};

// Zero rows [numbin_c, numbin_c + kLevelPad) of every level >= 3 array.
__global__ void k_zero_level_pads(float *l3, float *lvd, LevelPads P) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)kLevelPad * P.nslots;
    const int lvl = (int)blockIdx.y;
    if (i >= per || lvl >= P.nlev) return;
    float *base = lvl == 0 ? l3 : lvd + P.off[lvl];
    base[(int64_t)P.numbin[lvl] * P.nslots + i] = 0.f;
}

// Levels c0+1 .. c0+4 from level c0 by the pairwise-sum rebinning
// (IFCS:9477-9478): v_c[y] = v_{c-1}[2y] + v_{c-1}[2y+1], summed in FP64
// and stored as FP32 (as level 3).  One
// thread per (slot, 16 consecutive level-c0 values): it reads the 16
// values once and writes the 8 + 4 + 2 + 1 values they produce on the next
// four levels (levels past nlev, and values past each level's numbin, are
// not written).
struct RebinArgs {
    const float *src;      // level c0 (l3, or lvd + loff[c0])
    float *dst[4];         // lvd + loff[c0 + 1 + k]
    int numbin_src;
    int numbin[4];         // numbin of levels c0+1 .. c0+4 (0 = level absent)
    int nslots;
};

__global__ void k_corr_rebin4(RebinArgs R) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ngrp = (R.numbin_src + 15) / 16;
    if (i >= ngrp * R.nslots) return;
    const int64_t grp = i / R.nslots, s = i - grp * R.nslots;
    double v[16];
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const int64_t y = min(grp * 16 + k, (int64_t)R.numbin_src - 1);
        v[k] = (double)R.src[y * R.nslots + s];
    }
    int n = 16;
#pragma unroll
    for (int lv = 0; lv < 4; lv++) {
        n >>= 1;
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (k < n) v[k] = v[2 * k] + v[2 * k + 1];
        if (R.numbin[lv] > 0) {
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int64_t y = grp * n + k;
                if (k < n && y < R.numbin[lv]) R.dst[lv][y * R.nslots + s] = (float)v[k];
            }
        }
    }
}

// Values [ybeg, yend) of one level; yend - ybeg and ybeg are multiples of G
// and of the block length B, so every block of the blocked variance lies in
// one call.  The window holds the stored FP32 values; each group's 16
// products per delay are summed in FP32 (about 1e-7 relative), the block
// and chunk sums in FP64.
template <int G, bool SAME>
__device__ void high_level(const LevelCtx &L, const double *alpha, const double *beta, int B, int pnum,
                           int ybeg, int yend, double *cross, double *a1, double *a2, float *gls_row, int gls_ld) {
    constexpr int WN = G + 16;
    float w[WN], ga[G];
    double sab[8], head[8];
    double sa = 0.0, sbk = 0.0;   // block sums of a_y and of b_y over the block's products
#pragma unroll
    for (int i = 0; i < 8; i++) {
        sab[i] = head[i] = 0.0;
        cross[i] = a1[i] = a2[i] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < WN; k++) w[k] = L.val(L.sb, ybeg + k);
    if (!SAME) {
#pragma unroll
        for (int k = 0; k < G; k++) ga[k] = L.val(L.sa, ybeg + k);
    }
    for (int y0 = ybeg; y0 < yend; y0 += G) {
        // the next group's values, in flight while this group is reduced
        // (past yend they are values of the next chunk or zero: in bounds)
        float nb[G], na[SAME ? 1 : G];
#pragma unroll
        for (int k = 0; k < G; k++) {
            nb[k] = L.val(L.sb, y0 + WN + k);
            if (!SAME) na[k] = L.val(L.sa, y0 + G + k);
        }
        if (SAME) {
#pragma unroll
            for (int k = 0; k < G; k++) ga[k] = w[k];
        }
        if ((y0 & (B - 1)) == 0) {
            // block start (blocks hold whole groups: B >= G): head[i] = sum of
            // the block's first 9 + i b-values
            double pre = 0.0;
#pragma unroll
            for (int k = 0; k < 9; k++) pre += (double)w[k];
#pragma unroll
            for (int i = 0; i < 8; i++) {
                if (i > 0) pre += (double)w[8 + i];
                head[i] = pre;
            }
        }
        float gs = 0.f, gb = 0.f;
#pragma unroll
        for (int k = 0; k < G; k++) {
            gs += ga[k];
            gb += w[k];
        }
        sa += (double)gs;
        sbk += (double)gb;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            float sv = 0.f;
#pragma unroll
            for (int k = 0; k < G; k++) sv = fmaf(ga[k], w[k + 9 + i], sv);
            sab[i] += (double)sv;
        }
        const int ye = y0 + G;
        if ((ye & (B - 1)) == 0) {
            const bool inblk = (ye / B - 1) < pnum;
            // sum_{y in block} b_{y+d} = block sum - head(d) + tail(d), the tail
            // being the first d b-values after the block (w[G..G+15])
            double tail = 0.0;
#pragma unroll
            for (int k = 0; k < 9; k++) tail += (double)w[G + k];
#pragma unroll
            for (int i = 0; i < 8; i++) {
                if (i > 0) tail += (double)w[G + 8 + i];
                const double sbd = sbk - head[i] + tail;
                double S = sab[i] - beta[i] * sa - alpha[i] * sbd + (double)B * alpha[i] * beta[i];
                if (inblk) {
                    a1[i] += S;
                    a2[i] += S * S;
                    if (gls_row) gls_row[(int64_t)(ye / B - 1) * gls_ld + i] = (float)S;
                }
                cross[i] += sab[i];
                sab[i] = 0.0;
            }
            sa = 0.0;
            sbk = 0.0;
        }
#pragma unroll
        for (int k = 0; k < 16; k++) w[k] = w[k + G];
#pragma unroll
        for (int k = 0; k < G; k++) {
            w[16 + k] = nb[k];
            if (!SAME) ga[k] = na[k];
        }
    }
#pragma unroll
    for (int i = 0; i < 8; i++) cross[i] += sab[i];
}

template <bool SAME>
__device__ __forceinline__ void high_dispatch(const MainArgs &A, int c, int G, int sa, int sb, const double *alpha,
                                              const double *beta, int B, int pnum, int ybeg, int yend, double *cross,
                                              double *a1, double *a2, float *grow, int ld) {
    const CorrTables &ct = *A.ctp;
    const LevelCtx L{c == 3 ? A.l3 : A.lvd + A.loff[c], A.g.nslots, sa, sb, c, ct.numbin[c]};
    switch (G) {
        case 16: high_level<16, SAME>(L, alpha, beta, B, pnum, ybeg, yend, cross, a1, a2, grow, ld); break;
        case 8: high_level<8, SAME>(L, alpha, beta, B, pnum, ybeg, yend, cross, a1, a2, grow, ld); break;
        case 4: high_level<4, SAME>(L, alpha, beta, B, pnum, ybeg, yend, cross, a1, a2, grow, ld); break;
        case 2: high_level<2, SAME>(L, alpha, beta, B, pnum, ybeg, yend, cross, a1, a2, grow, ld); break;
        default: high_level<1, SAME>(L, alpha, beta, B, pnum, ybeg, yend, cross, a1, a2, grow, ld); break;
    }
}

// One thread per (pair, level >= 3, time chunk of the main kernel).  The
// ACF pairs and the CCF pairs fill separate workgroups, so whether a and b
// are one series is uniform per wave and an ACF loads its series once.
// With both kinds, workgroup x = 16 g + r takes the ACF of pixel block
// 8 g + r (r < 8) or the CCF of pixel block 8 g + r - 8: the two workgroups
// of a pixel block are 8 apart, so they run on the same XCD at about the
// same time and the CCF's series (the ACF's own, and the neighbours' one
// slot further) come from that XCD's L2.  With one kind, x is the pixel block.
__global__ void __launch_bounds__(64, 2) k_corr_high(MainArgs A, double *hpart /* [chunk][pair][nhigh][3] */,
                                                  int nhigh, int both) {
    const int bx = (int)blockIdx.x;
    const bool acf_block = both ? (bx & 15) < 8 : A.pd.acf_first != 0;
    const int pblk = both ? 8 * (bx >> 4) + (bx & 7) : bx;
    const int p = pblk * 64 + threadIdx.x;
    const int pidx = acf_block ? p : (A.pd.acf_first ? A.pd.npix : 0) + p;
    const int c = 3 + blockIdx.y;
    const int chunk = blockIdx.z;
    const CorrTables &ct = *A.ctp;
    if (c >= ct.nlevels) return;   // uniform per workgroup
    const bool live = p < A.pd.npix && pidx < A.pd.npairs;
    const int indp = live ? A.ind[pidx] : 0x7fff;
    // group size: the largest power of two up to 16 that divides the block
    // length 2^(ind - c) of every pair of the wave (blocks hold whole
    // groups); fewer, longer groups carry less per-group overhead
    int imin = indp;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) imin = min(imin, __shfl_xor(imin, off, 64));
    if (!live) return;
    int sa, sb;
    pair_series(A.pd, pidx, sa, sb);
    const int ka = A.g.slot_of[sa], kb = A.g.slot_of[sb];   // coarse levels are stored per slot
    const double *Fa = A.F + (int64_t)ka * ct.npos, *Fb = A.F + (int64_t)kb * ct.npos;
    const double mua = level_shift(A, sa, c);
    const double mub = level_shift(A, sb, c);
    const int x0 = ct.first_chan[c];
    double alpha[8], beta[8];
    for (int i = 0; i < 8; i++) {
        int x = x0 + i;
        double n = (double)(ct.numbin[c] - ct.del[x]);
        alpha[i] = Fa[ct.pos_a_end[x]] / n - mua;
        beta[i] = (Fb[ct.pos_b_end[x]] - Fb[ct.pos_b_beg[x]]) / n - mub;
    }
    const int B = 1 << (indp - c);
    const int pnum = (int)floor(ct.mtab_last / pow(2.0, fmax(indp - ct.log2_samp_last, 0.0)));
    double cross[8], a1[8], a2[8];
    const int len = A.chunk_len >> c;
    const int ybeg = chunk * len;
    const int yend = (chunk == A.nchunk - 1) ? ct.numbin[c] : ybeg + len;
    float *grow = A.gls ? A.gls + (int64_t)pidx * A.gls_nblk * (ct.chanum - 1) + (x0 - 1) : nullptr;
    const int G = min(16, 1 << max(imin - c, 0));
    const int ld = ct.chanum - 1;
    // (wave-uniform: a workgroup holds only ACF or only CCF pairs)
    if (acf_block && A.pd.acf_first)
        high_dispatch<true>(A, c, G, ka, kb, alpha, beta, B, pnum, ybeg, yend, cross, a1, a2, grow, ld);
    else
        high_dispatch<false>(A, c, G, ka, kb, alpha, beta, B, pnum, ybeg, yend, cross, a1, a2, grow, ld);
    double *out = hpart + (((int64_t)chunk * A.pd.npairs + pidx) * nhigh + (x0 - kMainChan)) * 3;
    for (int i = 0; i < 8; i++) {
        out[i * 3 + 0] = cross[i];
        out[i * 3 + 1] = a1[i];
        out[i * 3 + 2] = a2[i];
    }
}

// G and blocked variance of every channel from the accumulated sums.
__global__ void k_corr_finalize(MainArgs A, const double *hpart, int nhigh, double *corav, double *bvar) {
    const int pidx = blockIdx.x;
    const int x = threadIdx.x;
    const CorrTables &ct = *A.ctp;
    if (pidx >= A.pd.npairs || x >= ct.chanum) return;
    int sa, sb;
    pair_series(A.pd, pidx, sa, sb);
    const int c = ct.level[x];
    const double *Fa = A.F + (int64_t)A.g.slot_of[sa] * ct.npos, *Fb = A.F + (int64_t)A.g.slot_of[sb] * ct.npos;
    const double n = (double)(ct.numbin[c] - ct.del[x]);
    const double am = Fa[ct.pos_a_end[x]] / n;
    const double bm = (Fb[ct.pos_b_end[x]] - Fb[ct.pos_b_beg[x]]) / n;
    const double mua = level_shift(A, sa, c);
    const double mub = level_shift(A, sb, c);
    const double alpha = (double)(float)(am - mua), beta = (double)(float)(bm - mub);
    double cross = 0.0, a1 = 0.0, a2 = 0.0;
    if (x < kMainChan) {
#pragma unroll 4
        for (int k = 0; k < A.nchunk; k++) {
            const float *o = A.part + (((int64_t)k * A.pd.npairs + pidx) * kMainChan + x) * 3;
            cross += (double)o[0];
            a1 += (double)o[1];
            a2 += (double)o[2];
        }
    } else {
#pragma unroll 4
        for (int k = 0; k < A.nchunk; k++) {
            const double *o = hpart + (((int64_t)k * A.pd.npairs + pidx) * nhigh + (x - kMainChan)) * 3;
            cross += o[0];
            a1 += o[1];
            a2 += o[2];
        }
    }
    const int indp = A.ind[pidx];
    const double B = (double)(1 << (indp - c));
    const double pn = floor(ct.mtab_last / pow(2.0, fmax(indp - ct.log2_samp_last, 0.0)));
    // sum_y (a_y - abar)(b_{y+d} - bbar) = sum a~ b~ - n alpha beta, with the
    // alpha, beta values (FP32-rounded) used inside the block sums
    const double tot = cross - n * (am - mua) * (bm - mub);
    corav[(int64_t)pidx * ct.chanum + x] = tot / (n * am * bm);
    (void)alpha;
    (void)beta;
    const double sp = a1 / B, sp2 = a2 / (B * B), ab = am * bm;
    bvar[(int64_t)pidx * ct.chanum + x] = (sp2 / pn - (sp / pn) * (sp / pn)) / ((pn - 1) * (ab * ab));
}

bool correlate_fast_supported(const CorrTables &ct) {
    // level layout of p = 16 with at least levels 0..3 and blocks >= 16
    // values on level 2 (ind >= q - 1 >= 6); blocks never longer than the
    // 2^(blocknum-1)-frame chunks (ind <= blocknum - 1 needs blocknum >= q)
    return ct.p16 && ct.nlevels >= 7 && ct.nlevels <= 12 && ct.q == ct.nlevels && ct.numbin[3] >= 32 &&
           ct.blocknum >= ct.q;
}

int corr_fast_chunk_len(const CorrTables &ct, int npairs) {
    // one chunk per largest blocking block: boundaries are multiples of
    // every block size and a chunk holds at most 32 blocks of 2^ind frames
    (void)npairs;
    return 1 << (ct.blocknum - 1);
}

void launch_correlate_fast(const SeriesGeom &g, const BleachDesc &bd, const CorrTables &ct, const CorrTables *dct,
                           const PairDesc &pd, const double *F, const float *X, const int32_t *ind, float *part,
                           double *hpart, float *l3, float *lvd, int chunk_len, double *corav, double *bvar,
                           hipStream_t st, hipEvent_t ev_main_begin, hipEvent_t ev_main_end, float *gls_blk,
                           int gls_nblk, const int32_t *xexp) {
    MainArgs A;
    A.g = g;
    A.bd = bd;
    A.ctp = dct;
    A.pd = pd;
    A.F = F;
    A.X = X;
    A.ind = ind;
    A.chunk_len = chunk_len;
    A.nchunk = (g.T + chunk_len - 1) / chunk_len;
    A.has_acf = pd.acf_first;
    A.has_ccf = pd.has_ccf;
    A.part = part;
    A.l3 = l3;
    A.lvd = lvd;
    A.gls = gls_blk;
    A.gls_nblk = gls_nblk;
    int64_t off = 0;
    for (int c = 0; c < kMaxLevel; c++) {
        A.loff[c] = off;
        if (c >= 4 && c < ct.nlevels) off += ((int64_t)ct.numbin[c] + kLevelPad) * g.nslots;
    }
    const int tiles = (g.npix + 63) / 64;
    const int threads = pd.has_ccf ? 768 : 384;
    if (ev_main_begin) (void)hipEventRecord(ev_main_begin, st);
    if (xexp)
        launch_corr_mfma(g, bd, dct, pd, F, X, ind, xexp, chunk_len, A.nchunk, part, gls_blk, gls_nblk, st);
    else
        hipLaunchKernelGGL(k_corr_main, dim3(tiles, A.nchunk), dim3(threads), 0, st, A);
    if (ev_main_end) (void)hipEventRecord(ev_main_end, st);
    for (int c0 = 3; c0 + 1 < ct.nlevels; c0 += 4) {
        RebinArgs R{};
        R.src = c0 == 3 ? l3 : lvd + A.loff[c0];
        R.numbin_src = ct.numbin[c0];
        for (int k = 0; k < 4; k++) {
            const int c = c0 + 1 + k;
            R.dst[k] = c < ct.nlevels ? lvd + A.loff[c] : nullptr;
            R.numbin[k] = c < ct.nlevels ? ct.numbin[c] : 0;
        }
        R.nslots = g.nslots;
        const int64_t n = (int64_t)((ct.numbin[c0] + 15) / 16) * g.nslots;
        hipLaunchKernelGGL(k_corr_rebin4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, R);
    }
    LevelPads P{};
    P.nlev = ct.nlevels - 3;
    P.nslots = g.nslots;
    for (int k = 0; k < P.nlev; k++) {
        P.numbin[k] = ct.numbin[3 + k];
        P.off[k] = A.loff[3 + k];
    }
    if (P.nlev > 0)
        hipLaunchKernelGGL(k_zero_level_pads, dim3((unsigned)(((int64_t)kLevelPad * g.nslots + 255) / 256), P.nlev),
                           dim3(256), 0, st, l3, lvd, P);
    const int nhigh = ct.chanum - kMainChan;
    const int nbp = (pd.npix + 63) / 64;   // workgroups per pair type
    const int both = pd.acf_first && pd.has_ccf;
    const int nb_all = both ? 16 * ((nbp + 7) / 8) : nbp;
    hipLaunchKernelGGL(k_corr_high, dim3(nb_all, ct.nlevels - 3, A.nchunk), dim3(64), 0, st, A, hpart, nhigh, both);
    hipLaunchKernelGGL(k_corr_finalize, dim3(pd.npairs), dim3(128), 0, st, A, (const double *)hpart, nhigh, corav,
                       bvar);
}

}  // namespace imfcs
