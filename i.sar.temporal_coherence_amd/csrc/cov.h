/* SHP-masked covariance and coherence matrix of one pixel (tcoh.cl:567-
 * 639): C_ij = sum over accepted candidates of s_i conj(s_j), then
 * T_ij = C_ij / sqrt(C_ii C_jj) with the optional bias correction of the
 * magnitudes, and the pair coherences |T_ij|. T (i < j) goes to LDS,
 * packed row by row ((0,1),(0,2)...,(1,2)...), for the linking stage; the
 * pair planes go to global memory. Two arithmetic forms:
 *
 * plain (cov_plain_group): FP32 on the vector ALUs, one pixel per group
 *   of G lanes. The upper triangle (with diagonal) is cut into NBLK x NBLK
 *   blocks of B x B; lane l owns upper block l and keeps its B*B complex
 *   sums in registers; per accepted candidate (dy-major, the reference's
 *   summation order) it loads the B row and B column samples it needs.
 *
 * mixed (cov_mfma_pixel): the whole wavefront on one pixel, as a real
 *   Gram matrix. With R the 2N x L matrix [Re s_k ; Im s_k] of the L
 *   accepted candidates, G = R R^T, C_re = G_rr + G_ii and
 *   C_im[u][v] = G_ri[v][u] - G_ri[u][v]. Each sample is split as
 *   hi + lo (FP16: x = hi + lo + O(2^-22 x); BF16: O(2^-18 x)) and G is
 *   accumulated in FP32 from hi.hi + hi.lo + lo.hi with
 *   v_mfma_f32_16x16x32_{f16,bf16}. The accepted candidates are compacted
 *   into an LDS list padded to a multiple of 32 with a zero pixel; A and B
 *   operands share one lane layout, so the candidate order inside a
 *   32-candidate chunk does not matter. Dates are padded to NB with zeros.
 */
#ifndef TCOH_COV_H
#define TCOH_COV_H

#include "common.h"

/* Normalised and (optionally) debiased T_ij from the sums (tcoh.cl:610-
   636); den = sqrt(C_ii) sqrt(C_jj). */
__device__ __forceinline__ float2 coherence_entry(float cre, float cim,
                                                  float den, int bias,
                                                  int looks)
{
    float re = 0.0f, im = 0.0f;

    if (den > 0.0f) {
        re = cre / den;
        im = cim / den;
        if (bias && looks > 1) {
            const float m2 = re * re + im * im;
            const float m2c = (looks * m2 - 1.0f) / (looks - 1.0f);
            const float f = (m2 > 0.0f && m2c > 0.0f) ? sqrtf(m2c / m2) : 0.0f;

            re *= f;
            im *= f;
        }
    }
    return make_float2(re, im);
}

/* Pair plane of T_ij, or -1 if (i, j) is not an output pair. */
__device__ __forceinline__ int pair_plane(const KParams &kp, int i, int j)
{
    if (!kp.npairs)
        return -1;
    return kp.pairs_all ? upper_index(kp.n, i, j) : (j == i + 1 ? i : -1);
}

template <int G> struct PlainShape {
    static constexpr int NBLK = G == 8 ? 3 : (G == 16 ? 5 : 10);
};

/* Plain form, one pixel per group. diag: LDS scratch of N floats. */
template <int N, int G>
__device__ __forceinline__ void cov_plain_group(const KParams &kp,
                                                const float2 *__restrict__ slc,
                                                const uint32_t *__restrict__ mask,
                                                int p, int cnt, bool est,
                                                bool live, int lg, float *diag,
                                                float2 *tl,
                                                float *__restrict__ pairs)
{
    constexpr int NBLK = PlainShape<G>::NBLK;
    constexpr int B = (N + NBLK - 1) / NBLK;
    const size_t plane = (size_t)kp.rows * kp.cols;
    const int y = p / kp.cols, x = p - (p / kp.cols) * kp.cols;
    int bi = 0, bj = 0;
    float cr[B][B], ci[B][B];

    /* Upper block number lg -> (bi, bj), bi <= bj. */
    {
        int t = lg;

        while (bi < NBLK && t >= NBLK - bi) {
            t -= NBLK - bi;
            bi++;
        }
        bj = bi + t;
    }
    const bool own = lg < NBLK * (NBLK + 1) / 2;
    const int r0 = bi * B, c0 = bj * B;

#pragma unroll
    for (int r = 0; r < B; r++)
#pragma unroll
        for (int c = 0; c < B; c++)
            cr[r][c] = ci[r][c] = 0.0f;

    if (est && own) {
        const uint32_t *mk = mask + (size_t)p * kp.words;
        int bit = 0;

        for (int dy = 0; dy < kp.wa; dy++) {
            for (int dx = 0; dx < kp.wr; dx++, bit++) {
                if (!((mk[bit >> 5] >> (bit & 31)) & 1u))
                    continue;
                const float2 *s = slc + ((size_t)(y + dy) * kp.w + x + dx) * N;
                float2 sr[B], sc[B];

#pragma unroll
                for (int r = 0; r < B; r++) {
                    sr[r] = r0 + r < N ? s[r0 + r] : make_float2(0.f, 0.f);
                    sc[r] = c0 + r < N ? s[c0 + r] : make_float2(0.f, 0.f);
                }
#pragma unroll
                for (int r = 0; r < B; r++)
#pragma unroll
                    for (int c = 0; c < B; c++) {
                        cr[r][c] += sr[r].x * sc[c].x + sr[r].y * sc[c].y;
                        ci[r][c] += sr[r].y * sc[c].x - sr[r].x * sc[c].y;
                    }
            }
        }
    }
    if (own && bi == bj) {
#pragma unroll
        for (int r = 0; r < B; r++)
            if (r0 + r < N)
                diag[r0 + r] = cr[r][r];
    }
    wave_sync();
    if (!live || !own)
        return;
#pragma unroll
    for (int r = 0; r < B; r++) {
        const int i = r0 + r;

#pragma unroll
        for (int c = 0; c < B; c++) {
            const int j = c0 + c;

            if (i >= N || j >= N || j <= i)
                continue;
            const int pidx = pair_plane(kp, i, j);

            if (!est) {
                if (pidx >= 0)
                    pairs[(size_t)pidx * plane + p] = __builtin_nanf("");
                continue;
            }
            const float2 t = coherence_entry(
                cr[r][c], ci[r][c], sqrtf(diag[i]) * sqrtf(diag[j]), kp.bias, cnt);

            tl[upper_index(N, i, j)] = t;
            if (pidx >= 0)
                pairs[(size_t)pidx * plane + p] = hypotf(t.x, t.y);
        }
    }
}

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef unsigned int uint4_t __attribute__((ext_vector_type(4)));

template <int BF16>
__device__ __forceinline__ float4_t mfma(uint4_t a, uint4_t b, float4_t c)
{
    if (BF16)
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            __builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b), c,
            0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(
            __builtin_bit_cast(half8_t, a), __builtin_bit_cast(half8_t, b), c, 0,
            0, 0);
}

/* Operands of one 16-date block for the lane's 8 candidates: low halves
   (hi parts) and high halves (lo parts) of the re and im words. */
struct Opnd {
    uint4_t rh, rl, ih, il;
};

__device__ __forceinline__ Opnd make_opnd(const uint2 (&v)[8])
{
    Opnd o;

#pragma unroll
    for (int q = 0; q < 4; q++) {
        o.rh[q] = __builtin_amdgcn_perm(v[2 * q + 1].x, v[2 * q].x, 0x05040100u);
        o.rl[q] = __builtin_amdgcn_perm(v[2 * q + 1].x, v[2 * q].x, 0x07060302u);
        o.ih[q] = __builtin_amdgcn_perm(v[2 * q + 1].y, v[2 * q].y, 0x05040100u);
        o.il[q] = __builtin_amdgcn_perm(v[2 * q + 1].y, v[2 * q].y, 0x07060302u);
    }
    return o;
}

/* LDS words the mixed form needs: candidate list, then the C_re and G_ri
   tiles over it. */
template <int NB> struct MfmaLds {
    static constexpr int LIST = TCOH_MAX_WINDOW + 32;
    static constexpr int TILES = 2 * NB * (NB + 1);
    static constexpr int WORDS = LIST > TILES ? LIST : TILES;
};

/* Mixed form, whole wavefront on pixel p (uniform). hl[pix][date] =
   (re, im) words of hi | lo << 16; pixel npad is zero. */
template <int N, int NB, int BF16>
__device__ __forceinline__ void cov_mfma_pixel(const KParams &kp,
                                               const uint2 *__restrict__ hl,
                                               const uint32_t *__restrict__ mask,
                                               int p, int cnt, float *lds,
                                               float2 *tl,
                                               float *__restrict__ pairs)
{
    constexpr int NBK = NB / 16;             /* 16-date blocks */
    constexpr int NUP = NBK * (NBK + 1) / 2; /* upper block pairs */
    constexpr int NT = N * (N - 1) / 2;
    int *list = reinterpret_cast<int *>(lds);
    float(*gre)[NB + 1] = reinterpret_cast<float(*)[NB + 1]>(lds);
    float(*gri)[NB + 1] = reinterpret_cast<float(*)[NB + 1]>(lds + NB * (NB + 1));
    const int lane = threadIdx.x;
    const int y = p / kp.cols, x = p - (p / kp.cols) * kp.cols;
    const size_t plane = (size_t)kp.rows * kp.cols;

    if (cnt < kp.min_looks) {
        if (kp.npairs)
            for (int t = lane; t < kp.npairs; t += WAVE)
                pairs[(size_t)t * plane + p] = __builtin_nanf("");
        return;
    }

    /* Compact the accepted candidates of the window. */
    const uint32_t *mk = mask + (size_t)p * kp.words;
    const int kw = kp.wa * kp.wr;

    wave_sync();
    for (int b = lane; b < kw; b += WAVE) {
        const int wd = b >> 5;
        const uint32_t word = mk[wd];

        if ((word >> (b & 31)) & 1u) {
            int rank = __popc(word & ((1u << (b & 31)) - 1u));

            for (int t = 0; t < wd; t++)
                rank += __popc(mk[t]);
            const int dy = b / kp.wr, dx = b - dy * kp.wr;

            list[rank] = (y + dy) * kp.w + x + dx;
        }
    }
    const int lpad = (cnt + 31) & ~31;
    for (int t = cnt + lane; t < lpad; t += WAVE)
        list[t] = kp.npad;
    wave_sync();

    const int i16 = lane & 15, g = lane >> 4;
    float4_t acc_rr[NUP], acc_ii[NUP], acc_ri[NBK * NBK];

#pragma unroll
    for (int t = 0; t < NUP; t++)
        acc_rr[t] = acc_ii[t] = float4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < NBK * NBK; t++)
        acc_ri[t] = float4_t{0.f, 0.f, 0.f, 0.f};

    /* Gathers run two chunks ahead of the MFMAs. */
    uint2 cur[NBK][8], nxt[NBK][8], nx2[NBK][8];
    auto fetch = [&](int c0, uint2(&v)[NBK][8]) {
        int idx[8];

#pragma unroll
        for (int e = 0; e < 8; e++)
            idx[e] = list[c0 + g * 8 + e];
#pragma unroll
        for (int bk = 0; bk < NBK; bk++) {
            const int d = bk * 16 + i16;

#pragma unroll
            for (int e = 0; e < 8; e++)
                v[bk][e] = d < N ? hl[(size_t)idx[e] * N + d] : make_uint2(0u, 0u);
        }
    };

    fetch(0, cur);
    if (32 < lpad)
        fetch(32, nxt);
    for (int c0 = 0; c0 < lpad; c0 += 32) {
        if (c0 + 64 < lpad)
            fetch(c0 + 64, nx2);
        Opnd op[NBK];

#pragma unroll
        for (int bk = 0; bk < NBK; bk++)
            op[bk] = make_opnd(cur[bk]);
#pragma unroll
        for (int bk = 0; bk < NBK; bk++)
#pragma unroll
            for (int e = 0; e < 8; e++) {
                cur[bk][e] = nxt[bk][e];
                nxt[bk][e] = nx2[bk][e];
            }
        int t = 0;
#pragma unroll
        for (int a = 0; a < NBK; a++)
#pragma unroll
            for (int b = a; b < NBK; b++, t++) {
                acc_rr[t] = mfma<BF16>(op[a].rh, op[b].rh, acc_rr[t]);
                acc_rr[t] = mfma<BF16>(op[a].rh, op[b].rl, acc_rr[t]);
                acc_rr[t] = mfma<BF16>(op[a].rl, op[b].rh, acc_rr[t]);
                acc_ii[t] = mfma<BF16>(op[a].ih, op[b].ih, acc_ii[t]);
                acc_ii[t] = mfma<BF16>(op[a].ih, op[b].il, acc_ii[t]);
                acc_ii[t] = mfma<BF16>(op[a].il, op[b].ih, acc_ii[t]);
            }
#pragma unroll
        for (int a = 0; a < NBK; a++)
#pragma unroll
            for (int b = 0; b < NBK; b++) {
                float4_t &r = acc_ri[a * NBK + b];

                r = mfma<BF16>(op[a].rh, op[b].ih, r);
                r = mfma<BF16>(op[a].rh, op[b].il, r);
                r = mfma<BF16>(op[a].rl, op[b].ih, r);
            }
    }

    /* D[4g + r][i16] of each tile to LDS, over the dead list. */
    wave_sync();
    {
        int t = 0;

#pragma unroll
        for (int a = 0; a < NBK; a++)
#pragma unroll
            for (int b = a; b < NBK; b++, t++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    gre[a * 16 + 4 * g + r][b * 16 + i16] =
                        acc_rr[t][r] + acc_ii[t][r];
#pragma unroll
        for (int a = 0; a < NBK; a++)
#pragma unroll
            for (int b = 0; b < NBK; b++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    gri[a * 16 + 4 * g + r][b * 16 + i16] =
                        acc_ri[a * NBK + b][r];
    }
    wave_sync();

    for (int e = lane; e < NT; e += WAVE) {
        int i = 0, rem = e;

        while (rem >= N - 1 - i) {
            rem -= N - 1 - i;
            i++;
        }
        const int j = i + 1 + rem;
        const float2 t = coherence_entry(gre[i][j], gri[j][i] - gri[i][j],
                                         sqrtf(gre[i][i]) * sqrtf(gre[j][j]),
                                         kp.bias, cnt);
        const int pidx = pair_plane(kp, i, j);

        tl[e] = t;
        if (pidx >= 0)
            pairs[(size_t)pidx * plane + p] = hypotf(t.x, t.y);
    }
    wave_sync();
}

#endif
