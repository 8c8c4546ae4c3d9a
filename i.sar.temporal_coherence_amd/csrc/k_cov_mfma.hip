/* SHP-masked covariance on the CDNA4 matrix cores (the mixed form).
 *
 * Same result as the plain form (tcoh.cl:567-639): C = sum over accepted
 * candidates of s s^H, then T and the pair coherences. Written as a real
 * Gram matrix: with R the 2N x L matrix [Re s_k ; Im s_k] of the L
 * accepted candidates,
 *     G = R R^T,  C_re = G_rr + G_ii,  C_im[u][v] = G_ri[v][u] - G_ri[u][v]
 * (G_rr = Re Re^T, G_ii = Im Im^T, G_ri = Re Im^T).
 *
 * Each sample x is split as x = hi + lo + O(2^-22 x) with hi, lo FP16
 * (or 2^-18 with BF16), and G is accumulated in FP32 from the three
 * products hi.hi + hi.lo + lo.hi with v_mfma_f32_16x16x32_{f16,bf16}.
 * FP16 needs a common power-of-two scale of the strip (T does not depend
 * on it).
 *
 * One wavefront per pixel. The accepted candidates are compacted into an
 * LDS list (window order), padded to a multiple of 32 with a zero pixel;
 * every 32-candidate chunk feeds, per 16-date block pair, MFMA tiles
 * whose operand lane layout is the same for A and B, so the candidate
 * order inside a chunk does not matter. Dates are padded to NB (16, 32
 * or 64) with zeros.
 */

#include "common.h"

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef unsigned int uint4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t f16_bits(float x)
{
    const _Float16 h = (_Float16)x;
    unsigned short u;

    __builtin_memcpy(&u, &h, 2);
    return u;
}

__device__ __forceinline__ float f16_value(uint32_t u)
{
    const unsigned short s = (unsigned short)u;
    _Float16 h;

    __builtin_memcpy(&h, &s, 2);
    return (float)h;
}

/* Round to nearest even BF16 (finite inputs). */
__device__ __forceinline__ uint32_t bf16_bits(float x)
{
    const uint32_t b = __float_as_uint(x);

    return (b + 0x7fffu + ((b >> 16) & 1u)) >> 16;
}

__device__ __forceinline__ float bf16_value(uint32_t u)
{
    return __uint_as_float(u << 16);
}

/* hl[pix][date] = (re, im), each a 32-bit word hi | lo << 16. Pixel npad
   is all zeros. */
__global__ __launch_bounds__(256) void k_split(KParams kp,
                                              const float2 *__restrict__ slc,
                                              uint2 *__restrict__ hl, int bf16,
                                              float scale)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t total = ((size_t)kp.npad + 1) * kp.n;

    if (i >= total)
        return;
    float2 v = i < (size_t)kp.npad * kp.n ? slc[i] : make_float2(0.f, 0.f);
    uint32_t w[2];
    const float c[2] = {v.x * scale, v.y * scale};

#pragma unroll
    for (int t = 0; t < 2; t++) {
        uint32_t h, l;

        if (bf16) {
            h = bf16_bits(c[t]);
            l = bf16_bits(c[t] - bf16_value(h));
        }
        else {
            h = f16_bits(c[t]);
            l = f16_bits(c[t] - f16_value(h));
        }
        w[t] = h | (l << 16);
    }
    hl[i] = make_uint2(w[0], w[1]);
}

/* Largest |re|, |im| of the strip as float bits (non-negative floats
   order like their bit patterns). */
__global__ __launch_bounds__(256) void k_absmax(const float2 *__restrict__ slc,
                                               size_t n,
                                               unsigned int *__restrict__ out)
{
    unsigned int m = 0;

    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (size_t)gridDim.x * 256) {
        const float2 v = slc[i];
        const float a = fmaxf(fabsf(v.x), fabsf(v.y));

        if (isfinite(a))
            m = max(m, __float_as_uint(a));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        m = max(m, (unsigned int)__shfl_xor((int)m, o, 64));
    if ((threadIdx.x & 63) == 0)
        atomicMax(out, m);
}

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

template <int NB, int BF16>
__global__ __launch_bounds__(64) void k_cov_mfma(KParams kp,
                                                const uint2 *__restrict__ hl,
                                                const uint32_t *__restrict__ mask,
                                                const int *__restrict__ count,
                                                float2 *__restrict__ tbuf,
                                                float *__restrict__ pairs,
                                                int pix0, int npix)
{
    constexpr int NBK = NB / 16;               /* 16-date blocks */
    constexpr int NUP = NBK * (NBK + 1) / 2;   /* upper block pairs */
    __shared__ int list[TCOH_MAX_WINDOW + 32];
    __shared__ float gre[NB][NB + 1];          /* C_re, upper */
    __shared__ float gri[NB][NB + 1];          /* G_ri = Re Im^T */
    const int lane = threadIdx.x;
    const int q = blockIdx.x;

    if (q >= npix)
        return;
    const int p = pix0 + q;
    const int n = kp.n;
    const int y = p / kp.cols, x = p - (p / kp.cols) * kp.cols;
    const int cnt = count[p];
    const bool est = cnt >= kp.min_looks;
    const size_t plane = (size_t)kp.rows * kp.cols;
    const int ntri = n * (n - 1) / 2;

    if (!est) {
        if (kp.npairs) {
            const int np = kp.npairs;

            for (int t = lane; t < np; t += WAVE)
                pairs[(size_t)t * plane + p] = __builtin_nanf("");
        }
        return;
    }

    /* Compact the accepted candidates of the window. */
    const uint32_t *mk = mask + (size_t)p * kp.words;
    const int kw = kp.wa * kp.wr;
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

    for (int c0 = 0; c0 < lpad; c0 += 32) {
        int idx[8];

#pragma unroll
        for (int e = 0; e < 8; e++)
            idx[e] = list[c0 + g * 8 + e];
        Opnd op[NBK];

#pragma unroll
        for (int bk = 0; bk < NBK; bk++) {
            const int d = bk * 16 + i16;
            uint2 v[8];

#pragma unroll
            for (int e = 0; e < 8; e++)
                v[e] = d < n ? hl[(size_t)idx[e] * n + d] : make_uint2(0u, 0u);
            op[bk] = make_opnd(v);
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

    /* D[4g + r][i16] of each tile to LDS. */
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

    /* T_ij = C_ij / sqrt(C_ii C_jj), bias, pairs (tcoh.cl:610-639). */
    float2 *tp = tbuf + (size_t)q * ntri;
    for (int e = lane; e < ntri; e += WAVE) {
        int i = 0, rem = e;

        while (rem >= n - 1 - i) {
            rem -= n - 1 - i;
            i++;
        }
        const int j = i + 1 + rem;
        const float den = sqrtf(gre[i][i]) * sqrtf(gre[j][j]);
        float re = 0.0f, im = 0.0f;

        if (den > 0.0f) {
            re = gre[i][j] / den;
            im = (gri[j][i] - gri[i][j]) / den;
            if (kp.bias && cnt > 1) {
                const float m2 = re * re + im * im;
                const float m2c = (cnt * m2 - 1.0f) / (cnt - 1.0f);
                const float f =
                    (m2 > 0.0f && m2c > 0.0f) ? sqrtf(m2c / m2) : 0.0f;

                re *= f;
                im *= f;
            }
        }
        tp[e] = make_float2(re, im);
        if (kp.npairs) {
            const int pidx = kp.pairs_all ? e : (j == i + 1 ? i : -1);

            if (pidx >= 0)
                pairs[(size_t)pidx * plane + p] = hypotf(re, im);
        }
    }
}

hipError_t launch_split(const KParams &kp, const float2 *slc, uint32_t *hl,
                        int bf16, float scale, hipStream_t s)
{
    const size_t total = ((size_t)kp.npad + 1) * kp.n;

    hipLaunchKernelGGL(k_split, dim3((unsigned)((total + 255) / 256)), dim3(256),
                       0, s, kp, slc, (uint2 *)hl, bf16, scale);
    return hipGetLastError();
}

__global__ void k_zero_u32(unsigned int *out) { *out = 0u; }

hipError_t launch_absmax(const float2 *slc, size_t n, unsigned int *out,
                         hipStream_t s)
{
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, out);
    hipLaunchKernelGGL(k_absmax, dim3(1024), dim3(256), 0, s, slc, n, out);
    return hipGetLastError();
}

template <int NB>
static void launch_nb(const KParams &kp, const uint2 *hl, const uint32_t *mask,
                      const int *count, float2 *tbuf, float *pairs, int pix0,
                      int npix, int bf16, hipStream_t s)
{
    if (bf16)
        hipLaunchKernelGGL((k_cov_mfma<NB, 1>), dim3(npix), dim3(WAVE), 0, s, kp,
                           hl, mask, count, tbuf, pairs, pix0, npix);
    else
        hipLaunchKernelGGL((k_cov_mfma<NB, 0>), dim3(npix), dim3(WAVE), 0, s, kp,
                           hl, mask, count, tbuf, pairs, pix0, npix);
}

hipError_t launch_cov_mfma(const KParams &kp, const uint32_t *hl,
                           const uint32_t *mask, const int *count,
                           float2 *tbuf, float *pairs, int pix0, int npix,
                           int bf16, hipStream_t s)
{
    const uint2 *h = (const uint2 *)hl;

    if (kp.n <= 16)
        launch_nb<16>(kp, h, mask, count, tbuf, pairs, pix0, npix, bf16, s);
    else if (kp.n <= 32)
        launch_nb<32>(kp, h, mask, count, tbuf, pairs, pix0, npix, bf16, s);
    else
        launch_nb<64>(kp, h, mask, count, tbuf, pairs, pix0, npix, bf16, s);
    return hipGetLastError();
}
