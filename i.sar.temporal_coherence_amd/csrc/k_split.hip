/* Preparation of the mixed-precision covariance: every sample component x
 * becomes a 32-bit word hi | lo << 16 with hi = round16(x), lo =
 * round16(x - hi), FP16 or BF16, round to nearest even. FP16 scales each
 * pixel by its own power of two first, the one that puts the pixel's
 * largest component magnitude in [2^13, 2^14) (hi parts then fit FP16,
 * with 2^-22 of relative accuracy for hi + lo), and stores its exponent;
 * the covariance rescales every window to its largest candidate exactly
 * (cov.h), so a pixel's T depends only on its window's samples, as in the
 * reference, whatever else the strip holds. BF16 has the FP32 range and
 * takes no scale. Each pixel takes NB = shp_bucket(N) date slots (slots
 * N .. NB-1 zero), so that the covariance gathers whole, aligned 16-date
 * blocks without guards; one zero pixel is appended (index npad) for the
 * padding of candidate lists. */

#include "common.h"

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

/* One thread per sample slot; the NB threads of a pixel are adjacent
   lanes of one wavefront (NB divides 64), which reduce the pixel's
   largest component magnitude among themselves. */
__global__ __launch_bounds__(256) void k_split(KParams kp,
                                              const float2 *__restrict__ slc,
                                              uint2 *__restrict__ hl,
                                              short *__restrict__ ex, int bf16)
{
    const int nb = shp_bucket(kp.n);
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t total = ((size_t)kp.npad + 1) * nb;
    const size_t pix = i / nb;
    const int d = (int)(i - pix * nb);
    const bool in = i < total && pix < (size_t)kp.npad && d < kp.n;
    const float2 v = in ? slc[pix * kp.n + d] : make_float2(0.f, 0.f);
    float scale = 1.0f;

    if (!bf16) {
        float m = fmaxf(fabsf(v.x), fabsf(v.y));

        if (!isfinite(m))
            m = 0.0f;
        for (int o = nb / 2; o > 0; o >>= 1)
            m = fmaxf(m, __shfl_xor(m, o, nb));
        int e = SPLIT_EX_ZERO;

        if (m > 0.0f) {
            frexpf(m, &e);
            /* (a scale above 2^127 would overflow: such tiny pixels keep
               their largest component below 2^13) */
            e = max(e, -113);
            scale = ldexpf(1.0f, 14 - e);
        }
        if (d == 0 && i < total)
            ex[pix] = (short)e;
    }
    if (i >= total)
        return;
    const float c[2] = {v.x * scale, v.y * scale};
    uint32_t w[2];

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

hipError_t launch_split(const KParams &kp, const float2 *slc, uint32_t *hl, short *ex,
                        int bf16, hipStream_t s)
{
    const size_t total = ((size_t)kp.npad + 1) * shp_bucket(kp.n);

    hipLaunchKernelGGL(k_split, dim3((unsigned)((total + 255) / 256)), dim3(256),
                       0, s, kp, slc, (uint2 *)hl, ex, bf16);
    return hipGetLastError();
}
