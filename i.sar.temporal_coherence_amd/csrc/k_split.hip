/* Preparation of the mixed-precision covariance: every sample component x
 * (after a common power-of-two scale of the strip) becomes a 32-bit word
 * hi | lo << 16 with hi = round16(x), lo = round16(x - hi), FP16 or BF16,
 * round to nearest even. Each pixel takes NB = shp_bucket(N) date slots
 * (slots N .. NB-1 zero), so that the covariance gathers whole, aligned
 * 16-date blocks without guards; one zero pixel is appended (index npad)
 * for the padding of candidate lists. */

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

__global__ __launch_bounds__(256) void k_split(KParams kp,
                                              const float2 *__restrict__ slc,
                                              uint2 *__restrict__ hl, int bf16,
                                              const unsigned int *__restrict__ absmax)
{
    const int nb = shp_bucket(kp.n);
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t total = ((size_t)kp.npad + 1) * nb;

    if (i >= total)
        return;
    /* FP16: the power of two that puts the strip's largest amplitude in
       [2^13, 2^14), so every component's hi part fits FP16 and products
       and sums of up to 1024 candidates stay far below FP32 overflow; T is
       invariant to the common factor. BF16 keeps the FP32 range. */
    float scale = 1.0f;

    if (absmax) {
        const float mx = __uint_as_float(*absmax);

        if (mx > 0.0f && isfinite(mx)) {
            int ex;

            frexpf(mx, &ex);
            scale = ldexpf(1.0f, 14 - ex);
        }
    }
    const size_t pix = i / nb;
    const int d = (int)(i - pix * nb);
    const float2 v = (pix < (size_t)kp.npad && d < kp.n) ? slc[pix * kp.n + d]
                                                          : make_float2(0.f, 0.f);
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

hipError_t launch_split(const KParams &kp, const float2 *slc, uint32_t *hl,
                        int bf16, const unsigned int *absmax, hipStream_t s)
{
    const size_t total = ((size_t)kp.npad + 1) * shp_bucket(kp.n);

    hipLaunchKernelGGL(k_split, dim3((unsigned)((total + 255) / 256)), dim3(256),
                       0, s, kp, slc, (uint2 *)hl, bf16, absmax);
    return hipGetLastError();
}

