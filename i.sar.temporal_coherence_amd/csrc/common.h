/* Shared definitions of the gfx950 phase-linking kernels. */
#ifndef TCOH_COMMON_H
#define TCOH_COMMON_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

/* Settings of one handle as the kernels read them (the reference bakes
   them in with -D, ocl.c:222-229; here they are runtime arguments so one
   code object serves every stack of the same depth bucket). */
struct KParams {
    int n;                 /* dates */
    int wa, wr;            /* window, lines x samples (odd) */
    int w;                 /* padded strip width = cols + wr - 1 */
    int rows, cols;        /* output pixels of the strip */
    int npad;              /* padded_rows * w */
    int test;              /* 0 KS, 1 AD, 2 tlog */
    int emi, bias;
    int min_looks, ks_max;
    int npairs, pairs_all;
    int words;             /* 32-bit mask words per output pixel */
    float ad_norm, ad_sigma, ad_crit, tlog_crit;
};

/* Largest window the SHP mask supports (candidates per pixel). */
#define TCOH_MAX_WINDOW 1024

/* Lanes of one work-group: every kernel runs one wavefront per group. */
#define WAVE 64

/* Order LDS traffic between the lanes of the (single) wavefront of a
   work-group: the DS instructions of one wavefront execute in issue order,
   so a compiler-level fence is all that is needed, and unlike s_barrier it
   may sit in code where the lanes of different pixels diverge. */
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

/* conj(a) * b */
__device__ __forceinline__ float2 cmulc(float2 a, float2 b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

/* Reductions over the G consecutive lanes of a group (G a power of two
   dividing 64). */
template <int G> __device__ __forceinline__ float gsum(float v)
{
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1)
        v += __shfl_xor(v, o, G);
    return v;
}

template <int G> __device__ __forceinline__ double gsumd(double v)
{
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1)
        v += __shfl_xor(v, o, G);
    return v;
}

template <int G> __device__ __forceinline__ float gmax(float v)
{
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1)
        v = fmaxf(v, __shfl_xor(v, o, G));
    return v;
}

template <int G> __device__ __forceinline__ float gmin(float v)
{
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1)
        v = fminf(v, __shfl_xor(v, o, G));
    return v;
}

template <int G> __device__ __forceinline__ int gmaxi(int v)
{
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1)
        v = max(v, __shfl_xor(v, o, G));
    return v;
}

template <int G> __device__ __forceinline__ unsigned long long
gor64(unsigned long long v)
{
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1)
        v |= __shfl_xor(v, o, G);
    return v;
}

/* Offset of T_ij (i < j) in the packed strict upper triangle, row by row:
   (0,1),(0,2)...(0,N-1),(1,2)... -- also the pair order of pairs_mode=all
   (tcoh.cl:19-23). */
__host__ __device__ __forceinline__ int upper_index(int n, int i, int j)
{
    return i * n - i * (i + 1) / 2 + (j - i - 1);
}

/* Kernel launchers (defined in the .hip files). Each returns a HIP error
   code; stream is the handle's stream. */
struct LaunchCtx {
    hipStream_t stream;
};

hipError_t launch_prep(const KParams &kp, const float *amp, float *ampT,
                       const unsigned char *valid, float2 *lstat,
                       hipStream_t s);
hipError_t launch_shp(const KParams &kp, const float *ampT,
                      const unsigned char *valid, const float2 *lstat,
                      const float *adtab, uint32_t *mask, int *count,
                      hipStream_t s);
hipError_t launch_cov_plain(const KParams &kp, const float2 *slc,
                            const uint32_t *mask, const int *count,
                            float2 *tbuf, float *pairs, int pix0, int npix,
                            hipStream_t s);
hipError_t launch_split(const KParams &kp, const float2 *slc, uint32_t *hl,
                        int bf16, float scale, hipStream_t s);
hipError_t launch_cov_mfma(const KParams &kp, const uint32_t *hl,
                           const uint32_t *mask, const int *count,
                           float2 *tbuf, float *pairs, int pix0, int npix,
                           int bf16, hipStream_t s);
template <int N>
hipError_t launch_eig_n(const KParams &kp, const float2 *tbuf,
                        const int *count, float *coh, int pix0, int npix,
                        hipStream_t s);
hipError_t launch_absmax(const float2 *slc, size_t n, unsigned int *out,
                         hipStream_t s);

#endif
