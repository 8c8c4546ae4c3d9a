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
    float *phase;          /* optional: linked phase of every date, planar
                              [N][rows][cols] (radians, NaN where gamma is) */
    int shp_stripe;        /* XCD stripe widths of the SHP and linking */
    int link_stripe;       /* launches, in tiles (0: an eighth of a row) */
};

/* Largest window the SHP mask supports (candidates per pixel). */
#define TCOH_MAX_WINDOW 1024

/* Largest span (bytes) of the split-sample rows one linking launch
   gathers from, its band's window rows: the gathers address them with
   32-bit offsets below the candidate lists' pad value (0xFFFF0000). */
#define LIST_SPAN_MAX 0xFFF00000u

/* Lanes of one work-group: every kernel runs one wavefront per group. */
#define WAVE 64

typedef float float4_t __attribute__((ext_vector_type(4)));

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

/* The same value, but held in a vector register: offsets derived from it
   are computed per lane instead of being hoisted into scalar registers
   (which a fully unrolled loop over dates would otherwise spill). The
   empty asm only fixes the register class; it emits no instruction. */
__device__ __forceinline__ int in_vgpr(int x)
{
    asm volatile("" : "+v"(x));
    return x;
}

/* Date planes of the planar amplitude copy: the SHP kernel's depth
   bucket, so that loads of planes n..NB-1 stay in bounds. */
__host__ __device__ __forceinline__ int shp_bucket(int n)
{
    return n <= 16 ? 16 : (n <= 32 ? 32 : 64);
}

/* Hardware reciprocal (1 ulp). Used where a quotient only feeds a
   result checked to a tolerance, never in a decision that must match the
   reference bit for bit. */
__device__ __forceinline__ float rcpf(float x)
{
    return __builtin_amdgcn_rcpf(x);
}

/* Hardware square root (1 ulp, one instruction), where the result only
   has to meet a tolerance; sqrtf is the correctly rounded one, a sequence
   of some twenty instructions under hipcc's defaults. */
__device__ __forceinline__ float sqrt_fast(float x)
{
    return __builtin_amdgcn_sqrtf(x);
}

template <bool EXACT> __device__ __forceinline__ float qsqrt(float x)
{
    return EXACT ? sqrtf(x) : sqrt_fast(x);
}

/* a / b: correctly rounded when EXACT (the reference's division), else
   through the hardware reciprocal. */
template <bool EXACT> __device__ __forceinline__ float qdiv(float a, float b)
{
    return EXACT ? a / b : a * rcpf(b);
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
   dividing 64). For G <= 16 the group is inside one 16-lane DPP row and
   the all-reduce is four DPP moves: quad_perm [1,0,3,2] (xor 1),
   quad_perm [2,3,0,1] (xor 2), row_half_mirror (lane i <-> 7-i of each
   8) and row_mirror (lane i <-> 15-i); every lane ends with the total.
   All lanes of the wavefront must be active. */
#define DPP_XOR1 0xB1
#define DPP_XOR2 0x4E
#define DPP_HALF_MIRROR 0x141
#define DPP_MIRROR 0x140

template <int CTRL> __device__ __forceinline__ float dpp_f(float v)
{
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL,
                                           0xF, 0xF, false));
}

template <int CTRL> __device__ __forceinline__ int dpp_i(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

/* A whole-wavefront group (G = 64): each 16-lane row is reduced with the
   four DPP steps above, then the four row results are read into scalar
   registers (v_readlane) and combined there; every lane gets the total,
   and no LDS round trip (ds_bpermute) is on the path. */
__device__ __forceinline__ float rows4(float v, int r)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), r));
}

template <int G> __device__ __forceinline__ float gsum(float v)
{
    if (G == 64) {
        v += dpp_f<DPP_XOR1>(v);
        v += dpp_f<DPP_XOR2>(v);
        v += dpp_f<DPP_HALF_MIRROR>(v);
        v += dpp_f<DPP_MIRROR>(v);
        return (rows4(v, 0) + rows4(v, 16)) + (rows4(v, 32) + rows4(v, 48));
    }
    if (G <= 16) {
        if (G >= 2)
            v += dpp_f<DPP_XOR1>(v);
        if (G >= 4)
            v += dpp_f<DPP_XOR2>(v);
        if (G >= 8)
            v += dpp_f<DPP_HALF_MIRROR>(v);
        if (G >= 16)
            v += dpp_f<DPP_MIRROR>(v);
        return v;
    }
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
    if (G == 64) {
        v = fmaxf(v, dpp_f<DPP_XOR1>(v));
        v = fmaxf(v, dpp_f<DPP_XOR2>(v));
        v = fmaxf(v, dpp_f<DPP_HALF_MIRROR>(v));
        v = fmaxf(v, dpp_f<DPP_MIRROR>(v));
        return fmaxf(fmaxf(rows4(v, 0), rows4(v, 16)), fmaxf(rows4(v, 32), rows4(v, 48)));
    }
    if (G <= 16) {
        if (G >= 2)
            v = fmaxf(v, dpp_f<DPP_XOR1>(v));
        if (G >= 4)
            v = fmaxf(v, dpp_f<DPP_XOR2>(v));
        if (G >= 8)
            v = fmaxf(v, dpp_f<DPP_HALF_MIRROR>(v));
        if (G >= 16)
            v = fmaxf(v, dpp_f<DPP_MIRROR>(v));
        return v;
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1)
        v = fmaxf(v, __shfl_xor(v, o, G));
    return v;
}

template <int G> __device__ __forceinline__ float gmin(float v)
{
    if (G == 64) {
        v = fminf(v, dpp_f<DPP_XOR1>(v));
        v = fminf(v, dpp_f<DPP_XOR2>(v));
        v = fminf(v, dpp_f<DPP_HALF_MIRROR>(v));
        v = fminf(v, dpp_f<DPP_MIRROR>(v));
        return fminf(fminf(rows4(v, 0), rows4(v, 16)), fminf(rows4(v, 32), rows4(v, 48)));
    }
    if (G <= 16) {
        if (G >= 2)
            v = fminf(v, dpp_f<DPP_XOR1>(v));
        if (G >= 4)
            v = fminf(v, dpp_f<DPP_XOR2>(v));
        if (G >= 8)
            v = fminf(v, dpp_f<DPP_HALF_MIRROR>(v));
        if (G >= 16)
            v = fminf(v, dpp_f<DPP_MIRROR>(v));
        return v;
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1)
        v = fminf(v, __shfl_xor(v, o, G));
    return v;
}

template <int G> __device__ __forceinline__ int gmaxi(int v)
{
    if (G == 64) {
        v = max(v, dpp_i<DPP_XOR1>(v));
        v = max(v, dpp_i<DPP_XOR2>(v));
        v = max(v, dpp_i<DPP_HALF_MIRROR>(v));
        v = max(v, dpp_i<DPP_MIRROR>(v));
        return max(max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
                   max(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
    }
    if (G <= 16) {
        if (G >= 2)
            v = max(v, dpp_i<DPP_XOR1>(v));
        if (G >= 4)
            v = max(v, dpp_i<DPP_XOR2>(v));
        if (G >= 8)
            v = max(v, dpp_i<DPP_HALF_MIRROR>(v));
        if (G >= 16)
            v = max(v, dpp_i<DPP_MIRROR>(v));
        return v;
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1)
        v = max(v, __shfl_xor(v, o, G));
    return v;
}

template <int G> __device__ __forceinline__ int gori(int v)
{
    if (G == 64) {
        v |= dpp_i<DPP_XOR1>(v);
        v |= dpp_i<DPP_XOR2>(v);
        v |= dpp_i<DPP_HALF_MIRROR>(v);
        v |= dpp_i<DPP_MIRROR>(v);
        return (__builtin_amdgcn_readlane(v, 0) | __builtin_amdgcn_readlane(v, 16)) |
               (__builtin_amdgcn_readlane(v, 32) | __builtin_amdgcn_readlane(v, 48));
    }
    if (G <= 16) {
        if (G >= 2)
            v |= dpp_i<DPP_XOR1>(v);
        if (G >= 4)
            v |= dpp_i<DPP_XOR2>(v);
        if (G >= 8)
            v |= dpp_i<DPP_HALF_MIRROR>(v);
        if (G >= 16)
            v |= dpp_i<DPP_MIRROR>(v);
        return v;
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1)
        v |= __shfl_xor(v, o, G);
    return v;
}

template <int G> __device__ __forceinline__ unsigned long long
gor64(unsigned long long v)
{
    const unsigned int lo = (unsigned int)gori<G>((int)(unsigned int)v);
    const unsigned int hi = (unsigned int)gori<G>((int)(unsigned int)(v >> 32));

    return ((unsigned long long)hi << 32) | lo;
}

/* The value x of lane l (0 <= l < G, a constant once the caller's loops
   are unrolled) of the caller's group, in every lane of the group, without
   LDS: G = 16 is one DPP row and row_newbcast:l (DPP control 0x150 + l)
   broadcasts its lane l; G = 8 broadcasts lanes l and 8 + l of the row and
   each half-row keeps its own; G = 64 reads the lane into a scalar
   register. All lanes of the wavefront must be active. */
template <int CTRL> __device__ __forceinline__ float dpp_mov(float v)
{
    /* every lane of the row gets a value (no old operand to keep) */
    return __builtin_bit_cast(
        float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF,
                                        0xF, true));
}

#define TCOH_ROW_BCAST(L)                                                      \
    case L:                                                                    \
        return dpp_mov<0x150 + L>(x)

__device__ __forceinline__ float row_bcast(float x, int l)
{
    switch (l) {
        TCOH_ROW_BCAST(0); TCOH_ROW_BCAST(1); TCOH_ROW_BCAST(2); TCOH_ROW_BCAST(3);
        TCOH_ROW_BCAST(4); TCOH_ROW_BCAST(5); TCOH_ROW_BCAST(6); TCOH_ROW_BCAST(7);
        TCOH_ROW_BCAST(8); TCOH_ROW_BCAST(9); TCOH_ROW_BCAST(10); TCOH_ROW_BCAST(11);
        TCOH_ROW_BCAST(12); TCOH_ROW_BCAST(13); TCOH_ROW_BCAST(14);
    default:
        return dpp_mov<0x150 + 15>(x);
    }
}
#undef TCOH_ROW_BCAST

template <int G> __device__ __forceinline__ float gbcast(float x, int l)
{
    static_assert(G == 8 || G == 16 || G == 64, "group of 8, 16 or 64 lanes");
    if (G == 16)
        return row_bcast(x, l);
    if (G == 8) {
        const float a = row_bcast(x, l), b = row_bcast(x, 8 + l);

        return (threadIdx.x & 8) ? b : a;
    }
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
}

template <int G> __device__ __forceinline__ float2 gbcast2(float2 x, int l)
{
    return make_float2(gbcast<G>(x.x, l), gbcast<G>(x.y, l));
}

/* XCD-aware tiling of a band of R rows x nbx tiles (a tile = TW
   consecutive pixels of one row). The hardware deals work-groups to the 8
   XCDs round-robin (id % 8); with a grid of xcd_grid() groups, group id
   maps to logical tile x*per + id/8 of XCD x, and logical tiles run
   stripe by stripe (stripes sb tiles wide, top to bottom), so each XCD's
   L2 holds the windows of a few neighbouring stripes. sb is the launch's
   tuned stripe width (KParams shp_stripe / link_stripe); 0 takes an eighth
   of the row, one stripe per XCD; nbx makes every XCD a band of rows.
   Returns false for the padding groups. */
#define NXCD 8

__host__ __device__ __forceinline__ int xcd_stripe(int nbx, int want)
{
    const int sb = want > 0 ? want : (nbx + NXCD - 1) / NXCD;

    return sb < nbx ? sb : nbx;
}

__host__ __device__ __forceinline__ int xcd_logical(int R, int nbx, int want)
{
    const int sb = xcd_stripe(nbx, want);

    return ((nbx + sb - 1) / sb) * R * sb;
}

__host__ __forceinline__ int xcd_grid(int R, int nbx, int want)
{
    const int per = (xcd_logical(R, nbx, want) + NXCD - 1) / NXCD;

    return per * NXCD;
}

__device__ __forceinline__ bool xcd_tile(int bid, int R, int nbx, int want, int *row,
                                         int *tx)
{
    const int nl = xcd_logical(R, nbx, want);
    const int per = (nl + NXCD - 1) / NXCD;
    const int l = (bid % NXCD) * per + bid / NXCD;
    const int sb = xcd_stripe(nbx, want);

    if (l >= nl)
        return false;
    const int s = l / (R * sb), rem = l - s * (R * sb);

    *row = rem / sb;
    *tx = s * sb + rem % sb;
    return *tx < nbx;
}

/* Offset of T_ij (i < j) in the packed strict upper triangle, row by row:
   (0,1),(0,2)...(0,N-1),(1,2)... -- also the pair order of pairs_mode=all
   (tcoh.cl:19-23). */
__host__ __device__ __forceinline__ int upper_index(int n, int i, int j)
{
    return i * n - i * (i + 1) / 2 + (j - i - 1);
}

/* Kernel launchers (defined in the .hip files). Each returns a HIP error
   code; s is the handle's stream. */
hipError_t launch_prep(const KParams &kp, const float *amp, float *ampT,
                       const unsigned char *valid, float2 *lstat, hipStream_t s);
hipError_t launch_shp(const KParams &kp, const float *ampT,
                      const unsigned char *valid, const float2 *lstat,
                      const float *adtab, uint32_t *mask, int *count,
                      hipStream_t s);
/* The mixed forms' samples as hi | lo << 16 words (uint2 per sample, NB
   date slots per pixel, one zero pixel appended); FP16 with each pixel
   scaled by its own power of two, whose exponent goes to ex (one short a
   pixel, SPLIT_EX_ZERO for a pixel without non-zero finite samples). */
#define SPLIT_EX_ZERO (-32768)
hipError_t launch_split(const KParams &kp, const float2 *slc, uint32_t *hl, short *ex,
                        int bf16, hipStream_t s);
/* Strip preparation (main.c:52-127) on the device. */
struct StripArgs {
    const float *i, *q;       /* date planes [n][nrows][ncols], NaN null */
    const double *heights;    /* [nrows][ncols], NaN null; or none */
    const double *grid;       /* phase grid [node][height][date]; or none */
    float2 *slc;              /* [padded pixel][n] */
    float *amp;               /* [padded pixel][n], increasing */
    unsigned char *valid;     /* [padded pixel] */
    int n, nrows, ncols;      /* dates, region size */
    int row0, prow, halo, pc; /* first padded row (plane rows), padded rows,
                                 halo, padded width */
    int plane_row0;           /* region row of plane row 0 (the planes may
                                 hold a band of the region's rows) */
    int grid_cols, step, nh;  /* phase grid */
    double hx[3], h0;         /* grid heights, height without elevation */
};
hipError_t launch_strip(const StripArgs &sa, hipStream_t s);
/* EMI state of one band (band pixel bp = output pixel - first pixel of
   the band; capacity: the band's pixels). k_link stores T of every
   estimated pixel (NT float2 each at t) and appends the pixels whose |T|
   fails the Cholesky test to queue 1 (pix, n[0] entries); the first queue
   kernel inverts theirs by Gauss-Jordan and appends those beyond its
   condition bound to queue 2 (next, n[1] entries) for the Jacobi. The
   linking kernels append the pixels whose FP32 eigenvector needs the FP64
   refinement to queue 3 (ref, n[2] entries: bp | kind << 30, kind 1 for a
   Cholesky inverse, 2 for a Gauss-Jordan one), with their eigenvector (N
   float2 at z). n is cleared before each band. */
struct EmiQueue {
    float2 *t;
    float2 *z;
    int *ref;
    int *pix;
    int *n;
    int *next;
};
#define REFINE_KIND_SHIFT 30
/* Fused covariance + linking of output rows row0 .. row0+nrows-1
   (precision: 0 plain, 1 split FP16, 2 split BF16; EMI always plain with
   FP64 sums), one instantiation per stack depth N. EMI appends the pixels
   whose |T| fails the Cholesky test to queue 1 instead of linking them. */
template <int N>
hipError_t launch_link_n(const KParams &kp, const float2 *slc, const uint2 *hl,
                         const short *ex, const uint32_t *mask, const int *count,
                         float *coh, float *pairs, int row0, int nrows, int precision,
                         const EmiQueue &q, hipStream_t s);
/* The EMI kernels of a band after launch_link_n on the same stream (the
   two queue kernels, then the FP64 refinement): pbase is the band's first
   output pixel, cap its pixels; they read the queue counts on the device. */
template <int N>
hipError_t launch_emi_queue_n(const KParams &kp, const EmiQueue &q, int pbase, int cap,
                              float *coh, hipStream_t s);
/* The AD test's term table (2N(N+1) floats) on the device. */
hipError_t launch_adtab(int n, float *q, hipStream_t s);
/* Clear n 32-bit words (a kernel, not a memset: capturable in a graph). */
hipError_t launch_zero_words(uint32_t *p, int n, hipStream_t s);

#endif
