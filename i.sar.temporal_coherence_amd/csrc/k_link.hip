/* Fused covariance + phase linking of a band of output rows (tcoh.cl:
 * 567-744): each work-group (one wavefront) takes PPW = 64/G horizontally
 * adjacent pixels, builds their coherence matrices T in LDS (cov.h, plain
 * or mixed precision) and links them with G lanes per pixel (link.h), so T
 * never leaves the compute unit. The LDS is sized per launch from the
 * window (a 21x7 window at N=30 takes under 20 KB per group). Groups are
 * tiled XCD by XCD over column stripes (common.h) so that the windows of
 * one stripe stay in that XCD's L2. Instantiated once per stack depth N
 * (one translation unit each).
 */

#include "cov.h"
#include "link.h"
#include "refine.h"

template <int N> struct LinkShape {
    static constexpr int G = N <= 16 ? 8 : (N <= 32 ? 16 : 64);
    static constexpr int PPW = WAVE / G;
    static constexpr int NT = N * (N - 1) / 2;
    static constexpr int NB = N <= 16 ? 16 : (N <= 32 ? 32 : 64);
    /* Dynamic LDS of a group: T of each pixel (alive through both stages),
       then a scratch area the covariance and the linking take in turn,
       then (EMI) one inversion area per pixel, of the member of EmiLds the
       kernel uses (k_link: the Cholesky factor; the queue kernels: the
       Gauss-Jordan or the Jacobi state). */
    static constexpr int TL_BYTES = (PPW * NT * 8 + 15) & ~15;
    static constexpr int EIG_BYTES = PPW * (int)sizeof(EigLds<N>);
    static constexpr int CH_STRIDE = ((int)sizeof(EmiLds<N>::ch) + 15) & ~15;
    static constexpr int GJ_STRIDE = ((int)sizeof(EmiLds<N>::gj) + 15) & ~15;
    static constexpr int JAC_STRIDE = ((int)sizeof(EmiLds<N>::jac) + 15) & ~15;
};

template <int N> static int link_scratch_bytes(const KParams &kp, int precision)
{
    typedef LinkShape<N> S;
    /* EMI always takes the plain form with FP64 sums; the plain form
       keeps its diagonal sums and the pixels' mask words in LDS */
    const int cov = kp.emi ? S::PPW * (N * 8 + kp.words * 4)
                    : precision ? mfma_scratch_bytes(S::PPW, kp.wa * kp.wr, kp.words, S::NB)
                                : S::PPW * (N * 4 + kp.words * 4);

    return ((cov > S::EIG_BYTES ? cov : S::EIG_BYTES) + 15) & ~15;
}

/* Append the wavefront's deferred pixels (lane lg == 0 of each group with
   deferred set) to a queue: returns the slot of the caller's group (valid
   where deferred). All lanes of the wavefront must be active. */
template <int G>
__device__ __forceinline__ int queue_slot(int *qn, bool lead_deferred, int gi)
{
    const unsigned long long b = __ballot(lead_deferred);
    int base = 0;

    if (b == 0ull)
        return 0;
    if (threadIdx.x == 0)
        base = atomicAdd(qn, __popcll(b));
    base = __shfl(base, 0);
    return base + __popcll(b & ((1ull << (gi * G)) - 1ull));
}

/* The linked phases arg(z_c) of the group's pixel p from its phasors in
   LDS (as link_group leaves them), when the caller asked for them. */
template <int N, int G>
__device__ __forceinline__ void store_phases(const KParams &kp, const EigLds<N> &sh, int p,
                                             int lg, bool linked)
{
    if (!kp.phase)
        return;
    const size_t plane = (size_t)kp.rows * kp.cols;

    for (int c = lg; c < N; c += G) {
        const float2 z = sh.z[c];

        kp.phase[(size_t)c * plane + p] = linked ? atan2f(z.y, z.x) : __builtin_nanf("");
    }
}

/* cov: 0 plain FP32, 1 split FP16 MFMA, 2 split BF16 MFMA (uniform per
   launch: one code object serves the three forms, so the linking code is
   compiled once per depth and estimator); EMI ignores it (plain, FP64
   sums). EMI pixels whose |T| fails the Cholesky pivot test go to queue 1
   (T, NT float2 each, and the output pixel), in place of a result.
   EVD kernels run two wavefronts per SIMD (256 registers each) at every
   depth: above 32 dates that spills (1.4 KB a lane at N=60) but the second
   wavefront hides more than the spills cost (N=45: 49.7 -> 32.8 ms, N=60:
   33.9 -> 29.8 ms on 768^2 and 512^2 frames); EMI keeps one (512). */
template <int N, int EMI>
__global__ __launch_bounds__(64, EMI && N > 32 ? 1 : 2) void k_link(KParams kp,
                                               const float2 *__restrict__ slc,
                                               const uint2 *__restrict__ hl,
                                               const short *__restrict__ ex,
                                               const uint32_t *__restrict__ mask,
                                               const int *__restrict__ count,
                                               float *__restrict__ coh,
                                               float *__restrict__ pairs,
                                               int row0, int nrows,
                                               int scratch_bytes, int cov,
                                               EmiQueue q)
{
    typedef LinkShape<N> S;
    constexpr int G = S::G, PPW = S::PPW, NT = S::NT, NB = S::NB;
    /* EMI up to 32 dates: k_cov_emi has left T in the band's store */
    constexpr bool SPLIT = EMI && N <= 32;
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    float2 *tl = reinterpret_cast<float2 *>(lds);
    unsigned char *scratch = lds + (SPLIT ? 0 : S::TL_BYTES);
    const int gi = threadIdx.x / G, lg = threadIdx.x % G;
    EmiLds<N> *em = reinterpret_cast<EmiLds<N> *>(scratch + scratch_bytes +
                                                  (EMI ? gi * S::CH_STRIDE : 0));
    int r, tx;

    if (!xcd_tile(blockIdx.x, nrows, (kp.cols + PPW - 1) / PPW, kp.link_stripe, &r, &tx))
        return;
    const int x = tx * PPW + gi;
    const bool live = x < kp.cols;
    const int p = (row0 + r) * kp.cols + (live ? x : 0);
    const int cnt = live ? count[p] : -1;
    const bool est = live && cnt >= kp.min_looks;

    if constexpr (SPLIT) {
    }
    else if constexpr (EMI)
        cov_plain_group<N, G, double>(
            kp, slc, mask, p, cnt, est, live, lg, reinterpret_cast<double *>(scratch) + gi * N,
            reinterpret_cast<uint32_t *>(scratch + PPW * N * 8) + gi * kp.words, tl + gi * NT,
            pairs);
    else if (cov == 0)
        cov_plain_group<N, G, float>(
            kp, slc, mask, p, cnt, est, live, lg, reinterpret_cast<float *>(scratch) + gi * N,
            reinterpret_cast<uint32_t *>(scratch + PPW * N * 4) + gi * kp.words, tl + gi * NT,
            pairs);
    else if (cov == 1)
        cov_mfma_band<N, NB, PPW, 0>(kp, hl, ex, mask, count, row0 + r, tx * PPW, row0, nrows,
                                     scratch, tl, pairs);
    else
        cov_mfma_band<N, NB, PPW, 1>(kp, hl, ex, mask, count, row0 + r, tx * PPW, row0, nrows,
                                     scratch, tl, pairs);
    wave_sync();
    EigLds<N> *sh = reinterpret_cast<EigLds<N> *>(scratch);
    const int bp = r * kp.cols + (live ? x : 0); /* band pixel */
    bool deferred;

    if constexpr (EMI && !SPLIT) {
        /* T of the estimated pixel for the queues and the refinement */
        if (est) {
            float2 *dst = q.t + (size_t)bp * NT;

            for (int t = lg; t < NT; t += G)
                dst[t] = tl[gi * NT + t];
        }
    }
    /* (only live groups store an eigenvector: a group past the row's end
       has bp of the row's first pixel) */
    bool refine = false;
    const float gam = link_group<N, G, EMI, SPLIT ? INV_CHOLESKY_GLOBAL : INV_CHOLESKY>(
        kp, SPLIT ? q.t + (size_t)bp * NT : tl + gi * NT, est, lg, sh[gi], *em, deferred,
        EMI && live ? q.z + (size_t)bp * N : nullptr, EMI ? &refine : nullptr);

    if constexpr (EMI) {
        const int slot = queue_slot<G>(q.n, lg == 0 && deferred, gi);
        const bool fine = live && !deferred && refine && !__builtin_isnan(gam);
        const int s3 = queue_slot<G>(q.n + 2, lg == 0 && fine, gi);

        if (lg == 0 && live && deferred)
            q.pix[slot] = bp;
        if (lg == 0 && fine)
            q.ref[s3] = bp | (1 << REFINE_KIND_SHIFT);
    }
    if (live && lg == 0 && !deferred)
        coh[p] = gam;
    if (live && !deferred)
        store_phases<N, G>(kp, sh[gi], p, lg, !__builtin_isnan(gam));
}

/* EMI up to 32 dates, first of two kernels: the band's FP64 covariance
   (cov_plain_group<double>), T of each estimated pixel to the band's store
   q.t, and the pair maps. Apart, the FP64 sums and the linking each fit
   256 registers, and both kernels run two wavefronts per SIMD (together,
   one); k_link then reads T from q.t. */
template <int N>
__global__ __launch_bounds__(64, 2) void k_cov_emi(KParams kp, const float2 *__restrict__ slc,
                                                   const uint32_t *__restrict__ mask,
                                                   const int *__restrict__ count,
                                                   float *__restrict__ pairs, int row0,
                                                   int nrows, EmiQueue q)
{
    typedef LinkShape<N> S;
    constexpr int G = S::G, PPW = S::PPW, NT = S::NT;
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    float2 *tl = reinterpret_cast<float2 *>(lds);
    unsigned char *scratch = lds + S::TL_BYTES;
    const int gi = threadIdx.x / G, lg = threadIdx.x % G;
    int r, tx;

    if (!xcd_tile(blockIdx.x, nrows, (kp.cols + PPW - 1) / PPW, kp.link_stripe, &r, &tx))
        return;
    const int x = tx * PPW + gi;
    const bool live = x < kp.cols;
    const int p = (row0 + r) * kp.cols + (live ? x : 0);
    const int cnt = live ? count[p] : -1;
    const bool est = live && cnt >= kp.min_looks;

    cov_plain_group<N, G, double>(
        kp, slc, mask, p, cnt, est, live, lg, reinterpret_cast<double *>(scratch) + gi * N,
        reinterpret_cast<uint32_t *>(scratch + PPW * N * 8) + gi * kp.words, tl + gi * NT,
        pairs);
    wave_sync();
    if (est) {
        float2 *dst = q.t + ((size_t)r * kp.cols + x) * NT;

        for (int t = lg; t < NT; t += G)
            dst[t] = tl[gi * NT + t];
    }
}

/* Queue 1 (pixels whose |T| failed the Cholesky test), PPW to a
   wavefront, T read from the band's store in place: |T|^-1 by Gauss-Jordan
   where the condition bound shows it is the reference's pseudo-inverse,
   then the linking; the rest go on to queue 2. Persistent: a fixed grid
   strides over the count that k_link left in q.n[0], so that nothing waits
   for it on the host. */
template <int N>
__global__ __launch_bounds__(64, 1) void k_emi_queue(KParams kp, EmiQueue q, int pbase,
                                                     float *__restrict__ coh)
{
    typedef LinkShape<N> S;
    constexpr int G = S::G, PPW = S::PPW, NT = S::NT;
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    EigLds<N> *sh = reinterpret_cast<EigLds<N> *>(lds);
    const int gi = threadIdx.x / G, lg = threadIdx.x % G;
    EmiLds<N> *em = reinterpret_cast<EmiLds<N> *>(lds + S::EIG_BYTES + gi * S::GJ_STRIDE);
    const int total = q.n[0];

    for (int base = blockIdx.x * PPW; base < total; base += gridDim.x * PPW) {
        const int slot = base + gi;
        const bool live = slot < total;
        const int bp = live ? q.pix[slot] : 0;
        bool deferred, refine = false;
        const float gam = link_group<N, G, 1, INV_GAUSS_JORDAN>(
            kp, q.t + (size_t)bp * NT, live, lg, sh[gi], *em, deferred,
            live ? q.z + (size_t)bp * N : nullptr, &refine);
        const int s2 = queue_slot<G>(q.n + 1, lg == 0 && live && deferred, gi);
        const bool fine = live && !deferred && refine && !__builtin_isnan(gam);
        const int s3 = queue_slot<G>(q.n + 2, lg == 0 && fine, gi);

        if (live && lg == 0) {
            if (deferred)
                q.next[s2] = bp;
            else
                coh[pbase + bp] = gam;
        }
        if (lg == 0 && fine)
            q.ref[s3] = bp | (2 << REFINE_KIND_SHIFT);
        if (live && !deferred)
            store_phases<N, G>(kp, sh[gi], pbase + bp, lg, !__builtin_isnan(gam));
    }
}

/* Queue 2 (pixels whose |T| the condition bound leaves to the eigen
   pseudo-inverse): the FP64 Jacobi pseudo-inverse, then the linking (not
   refined: the refinement's inverse is not that pseudo-inverse). */
template <int N>
__global__ __launch_bounds__(64, 1) void k_emi_jacobi(KParams kp, EmiQueue q, int pbase,
                                                      float *__restrict__ coh)
{
    typedef LinkShape<N> S;
    constexpr int G = S::G, PPW = S::PPW, NT = S::NT;
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    EigLds<N> *sh = reinterpret_cast<EigLds<N> *>(lds);
    const int gi = threadIdx.x / G, lg = threadIdx.x % G;
    EmiLds<N> *em = reinterpret_cast<EmiLds<N> *>(lds + S::EIG_BYTES + gi * S::JAC_STRIDE);
    const int total = q.n[1];

    for (int base = blockIdx.x * PPW; base < total; base += gridDim.x * PPW) {
        const int k = base + gi;
        const bool live = k < total;
        const int bp = live ? q.next[k] : 0;
        bool deferred;
        const float gam = link_group<N, G, 1, INV_JACOBI>(kp, q.t + (size_t)bp * NT, live,
                                                          lg, sh[gi], *em, deferred);

        if (live && lg == 0)
            coh[pbase + bp] = gam;
        if (live)
            store_phases<N, G>(kp, sh[gi], pbase + bp, lg, !__builtin_isnan(gam));
    }
}

/* The FP64 refinement (refine.h) of the band's pixels in queue 3 (linked,
   |T|^-1 the inverse, a second eigenvalue of M close to the smallest): one
   inverse iteration from the FP32 eigenvector, phasors and gamma in FP64,
   which replace the FP32 results. G = 32 lanes a pixel (two pixels a
   wavefront) up to 32 dates, 64 above; a fixed grid strides over the
   queue. */
template <int N> struct RefineShape {
    static constexpr int G = N <= 32 ? 32 : 64;
    static constexpr int PPW = WAVE / G;
    static constexpr int STRIDE = ((int)sizeof(RefineLds<N>) + 15) & ~15;
};

template <int N>
__global__ __launch_bounds__(64, 1) void k_emi_refine(KParams kp, EmiQueue q, int pbase,
                                                      float *__restrict__ coh)
{
    typedef RefineShape<N> R;
    constexpr int G = R::G, PPW = R::PPW, NT = N * (N - 1) / 2;
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    const int gi = threadIdx.x / G, lg = threadIdx.x % G;
    RefineLds<N> &sh = *reinterpret_cast<RefineLds<N> *>(lds + gi * R::STRIDE);
    const size_t plane = (size_t)kp.rows * kp.cols;
    const int total = q.n[2];

    for (int base = blockIdx.x * PPW; base < total; base += gridDim.x * PPW) {
        const int k = base + gi;
        const int e = k < total ? q.ref[k] : 0;
        const int how = e >> REFINE_KIND_SHIFT;
        const int bp = e & ((1 << REFINE_KIND_SHIFT) - 1);
        double g;

        if (how == 0)
            continue;
        if (!refine_pixel<N, G>(q.t + (size_t)bp * NT, q.z + (size_t)bp * N, lg, sh,
                                how == 2, g))
            continue;
        if (lg == 0)
            coh[pbase + bp] = (float)g;
        if (kp.phase && lg < N) {
            const double2 u = sh.z[lg];

            kp.phase[(size_t)lg * plane + pbase + bp] = (float)atan2(u.y, u.x);
        }
    }
}

static hipError_t lds_limit(const void *fn, size_t lds)
{
    if (lds <= 65536)
        return hipSuccess;
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

/* Groups of the persistent queue kernels: enough for the whole chip, no
   more than the queue can use. */
static unsigned queue_grid(int cap, int ppw, int groups)
{
    const int need = (cap + ppw - 1) / ppw;

    return (unsigned)(need < groups ? (need > 0 ? need : 1) : groups);
}

template <int N>
hipError_t launch_emi_queue_n(const KParams &kp, const EmiQueue &q, int pbase, int cap,
                              float *coh, hipStream_t s)
{
    typedef LinkShape<N> S;
    typedef RefineShape<N> R;
    const size_t lq = (size_t)S::EIG_BYTES + (size_t)S::PPW * S::GJ_STRIDE;
    const size_t lj = (size_t)S::EIG_BYTES + (size_t)S::PPW * S::JAC_STRIDE;
    const size_t lr = (size_t)R::PPW * R::STRIDE;
    hipError_t e;

    if ((e = lds_limit(reinterpret_cast<const void *>(&k_emi_queue<N>), lq)) != hipSuccess)
        return e;
    if ((e = lds_limit(reinterpret_cast<const void *>(&k_emi_jacobi<N>), lj)) != hipSuccess)
        return e;
    if ((e = lds_limit(reinterpret_cast<const void *>(&k_emi_refine<N>), lr)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_emi_queue<N>, dim3(queue_grid(cap, S::PPW, 4096)), dim3(WAVE), lq,
                       s, kp, q, pbase, coh);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_emi_jacobi<N>, dim3(queue_grid(cap, S::PPW, 1024)), dim3(WAVE), lj,
                       s, kp, q, pbase, coh);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_emi_refine<N>, dim3(queue_grid(cap, R::PPW, 4096)), dim3(WAVE), lr,
                       s, kp, q, pbase, coh);
    return hipGetLastError();
}

template <int N>
hipError_t launch_link_n(const KParams &kp, const float2 *slc, const uint2 *hl,
                         const short *ex, const uint32_t *mask, const int *count,
                         float *coh, float *pairs, int row0, int nrows, int precision,
                         const EmiQueue &q, hipStream_t s)
{
    typedef LinkShape<N> S;
    const dim3 grid(xcd_grid(nrows, (kp.cols + S::PPW - 1) / S::PPW, kp.link_stripe));
    const int scratch = link_scratch_bytes<N>(kp, precision);
    const bool split = kp.emi && N <= 32;
    /* the split EMI kernel keeps no T in LDS: its eigen area and the
       Cholesky factors only */
    const size_t lds = (split ? 0 : (size_t)S::TL_BYTES) + scratch +
                       (kp.emi ? S::PPW * S::CH_STRIDE : 0);

    if (kp.emi && !(q.t && q.z && q.ref && q.pix && q.n && q.next))
        return hipErrorInvalidValue;
    if constexpr (N <= 32) {
        if (split) {
            const size_t lc = (size_t)S::TL_BYTES + ((S::PPW * (N * 8 + kp.words * 4) + 15) & ~15);
            const hipError_t e = lds_limit(reinterpret_cast<const void *>(&k_cov_emi<N>), lc);

            if (e != hipSuccess)
                return e;
            hipLaunchKernelGGL(k_cov_emi<N>, grid, dim3(WAVE), lc, s, kp, slc, mask, count,
                               pairs, row0, nrows, q);
            if (hipGetLastError() != hipSuccess)
                return hipErrorLaunchFailure;
        }
    }
#define TCOH_LINK_LAUNCH(EMI)                                                   \
    do {                                                                        \
        const hipError_t e =                                                    \
            lds_limit(reinterpret_cast<const void *>(&k_link<N, EMI>), lds);    \
        if (e != hipSuccess)                                                    \
            return e;                                                           \
        hipLaunchKernelGGL((k_link<N, EMI>), grid, dim3(WAVE), lds, s, kp, slc, \
                           hl, ex, mask, count, coh, pairs, row0, nrows,        \
                           scratch, precision, q);                              \
    } while (0)
    if (kp.emi)
        TCOH_LINK_LAUNCH(1);
    else
        TCOH_LINK_LAUNCH(0);
#undef TCOH_LINK_LAUNCH
    return hipGetLastError();
}

/* One translation unit per stack depth (the build compiles this file once
   for every N, in parallel). */
template hipError_t launch_link_n<TCOH_EIG_N>(const KParams &, const float2 *,
                                              const uint2 *, const short *,
                                              const uint32_t *, const int *, float *,
                                              float *, int, int, int, const EmiQueue &,
                                              hipStream_t);
template hipError_t launch_emi_queue_n<TCOH_EIG_N>(const KParams &, const EmiQueue &, int,
                                                   int, float *, hipStream_t);
