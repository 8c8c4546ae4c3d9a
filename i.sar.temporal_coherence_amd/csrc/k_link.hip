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

template <int N> struct LinkShape {
    static constexpr int G = N <= 16 ? 8 : (N <= 32 ? 16 : 64);
    static constexpr int PPW = WAVE / G;
    static constexpr int NT = N * (N - 1) / 2;
    static constexpr int NB = N <= 16 ? 16 : (N <= 32 ? 32 : 64);
    /* Dynamic LDS of a group: T of each pixel (alive through both stages),
       then a scratch area the covariance and the linking take in turn,
       then (EMI) the inversion areas. */
    static constexpr int TL_BYTES = (PPW * NT * 8 + 15) & ~15;
    static constexpr int EIG_BYTES = PPW * (int)sizeof(EigLds<N>);
    static constexpr int EMI_BYTES = PPW * (int)sizeof(EmiLds<N>);
    /* A group that leaves its Jacobi pixels to the queue uses only the
       Cholesky part of EmiLds: smaller strides let one more group share a
       compute unit's LDS (at N=30, 35 KB a group instead of 51 KB: 4 groups
       per CU, one per SIMD, instead of 3). */
    static constexpr int CH_STRIDE = ((int)sizeof(EmiLds<N>::ch) + 15) & ~15;
    static constexpr int CH_BYTES = PPW * CH_STRIDE;
};

template <int N> static int link_scratch_bytes(const KParams &kp, int precision)
{
    typedef LinkShape<N> S;
    const int cov = precision ? mfma_scratch_bytes(S::PPW, kp.wa * kp.wr, kp.words, S::NB)
                              : S::PPW * N * 4;

    return ((cov > S::EIG_BYTES ? cov : S::EIG_BYTES) + 15) & ~15;
}

/* cov: 0 plain FP32, 1 split FP16 MFMA, 2 split BF16 MFMA (uniform per
   launch: one code object serves the three forms, so the linking code is
   compiled once per depth and estimator). */
/* EMI pixels whose |T| fails the Cholesky pivot test, deferred to
   k_emi_queue: T (NT float2 each), the output pixel, and the count. */
struct EmiQueue {
    float2 *t;
    int *pix;
    int *n;
};

template <int N, int EMI>
__global__ __launch_bounds__(64, (EMI || N > 32) ? 1 : 2) void k_link(KParams kp,
                                               const float2 *__restrict__ slc,
                                               const uint2 *__restrict__ hl,
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
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    float2 *tl = reinterpret_cast<float2 *>(lds);
    unsigned char *scratch = lds + S::TL_BYTES;
    /* With a queue, a group whose |T| needs the Jacobi pseudo-inverse
       leaves it to k_emi_queue, where every group of a wavefront needs it:
       here one such group would hold up the wavefront's other three. */
    const bool defer = EMI && q.t != nullptr;
    const int gi = threadIdx.x / G, lg = threadIdx.x % G;
    EmiLds<N> *em = reinterpret_cast<EmiLds<N> *>(
        scratch + scratch_bytes +
        (EMI ? gi * (defer ? S::CH_STRIDE : (int)sizeof(EmiLds<N>)) : 0));
    int r, tx;

    if (!xcd_tile(blockIdx.x, nrows, (kp.cols + PPW - 1) / PPW, &r, &tx))
        return;
    const int x = tx * PPW + gi;
    const bool live = x < kp.cols;
    const int p = (row0 + r) * kp.cols + (live ? x : 0);
    const int cnt = live ? count[p] : -1;
    const bool est = live && cnt >= kp.min_looks;

    if (cov == 0)
        cov_plain_group<N, G>(kp, slc, mask, p, cnt, est, live, lg,
                              reinterpret_cast<float *>(scratch) + gi * N,
                              tl + gi * NT, pairs);
    else if (cov == 1)
        cov_mfma_band<N, NB, PPW, 0>(kp, hl, mask, count, row0 + r, tx * PPW,
                                     scratch, tl, pairs);
    else
        cov_mfma_band<N, NB, PPW, 1>(kp, hl, mask, count, row0 + r, tx * PPW,
                                     scratch, tl, pairs);
    wave_sync();
    EigLds<N> *sh = reinterpret_cast<EigLds<N> *>(scratch);
    bool deferred;
    const float gam = link_group<N, G, EMI>(kp, tl + gi * NT, est, lg, sh[gi], *em,
                                            !defer, deferred);

    if (defer) {
        const unsigned long long b = __ballot(lg == 0 && deferred);

        if (b) {
            int base = 0;

            if (threadIdx.x == 0)
                base = atomicAdd(q.n, __popcll(b));
            base = __shfl(base, 0);
            if (deferred) {
                const int slot =
                    base + __popcll(b & ((1ull << (gi * G)) - 1ull));
                float2 *dst = q.t + (size_t)slot * NT;

                for (int t = lg; t < NT; t += G)
                    dst[t] = tl[gi * NT + t];
                if (lg == 0)
                    q.pix[slot] = p;
            }
        }
    }
    if (live && lg == 0 && !deferred)
        coh[p] = gam;
}

/* The deferred EMI pixels, PPW to a wavefront: T read from the queue in
   place (no LDS copy, so that four groups share a compute unit's LDS), the
   Jacobi pseudo-inverse of |T| (the Cholesky test fails again, as it did in
   k_link), then the linking as in k_link. */
template <int N>
__global__ __launch_bounds__(64, 1) void k_emi_queue(KParams kp,
                                                     const float2 *__restrict__ qt,
                                                     const int *__restrict__ qpix,
                                                     int qcount,
                                                     float *__restrict__ coh)
{
    typedef LinkShape<N> S;
    constexpr int G = S::G, PPW = S::PPW, NT = S::NT;
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    EigLds<N> *sh = reinterpret_cast<EigLds<N> *>(lds);
    EmiLds<N> *ems = reinterpret_cast<EmiLds<N> *>(lds + S::EIG_BYTES);
    const int gi = threadIdx.x / G, lg = threadIdx.x % G;
    const int slot = blockIdx.x * PPW + gi;
    const bool live = slot < qcount;
    /* a group past the end reads nothing (not estimated) */
    const float2 *tq = qt + (size_t)(live ? slot : 0) * NT;
    bool deferred;
    const float gam = link_group<N, G, 1>(kp, tq, live, lg, sh[gi], ems[gi], true,
                                          deferred);

    if (live && lg == 0)
        coh[qpix[slot]] = gam;
}

template <int N>
hipError_t launch_emi_queue_n(const KParams &kp, const float2 *qt, const int *qpix,
                              int qcount, float *coh, hipStream_t s)
{
    typedef LinkShape<N> S;
    const size_t lds = (size_t)S::EIG_BYTES + S::EMI_BYTES;

    if (qcount <= 0)
        return hipSuccess;
    if (lds > 65536) {
        const hipError_t e = hipFuncSetAttribute(
            reinterpret_cast<const void *>(&k_emi_queue<N>),
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess)
            return e;
    }
    hipLaunchKernelGGL(k_emi_queue<N>, dim3((qcount + S::PPW - 1) / S::PPW),
                       dim3(WAVE), lds, s, kp, qt, qpix, qcount, coh);
    return hipGetLastError();
}

template <int N>
hipError_t launch_link_n(const KParams &kp, const float2 *slc, const uint2 *hl,
                         const uint32_t *mask, const int *count, float *coh,
                         float *pairs, int row0, int nrows, int precision,
                         float2 *qt, int *qpix, int *qn, hipStream_t s)
{
    const EmiQueue q = {qt, qpix, qn};

    typedef LinkShape<N> S;
    const dim3 grid(xcd_grid(nrows, (kp.cols + S::PPW - 1) / S::PPW));
    const int scratch = link_scratch_bytes<N>(kp, precision);
    const size_t lds = (size_t)S::TL_BYTES + scratch +
                       (kp.emi ? (qt ? S::CH_BYTES : S::EMI_BYTES) : 0);

#define TCOH_LINK_LAUNCH(EMI)                                                   \
    do {                                                                        \
        if (lds > 65536) {                                                      \
            const hipError_t e = hipFuncSetAttribute(                           \
                reinterpret_cast<const void *>(&k_link<N, EMI>),                \
                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);          \
            if (e != hipSuccess)                                                \
                return e;                                                       \
        }                                                                       \
        hipLaunchKernelGGL((k_link<N, EMI>), grid, dim3(WAVE), lds, s, kp, slc, \
                           hl, mask, count, coh, pairs, row0, nrows, scratch,   \
                           precision, q);                                       \
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
                                              const uint2 *, const uint32_t *,
                                              const int *, float *, float *, int,
                                              int, int, float2 *, int *, int *,
                                              hipStream_t);
template hipError_t launch_emi_queue_n<TCOH_EIG_N>(const KParams &, const float2 *,
                                                   const int *, int, float *,
                                                   hipStream_t);
