/* Fused covariance + phase linking of a band of output rows (tcoh.cl:
 * 567-744): each work-group (one wavefront) takes PPW = 64/G horizontally
 * adjacent pixels, builds their coherence matrices T in LDS (cov.h, plain
 * or mixed precision) and links them with G lanes per pixel (link.h), so T
 * never leaves the compute unit. Groups are tiled XCD by XCD over column
 * stripes (common.h) so that the windows of one stripe stay in that XCD's
 * L2. Instantiated once per stack depth N (one translation unit each).
 */

#include "cov.h"
#include "link.h"

template <int N> struct LinkShape {
    static constexpr int G = N <= 16 ? 8 : (N <= 32 ? 16 : 64);
    static constexpr int PPW = WAVE / G;
    static constexpr int NT = N * (N - 1) / 2;
    static constexpr int NB = N <= 16 ? 16 : (N <= 32 ? 32 : 64);
};

/* COV: 0 plain FP32, 1 split FP16 MFMA, 2 split BF16 MFMA. */
template <int N, int COV, int EMI>
__global__ __launch_bounds__(64, 2) void k_link(KParams kp,
                                               const float2 *__restrict__ slc,
                                               const uint2 *__restrict__ hl,
                                               const uint32_t *__restrict__ mask,
                                               const int *__restrict__ count,
                                               float *__restrict__ coh,
                                               float *__restrict__ pairs,
                                               int row0, int nrows)
{
    typedef LinkShape<N> S;
    constexpr int G = S::G, PPW = S::PPW, NT = S::NT, NB = S::NB;
    constexpr int WC = COV ? MfmaLds<NB>::WORDS : PPW * N;
    constexpr int WE = (int)(PPW * sizeof(EigLds<N>) / 4);
    /* T of each group's pixel, alive through both stages; the covariance
       scratch and the linking scratch share the rest. */
    __shared__ float2 tl[PPW][NT];
    __shared__ __attribute__((aligned(16))) float scratch[WC > WE ? WC : WE];
    __shared__ EmiLds<N> ems[EMI ? PPW : 1];
    const int gi = threadIdx.x / G, lg = threadIdx.x % G;
    int r, tx;

    if (!xcd_tile(blockIdx.x, nrows, (kp.cols + PPW - 1) / PPW, &r, &tx))
        return;
    const int x = tx * PPW + gi;
    const bool live = x < kp.cols;
    const int p = (row0 + r) * kp.cols + (live ? x : 0);
    const int cnt = live ? count[p] : -1;
    const bool est = live && cnt >= kp.min_looks;

    if (COV == 0) {
        cov_plain_group<N, G>(kp, slc, mask, p, cnt, est, live, lg,
                              scratch + gi * N, tl[gi], pairs);
    }
    else {
        for (int g = 0; g < PPW; g++) {
            const int xg = tx * PPW + g;

            if (xg >= kp.cols)
                break;
            const int pg = (row0 + r) * kp.cols + xg;

            cov_mfma_pixel<N, NB, COV == 2>(kp, hl, mask, pg, count[pg], scratch,
                                            tl[g], pairs);
        }
    }
    wave_sync();
    EigLds<N> *sh = reinterpret_cast<EigLds<N> *>(scratch);
    const float gam = link_group<N, G, EMI>(kp, tl[gi], est, lg, sh[gi],
                                            ems[EMI ? gi : 0]);

    if (live && lg == 0)
        coh[p] = gam;
}

template <int N>
hipError_t launch_link_n(const KParams &kp, const float2 *slc, const uint2 *hl,
                         const uint32_t *mask, const int *count, float *coh,
                         float *pairs, int row0, int nrows, int precision,
                         hipStream_t s)
{
    typedef LinkShape<N> S;
    const dim3 grid(xcd_grid(nrows, (kp.cols + S::PPW - 1) / S::PPW));

#define TCOH_LINK_LAUNCH(COV, EMI)                                           \
    hipLaunchKernelGGL((k_link<N, COV, EMI>), grid, dim3(WAVE), 0, s, kp,    \
                       slc, hl, mask, count, coh, pairs, row0, nrows)
    if (kp.emi) {
        if (precision == 0)
            TCOH_LINK_LAUNCH(0, 1);
        else if (precision == 1)
            TCOH_LINK_LAUNCH(1, 1);
        else
            TCOH_LINK_LAUNCH(2, 1);
    }
    else {
        if (precision == 0)
            TCOH_LINK_LAUNCH(0, 0);
        else if (precision == 1)
            TCOH_LINK_LAUNCH(1, 0);
        else
            TCOH_LINK_LAUNCH(2, 0);
    }
#undef TCOH_LINK_LAUNCH
    return hipGetLastError();
}

/* One translation unit per stack depth (the build compiles this file once
   for every N, in parallel). */
template hipError_t launch_link_n<TCOH_EIG_N>(const KParams &, const float2 *,
                                              const uint2 *, const uint32_t *,
                                              const int *, float *, float *, int,
                                              int, int, hipStream_t);
