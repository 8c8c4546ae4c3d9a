/* Statistically homogeneous pixel selection on gfx950.
 *
 * One lane per output pixel. The strip's sorted amplitudes are first
 * transposed to date-planar order (ampT[k][pixel]) so that the lanes of a
 * wavefront, which hold consecutive pixels, read consecutive addresses for
 * every candidate. Each lane emits the accepted candidates of its window
 * as a bit mask (dy-major, bit dy*WR+dx, as the reference's candidate
 * order tcoh.cl:572-576) and the SHP count (tcoh.cl:603).
 *
 * KS (tcoh.cl:34-54). The reference walks the merge of the two sorted
 * samples, centre first on ties, and accepts when the largest |ia - ib|
 * stays <= KS_MAX. After a centre sample a_i is taken, ib counts the
 * candidate samples strictly below a_i; after a candidate sample b_j,
 * ia counts the centre samples <= b_j. With K = KS_MAX the walk therefore
 * stays within K exactly when, for every i >= K,
 *     b[i-K] < a[i]   and   a[i-K] <= b[i],
 * which the lanes test without branches (states after one sample is used
 * up only shrink |ia - ib|). Same float comparisons, so the decision is
 * bit-for-bit the reference's.
 *
 * AD (tcoh.cl:56-75). The same merge gives m_j, the centre samples among
 * the first j merged; the kernel adds q(j, m_j) = (float)(2Nm - jN)^2 /
 * (float)(j(2N - j)) in j order in FP32, with q taken from a table the
 * host fills with the same two IEEE roundings, then tests
 * (AD_NORM * s - 1) / AD_SIGMA <= AD_CRIT without contraction.
 *
 * tlog (tcoh.cl:77-95, 555-565). The log-amplitude mean and unbiased
 * variance of every pixel are computed once in the reference's summation
 * order; log is the double logarithm rounded to float.
 */

#include "common.h"

#pragma clang fp contract(off)

__device__ __forceinline__ float logf_rounded(float x)
{
    return (float)log((double)x);
}

/* ampT[k][pix] = amp[pix][k]; lstat[pix] = (mean, var) of log amplitudes
   of valid pixels when the t-test is used. */
__global__ __launch_bounds__(256) void k_prep(KParams kp,
                                             const float *__restrict__ amp,
                                             float *__restrict__ ampT,
                                             const unsigned char *__restrict__ valid,
                                             float2 *__restrict__ lstat)
{
    const int pix = blockIdx.x * 256 + threadIdx.x;
    const int n = kp.n;

    if (pix >= kp.npad)
        return;
    for (int k = 0; k < n; k++)
        ampT[(size_t)k * kp.npad + pix] = amp[(size_t)pix * n + k];
    if (kp.test == 2) {
        float mean = 0.0f, var = 0.0f;

        if (valid[pix]) {
            for (int k = 0; k < n; k++)
                mean += logf_rounded(amp[(size_t)pix * n + k]);
            mean /= n;
            for (int k = 0; k < n; k++) {
                const float d = logf_rounded(amp[(size_t)pix * n + k]) - mean;

                var += d * d;
            }
            var /= n - 1;
        }
        lstat[pix] = make_float2(mean, var);
    }
}

template <int NB>
__device__ __forceinline__ bool ks_ok(const float (&a)[NB], int n, int kmax,
                                      const float *__restrict__ ampT,
                                      int npad, int cand)
{
    bool ok = kmax >= 1;

#pragma unroll
    for (int i = 0; i < NB; i++)
        if (i >= kmax && i < n)
            ok &= ampT[(size_t)(i - kmax) * npad + cand] < a[i];
#pragma unroll
    for (int j = 0; j < NB; j++)
        if (j + kmax < n)
            ok &= a[j] <= ampT[(size_t)(j + kmax) * npad + cand];
    return ok;
}

template <int NB, int TEST>
__global__ __launch_bounds__(64) void k_shp(KParams kp,
                                           const float *__restrict__ ampT,
                                           const unsigned char *__restrict__ valid,
                                           const float2 *__restrict__ lstat,
                                           const float *__restrict__ adtab,
                                           uint32_t *__restrict__ mask,
                                           int *__restrict__ count)
{
    /* AD walks the centre sample with a data-dependent index. */
    __shared__ float acol[TEST == 1 ? NB : 1][WAVE];
    const int out = blockIdx.x * WAVE + threadIdx.x;
    const int n = kp.n;
    const int y = out / kp.cols, x = out - (out / kp.cols) * kp.cols;
    const int centre = (y + kp.wa / 2) * kp.w + x + kp.wr / 2;
    uint32_t *mk = mask + (size_t)out * kp.words;
    float a[NB];
    float mean_a = 0.0f, var_a = 0.0f;
    uint32_t word = 0;
    int bit = 0, wi = 0, looks = 0;

    if (out >= kp.rows * kp.cols)
        return;
    if (!valid[centre]) {
        count[out] = -1;
        for (int i = 0; i < kp.words; i++)
            mk[i] = 0u;
        return;
    }
#pragma unroll
    for (int i = 0; i < NB; i++)
        a[i] = i < n ? ampT[(size_t)i * kp.npad + centre] : 0.0f;
    if (TEST == 1) {
#pragma unroll
        for (int i = 0; i < NB; i++)
            if (i < n)
                acol[i][threadIdx.x] = a[i];
    }
    if (TEST == 2) {
        const float2 st = lstat[centre];

        mean_a = st.x;
        var_a = st.y;
    }

    for (int dy = 0; dy < kp.wa; dy++) {
        const int row = (y + dy) * kp.w + x;

        for (int dx = 0; dx < kp.wr; dx++) {
            const int cand = row + dx;
            bool ok = valid[cand] != 0;

            if (ok) {
                if (TEST == 0) {
                    ok = ks_ok<NB>(a, n, kp.ks_max, ampT, kp.npad, cand);
                }
                else if (TEST == 1) {
                    int ia = 0, ib = 0, m = 0;
                    float s = 0.0f;

                    for (int j = 1; j < 2 * n; j++) {
                        bool take_a;

                        if (ia >= n)
                            take_a = false;
                        else if (ib >= n)
                            take_a = true;
                        else
                            take_a = acol[ia][threadIdx.x] <=
                                     ampT[(size_t)ib * kp.npad + cand];
                        m += take_a;
                        ia += take_a;
                        ib += !take_a;
                        s = s + adtab[j * (n + 1) + m];
                    }
                    ok = __fdiv_rn(__fsub_rn(__fmul_rn(kp.ad_norm, s), 1.0f),
                                   kp.ad_sigma) <= kp.ad_crit;
                }
                else {
                    const float2 st = lstat[cand];
                    const float se = __fsqrt_rn(__fdiv_rn(var_a + st.y, (float)n));

                    ok = !(se > 0.0f) ||
                         __fdiv_rn(fabsf(mean_a - st.x), se) <= kp.tlog_crit;
                }
            }
            word |= (uint32_t)ok << bit;
            looks += ok;
            if (++bit == 32) {
                mk[wi++] = word;
                word = 0;
                bit = 0;
            }
        }
    }
    if (bit)
        mk[wi] = word;
    count[out] = looks;
}

hipError_t launch_prep(const KParams &kp, const float *amp, float *ampT,
                       const unsigned char *valid, float2 *lstat,
                       hipStream_t s)
{
    hipLaunchKernelGGL(k_prep, dim3((kp.npad + 255) / 256), dim3(256), 0, s,
                       kp, amp, ampT, valid, lstat);
    return hipGetLastError();
}

template <int NB>
static void launch_shp_nb(const KParams &kp, const float *ampT,
                          const unsigned char *valid, const float2 *lstat,
                          const float *adtab, uint32_t *mask, int *count,
                          hipStream_t s)
{
    const dim3 grid((kp.rows * kp.cols + WAVE - 1) / WAVE);

    if (kp.test == 0)
        hipLaunchKernelGGL((k_shp<NB, 0>), grid, dim3(WAVE), 0, s, kp, ampT,
                           valid, lstat, adtab, mask, count);
    else if (kp.test == 1)
        hipLaunchKernelGGL((k_shp<NB, 1>), grid, dim3(WAVE), 0, s, kp, ampT,
                           valid, lstat, adtab, mask, count);
    else
        hipLaunchKernelGGL((k_shp<NB, 2>), grid, dim3(WAVE), 0, s, kp, ampT,
                           valid, lstat, adtab, mask, count);
}

hipError_t launch_shp(const KParams &kp, const float *ampT,
                      const unsigned char *valid, const float2 *lstat,
                      const float *adtab, uint32_t *mask, int *count,
                      hipStream_t s)
{
    if (kp.n <= 16)
        launch_shp_nb<16>(kp, ampT, valid, lstat, adtab, mask, count, s);
    else if (kp.n <= 32)
        launch_shp_nb<32>(kp, ampT, valid, lstat, adtab, mask, count, s);
    else
        launch_shp_nb<64>(kp, ampT, valid, lstat, adtab, mask, count, s);
    return hipGetLastError();
}
