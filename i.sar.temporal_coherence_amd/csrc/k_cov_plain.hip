/* SHP-masked covariance, FP32 on the vector ALUs (the plain form).
 *
 * C_ij = sum over accepted candidates of s_i conj(s_j) (tcoh.cl:567-602),
 * then T_ij = C_ij / sqrt(C_ii C_jj) with the optional bias correction of
 * the magnitudes (tcoh.cl:610-636) and the pair coherences |T_ij|
 * (tcoh.cl:515-527, 637-639).
 *
 * Layout: G lanes per pixel. The upper triangle (with diagonal) of the
 * N x N matrix is cut into NBLK x NBLK blocks of B x B; lane l of a group
 * owns upper block l (NBLK(NBLK+1)/2 <= G), keeps its B*B complex sums in
 * registers and, per accepted candidate, loads the B row samples and B
 * column samples it needs (L1/L2 hits: the 4 pixels of a wavefront share
 * most of their windows). Candidates are visited in the reference's
 * dy-major order, so every sum has the reference's summation order.
 *
 * Output: T_ij (i < j) packed row by row per pixel into the batch scratch,
 * and the pair planes. Pixels with fewer than MIN_LOOKS SHPs (or an
 * invalid centre) get NaN pairs and no T.
 */

#include "common.h"

template <int B, int NBLK>
__global__ __launch_bounds__(64) void k_cov_plain(KParams kp,
                                                  const float2 *__restrict__ slc,
                                                  const uint32_t *__restrict__ mask,
                                                  const int *__restrict__ count,
                                                  float2 *__restrict__ tbuf,
                                                  float *__restrict__ pairs,
                                                  int pix0, int npix)
{
    constexpr int G = NBLK <= 5 ? 16 : 64;
    constexpr int PPW = WAVE / G;
    constexpr int NMAX = B * NBLK;
    __shared__ float diag[PPW][NMAX];
    const int gi = threadIdx.x / G, lg = threadIdx.x % G;
    const int q = blockIdx.x * PPW + gi;
    const bool live = q < npix;
    const int p = pix0 + (live ? q : 0);
    const int n = kp.n;
    const int y = p / kp.cols, x = p - (p / kp.cols) * kp.cols;
    const int cnt = live ? count[p] : -1;
    const bool est = cnt >= kp.min_looks;
    const int ntri = n * (n - 1) / 2;
    const size_t plane = (size_t)kp.rows * kp.cols;
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
                const float2 *s =
                    slc + ((size_t)(y + dy) * kp.w + x + dx) * n;
                float2 sr[B], sc[B];

#pragma unroll
                for (int r = 0; r < B; r++) {
                    sr[r] = r0 + r < n ? s[r0 + r] : make_float2(0.f, 0.f);
                    sc[r] = c0 + r < n ? s[c0 + r] : make_float2(0.f, 0.f);
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
            if (r0 + r < n)
                diag[gi][r0 + r] = cr[r][r];
    }
    wave_sync();
    if (!live || !own)
        return;

    float2 *tp = tbuf + (size_t)q * ntri;
#pragma unroll
    for (int r = 0; r < B; r++) {
        const int i = r0 + r;

#pragma unroll
        for (int c = 0; c < B; c++) {
            const int j = c0 + c;
            float re = 0.0f, im = 0.0f;
            int pidx = -1;

            if (i >= n || j >= n || j <= i)
                continue;
            if (kp.npairs)
                pidx = kp.pairs_all ? upper_index(n, i, j) : (j == i + 1 ? i : -1);
            if (!est) {
                if (pidx >= 0)
                    pairs[(size_t)pidx * plane + p] = __builtin_nanf("");
                continue;
            }
            {
                const float den = sqrtf(diag[gi][i]) * sqrtf(diag[gi][j]);

                if (den > 0.0f) {
                    re = cr[r][c] / den;
                    im = ci[r][c] / den;
                    if (kp.bias && cnt > 1) {
                        const float m2 = re * re + im * im;
                        const float m2c = (cnt * m2 - 1.0f) / (cnt - 1.0f);
                        const float f =
                            (m2 > 0.0f && m2c > 0.0f) ? sqrtf(m2c / m2) : 0.0f;

                        re *= f;
                        im *= f;
                    }
                }
            }
            tp[upper_index(n, i, j)] = make_float2(re, im);
            if (pidx >= 0)
                pairs[(size_t)pidx * plane + p] = hypotf(re, im);
        }
    }
}

template <int B, int NBLK>
static void launch_b(const KParams &kp, const float2 *slc,
                     const uint32_t *mask, const int *count, float2 *tbuf,
                     float *pairs, int pix0, int npix, hipStream_t s)
{
    constexpr int G = NBLK <= 5 ? 16 : 64;
    constexpr int PPW = WAVE / G;

    hipLaunchKernelGGL((k_cov_plain<B, NBLK>), dim3((npix + PPW - 1) / PPW),
                       dim3(WAVE), 0, s, kp, slc, mask, count, tbuf, pairs,
                       pix0, npix);
}

hipError_t launch_cov_plain(const KParams &kp, const float2 *slc,
                            const uint32_t *mask, const int *count,
                            float2 *tbuf, float *pairs, int pix0, int npix,
                            hipStream_t s)
{
    const int n = kp.n;

    if (n <= 10)
        launch_b<2, 5>(kp, slc, mask, count, tbuf, pairs, pix0, npix, s);
    else if (n <= 15)
        launch_b<3, 5>(kp, slc, mask, count, tbuf, pairs, pix0, npix, s);
    else if (n <= 20)
        launch_b<4, 5>(kp, slc, mask, count, tbuf, pairs, pix0, npix, s);
    else if (n <= 25)
        launch_b<5, 5>(kp, slc, mask, count, tbuf, pairs, pix0, npix, s);
    else if (n <= 30)
        launch_b<6, 5>(kp, slc, mask, count, tbuf, pairs, pix0, npix, s);
    else if (n <= 35)
        launch_b<7, 5>(kp, slc, mask, count, tbuf, pairs, pix0, npix, s);
    else if (n <= 40)
        launch_b<8, 5>(kp, slc, mask, count, tbuf, pairs, pix0, npix, s);
    else if (n <= 50)
        launch_b<5, 10>(kp, slc, mask, count, tbuf, pairs, pix0, npix, s);
    else if (n <= 60)
        launch_b<6, 10>(kp, slc, mask, count, tbuf, pairs, pix0, npix, s);
    else
        launch_b<7, 10>(kp, slc, mask, count, tbuf, pairs, pix0, npix, s);
    return hipGetLastError();
}
