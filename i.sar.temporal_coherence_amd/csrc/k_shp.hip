/* Statistically homogeneous pixel selection on gfx950.
 *
 * One lane per output pixel. The strip's sorted amplitudes are first
 * transposed to planes of four dates (ampT[k/4][pixel][k%4], a float4 a
 * pixel and plane) so that the lanes of a wavefront, which hold
 * consecutive pixels, read one contiguous 1 KB block per four dates of
 * every candidate: one 16-byte load and one address step per four dates. Each lane emits the accepted candidates of its window
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

/* ampT[k/4][pix][k%4] = amp[pix][k] (dates n..NB-1 zero); lstat[pix] =
   (mean, var) of log amplitudes
   of valid pixels when the t-test is used. A block stages its PREP_PX
   pixels' amplitudes (contiguous in amp) in LDS with coalesced reads, then
   writes the date planes; a row of the tile is padded to n + 1 words so
   that the per-pixel reads do not collide in LDS banks. The tile is sized
   from n at launch (a tile for 64 dates took 33 KB a block and held a CU
   to 4 blocks, too few loads in flight for this copy). */
#define PREP_PX 128

__global__ __launch_bounds__(PREP_PX) void k_prep(KParams kp,
                                                 const float *__restrict__ amp,
                                                 float *__restrict__ ampT,
                                                 const unsigned char *__restrict__ valid,
                                                 float2 *__restrict__ lstat)
{
    extern __shared__ float tile[]; /* PREP_PX x (n + 1) */
    const int n = kp.n, n1 = n + 1;
    const int p0 = blockIdx.x * PREP_PX;
    const int np = min(PREP_PX, kp.npad - p0);
    const float *src = amp + (size_t)p0 * n;
    const float inv_n = 1.0f / (float)n;

    /* t < 128 * 64: (t + 0.5) / n is at least 0.5 / n away from an
       integer, far more than the float product's error */
    for (int t = threadIdx.x; t < np * n; t += PREP_PX) {
        const int pp = (int)(((float)t + 0.5f) * inv_n);

        tile[pp * n1 + (t - pp * n)] = src[t];
    }
    __syncthreads();
    if ((int)threadIdx.x >= np)
        return;
    const int pix = p0 + threadIdx.x;
    const float *a = tile + threadIdx.x * n1;

    float4 *quad = reinterpret_cast<float4 *>(ampT);

    for (int k = 0; k < shp_bucket(n); k += 4)
        quad[(size_t)(k >> 2) * kp.npad + pix] =
            make_float4(k < n ? a[k] : 0.0f, k + 1 < n ? a[k + 1] : 0.0f,
                        k + 2 < n ? a[k + 2] : 0.0f, k + 3 < n ? a[k + 3] : 0.0f);
    if (kp.test == 2) {
        float mean = 0.0f, var = 0.0f;

        if (valid[pix]) {
            for (int k = 0; k < n; k++)
                mean += logf_rounded(a[k]);
            mean /= n;
            for (int k = 0; k < n; k++) {
                const float d = logf_rounded(a[k]) - mean;

                var += d * d;
            }
            var /= n - 1;
        }
        lstat[pix] = make_float2(mean, var);
    }
}

/* With ahi[j] = a[j+K] and alo[j] = a[j-K] (the centre shifted by K,
   loaded once per pixel): accept iff b[j] < ahi[j] for j < N-K and
   alo[j] <= b[j] for j >= K. All N candidate samples are loaded
   unconditionally (clamped index) so the loads issue back to back. */
/* The NB dates of pixel pix (n..NB-1 zero) from the planes of four:
   quad q at (q * npad + pix) float4s; npv is npad in a vector register,
   so the plane offsets are per-lane adds, one per four dates. */
template <int NB>
__device__ __forceinline__ void load_dates(float (&b)[NB], const float *__restrict__ ampT,
                                           int pix, int npv)
{
    const float4 *quad = reinterpret_cast<const float4 *>(ampT);
    int idx = pix;

#pragma unroll
    for (int q = 0; q < NB / 4; q++) {
        const float4 v = quad[idx];

        b[4 * q] = v.x;
        b[4 * q + 1] = v.y;
        b[4 * q + 2] = v.z;
        b[4 * q + 3] = v.w;
        idx += npv;
    }
}

/* Date d of pixel pix (one float of its plane of four). */
__device__ __forceinline__ float load_date(const float *__restrict__ ampT, int pix, int d,
                                           int npv)
{
    return ampT[((size_t)(d >> 2) * npv + pix) * 4 + (d & 3)];
}

/* The KS decision from a candidate's NB dates b (loaded by load_dates:
   ampT holds NB dates a pixel, n..NB-1 zero, so every load is in bounds). */
template <int NB>
__device__ __forceinline__ bool ks_test(const float (&ahi)[NB], const float (&alo)[NB],
                                        int kmax, const float (&b)[NB])
{
    bool ok = kmax >= 1;

    /* Dates without a condition carry NaN thresholds: !(b >= NaN) and
       !(NaN > b) hold, and for numbers they are b < ahi and alo <= b
       (amplitudes of valid pixels are never NaN). The tests of a plane of
       four are folded in before the next plane's (a scheduling barrier):
       hoisted together, their 64 lane masks outgrew the scalar registers. */
#pragma unroll
    for (int j = 0; j < NB; j++) {
        ok &= !(b[j] >= ahi[j]) & !(alo[j] > b[j]);
        if (j % 4 == 3)
            __builtin_amdgcn_sched_barrier(0);
    }
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
    /* AD walks the centre's and the candidate's samples with data-dependent
       indices: both sit in LDS, one column per lane (the candidate's loaded
       in one burst of independent loads), so that a step of the walk waits
       on LDS rather than on a global load. */
    __shared__ float acol[TEST == 1 ? NB : 1][WAVE];
    __shared__ float bcol[TEST == 1 ? NB : 1][WAVE];
    const int n = kp.n;
    int y, tx;

    if (!xcd_tile(blockIdx.x, kp.rows, (kp.cols + WAVE - 1) / WAVE, kp.shp_stripe, &y, &tx))
        return;
    const int x = tx * WAVE + threadIdx.x;

    if (x >= kp.cols)
        return;
    const int out = y * kp.cols + x;
    const int centre = (y + kp.wa / 2) * kp.w + x + kp.wr / 2;
    uint32_t *mk = mask + (size_t)out * kp.words;
    float a[NB], ahi[NB], alo[NB];
    float mean_a = 0.0f, var_a = 0.0f;
    uint32_t word = 0;
    int bit = 0, wi = 0, looks = 0;

    if (!valid[centre]) {
        count[out] = -1;
        for (int i = 0; i < kp.words; i++)
            mk[i] = 0u;
        return;
    }
    const int npv = in_vgpr(kp.npad);

    load_dates<NB>(a, ampT, centre, npv);

    if (TEST == 0) {
        /* ahi[j] = a[j+K] for j < N-K, alo[j] = a[j-K] for K <= j < N,
           NaN where the reference walk has no such condition */
        const int k = kp.ks_max;
        const float nan = __builtin_nanf("");

#pragma unroll
        for (int j = 0; j < NB; j++) {
            const float h = load_date(ampT, centre, min(j + k, n - 1), npv);
            const float l = load_date(ampT, centre, min(max(j - k, 0), n - 1), npv);

            ahi[j] = j + k < n ? h : nan;
            alo[j] = (j >= k && j < n) ? l : nan;
        }
    }
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

    /* candidate bit k (dy-major) of the mask and the count */
    auto emit = [&](bool ok) {
        word |= (uint32_t)ok << bit;
        looks += ok;
        if (++bit == 32) {
            mk[wi++] = word;
            word = 0;
            bit = 0;
        }
    };

    /* KS up to 32 date slots: two candidates an iteration, both loaded
       before either is tested, so that the second's loads overlap the
       first's compares (every load is in bounds: both are tested
       unconditionally). At 64 slots the second candidate's registers
       would spill (N=60: 47 -> 73 ms a frame); one at a time there. */
    constexpr bool KS_PAIRS = TEST == 0 && NB <= 32;

    if (KS_PAIRS) {
        const int kw = kp.wa * kp.wr;
        int dy = 0, dx = 0;

        for (int kk = 0; kk < kw; kk += 2) {
            const int c0 = (y + dy) * kp.w + x + dx;

            if (++dx == kp.wr) {
                dx = 0;
                dy++;
            }
            const bool two = kk + 1 < kw;
            const int c1 = two ? (y + dy) * kp.w + x + dx : c0;

            if (two && ++dx == kp.wr) {
                dx = 0;
                dy++;
            }
            float b0[NB], b1[NB];

            load_dates<NB>(b0, ampT, c0, npv);
            load_dates<NB>(b1, ampT, c1, npv);
            const bool v0 = valid[c0] != 0, v1 = valid[c1] != 0;

            emit(ks_test<NB>(ahi, alo, kp.ks_max, b0) & v0);
            if (two)
                emit(ks_test<NB>(ahi, alo, kp.ks_max, b1) & v1);
        }
    }
    for (int dy = 0; !KS_PAIRS && dy < kp.wa; dy++) {
        const int row = (y + dy) * kp.w + x;

        for (int dx = 0; dx < kp.wr; dx++) {
            const int cand = row + dx;
            bool ok = valid[cand] != 0;

            if (TEST == 0) {
                float b[NB];

                load_dates<NB>(b, ampT, cand, npv);
                ok = ks_test<NB>(ahi, alo, kp.ks_max, b) & ok;
            }
            else if (ok) {
                if (TEST == 1) {
                    int ia = 0, ib = 0, m = 0;
                    float s = 0.0f;
                    float b[NB];

                    load_dates<NB>(b, ampT, cand, npv);
#pragma unroll
                    for (int i = 0; i < NB; i++)
                        if (i < n)
                            bcol[i][threadIdx.x] = b[i];
                    /* the current sample of each side; the side that
                       advances is reloaded by one LDS read from the
                       selected column (no branch) */
                    float va = acol[0][threadIdx.x], vb = bcol[0][threadIdx.x];

                    /* One step of the merge; returns the step's table
                       term, whose load is left in flight. */
                    auto step = [&](int j) -> float {
                        const bool take_a = ia < n && (ib >= n || va <= vb);

                        m += take_a;
                        ia += take_a;
                        ib += !take_a;
                        const float *src = take_a ? &acol[min(ia, n - 1)][threadIdx.x]
                                                  : &bcol[min(ib, n - 1)][threadIdx.x];
                        const float v = *src;

                        va = take_a ? v : va;
                        vb = take_a ? vb : v;
                        return adtab[j * (n + 1) + m];
                    };
                    /* Eight steps issue their table loads before the first
                       of their terms is added, so the loads overlap the
                       walk's LDS chain; the terms are still added one by
                       one in j order. */
                    int j = 1;

                    for (; j + 7 < 2 * n; j += 8) {
                        float t[8];

#pragma unroll
                        for (int u = 0; u < 8; u++)
                            t[u] = step(j + u);
#pragma unroll
                        for (int u = 0; u < 8; u++)
                            s = s + t[u];
                    }
                    for (; j < 2 * n; j++)
                        s = s + step(j);
                    /* plain operators: contraction is off in this file
                       (hipcc's __f*_rn helpers are contractable operators
                       and a native square root) */
                    ok = (kp.ad_norm * s - 1.0f) / kp.ad_sigma <= kp.ad_crit;
                }
                else {
                    const float2 st = lstat[cand];
                    /* sqrtf and / are correctly rounded (hipcc's default
                       -fhip-fp32-correctly-rounded-divide-sqrt) */
                    const float se = sqrtf((var_a + st.y) / (float)n);

                    ok = !(se > 0.0f) || fabsf(mean_a - st.x) / se <= kp.tlog_crit;
                }
            }
            emit(ok);
        }
    }
    if (bit)
        mk[wi] = word;
    count[out] = looks;
}

/* The AD terms q(j, m) = (float)(2Nm - jN)^2 / (float)(j(2N - j)) of the
   SHP walk, rounded twice as tcoh.cl:71-72 rounds them (the square, then
   the correctly rounded quotient; no contraction in this file), for j <
   2N and m <= N (row j = 0 is zero). */
__global__ void k_adtab(int n, float *q)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;

    if (i >= 2 * n * (n + 1))
        return;
    const int j = i / (n + 1), m = i - j * (n + 1);

    if (j == 0) {
        q[i] = 0.0f;
        return;
    }
    const float num = (float)(2 * n * m - j * n);
    const float den = (float)(j * (2 * n - j));
    const float sq = num * num;

    q[i] = sq / den;
}

hipError_t launch_adtab(int n, float *q, hipStream_t s)
{
    const int total = 2 * n * (n + 1);

    hipLaunchKernelGGL(k_adtab, dim3((total + 255) / 256), dim3(256), 0, s, n, q);
    return hipGetLastError();
}

__global__ void k_zero_words(uint32_t *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;

    if (i < n)
        p[i] = 0u;
}

hipError_t launch_zero_words(uint32_t *p, int n, hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_zero_words, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_prep(const KParams &kp, const float *amp, float *ampT,
                       const unsigned char *valid, float2 *lstat, hipStream_t s)
{
    hipLaunchKernelGGL(k_prep, dim3((kp.npad + PREP_PX - 1) / PREP_PX), dim3(PREP_PX),
                       (size_t)PREP_PX * (kp.n + 1) * sizeof(float), s, kp, amp, ampT, valid,
                       lstat);
    return hipGetLastError();
}

template <int NB>
static void launch_shp_nb(const KParams &kp, const float *ampT,
                          const unsigned char *valid, const float2 *lstat,
                          const float *adtab, uint32_t *mask, int *count,
                          hipStream_t s)
{
    const dim3 grid(xcd_grid(kp.rows, (kp.cols + WAVE - 1) / WAVE, kp.shp_stripe));

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
