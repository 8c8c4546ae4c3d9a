/* Phase linking and temporal coherence on gfx950 (tcoh.cl:641-744):
 * device functions of the linking stage of k_link.
 *
 * G lanes per pixel; lane l of a group owns the columns c = l + m*G
 * (m < CPL) of the pixel's N x N Hermitian matrix and keeps them, both
 * triangles, in registers. The coherence matrix T comes from the group's
 * LDS (packed strict upper triangle); row-wise data every lane needs
 * (Householder vectors, the tridiagonal, the phasors) passes through a
 * small LDS area of the group.
 *
 *  connectivity (tcoh.cl:494-513)   bit-mask breadth-first search
 *  EVD (tcoh.cl:649-656)            eigenvector of the largest eigenvalue
 *  EMI (tcoh.cl:657-694)            |T|^-1 by Cholesky in FP64 (pivot
 *                                   floor 1e-5 as the reference); else the
 *                                   eigen pseudo-inverse (relative floor
 *                                   1e-6), in FP64: a Gauss-Jordan
 *                                   inverse where a condition bound shows
 *                                   that the floor drops nothing, else a
 *                                   cyclic Jacobi; then the smallest
 *                                   eigenvalue of |T|^-1 o T
 *  eigenvector (tcoh.cl:201-444)    Householder reduction to a real
 *                                   tridiagonal (zhetd2, lower), then the
 *                                   eigenvalue of the wanted rank by
 *                                   G-way multisection of Sturm counts,
 *                                   inverse iteration (3 steps, pivoted
 *                                   LU) and the Householder back-transform
 *  phasors, gamma (tcoh.cl:696-744)
 *
 * A work-group is one wavefront. Groups of the same wavefront may take
 * different branches (Cholesky or Jacobi, early convergence), so lanes
 * order their LDS traffic with wavefront-scope fences, never s_barrier.
 */

#ifndef TCOH_EIG_H
#define TCOH_EIG_H

#include "common.h"

#define PHASOR_FLOOR 1.0e-3f
#define CHOL_FLOOR 1.0e-5
#define EIG_FLOOR 1.0e-6f
#define MAX_SWEEPS 40

/* Keep the compiler from hoisting the LDS loads of later steps of an
   unrolled sequence above this point (it would hold them all in registers
   and spill them): a compiler-level memory barrier, which emits nothing,
   and a scheduling barrier. */
#define TCOH_STEP_FENCE()                                                      \
    do {                                                                       \
        asm volatile("" ::: "memory");                                         \
        __builtin_amdgcn_sched_barrier(0);                                     \
    } while (0)

/* Per-pixel LDS of the eigen solve: the tridiagonal (d, e) and tau of the
   reduction, the inverse iteration's LU factors (u0 holds 1/pivot) and
   result y, the phasors z. In groups of 64 lanes (depths above 32) the
   reduction broadcasts its row operands through vw[r] = (v_r, w_r) (one
   16-byte word a row), which the inverse iteration's arrays replace. */
template <int N> struct EigLds {
    union {
        float4 vw[(N + 3) & ~3];
        struct {
            float u0[N], u1[N], u2[N], l[N], y[N];
            int piv[N];
        };
    };
    float2 tau[N];
    float d[N], e[N], e2[N]; /* e2 = e * e, rounded as the Sturm count's */
    float2 z[N];
};

/* Inversion state of one pixel's |T| (EMI), FP64, one area per group; each
   kernel reserves only the member it uses:
   ch   Cholesky factor (lower, column-major packed), reciprocal diagonal,
        the pivot flag and, above 32 dates, W = L^-1
   gj   Gauss-Jordan exchange: the pivot row of the current step, the
        pivot row of every step and |T|^-1 (float, both triangles)
   jac  cyclic Jacobi eigendecomposition: A, V (row-major N x N) and the
        reciprocal eigenvalues */
template <int N> union __attribute__((aligned(16))) EmiLds {
    struct {
        double lc[N * (N + 1) / 2];
        double rd[N];
        double flag;
        double w[N > 32 ? N * (N + 1) / 2 : 1]; /* L^-1 (N > 32) */
    } ch;
    struct {
        double prow[N];
        int pk[N];
        float inv[N * N];
    } gj;
    struct {
        double a[N * N];
        double v[N * N];
        double il[N];
    } jac;
};

template <int N> __device__ __forceinline__ int lpos(int i, int j)
{
    /* L[i][j], i >= j, column-major packed. */
    return j * N - j * (j - 1) / 2 + (i - j);
}

/* The group's copy of the pixel's N x N Hermitian matrix: lane lg holds
   the columns c = lg + m*G (m < CPL), both triangles, real and imaginary
   parts apart, rows in groups of four (one float4 per group), the layout of
   the accumulator of v_mfma_f32_4x4x1f32: in each block of four adjacent
   lanes, that instruction's result register i of lane j is (its A operand
   in the block's lane i) x (its B operand in lane j) + C. One instruction
   thus adds a rank-1 term to four rows of every lane's column. Rows N ..
   4*NG-1 are padding. */
template <int N, int CPL> struct HMat {
    static constexpr int NG = (N + 3) / 4;
    float4_t re[CPL][NG], im[CPL][NG];

    __device__ __forceinline__ float2 get(int m, int r) const
    {
        return make_float2(re[m][r >> 2][r & 3], im[m][r >> 2][r & 3]);
    }
    __device__ __forceinline__ void set(int m, int r, float2 v)
    {
        re[m][r >> 2][r & 3] = v.x;
        im[m][r >> 2][r & 3] = v.y;
    }
};

/* c + a_row x b_col on the 4 x 4 blocks (FP32 matrix cores). */
__device__ __forceinline__ float4_t rank1x4(float a, float b, float4_t c)
{
    return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0);
}

/* One row of the Householder matrix-vector product for G = 16, with the
   row operand broadcast inside the multiply-adds: VOP2 v_fmac_f32 with
   DPP row_newbcast:L reads lane L of the row as src0. The leading s_nop 1
   gives the two wait states a DPP read needs after a VALU write of its
   source, and one a VALU read needs after a transcendental write; nothing
   inside the block writes a broadcast source. */
#define TCOH_FMAC_BC(D, S, B, L)                                               \
    "v_fmac_f32_dpp " D ", " S ", " B " row_newbcast:" L                        \
    " row_mask:0xf bank_mask:0xf\n"
#define TCOH_MUL_BC(D, S, B, L)                                                \
    "v_mul_f32_dpp " D ", " S ", " B " row_newbcast:" L                         \
    " row_mask:0xf bank_mask:0xf\n"

template <int L> struct RowOps {
    /* The first row of the partial sums: written, not accumulated. */
    static __device__ __forceinline__ void matvec1_first(float &sr, float &sia,
                                                         float &sib, float vx,
                                                         float vy, float ax, float ay)
    {
        asm volatile("s_nop 1\n"
                     TCOH_MUL_BC("%0", "%3", "%5", "%7")
                     TCOH_FMAC_BC("%0", "%4", "%6", "%7")
                     TCOH_MUL_BC("%1", "%4", "%5", "%7")
                     TCOH_MUL_BC("%2", "%3", "%6", "%7")
                     : "=&v"(sr), "=&v"(sia), "=&v"(sib)
                     : "v"(vx), "v"(vy), "v"(ax), "v"(ay), "n"(L));
    }
    /* w partial sums: sr += vx*ax + vy*ay, sia += vy*ax, sib += vx*ay */
    static __device__ __forceinline__ void matvec1(float &sr, float &sia, float &sib,
                                                   float vx, float vy, float ax,
                                                   float ay)
    {
        asm volatile("s_nop 1\n"
                     TCOH_FMAC_BC("%0", "%3", "%5", "%7")
                     TCOH_FMAC_BC("%0", "%4", "%6", "%7")
                     TCOH_FMAC_BC("%1", "%4", "%5", "%7")
                     TCOH_FMAC_BC("%2", "%3", "%6", "%7")
                     : "+v"(sr), "+v"(sia), "+v"(sib)
                     : "v"(vx), "v"(vy), "v"(ax), "v"(ay), "n"(L));
    }
};

/* One row r > k of a 64-lane group's matrix-vector product, with v_r read
   from LDS as t.xy. */
__device__ __forceinline__ void matvec_row64(bool live, float4 t, float2 a, float &sr,
                                             float &sia, float &sib)
{
    if (!live)
        return;
    sr = fmaf(t.x, a.x, sr);
    sr = fmaf(t.y, a.y, sr);
    sia = fmaf(t.y, a.x, sia);
    sib = fmaf(t.x, a.y, sib);
}

#define TCOH_ROW_CASES(F)                                                      \
    F(0) F(1) F(2) F(3) F(4) F(5) F(6) F(7) F(8) F(9) F(10) F(11) F(12) F(13)  \
    F(14) F(15)

__device__ __forceinline__ void row_matvec1(int l, bool first, float &sr, float &sia,
                                            float &sib, float2 v, float2 a)
{
#define TCOH_CASE(L)                                                           \
    case L:                                                                    \
        if (first)                                                             \
            RowOps<L>::matvec1_first(sr, sia, sib, v.x, v.y, a.x, a.y);        \
        else                                                                   \
            RowOps<L>::matvec1(sr, sia, sib, v.x, v.y, a.x, a.y);              \
        break;
    switch (l) { TCOH_ROW_CASES(TCOH_CASE) }
#undef TCOH_CASE
}

/* Householder reduction of the group's matrix A to the real tridiagonal
   (d, e) in LDS; the Householder vector of step k is left in row k of the
   columns c > k, tau_k in LDS.

   Slot m holds columns m*G .. m*G+G-1; a slot whose columns are all
   finished at step k is skipped at compile time. Finished columns inside
   a live slot (c <= k) are updated like the others: their rows below k
   are never read again (this layout keeps Householder vectors in rows,
   not columns), so only the reductions are masked.

   Per step: the matrix-vector product w = tau A v on the vector ALUs, the
   row operand v_j of each row reaching the other lanes as a lane broadcast
   inside the multiply-adds (G = 16), a broadcast value (G = 8) or an LDS
   read (G = 64); then the rank-2 update A -= v w^H + w v^H on the matrix
   cores, eight 4 x 4 rank-1 products for every four rows of a slot (real
   part: v_r.x(-w_c.x) + v_r.y(-w_c.y) + w_r.x(-v_c.x) + w_r.y(-v_c.y);
   imaginary part: v_r.y(-w_c.x) + v_r.x w_c.y + w_r.y(-v_c.x) + w_r.x v_c.y),
   with (v_r, w_r) of the block lane's row read from LDS, zero for rows <= k
   and padding rows, so that a group of four rows is updated whole. */
template <int N, int G, int CPL, bool EXACT>
__device__ __forceinline__ void tridiagonalize(HMat<N, CPL> &A, EigLds<N> &sh, int lg)
{
    constexpr int NG = HMat<N, CPL>::NG;
    const int q4 = lg & 3;

#pragma unroll
    for (int k = 0; k < N - 1; k++) {
        float part = 0.0f;

#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const int c = lg + m * G;

            if (min(m * G + G - 1, N - 1) < k)
                continue;
            const float2 a = A.get(m, k);

            if (c == k)
                sh.d[k] = a.x;
            const float x2 = fmaf(a.x, a.x, a.y * a.y);

            part += (c >= k + 2 && c < N) ? x2 : 0.0f;
        }
        const float xn = gsum<G>(part);
        /* alpha = A[k+1][k] = conj(A[k][k+1]), row k of column k + 1 */
        const float2 ak = gbcast2<G>(A.get((k + 1) / G, k), (k + 1) % G);
        const float alr = ak.x, ali = -ak.y;
        float beta, tr_ = 0.0f, ti_ = 0.0f, sr = 0.0f, si = 0.0f;

        if (xn == 0.0f && ali == 0.0f)
            beta = alr;
        else {
            beta = -copysignf(qsqrt<EXACT>(alr * alr + ali * ali + xn), alr);
            tr_ = qdiv<EXACT>(beta - alr, beta);
            ti_ = qdiv<EXACT>(-ali, beta);
            {
                const float dr = alr - beta;
                const float den = dr * dr + ali * ali;

                sr = qdiv<EXACT>(dr, den);
                si = qdiv<EXACT>(-ali, den);
            }
        }
        float2 vc[CPL];
        float act[CPL];

#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const int c = lg + m * G;

            vc[m] = make_float2(0.f, 0.f);
            act[m] = 0.0f;
            if (min(m * G + G - 1, N - 1) < k + 1)
                continue;
            /* x_c = A[c][k] = conj(A[k][c]), v_c = x_c / (alpha - beta) */
            const float2 a = A.get(m, k);
            const float xr = a.x, xi = -a.y;
            const float2 xs = make_float2(fmaf(xr, sr, -xi * si),
                                          fmaf(xr, si, xi * sr));
            const bool live = c >= k + 1 && c < N;

            vc[m] = c == k + 1 ? make_float2(1.f, 0.f) : (live ? xs : vc[m]);
            if (c >= k + 2)
                A.set(m, k, xs);
            act[m] = live ? 1.0f : 0.0f;
        }
        if (lg == 0) {
            sh.e[k] = beta;
            sh.e2[k] = beta * beta;
            sh.tau[k] = make_float2(tr_, ti_);
        }

        /* w_c = tau sum_j conj(A[j][c]) v_j (row c is conj of column c);
           constant trip counts (predicated) keep every index of A static
           once the step loop is unrolled. */
        /* Every product is (broadcast operand) x (lane's own operand)
           accumulated without a sign, so that the broadcast can ride on the
           multiply-add itself; the imaginary part keeps its two sums apart. */
        float sr2[CPL], sia[CPL], sib[CPL];

#pragma unroll
        for (int m = 0; m < CPL; m++) {
            sr2[m] = 0.0f;
            sia[m] = 0.0f;
            sib[m] = 0.0f;
        }
        if constexpr (G == 64) {
            /* lane c publishes v_c; rows are read back as LDS broadcasts
               (a wavefront-wide readlane per value would cost a VALU
               instruction each and tie up scalar registers) */
            const int c = lg;

            if (c >= k + 1 && c < N) {
                sh.vw[c].x = vc[0].x;
                sh.vw[c].y = vc[0].y;
            }
            wave_sync();
            /* four rows in flight: rows j + 4 .. j + 7 load while rows
               j .. j + 3 compute (named values carried across the unrolled
               loop, so that no array is indexed before unrolling) */
            float4 c0, c1, c2, c3;

            c0 = (0 > k && 0 < N) ? sh.vw[0] : make_float4(0.f, 0.f, 0.f, 0.f);
            c1 = (1 > k && 1 < N) ? sh.vw[1] : make_float4(0.f, 0.f, 0.f, 0.f);
            c2 = (2 > k && 2 < N) ? sh.vw[2] : make_float4(0.f, 0.f, 0.f, 0.f);
            c3 = (3 > k && 3 < N) ? sh.vw[3] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int j = 0; j < N; j += 4) {
                const float4 n0 = (j + 4 > k && j + 4 < N) ? sh.vw[j + 4] : c0;
                const float4 n1 = (j + 5 > k && j + 5 < N) ? sh.vw[j + 5] : c1;
                const float4 n2 = (j + 6 > k && j + 6 < N) ? sh.vw[j + 6] : c2;
                const float4 n3 = (j + 7 > k && j + 7 < N) ? sh.vw[j + 7] : c3;

                matvec_row64(j > k && j < N, c0, A.get(0, j < N ? j : 0), sr2[0],
                             sia[0], sib[0]);
                matvec_row64(j + 1 > k && j + 1 < N, c1, A.get(0, j + 1 < N ? j + 1 : 0),
                             sr2[0], sia[0], sib[0]);
                matvec_row64(j + 2 > k && j + 2 < N, c2, A.get(0, j + 2 < N ? j + 2 : 0),
                             sr2[0], sia[0], sib[0]);
                matvec_row64(j + 3 > k && j + 3 < N, c3, A.get(0, j + 3 < N ? j + 3 : 0),
                             sr2[0], sia[0], sib[0]);
                TCOH_STEP_FENCE();
                c0 = n0;
                c1 = n1;
                c2 = n2;
                c3 = n3;
            }
        }
#pragma unroll
        for (int j = 0; j < N; j++) {
            if (G == 64 || j <= k)
                continue;
            if (G == 16) {
#pragma unroll
                for (int m = 0; m < CPL; m++) {
                    if (min(m * G + G - 1, N - 1) < k + 1)
                        continue;
                    row_matvec1(j % G, j == k + 1, sr2[m], sia[m], sib[m], vc[j / G],
                                A.get(m, j));
                }
                continue;
            }
            const float2 v = gbcast2<G>(vc[j / G], j % G);

#pragma unroll
            for (int m = 0; m < CPL; m++) {
                if (min(m * G + G - 1, N - 1) < k + 1)
                    continue;
                const float2 a = A.get(m, j);

                sr2[m] = fmaf(v.x, a.x, sr2[m]);
                sr2[m] = fmaf(v.y, a.y, sr2[m]);
                sia[m] = fmaf(v.y, a.x, sia[m]);
                sib[m] = fmaf(v.x, a.y, sib[m]);
            }
        }
        float2 wc[CPL];
        float pr = 0.0f, pi = 0.0f;

#pragma unroll
        for (int m = 0; m < CPL; m++) {
            wc[m] = make_float2(0.f, 0.f);
            if (min(m * G + G - 1, N - 1) < k + 1)
                continue;
            const float si2 = sia[m] - sib[m];

            wc[m] = make_float2(fmaf(tr_, sr2[m], -ti_ * si2),
                                fmaf(tr_, si2, ti_ * sr2[m]));
            pr = fmaf(act[m], fmaf(wc[m].x, vc[m].x, wc[m].y * vc[m].y), pr);
            pi = fmaf(act[m], fmaf(wc[m].x, vc[m].y, -wc[m].y * vc[m].x), pi);
        }
        pr = gsum<G>(pr);
        pi = gsum<G>(pi);
        {
            const float hr = -0.5f * fmaf(tr_, pr, -ti_ * pi);
            const float hi = -0.5f * fmaf(tr_, pi, ti_ * pr);

#pragma unroll
            for (int m = 0; m < CPL; m++) {
                if (min(m * G + G - 1, N - 1) < k + 1)
                    continue;
                wc[m].x = fmaf(hr, vc[m].x, fmaf(-hi, vc[m].y, wc[m].x));
                wc[m].y = fmaf(hr, vc[m].y, fmaf(hi, vc[m].x, wc[m].y));
            }
        }

        /* Row operands (v_r, w_r) for the update: every row of the groups
           the update touches, zero for finished columns (rows <= k) and
           padding (rows >= N). */
#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const int c = lg + m * G;
            const bool lv = c >= k + 1 && c < N;

            if (m * G < 4 * NG && c < 4 * NG)
                sh.vw[c] = lv ? make_float4(vc[m].x, vc[m].y, wc[m].x, wc[m].y)
                              : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        wave_sync();
        /* A[r][c] -= v_r conj(w_c) + w_r conj(v_c), r > k */
#pragma unroll
        for (int g0 = 0; g0 < NG; g0++) {
            if (4 * g0 + 3 <= k)
                continue;
            const float4 t = sh.vw[4 * g0 + q4];

#pragma unroll
            for (int m = 0; m < CPL; m++) {
                if (min(m * G + G - 1, N - 1) < k + 1)
                    continue;
                float4_t ar = A.re[m][g0], ai = A.im[m][g0];

                ar = rank1x4(t.x, -wc[m].x, ar);
                ai = rank1x4(t.y, -wc[m].x, ai);
                ar = rank1x4(t.y, -wc[m].y, ar);
                ai = rank1x4(t.x, wc[m].y, ai);
                ar = rank1x4(t.z, -vc[m].x, ar);
                ai = rank1x4(t.w, -vc[m].x, ai);
                ar = rank1x4(t.w, -vc[m].y, ar);
                ai = rank1x4(t.z, vc[m].y, ai);
                A.re[m][g0] = ar;
                A.im[m][g0] = ai;
            }
        }
        /* vw is written again only after every lane has read it */
        wave_sync();
    }
#pragma unroll
    for (int m = 0; m < CPL; m++)
        if (lg + m * G == N - 1)
            sh.d[N - 1] = A.get(m, N - 1).x;
    wave_sync();
}

/* Eigenvalue of rank idx (0 smallest) of the tridiagonal in LDS by G-way
   multisection of Sturm counts on the Gershgorin interval, with the
   reference's pivot guard (tcoh.cl:295-338). Returns lambda and norm.
   EXT (idx = N - 1 or 0, an extreme eigenvalue): a point is above lambda
   exactly when every pivot is negative (largest) or any is (smallest), so
   the count is a flag; and the search stops once the interval is within
   2^-20 of |lambda| (a few ulp; inverse iteration from such a shift
   converges as from the exact one), two rounds of 17x narrowing short of
   the float resolution at which the plain rule (no further progress)
   stops. Near lambda = 0 the relative test never holds and the plain rule
   ends the search. */
template <int N, int G, bool EXT>
__device__ __forceinline__ float tridiag_eigenvalue(const EigLds<N> &sh,
                                                    int idx, int lg,
                                                    float *norm)
{
    float lo = sh.d[0], hi = sh.d[0];

#pragma unroll
    for (int i = 0; i < N; i++) {
        const float r = (i > 0 ? fabsf(sh.e[i - 1]) : 0.0f) +
                        (i < N - 1 ? fabsf(sh.e[i]) : 0.0f);

        lo = fminf(lo, sh.d[i] - r);
        hi = fmaxf(hi, sh.d[i] + r);
    }
    *norm = fmaxf(fabsf(lo), fabsf(hi));
    const float pivmin = 1.0e-30f + 1.0e-7f * *norm * 1.0e-7f;

    for (int it = 0; it < 24; it++) {
        if (EXT && hi - lo <= 0x1p-20f * fmaxf(fabsf(lo), fabsf(hi)))
            break;
        const float step = (hi - lo) * (1.0f / (G + 1));
        const float x = fminf(lo + step * (float)(lg + 1), hi);
        float qv = sh.d[0] - x;
        int cnt = qv < 0.0f;
        bool allneg = qv < 0.0f, anyneg = qv < 0.0f;
        const bool top = idx != 0;

        /* The count only decides a side of x: the hardware reciprocal
           (1 ulp) is as good as a rounded division here. */
#pragma unroll
        for (int i = 1; i < N; i++) {
            if (fabsf(qv) < pivmin)
                qv = -pivmin;
            qv = sh.d[i] - x - sh.e2[i - 1] * __builtin_amdgcn_rcpf(qv);
            if (EXT && top)
                allneg &= qv < 0.0f;
            else if (EXT)
                anyneg |= qv < 0.0f;
            else
                cnt += qv < 0.0f;
        }
        const bool above = EXT ? (top ? allneg : anyneg) : cnt > idx;
        const float nlo = gmax<G>(above ? lo : x);
        const float nhi = gmin<G>(above ? x : hi);

        if (nlo <= lo && nhi >= hi)
            break;
        lo = nlo;
        hi = nhi;
    }
    return 0.5f * (lo + hi);
}

/* Inverse iteration of tcoh.cl:340-410; every lane of the group computes
   the same vector, y ends in LDS. Above 32 dates the factorisation and the
   substitutions stop every four steps (TCOH_STEP_FENCE) so that the
   compiler does not hoist every step's LDS operands ahead of the serial
   chain: with them all in registers the N=60 kernel spilled 1.2 KB a lane
   inside these loops (340 bytes with the fences; N=60 512^2: 28.2 ->
   19.5 ms). */
template <int N, bool EXACT>
__device__ __forceinline__ void tridiag_eigenvector(EigLds<N> &sh,
                                                    float lambda, float norm,
                                                    int lg)
{
    const float tiny = 1.0e-6f * fmaxf(norm, 1.0e-30f);
    float dd = sh.d[0] - lambda, s1 = N > 1 ? sh.e[0] : 0.0f;
    float y[N];

#pragma unroll
    for (int k = 0; k < N - 1; k++) {
        const float sub = sh.e[k], nd = sh.d[k + 1] - lambda;
        const float ns = k + 1 < N - 1 ? sh.e[k + 1] : 0.0f;
        float lk, a0, a1, a2;
        int pv;

        if (fabsf(dd) >= fabsf(sub)) {
            if (fabsf(dd) < tiny)
                dd = dd < 0.0f ? -tiny : tiny;
            lk = qdiv<EXACT>(sub, dd);
            pv = 0;
            a0 = dd;
            a1 = s1;
            a2 = 0.0f;
            dd = nd - lk * s1;
            s1 = ns;
        }
        else {
            lk = qdiv<EXACT>(dd, sub);
            pv = 1;
            a0 = sub;
            a1 = nd;
            a2 = ns;
            dd = s1 - lk * nd;
            s1 = -lk * ns;
        }
        sh.l[k] = lk;
        sh.piv[k] = pv;
        sh.u0[k] = EXACT ? a0 : rcpf(a0);
        sh.u1[k] = a1;
        sh.u2[k] = a2;
        if (N > 32 && k % 4 == 3)
            TCOH_STEP_FENCE();
    }
    if (fabsf(dd) < tiny)
        dd = dd < 0.0f ? -tiny : tiny;
    sh.u0[N - 1] = EXACT ? dd : rcpf(dd);
    sh.u1[N - 1] = 0.0f;
    sh.u2[N - 1] = 0.0f;

#pragma unroll
    for (int k = 0; k < N; k++)
        y[k] = 1.0f;
    for (int it = 0; it < 3; it++) {
        float nrm = 0.0f;

#pragma unroll
        for (int k = 0; k < N - 1; k++) {
            if (sh.piv[k]) {
                const float t = y[k];

                y[k] = y[k + 1];
                y[k + 1] = t;
            }
            y[k + 1] -= sh.l[k] * y[k];
            if (N > 32 && k % 4 == 3)
                TCOH_STEP_FENCE();
        }
#pragma unroll
        for (int k = N - 1; k >= 0; k--) {
            float v = y[k];

            if (k + 1 < N)
                v -= sh.u1[k] * y[k + 1];
            if (k + 2 < N)
                v -= sh.u2[k] * y[k + 2];
            y[k] = EXACT ? v / sh.u0[k] : v * sh.u0[k];
            if (N > 32 && k % 4 == 0)
                TCOH_STEP_FENCE();
        }
#pragma unroll
        for (int k = 0; k < N; k++)
            nrm = fmaxf(nrm, fabsf(y[k]));
        const float rn = rcpf(nrm);

#pragma unroll
        for (int k = 0; k < N; k++)
            y[k] = EXACT ? y[k] / nrm : y[k] * rn;
    }
    if (lg == 0) {
#pragma unroll
        for (int k = 0; k < N; k++)
            sh.y[k] = y[k];
    }
    wave_sync();
}

/* z = H_0 ... H_{N-2} y for the owned components (tcoh.cl:426-443). */
template <int N, int G, int CPL>
__device__ __forceinline__ void back_transform(const HMat<N, CPL> &A,
                                               const EigLds<N> &sh, int lg,
                                               float2 (&z)[CPL])
{
#pragma unroll
    for (int m = 0; m < CPL; m++) {
        const int c = lg + m * G;

        z[m] = make_float2(c < N ? sh.y[c < N ? c : 0] : 0.0f, 0.0f);
    }
#pragma unroll
    for (int k = N - 2; k >= 0; k--) {
        const float2 tau = sh.tau[k];
        float sr = 0.0f, si = 0.0f;
        float2 v[CPL];

#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const int c = lg + m * G;

            v[m] = make_float2(0.f, 0.f);
            /* a slot whose columns are all <= k + 1 has at most v = 1 at
               c = k + 1 (else none): known at compile time */
            if (min(m * G + G - 1, N - 1) < k + 1)
                continue;
            if (c == k + 1)
                v[m] = make_float2(1.f, 0.f);
            else if (c > k + 1 && c < N)
                v[m] = A.get(m, k);
            sr += v[m].x * z[m].x + v[m].y * z[m].y;
            si += v[m].x * z[m].y - v[m].y * z[m].x;
        }
        sr = gsum<G>(sr);
        si = gsum<G>(si);
        const float2 h = make_float2(tau.x * sr - tau.y * si,
                                     tau.x * si + tau.y * sr);

#pragma unroll
        for (int m = 0; m < CPL; m++) {
            if (min(m * G + G - 1, N - 1) < k + 1)
                continue;
            const float2 vh = cmul(v[m], h);

            z[m].x -= vh.x;
            z[m].y -= vh.y;
        }
    }
}

/* Eigenvector (owned components in z) of rank idx of the group's
   Hermitian matrix A (destroyed). */
template <int N, int G, int CPL, bool EXACT>
__device__ __forceinline__ void hermitian_eigenvector(HMat<N, CPL> &A,
                                                      EigLds<N> &sh, int idx,
                                                      int lg, float2 (&z)[CPL],
                                                      float *lam_out = nullptr,
                                                      float *norm_out = nullptr)
{
    float norm;

    tridiagonalize<N, G, CPL, EXACT>(A, sh, lg);
    /* both estimators take an extreme eigenvalue (EVD the largest, EMI the
       smallest) */
    const float lambda = tridiag_eigenvalue<N, G, true>(sh, idx, lg, &norm);
    tridiag_eigenvector<N, EXACT>(sh, lambda, norm, lg);
    back_transform<N, G, CPL>(A, sh, lg, z);
    if (lam_out) {
        *lam_out = lambda;
        *norm_out = norm;
    }
}

/* Whether the FP32 eigenvector of the smallest eigenvalue lambda of the
   tridiagonal in sh (||.|| <= norm) may be off by more than REFINE_TOL:
   its error is about eps_32 norm / gap, so it is when a second eigenvalue
   lies below lambda + eps_32 norm / REFINE_TOL (a Sturm count there). */
#define REFINE_TOL 1.0e-5f
template <int N>
__device__ __forceinline__ bool close_second_eigenvalue(const EigLds<N> &sh, float lambda,
                                                        float norm)
{
    const float x = lambda + (0x1p-24f / REFINE_TOL) * norm;
    const float pivmin = 1.0e-30f + 1.0e-7f * norm * 1.0e-7f;
    float qv = sh.d[0] - x;
    int cnt = qv < 0.0f;

#pragma unroll
    for (int i = 1; i < N; i++) {
        if (fabsf(qv) < pivmin)
            qv = -pivmin;
        qv = sh.d[i] - x - sh.e2[i - 1] / qv;
        cnt += qv < 0.0f;
    }
    return cnt >= 2;
}

/* T_rc of the pixel from the packed strict upper triangle. */
template <int N>
__device__ __forceinline__ float2 load_t(const float2 *__restrict__ tp, int r,
                                         int c)
{
    if (r == c)
        return make_float2(1.f, 0.f);
    if (r < c)
        return tp[upper_index(N, r, c)];
    const float2 t = tp[upper_index(N, c, r)];

    return make_float2(t.x, -t.y);
}

/* |T|^-1 into inv (owned columns) for EMI by Cholesky in FP64 with the
   reference's pivot floor (tcoh.cl:446-492). Returns false, with inv
   untouched, when a pivot is at or below the floor: the reference then
   takes the eigen pseudo-inverse (tcoh.cl:657-687), which the caller
   leaves to the queue kernels (gj_inverse, jacobi_pinv). */
template <int N, int G, int CPL>
__device__ __forceinline__ bool chol_inverse(const float2 *__restrict__ tp, bool est,
                                             EmiLds<N> &em, int lg, float (&inv)[CPL][N])
{
    bool ok = true;

    if constexpr (N > 32) {
        /* One column per lane (G = 64): L, L^-1 and the pivots in LDS with
           rolled loops (a register-resident factor of this size would hold
           N^2 doubles per lane in unrolled code). Lane c factors row c of
           |T| and then computes column c of W = L^-1; |T|^-1 = W^T W. */
        static_assert(CPL == 1, "one column per lane");
        double *l = em.ch.lc, *w = em.ch.w, *rd = em.ch.rd;
        const int c = lg;

        if (c < N)
            for (int j = 0; j <= c; j++) {
                const float2 t = est ? load_t<N>(tp, c, j)
                                     : make_float2(c == j ? 1.f : 0.f, 0.f);

                l[lpos<N>(c, j)] = (double)hypotf(t.x, t.y);
            }
        for (int j = 0; j < N; j++) {
            wave_sync();
            const double s = l[lpos<N>(j, j)];

            if (!(s > CHOL_FLOOR)) {
                ok = false;
                break;
            }
            const double ljj = sqrt(s), rl = 1.0 / ljj;

            wave_sync();
            if (c == j) {
                l[lpos<N>(j, j)] = ljj;
                rd[j] = rl;
            }
            if (c > j && c < N)
                l[lpos<N>(c, j)] *= rl;
            wave_sync();
            if (c > j && c < N) {
                const double lcj = l[lpos<N>(c, j)];

                for (int k = j + 1; k <= c; k++)
                    l[lpos<N>(c, k)] -= lcj * l[lpos<N>(k, j)];
            }
        }
        wave_sync();
        if (ok) {
            if (c < N) {
                w[lpos<N>(c, c)] = rd[c];
                for (int r = c + 1; r < N; r++) {
                    double sm = 0.0;

                    for (int k = c; k < r; k++)
                        sm += l[lpos<N>(r, k)] * w[lpos<N>(k, c)];
                    w[lpos<N>(r, c)] = -sm * rd[r];
                }
            }
            wave_sync();
#pragma unroll
            for (int r = 0; r < N; r++) {
                double sm = 0.0;

                if (c < N)
                    for (int k = max(r, c); k < N; k++)
                        sm += w[lpos<N>(k, r)] * w[lpos<N>(k, c)];
                inv[0][r] = (float)sm;
            }
        }
        wave_sync();
        if (ok)
            return true;
    }
    else {
    double g[CPL][N];

#pragma unroll
    for (int m = 0; m < CPL; m++) {
        const int c = lg + m * G;

#pragma unroll
        for (int r = 0; r < N; r++) {
            float2 t = make_float2(r == c ? 1.f : 0.f, 0.f);

            if (est && c < N)
                t = load_t<N>(tp, r, c);
            g[m][r] = (double)hypotf(t.x, t.y);
        }
    }
    /* Right-looking Cholesky in FP64 with the reference's pivot floor
       (tcoh.cl:446-492); column j of L and 1 / L_jj go to LDS. */
#pragma unroll
    for (int j = 0; j < N; j++) {
#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const int c = lg + m * G;

            if (c == j) {
                const double s = g[m][j];
                const double ljj = sqrt(s > 0.0 ? s : 1.0);
                const double rl = 1.0 / ljj;

                em.ch.flag = s > CHOL_FLOOR ? 1.0 : 0.0;
                em.ch.lc[lpos<N>(j, j)] = ljj;
                em.ch.rd[j] = rl;
                /* constant trip counts (predicated) keep every index of g
                   static once the outer loop is unrolled */
#pragma unroll
                for (int i = 0; i < N; i++)
                    if (i > j)
                        em.ch.lc[lpos<N>(i, j)] = g[m][i] * rl;
            }
        }
        wave_sync();
        ok = ok && em.ch.flag != 0.0;
#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const int c = lg + m * G;

            if (c > j && c < N) {
                const double lcj = em.ch.lc[lpos<N>(c, j)];

#pragma unroll
                for (int r = 0; r < N; r++)
                    if (r > j && r >= c)
                        g[m][r] -= em.ch.lc[lpos<N>(r, j)] * lcj;
            }
        }
        wave_sync();
        TCOH_STEP_FENCE();
    }
    if (ok) {
        /* Column c of the inverse: L y = e_c by columns of L (axpy form),
           then L^T x = y by dot products with the columns of L; all owned
           columns advance together, so each entry of L read from LDS is
           used at once by every slot. */
        double y[CPL][N];

#pragma unroll
        for (int m = 0; m < CPL; m++)
#pragma unroll
            for (int r = 0; r < N; r++)
                y[m][r] = r == lg + m * G ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < N; k++) {
            const double rk = em.ch.rd[k];

#pragma unroll
            for (int m = 0; m < CPL; m++)
                y[m][k] *= rk;
#pragma unroll
            for (int r = 0; r < N; r++) {
                if (r <= k)
                    continue;
                const double l = em.ch.lc[lpos<N>(r, k)];

#pragma unroll
                for (int m = 0; m < CPL; m++)
                    y[m][r] -= l * y[m][k];
            }
            TCOH_STEP_FENCE();
        }
#pragma unroll
        for (int r = N - 1; r >= 0; r--) {
            double sm[CPL];

#pragma unroll
            for (int m = 0; m < CPL; m++)
                sm[m] = y[m][r];
#pragma unroll
            for (int k = 0; k < N; k++) {
                if (k <= r)
                    continue;
                const double l = em.ch.lc[lpos<N>(k, r)];

#pragma unroll
                for (int m = 0; m < CPL; m++)
                    sm[m] -= l * y[m][k];
            }
            const double rr = em.ch.rd[r];

#pragma unroll
            for (int m = 0; m < CPL; m++)
                y[m][r] = sm[m] * rr;
            TCOH_STEP_FENCE();
        }
#pragma unroll
        for (int m = 0; m < CPL; m++)
#pragma unroll
            for (int r = 0; r < N; r++)
                inv[m][r] = lg + m * G < N ? (float)y[m][r] : 0.0f;
    }
    wave_sync();
    if (ok)
        return true;
    }
    return false;
}

template <int G> __device__ __forceinline__ double gmaxd(double v)
{
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1)
        v = fmax(v, __shfl_xor(v, o, G));
    return v;
}

template <int G> __device__ __forceinline__ int gmini(int v)
{
    return -gmaxi<G>(-v);
}

/* Condition bound below which the eigen pseudo-inverse of |T| with the
   reference's relative floor (tcoh.cl:669-677, EIG_FLOOR) is the inverse:
   for symmetric A, min |lambda| = 1 / ||A^-1||_2 >= 1 / ||A^-1||_F and
   max |lambda| <= ||A||_F, so ||A||_F ||A^-1||_F <= 1 / EIG_FLOOR drops no
   eigenvalue. Compared squared. */
#define PINV_IS_INV_KAPPA2 1.0e12

/* |T|^-1 into inv for a pixel whose Cholesky pivot test failed (|T| not
   safely positive definite, usually indefinite), by Gauss-Jordan exchange
   steps in FP64 with partial pivoting. Lane lg keeps rows r = lg + m G of
   the tableau in registers. Step k exchanges column k with the unused row
   p_k of largest |a_rk| (ties to the lowest row); that row is broadcast
   through LDS, and the tableau stays in place, so after N steps row p_i
   holds row i of the inverse in the column order p_0 .. p_{N-1}, which the
   owners scatter to LDS. Returns true when ||A||_F ||A^-1||_F is within
   the bound above (the reference's pseudo-inverse is then this inverse);
   false, for the eigen pseudo-inverse of jacobi_pinv, otherwise (also for
   a singular A, whose exchange divides by zero). */
template <int N, int G, int CPL>
__device__ __forceinline__ bool gj_inverse(const float2 *__restrict__ tp, bool est,
                                           EmiLds<N> &em, int lg, float (&inv)[CPL][N])
{
    double a[CPL][N];
    bool used[CPL];
    int step[CPL];
    double fa = 0.0;

#pragma unroll
    for (int m = 0; m < CPL; m++) {
        const int r = lg + m * G;

        used[m] = r >= N;
        step[m] = 0;
#pragma unroll
        for (int c = 0; c < N; c++) {
            float2 t = make_float2(r == c ? 1.f : 0.f, 0.f);

            if (est && r < N)
                t = load_t<N>(tp, r, c);
            a[m][c] = r < N ? (double)hypotf(t.x, t.y) : 0.0;
            fa = fma(a[m][c], a[m][c], fa);
        }
    }
    fa = gsumd<G>(fa);
#pragma unroll
    for (int k = 0; k < N; k++) {
        double bv = -1.0;
        int br = N;

#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const double v = fabs(a[m][k]);

            if (!used[m] && v > bv) {
                bv = v;
                br = lg + m * G;
            }
        }
        const double mx = gmaxd<G>(bv);
        const int p = gmini<G>(bv == mx ? br : N);

#pragma unroll
        for (int m = 0; m < CPL; m++)
            if (lg + m * G == p) {
#pragma unroll
                for (int c = 0; c < N; c++)
                    em.gj.prow[c] = a[m][c];
                used[m] = true;
                step[m] = k;
            }
        if (lg == 0)
            em.gj.pk[k] = p;
        wave_sync();
        const double rp = 1.0 / em.gj.prow[k];

#pragma unroll
        for (int m = 0; m < CPL; m++) {
            if (lg + m * G == p) {
#pragma unroll
                for (int c = 0; c < N; c++)
                    a[m][c] = c == k ? rp : -a[m][c] * rp;
            }
            else {
                const double f = a[m][k] * rp;

#pragma unroll
                for (int c = 0; c < N; c++)
                    a[m][c] = c == k ? f : fma(-f, em.gj.prow[c], a[m][c]);
            }
        }
        wave_sync();
        TCOH_STEP_FENCE();
    }
    double fi = 0.0;

#pragma unroll
    for (int m = 0; m < CPL; m++) {
        const int r = lg + m * G;

        if (r < N) {
            const int i = step[m];

#pragma unroll
            for (int c = 0; c < N; c++) {
                fi = fma(a[m][c], a[m][c], fi);
                em.gj.inv[i * N + em.gj.pk[c]] = (float)a[m][c];
            }
        }
    }
    fi = gsumd<G>(fi);
    wave_sync();
#pragma unroll
    for (int m = 0; m < CPL; m++) {
        const int c = lg + m * G;

#pragma unroll
        for (int r = 0; r < N; r++)
            inv[m][r] = c < N ? em.gj.inv[r * N + c] : 0.0f;
    }
    wave_sync();
    return fa * fi <= PINV_IS_INV_KAPPA2;
}

/* Rotations of the Jacobi below are skipped when |a_pq| is within this
   fraction of sqrt(|a_pp a_qq|) (negligible in FP64). */
#define JACOBI_TOL 1.0e-17

/* The reference's pseudo-inverse of |T| (tcoh.cl:657-687): its
   eigendecomposition, then V diag(1 / lambda) V^T with the eigenvalues
   below EIG_FLOOR of the largest magnitude left out. Here in FP64 by the
   cyclic Jacobi method (rows p < q in order), run to convergence: a sweep
   that finds every off-diagonal entry negligible (JACOBI_TOL) ends it,
   within the reference's sweep limit. Each rotation is the symmetric
   Schur rotation annihilating a_pq (Golub and Van Loan, Matrix
   Computations, 4th ed., 8.5.2), applied to the rows and columns k owned
   by the lanes of the group (k = lg, lg + G, ...) in LDS. Only pixels that
   gj_inverse could not settle come here (none in the benchmark scenes). */
template <int N, int G, int CPL>
__device__ __forceinline__ void jacobi_pinv(const float2 *__restrict__ tp, bool est,
                                            EmiLds<N> &em, int lg, float (&inv)[CPL][N])
{
    double *a = em.jac.a, *v = em.jac.v;

#pragma unroll 1
    for (int i = lg; i < N * N; i += G) {
        const int r = i / N, c = i - (i / N) * N;
        const float2 t = est ? load_t<N>(tp, r, c) : make_float2(r == c ? 1.f : 0.f, 0.f);

        a[i] = (double)hypotf(t.x, t.y);
        v[i] = r == c ? 1.0 : 0.0;
    }
    wave_sync();
#pragma unroll 1
    for (int sweep = 0; sweep < MAX_SWEEPS; sweep++) {
        bool rotated = false;

#pragma unroll 1
        for (int p = 0; p < N - 1; p++)
#pragma unroll 1
            for (int q = p + 1; q < N; q++) {
                const double apq = a[p * N + q], app = a[p * N + p], aqq = a[q * N + q];

                if (!(fabs(apq) > JACOBI_TOL * sqrt(fabs(app) * fabs(aqq))))
                    continue;
                const double tau = (aqq - app) / (2.0 * apq);
                const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(fma(tau, tau, 1.0)));
                const double c = 1.0 / sqrt(fma(t, t, 1.0)), s = t * c;

                wave_sync();
#pragma unroll 1
                for (int k = lg; k < N; k += G) {
                    if (k != p && k != q) {
                        const double akp = a[k * N + p], akq = a[k * N + q];
                        const double np_ = c * akp - s * akq, nq = s * akp + c * akq;

                        a[k * N + p] = np_;
                        a[p * N + k] = np_;
                        a[k * N + q] = nq;
                        a[q * N + k] = nq;
                    }
                    const double vkp = v[k * N + p], vkq = v[k * N + q];

                    v[k * N + p] = c * vkp - s * vkq;
                    v[k * N + q] = s * vkp + c * vkq;
                }
                if (lg == 0) {
                    a[p * N + p] = app - t * apq;
                    a[q * N + q] = aqq + t * apq;
                    a[p * N + q] = 0.0;
                    a[q * N + p] = 0.0;
                }
                wave_sync();
                rotated = true;
            }
        if (!rotated)
            break;
    }
    double lmax = 0.0;

    for (int k = 0; k < N; k++)
        lmax = fmax(lmax, fabs(a[k * N + k]));
    lmax *= (double)EIG_FLOOR;
    for (int k = lg; k < N; k += G) {
        const double lk = a[k * N + k];

        em.jac.il[k] = fabs(lk) < lmax ? 0.0 : 1.0 / lk;
    }
    wave_sync();
#pragma unroll
    for (int m = 0; m < CPL; m++) {
        const int c = lg + m * G;

#pragma unroll
        for (int r = 0; r < N; r++) {
            double sm = 0.0;

            if (c < N)
#pragma unroll 1
                for (int k = 0; k < N; k++)
                    sm = fma(v[r * N + k] * em.jac.il[k], v[c * N + k], sm);
            inv[m][r] = (float)sm;
        }
    }
    wave_sync();
}

#endif
