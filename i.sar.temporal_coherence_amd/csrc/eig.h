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
 *                                   eigen pseudo-inverse of a cyclic
 *                                   Jacobi in FP32 (relative floor 1e-6);
 *                                   then the smallest eigenvalue of
 *                                   |T|^-1 o T
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

/* vw[r] = (v_r, w_r) of the current Householder step, read as one
   16-byte word by the update. vw is dead once the reduction ends, and the
   inverse iteration's LU factors (u0 holds 1/pivot) and result y take its
   place. */
template <int N> struct EigLds {
    union {
        float4 vw[N];
        struct {
            float u0[N], u1[N], u2[N], l[N], y[N];
            int piv[N];
        };
    };
    float2 tau[N];
    float d[N], e[N];
    float2 z[N];
    float2 alpha;
};

/* Cholesky factor (FP64, lower, column-major packed, with the reciprocal
   diagonal) or, in the fallback, the Jacobi state (FP32): the rotations of
   a round, the matrix A and the eigenvectors V (NE x NE, NE = N rounded up
   to even; the padding row and column stay zero in A). */
template <int N> struct JacobiShape {
    static constexpr int NE = (N + 1) & ~1;
    static constexpr int NP = NE / 2;
};

template <int N> union __attribute__((aligned(16))) EmiLds {
    struct {
        double lc[N * (N + 1) / 2];
        double rd[N];
        double flag;
    } ch;
    struct {
        float4 rot[JacobiShape<N>::NP];
        int4 pq[JacobiShape<N>::NP]; /* p, q, rotation applied */
        float2 dd[JacobiShape<N>::NP];
        float il[JacobiShape<N>::NE];
        float a[JacobiShape<N>::NE * JacobiShape<N>::NE];
        float v[JacobiShape<N>::NE * JacobiShape<N>::NE];
    } jac;
};

template <int N> __device__ __forceinline__ int lpos(int i, int j)
{
    /* L[i][j], i >= j, column-major packed. */
    return j * N - j * (j - 1) / 2 + (i - j);
}

/* Householder reduction of the group's matrix (columns in A) to the real
   tridiagonal (d, e) in LDS; the Householder vector of step k is left in
   row k of the columns c > k (A[m][k]), tau_k in LDS.

   Slot m holds columns m*G .. m*G+G-1; a slot whose columns are all
   finished at step k is skipped at compile time. Finished columns inside
   a live slot (c <= k) are updated like the others: their rows below k
   are never read again (this layout keeps Householder vectors in rows,
   not columns), so only the reductions and LDS writes are masked. */
template <int N, int G, int CPL>
__device__ __forceinline__ void tridiagonalize(float2 (&A)[CPL][N],
                                               EigLds<N> &sh, int lg)
{
#pragma unroll
    for (int k = 0; k < N - 1; k++) {
        float part = 0.0f;

#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const int c = lg + m * G;

            if (min(m * G + G - 1, N - 1) < k)
                continue;
            if (c == k + 1)
                sh.alpha = make_float2(A[m][k].x, -A[m][k].y);
            if (c == k)
                sh.d[k] = A[m][k].x;
            const float x2 = fmaf(A[m][k].x, A[m][k].x, A[m][k].y * A[m][k].y);

            part += (c >= k + 2 && c < N) ? x2 : 0.0f;
        }
        const float xn = gsum<G>(part);
        wave_sync();
        const float alr = sh.alpha.x, ali = sh.alpha.y;
        float beta, tr_ = 0.0f, ti_ = 0.0f, sr = 0.0f, si = 0.0f;

        if (xn == 0.0f && ali == 0.0f)
            beta = alr;
        else {
            beta = -copysignf(sqrtf(alr * alr + ali * ali + xn), alr);
            const float rb = rcpf(beta);

            tr_ = (beta - alr) * rb;
            ti_ = -ali * rb;
            {
                const float dr = alr - beta;
                const float rd = rcpf(dr * dr + ali * ali);

                sr = dr * rd;
                si = -ali * rd;
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
            const float xr = A[m][k].x, xi = -A[m][k].y;
            const float2 xs = make_float2(fmaf(xr, sr, -xi * si),
                                          fmaf(xr, si, xi * sr));
            const bool live = c >= k + 1 && c < N;

            vc[m] = c == k + 1 ? make_float2(1.f, 0.f) : (live ? xs : vc[m]);
            if (c >= k + 2)
                A[m][k] = xs;
            act[m] = live ? 1.0f : 0.0f;
            if (live) {
                sh.vw[c].x = vc[m].x;
                sh.vw[c].y = vc[m].y;
            }
        }
        if (lg == 0) {
            sh.e[k] = beta;
            sh.tau[k] = make_float2(tr_, ti_);
        }
        wave_sync();

        /* w_c = tau sum_j conj(A[j][c]) v_j (row c is conj of column c). */
        float2 wc[CPL];
        float pr = 0.0f, pi = 0.0f;

#pragma unroll
        for (int m = 0; m < CPL; m++) {
            float sr2 = 0.0f, si2 = 0.0f;

            wc[m] = make_float2(0.f, 0.f);
            if (min(m * G + G - 1, N - 1) < k + 1)
                continue;
            /* constant trip counts (predicated) keep every index of A
               static once the step loop is unrolled */
#pragma unroll
            for (int j = 0; j < N; j++) {
                if (j <= k)
                    continue;
                const float4 vw = sh.vw[j];
                const float2 v = make_float2(vw.x, vw.y);

                sr2 = fmaf(A[m][j].x, v.x, sr2);
                sr2 = fmaf(A[m][j].y, v.y, sr2);
                si2 = fmaf(A[m][j].x, v.y, si2);
                si2 = fmaf(-A[m][j].y, v.x, si2);
            }
            wc[m] = make_float2(fmaf(tr_, sr2, -ti_ * si2), fmaf(tr_, si2, ti_ * sr2));
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
                const int c = lg + m * G;

                if (min(m * G + G - 1, N - 1) < k + 1)
                    continue;
                wc[m].x = fmaf(hr, vc[m].x, fmaf(-hi, vc[m].y, wc[m].x));
                wc[m].y = fmaf(hr, vc[m].y, fmaf(hi, vc[m].x, wc[m].y));
                if (c >= k + 1 && c < N) {
                    sh.vw[c].z = wc[m].x;
                    sh.vw[c].w = wc[m].y;
                }
            }
        }
        wave_sync();
        TCOH_STEP_FENCE();

        /* A[r][c] -= v_r conj(w_c) + w_r conj(v_c), r > k. */
#pragma unroll
        for (int m = 0; m < CPL; m++) {
            if (min(m * G + G - 1, N - 1) < k + 1)
                continue;
#pragma unroll
            for (int r = 0; r < N; r++) {
                if (r <= k)
                    continue;
                const float4 vw = sh.vw[r];
                const float2 v = make_float2(vw.x, vw.y), w = make_float2(vw.z, vw.w);
                float ax = A[m][r].x, ay = A[m][r].y;

                ax = fmaf(-v.x, wc[m].x, ax);
                ax = fmaf(-v.y, wc[m].y, ax);
                ax = fmaf(-w.x, vc[m].x, ax);
                ax = fmaf(-w.y, vc[m].y, ax);
                ay = fmaf(-v.y, wc[m].x, ay);
                ay = fmaf(v.x, wc[m].y, ay);
                ay = fmaf(-w.y, vc[m].x, ay);
                ay = fmaf(w.x, vc[m].y, ay);
                A[m][r] = make_float2(ax, ay);
            }
        }
    }
#pragma unroll
    for (int m = 0; m < CPL; m++)
        if (lg + m * G == N - 1)
            sh.d[N - 1] = A[m][N - 1].x;
    wave_sync();
}

/* Eigenvalue of rank idx (0 smallest) of the tridiagonal in LDS by G-way
   multisection of Sturm counts on the Gershgorin interval, with the
   reference's pivot guard (tcoh.cl:295-338). Returns lambda and norm. */
template <int N, int G>
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
        const float step = (hi - lo) * (1.0f / (G + 1));
        const float x = fminf(lo + step * (float)(lg + 1), hi);
        float qv = sh.d[0] - x;
        int cnt = qv < 0.0f;

        /* The count only decides a side of x: the hardware reciprocal
           (1 ulp) is as good as a rounded division here. */
#pragma unroll
        for (int i = 1; i < N; i++) {
            if (fabsf(qv) < pivmin)
                qv = -pivmin;
            qv = sh.d[i] - x - sh.e[i - 1] * sh.e[i - 1] *
                                   __builtin_amdgcn_rcpf(qv);
            cnt += qv < 0.0f;
        }
        const float nlo = gmax<G>(cnt <= idx ? x : lo);
        const float nhi = gmin<G>(cnt > idx ? x : hi);

        if (nlo <= lo && nhi >= hi)
            break;
        lo = nlo;
        hi = nhi;
    }
    return 0.5f * (lo + hi);
}

/* Inverse iteration of tcoh.cl:340-410; every lane of the group computes
   the same vector, y ends in LDS. */
template <int N>
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
            lk = sub * rcpf(dd);
            pv = 0;
            a0 = dd;
            a1 = s1;
            a2 = 0.0f;
            dd = nd - lk * s1;
            s1 = ns;
        }
        else {
            lk = dd * rcpf(sub);
            pv = 1;
            a0 = sub;
            a1 = nd;
            a2 = ns;
            dd = s1 - lk * nd;
            s1 = -lk * ns;
        }
        sh.l[k] = lk;
        sh.piv[k] = pv;
        sh.u0[k] = rcpf(a0);
        sh.u1[k] = a1;
        sh.u2[k] = a2;
    }
    if (fabsf(dd) < tiny)
        dd = dd < 0.0f ? -tiny : tiny;
    sh.u0[N - 1] = rcpf(dd);
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
        }
#pragma unroll
        for (int k = N - 1; k >= 0; k--) {
            float v = y[k];

            if (k + 1 < N)
                v -= sh.u1[k] * y[k + 1];
            if (k + 2 < N)
                v -= sh.u2[k] * y[k + 2];
            y[k] = v * sh.u0[k];
        }
#pragma unroll
        for (int k = 0; k < N; k++)
            nrm = fmaxf(nrm, fabsf(y[k]));
        const float rn = rcpf(nrm);

#pragma unroll
        for (int k = 0; k < N; k++)
            y[k] *= rn;
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
__device__ __forceinline__ void back_transform(const float2 (&A)[CPL][N],
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
            if (c == k + 1)
                v[m] = make_float2(1.f, 0.f);
            else if (c > k + 1 && c < N)
                v[m] = A[m][k];
            sr += v[m].x * z[m].x + v[m].y * z[m].y;
            si += v[m].x * z[m].y - v[m].y * z[m].x;
        }
        sr = gsum<G>(sr);
        si = gsum<G>(si);
        const float2 h = make_float2(tau.x * sr - tau.y * si,
                                     tau.x * si + tau.y * sr);

#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const float2 vh = cmul(v[m], h);

            z[m].x -= vh.x;
            z[m].y -= vh.y;
        }
    }
}

/* Eigenvector (owned components in z) of rank idx of the group's
   Hermitian matrix A (destroyed). */
template <int N, int G, int CPL>
__device__ __forceinline__ void hermitian_eigenvector(float2 (&A)[CPL][N],
                                                      EigLds<N> &sh, int idx,
                                                      int lg, float2 (&z)[CPL])
{
    float norm;

    tridiagonalize<N, G, CPL>(A, sh, lg);
    const float lambda = tridiag_eigenvalue<N, G>(sh, idx, lg, &norm);
    tridiag_eigenvector<N>(sh, lambda, norm, lg);
    back_transform<N, G, CPL>(A, sh, lg, z);
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

/* |T|^-1 into inv (owned columns) for EMI (tcoh.cl:657-687). */
template <int N, int G, int CPL>
__device__ __forceinline__ void abs_inverse(const float2 *__restrict__ tp, bool est,
                            EmiLds<N> &em, int lg, float (&inv)[CPL][N])
{
    double g[CPL][N];
    bool ok = true;

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
        return;

    /* Not safely positive definite: pseudo-inverse from a Jacobi
       eigendecomposition of |T| in FP32 with the reference's rotations,
       stopping rule and sweep limit (tcoh.cl:97-199, 663-687), in the
       parallel ordering: each round rotates NP disjoint pairs (round-robin
       tournament; a zero index pads odd N), one pair per lane, then updates
       the 2 x 2 blocks of A and the column pairs of V in parallel. |T| is
       real, so the phase step of the complex rotation is a sign. */
    typedef JacobiShape<N> J;
    constexpr int NE = J::NE, NP = J::NP, NB = NP * (NP + 1) / 2;
    float *a = em.jac.a, *v = em.jac.v;

    for (int i = lg; i < NE * NE; i += G) {
        const int r = i / NE, c = i - (i / NE) * NE;
        float2 t = make_float2(r == c ? 1.f : 0.f, 0.f);

        if (r < N && c < N && est)
            t = load_t<N>(tp, r, c);
        a[i] = (r < N && c < N) ? hypotf(t.x, t.y) : 0.0f;
        v[i] = r == c ? 1.0f : 0.0f;
    }
    wave_sync();
    for (int sweep = 0; sweep < MAX_SWEEPS; sweep++) {
        float off = 0.0f, dg = 0.0f;

        for (int i = lg; i < N * N; i += G) {
            const int r = i / N, c = i - (i / N) * N;
            const float x = a[r * NE + c];

            if (r == c)
                dg += x * x;
            else if (r < c)
                off += x * x;
        }
        off = gsum<G>(off);
        dg = gsum<G>(dg);
        if (off <= 1.0e-14f * dg)
            break;
        for (int round = 0; round < NE - 1; round++) {
            /* The rotation of pair lg: circle positions lg and NE-1-lg,
               position 0 fixed, the others advanced by the round. */
            if (lg < NP) {
                const int s2 = NE - 1 - lg;
                const int i1 = lg == 0 ? 0 : (lg - 1 + round) % (NE - 1) + 1;
                const int i2 = (s2 - 1 + round) % (NE - 1) + 1;
                const int p = min(i1, i2), q = max(i1, i2);
                const float apq = a[p * NE + q];
                const float r = fabsf(apq);
                float4 jr = make_float4(1.f, 0.f, 0.f, 1.f);
                float2 d = make_float2(0.f, 0.f);

                if (r != 0.0f) {
                    const float sg = apq < 0.0f ? -1.0f : 1.0f;
                    const float app = a[p * NE + p], aqq = a[q * NE + q];
                    const float theta = (aqq - app) / (2.0f * r);
                    float t;

                    if (fabsf(theta) > 1.0e18f)
                        t = 0.5f / theta;
                    else
                        t = (theta >= 0.0f ? 1.0f : -1.0f) /
                            (fabsf(theta) + sqrtf(theta * theta + 1.0f));
                    const float cs = 1.0f / sqrtf(t * t + 1.0f), sn = t * cs;

                    /* new col p = cs p - sn sg q, new col q = sn p + cs sg q */
                    jr = make_float4(cs, sn, -sn * sg, cs * sg);
                    d = make_float2(app - t * r, aqq + t * r);
                }
                em.jac.rot[lg] = jr;
                em.jac.dd[lg] = d;
                em.jac.pq[lg] = make_int4(p, q, r != 0.0f, 0);
            }
            wave_sync();
            /* A <- J^T A J, one upper 2 x 2 block (i, k) at a time. */
            {
                int i = 0, rem = lg;

                while (i < NP && rem >= NP - i) {
                    rem -= NP - i;
                    i++;
                }
                for (int b = lg; b < NB; b += G) {
                    const int k = i + rem;
                    const int4 pi = em.jac.pq[i], pk = em.jac.pq[k];

                    if (i == k) {
                        if (pi.z) {
                            const float2 d = em.jac.dd[i];

                            a[pi.x * NE + pi.x] = d.x;
                            a[pi.y * NE + pi.y] = d.y;
                            a[pi.x * NE + pi.y] = 0.0f;
                            a[pi.y * NE + pi.x] = 0.0f;
                        }
                    }
                    else if (pi.z || pk.z) {
                        const float4 ri = em.jac.rot[i], rk = em.jac.rot[k];
                        const float m00 = a[pi.x * NE + pk.x], m01 = a[pi.x * NE + pk.y];
                        const float m10 = a[pi.y * NE + pk.x], m11 = a[pi.y * NE + pk.y];
                        /* columns (pk, qk) times J_k, then rows (pi, qi) by J_i^T */
                        const float r00 = m00 * rk.x + m01 * rk.z, r01 = m00 * rk.y + m01 * rk.w;
                        const float r10 = m10 * rk.x + m11 * rk.z, r11 = m10 * rk.y + m11 * rk.w;
                        const float n00 = ri.x * r00 + ri.z * r10, n01 = ri.x * r01 + ri.z * r11;
                        const float n10 = ri.y * r00 + ri.w * r10, n11 = ri.y * r01 + ri.w * r11;

                        a[pi.x * NE + pk.x] = a[pk.x * NE + pi.x] = n00;
                        a[pi.x * NE + pk.y] = a[pk.y * NE + pi.x] = n01;
                        a[pi.y * NE + pk.x] = a[pk.x * NE + pi.y] = n10;
                        a[pi.y * NE + pk.y] = a[pk.y * NE + pi.y] = n11;
                    }
                    rem += G;
                    while (i < NP && rem >= NP - i) {
                        rem -= NP - i;
                        i++;
                    }
                }
            }
            /* V <- V J on the real rows. */
            for (int e = lg; e < N * NP; e += G) {
                const int r = e / NP, i = e - (e / NP) * NP;
                const int4 pi = em.jac.pq[i];

                if (pi.z) {
                    const float4 ri = em.jac.rot[i];
                    const float vp = v[r * NE + pi.x], vq = v[r * NE + pi.y];

                    v[r * NE + pi.x] = vp * ri.x + vq * ri.z;
                    v[r * NE + pi.y] = vp * ri.y + vq * ri.w;
                }
            }
            wave_sync();
        }
    }

    /* inv = V diag(1 / lambda) V^T with eigenvalues below 1e-6 of the
       largest magnitude dropped (tcoh.cl:669-677). */
    float lmax = 0.0f;

    for (int k = lg; k < N; k += G)
        lmax = fmaxf(lmax, fabsf(a[k * NE + k]));
    lmax = gmax<G>(lmax) * EIG_FLOOR;
    for (int k = lg; k < N; k += G) {
        const float lk = a[k * NE + k];

        em.jac.il[k] = fabsf(lk) < lmax ? 0.0f : 1.0f / lk;
    }
    wave_sync();
    {
        float w[CPL][N];

#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const int c = lg + m * G;

#pragma unroll
            for (int k = 0; k < N; k++)
                w[m][k] = c < N ? v[c * NE + k] * em.jac.il[k] : 0.0f;
        }
#pragma unroll
        for (int r = 0; r < N; r++) {
            float sm[CPL];

#pragma unroll
            for (int m = 0; m < CPL; m++)
                sm[m] = 0.0f;
#pragma unroll
            for (int k = 0; k < N; k++) {
                const float x = v[r * NE + k];

#pragma unroll
                for (int m = 0; m < CPL; m++)
                    sm[m] = fmaf(x, w[m][k], sm[m]);
            }
#pragma unroll
            for (int m = 0; m < CPL; m++)
                inv[m][r] = sm[m];
            TCOH_STEP_FENCE();
        }
    }
    wave_sync();
}

#endif
