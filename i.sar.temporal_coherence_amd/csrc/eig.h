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

/* Cholesky factor (FP64, lower, column by column) or, in the fallback,
   the Jacobi matrix and its eigenvectors (FP32). */
template <int N> union __attribute__((aligned(16))) EmiLds {
    double lc[N * (N + 1) / 2 + 1];
    struct {
        float a[N * N];
        float v[N * N];
        float lam[N];
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
#pragma unroll
            for (int j = k + 1; j < N; j++) {
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

        /* A[r][c] -= v_r conj(w_c) + w_r conj(v_c), r > k. */
#pragma unroll
        for (int m = 0; m < CPL; m++) {
            if (min(m * G + G - 1, N - 1) < k + 1)
                continue;
#pragma unroll
            for (int r = k + 1; r < N; r++) {
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

/* |T|^-1 into inv (owned columns) for EMI. */
template <int N, int G, int CPL>
__device__ void abs_inverse(const float2 *__restrict__ tp, bool est,
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
    /* Right-looking Cholesky; column j of L goes to LDS. */
#pragma unroll
    for (int j = 0; j < N; j++) {
#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const int c = lg + m * G;

            if (c == j) {
                const double s = g[m][j];
                const double ljj = sqrt(s > 0.0 ? s : 1.0);

                em.lc[N * (N + 1) / 2] = s > CHOL_FLOOR ? 1.0 : 0.0;
                em.lc[lpos<N>(j, j)] = ljj;
#pragma unroll
                for (int i = j + 1; i < N; i++)
                    em.lc[lpos<N>(i, j)] = g[m][i] / ljj;
            }
        }
        wave_sync();
        ok = ok && em.lc[N * (N + 1) / 2] != 0.0;
#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const int c = lg + m * G;

            if (c > j && c < N) {
                const double lcj = em.lc[lpos<N>(c, j)];

#pragma unroll
                for (int r = j + 1; r < N; r++)
                    if (r >= c)
                        g[m][r] -= em.lc[lpos<N>(r, j)] * lcj;
            }
        }
        wave_sync();
    }
    if (ok) {
        /* Column c of the inverse: L y = e_c, then L^T x = y. */
#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const int c = lg + m * G;
            double yv[N];

#pragma unroll
            for (int r = 0; r < N; r++) {
                double s = r == c ? 1.0 : 0.0;

#pragma unroll
                for (int k = 0; k < r; k++)
                    s -= em.lc[lpos<N>(r, k)] * yv[k];
                yv[r] = s / em.lc[lpos<N>(r, r)];
            }
#pragma unroll
            for (int r = N - 1; r >= 0; r--) {
                double s = yv[r];

#pragma unroll
                for (int k = r + 1; k < N; k++)
                    s -= em.lc[lpos<N>(k, r)] * yv[k];
                yv[r] = s / em.lc[lpos<N>(r, r)];
            }
#pragma unroll
            for (int r = 0; r < N; r++)
                inv[m][r] = c < N ? (float)yv[r] : 0.0f;
        }
    }
    wave_sync();
    if (ok)
        return;

    /* Not safely positive definite: pseudo-inverse from a cyclic Jacobi
       eigendecomposition of |T| in FP32 (tcoh.cl:97-199, 663-687); |T| is
       real, so the phase step of the complex rotation is a sign. */
    float *a = em.jac.a, *v = em.jac.v;

    for (int i = lg; i < N * N; i += G) {
        const int r = i / N, c = i % N;
        float2 t = make_float2(r == c ? 1.f : 0.f, 0.f);

        if (est)
            t = load_t<N>(tp, r, c);
        a[i] = hypotf(t.x, t.y);
        v[i] = r == c ? 1.0f : 0.0f;
    }
    wave_sync();
    for (int sweep = 0; sweep < MAX_SWEEPS; sweep++) {
        float off = 0.0f, dg = 0.0f;

        for (int p = 0; p < N; p++) {
            dg += a[p * N + p] * a[p * N + p];
            for (int q = p + 1; q < N; q++)
                off += a[p * N + q] * a[p * N + q];
        }
        if (off <= 1.0e-14f * dg)
            break;
        for (int p = 0; p < N - 1; p++) {
            for (int q = p + 1; q < N; q++) {
                const float apq = a[p * N + q];
                const float r = fabsf(apq);

                if (r == 0.0f)
                    continue;
                const float sg = apq < 0.0f ? -1.0f : 1.0f;
                const float theta = (a[q * N + q] - a[p * N + p]) / (2.0f * r);
                float t;

                if (fabsf(theta) > 1.0e18f)
                    t = 0.5f / theta;
                else
                    t = (theta >= 0.0f ? 1.0f : -1.0f) /
                        (fabsf(theta) + sqrtf(theta * theta + 1.0f));
                const float cs = 1.0f / sqrtf(t * t + 1.0f), sn = t * cs;
                const float app = a[p * N + p] - t * r;
                const float aqq = a[q * N + q] + t * r;
                wave_sync();
                for (int k = lg; k < N; k += G) {
                    if (k != p && k != q) {
                        const float kp = a[k * N + p], kq = sg * a[k * N + q];

                        a[k * N + p] = a[p * N + k] = cs * kp - sn * kq;
                        a[k * N + q] = a[q * N + k] = sn * kp + cs * kq;
                    }
                    const float vp = v[k * N + p], vq = sg * v[k * N + q];

                    v[k * N + p] = cs * vp - sn * vq;
                    v[k * N + q] = sn * vp + cs * vq;
                }
                wave_sync();
                if (lg == 0) {
                    a[p * N + p] = app;
                    a[q * N + q] = aqq;
                    a[p * N + q] = a[q * N + p] = 0.0f;
                }
                wave_sync();
            }
        }
    }
    {
        float lmax = 0.0f;

        for (int k = 0; k < N; k++)
            lmax = fmaxf(lmax, fabsf(a[k * N + k]));
        lmax *= EIG_FLOOR;
#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const int c = lg + m * G;

#pragma unroll
            for (int r = 0; r < N; r++) {
                float s = 0.0f;

                if (c < N)
                    for (int k = 0; k < N; k++) {
                        const float lk = a[k * N + k];
                        const float il = fabsf(lk) < lmax ? 0.0f : 1.0f / lk;

                        s += v[r * N + k] * v[c * N + k] * il;
                    }
                inv[m][r] = s;
            }
        }
    }
    wave_sync();
}


#endif
