/* Phase linking and temporal coherence on gfx950 (tcoh.cl:641-744).
 *
 * G lanes per pixel; lane l of a group owns the columns c = l + m*G
 * (m < CPL) of the pixel's N x N Hermitian matrix and keeps them, both
 * triangles, in registers. Row-wise data every lane needs (Householder
 * vectors, the tridiagonal, the phasors) passes through a small LDS area
 * of the group.
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

#include "common.h"

#define PHASOR_FLOOR 1.0e-3f
#define CHOL_FLOOR 1.0e-5
#define EIG_FLOOR 1.0e-6f
#define MAX_SWEEPS 40

template <int N> struct EigLds {
    float2 vb[N], wb[N];
    float2 tau[N];
    float d[N], e[N];
    float u0[N], u1[N], u2[N], l[N], y[N];
    int piv[N];
    float2 z[N];
    float2 alpha;
};

/* Cholesky factor (FP64, lower, column by column) or, in the fallback,
   the Jacobi matrix and its eigenvectors (FP32). */
template <int N> union EmiLds {
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
   row k of the columns c > k (A[m][k]), tau_k in LDS. */
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

            if (c == k + 1)
                sh.alpha = make_float2(A[m][k].x, -A[m][k].y);
            if (c == k)
                sh.d[k] = A[m][k].x;
            if (c >= k + 2 && c < N)
                part += A[m][k].x * A[m][k].x + A[m][k].y * A[m][k].y;
        }
        const float xn = gsum<G>(part);
        wave_sync();
        const float alr = sh.alpha.x, ali = sh.alpha.y;
        float beta, tr_ = 0.0f, ti_ = 0.0f, sr = 0.0f, si = 0.0f;

        if (xn == 0.0f && ali == 0.0f)
            beta = alr;
        else {
            const float dr0 = alr, di = ali;

            beta = -copysignf(sqrtf(alr * alr + ali * ali + xn), alr);
            tr_ = (beta - alr) / beta;
            ti_ = -ali / beta;
            {
                const float dr = dr0 - beta;
                const float den = dr * dr + di * di;

                sr = dr / den;
                si = -di / den;
            }
        }
        const float2 tau = make_float2(tr_, ti_);
        float2 vc[CPL];

#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const int c = lg + m * G;

            vc[m] = make_float2(0.f, 0.f);
            if (c == k + 1)
                vc[m] = make_float2(1.f, 0.f);
            if (c >= k + 2 && c < N) {
                /* x_c = A[c][k] = conj(A[k][c]), v_c = x_c / (alpha - beta) */
                const float2 xc = make_float2(A[m][k].x, -A[m][k].y);

                vc[m] = cmul(xc, make_float2(sr, si));
                A[m][k] = vc[m];
            }
            if (c >= k + 1 && c < N)
                sh.vb[c] = vc[m];
        }
        if (lg == 0) {
            sh.e[k] = beta;
            sh.tau[k] = tau;
        }
        wave_sync();

        /* w = tau A v (A Hermitian: row c is conj of column c). */
        float2 wc[CPL];
        float pr = 0.0f, pi = 0.0f;

#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const int c = lg + m * G;
            float sr2 = 0.0f, si2 = 0.0f;

#pragma unroll
            for (int j = k + 1; j < N; j++) {
                const float2 v = sh.vb[j];

                sr2 += A[m][j].x * v.x + A[m][j].y * v.y;
                si2 += A[m][j].x * v.y - A[m][j].y * v.x;
            }
            wc[m] = cmul(tau, make_float2(sr2, si2));
            if (c >= k + 1 && c < N) {
                pr += wc[m].x * vc[m].x + wc[m].y * vc[m].y;
                pi += wc[m].x * vc[m].y - wc[m].y * vc[m].x;
            }
        }
        pr = gsum<G>(pr);
        pi = gsum<G>(pi);
        {
            const float2 h = make_float2(-0.5f * (tr_ * pr - ti_ * pi),
                                         -0.5f * (tr_ * pi + ti_ * pr));

#pragma unroll
            for (int m = 0; m < CPL; m++) {
                const int c = lg + m * G;
                const float2 hv = cmul(h, vc[m]);

                wc[m].x += hv.x;
                wc[m].y += hv.y;
                if (c >= k + 1 && c < N)
                    sh.wb[c] = wc[m];
            }
        }
        wave_sync();

        /* A[r][c] -= v_r conj(w_c) + w_r conj(v_c), r, c > k. */
#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const int c = lg + m * G;

            if (c >= k + 1 && c < N) {
#pragma unroll
                for (int r = k + 1; r < N; r++) {
                    const float2 v = sh.vb[r], w = sh.wb[r];

                    A[m][r].x -= v.x * wc[m].x + v.y * wc[m].y +
                                 w.x * vc[m].x + w.y * vc[m].y;
                    A[m][r].y -= v.y * wc[m].x - v.x * wc[m].y +
                                 w.y * vc[m].x - w.x * vc[m].y;
                }
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

#pragma unroll
        for (int i = 1; i < N; i++) {
            if (fabsf(qv) < pivmin)
                qv = -pivmin;
            qv = sh.d[i] - x - sh.e[i - 1] * sh.e[i - 1] / qv;
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
            lk = sub / dd;
            pv = 0;
            a0 = dd;
            a1 = s1;
            a2 = 0.0f;
            dd = nd - lk * s1;
            s1 = ns;
        }
        else {
            lk = dd / sub;
            pv = 1;
            a0 = sub;
            a1 = nd;
            a2 = ns;
            dd = s1 - lk * nd;
            s1 = -lk * ns;
        }
        sh.l[k] = lk;
        sh.piv[k] = pv;
        sh.u0[k] = a0;
        sh.u1[k] = a1;
        sh.u2[k] = a2;
    }
    if (fabsf(dd) < tiny)
        dd = dd < 0.0f ? -tiny : tiny;
    sh.u0[N - 1] = dd;
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
            y[k] = v / sh.u0[k];
        }
#pragma unroll
        for (int k = 0; k < N; k++)
            nrm = fmaxf(nrm, fabsf(y[k]));
#pragma unroll
        for (int k = 0; k < N; k++)
            y[k] /= nrm;
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

template <int N, int G, int EMI>
__global__ __launch_bounds__(64) void k_eig(KParams kp,
                                           const float2 *__restrict__ tbuf,
                                           const int *__restrict__ count,
                                           float *__restrict__ coh, int pix0,
                                           int npix)
{
    constexpr int CPL = (N + G - 1) / G;
    constexpr int PPW = WAVE / G;
    constexpr int NT = N * (N - 1) / 2;
    __shared__ EigLds<N> shs[PPW];
    __shared__ EmiLds<N> ems[EMI ? PPW : 1];
    const int gi = threadIdx.x / G, lg = threadIdx.x % G;
    const int q = blockIdx.x * PPW + gi;
    const bool live = q < npix;
    const int p = pix0 + (live ? q : 0);
    EigLds<N> &sh = shs[gi];
    const int cnt = live ? count[p] : -1;
    bool est = live && cnt >= kp.min_looks;
    const float2 *tp = tbuf + (size_t)(live ? q : 0) * NT;
    float2 A[CPL][N];

#pragma unroll
    for (int m = 0; m < CPL; m++) {
        const int c = lg + m * G;

#pragma unroll
        for (int r = 0; r < N; r++) {
            A[m][r] = make_float2(r == c ? 1.f : 0.f, 0.f);
            if (est && c < N)
                A[m][r] = load_t<N>(tp, r, c);
        }
    }

    if (kp.bias) {
        unsigned long long adj[CPL];
        unsigned long long reach = 1ull;
        const unsigned long long all = N == 64 ? ~0ull : ((1ull << N) - 1);

#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const int c = lg + m * G;

            adj[m] = 0ull;
#pragma unroll
            for (int r = 0; r < N; r++)
                if (c < N && r != c && (A[m][r].x != 0.0f || A[m][r].y != 0.0f))
                    adj[m] |= 1ull << r;
        }
        for (int it = 0; it < N; it++) {
            unsigned long long add = 0ull;

#pragma unroll
            for (int m = 0; m < CPL; m++) {
                const int c = lg + m * G;

                if (c < N && ((reach >> c) & 1ull))
                    add |= adj[m];
            }
            const unsigned long long nr = reach | gor64<G>(add);

            if (nr == reach)
                break;
            reach = nr;
        }
        est = est && reach == all;
    }

    float2 z[CPL];

    if (!EMI) {
        hermitian_eigenvector<N, G, CPL>(A, sh, N - 1, lg, z);
    }
    else {
        float inv[CPL][N];

        abs_inverse<N, G, CPL>(tp, est, ems[EMI ? gi : 0], lg, inv);
        /* M = |T|^-1 o T, reloaded so that T need not stay in registers
           through the inversion. */
#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const int c = lg + m * G;

#pragma unroll
            for (int r = 0; r < N; r++) {
                float2 t = make_float2(r == c ? 1.f : 0.f, 0.f);

                if (est && c < N)
                    t = load_t<N>(tp, r, c);
                A[m][r] = make_float2(t.x * inv[m][r], t.y * inv[m][r]);
            }
        }
        hermitian_eigenvector<N, G, CPL>(A, sh, 0, lg, z);
    }

    /* Phasors relative to date N/2 or to the last strongest component. */
    float mags[CPL], lmax = 0.0f;

#pragma unroll
    for (int m = 0; m < CPL; m++) {
        const int c = lg + m * G;

        mags[m] = hypotf(z[m].x, z[m].y);
        if (c < N) {
            lmax = fmaxf(lmax, mags[m]);
            sh.z[c] = z[m];
        }
    }
    lmax = gmax<G>(lmax);
    wave_sync();
    int ref = N / 2;
    {
        const float2 zr = sh.z[N / 2];

        if (hypotf(zr.x, zr.y) <= PHASOR_FLOOR * lmax) {
            int last = -1;

#pragma unroll
            for (int m = 0; m < CPL; m++) {
                const int c = lg + m * G;

                if (c < N && mags[m] == lmax)
                    last = c;
            }
            ref = gmaxi<G>(last);
            if (ref < 0) /* only with non-finite magnitudes */
                ref = N / 2;
        }
    }
    const float2 zref = sh.z[ref];
    const float refm = hypotf(zref.x, zref.y);
    const float refr = zref.x / refm, refi = -zref.y / refm;
    const float floor_ = lmax * PHASOR_FLOOR;
    wave_sync();
#pragma unroll
    for (int m = 0; m < CPL; m++) {
        const int c = lg + m * G;
        const float re = z[m].x * refr - z[m].y * refi;
        const float im = z[m].x * refi + z[m].y * refr;
        const float mg = hypotf(re, im);

        z[m] = mg > floor_ ? make_float2(re / mg, im / mg) : make_float2(1.f, 0.f);
        if (c < N)
            sh.z[c] = z[m];
    }
    wave_sync();

    /* gamma: |sum_{i<j} unit(T_ij) conj(z_i) z_j| / (N(N-1)/2) */
    float gr = 0.0f, gim = 0.0f;

#pragma unroll
    for (int m = 0; m < CPL; m++) {
        const int c = lg + m * G;

        if (est && c < N) {
#pragma unroll
            for (int i = 0; i < N; i++) {
                if (i < c) {
                    const float2 t = tp[upper_index(N, i, c)];
                    const float mg = hypotf(t.x, t.y);
                    const float ur = mg > 0.0f ? t.x / mg : 1.0f;
                    const float ui = mg > 0.0f ? t.y / mg : 0.0f;
                    const float2 w = cmulc(sh.z[i], z[m]);

                    gr += ur * w.x - ui * w.y;
                    gim += ur * w.y + ui * w.x;
                }
            }
        }
    }
    gr = gsum<G>(gr);
    gim = gsum<G>(gim);
    if (live && lg == 0)
        coh[p] = est ? hypotf(gr, gim) / (0.5f * N * (N - 1))
                     : __builtin_nanf("");
}

template <int N>
hipError_t launch_eig_n(const KParams &kp, const float2 *tbuf,
                        const int *count, float *coh, int pix0, int npix,
                        hipStream_t s)
{
    constexpr int G = N <= 16 ? 8 : (N <= 32 ? 16 : 64);
    constexpr int PPW = WAVE / G;
    const dim3 grid((npix + PPW - 1) / PPW);

    if (kp.emi)
        hipLaunchKernelGGL((k_eig<N, G, 1>), grid, dim3(WAVE), 0, s, kp, tbuf,
                           count, coh, pix0, npix);
    else
        hipLaunchKernelGGL((k_eig<N, G, 0>), grid, dim3(WAVE), 0, s, kp, tbuf,
                           count, coh, pix0, npix);
    return hipGetLastError();
}

/* One translation unit per stack depth (the build compiles this file once
   for every N, in parallel). */
template hipError_t launch_eig_n<TCOH_EIG_N>(const KParams &, const float2 *,
                                             const int *, float *, int, int,
                                             hipStream_t);
