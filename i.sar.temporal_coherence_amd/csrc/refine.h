/* FP64 refinement of an EMI pixel's linked phases (tcoh.cl:657-744).
 *
 * The FP32 eigen solve of M = |T|^-1 o T (eig.h) has an absolute error of
 * the order of eps_32 ||M||, and the EMI matrix has ||M|| up to 1e5 where
 * |T| is near singular, against an eigenvalue gap that can be 0.1: its
 * eigenvector is then off by 1e-3 (gamma by 2e-4), as is the reference's
 * FP32 kernel. This step takes that eigenvector as the start of one inverse
 * iteration in FP64, (M - sigma I) y = z with sigma its Rayleigh quotient
 * (within the square of its error of the eigenvalue, so one step converges
 * to FP64 accuracy), and recomputes the phasors and gamma in FP64 from the
 * pixel's T. One pixel per group of G >= N lanes; lane r keeps row r.
 *
 *   |T|^-1      Gauss-Jordan exchange with partial pivoting (the inverse,
 *               for a Cholesky pixel and for a pixel whose condition bound
 *               showed the pseudo-inverse to be the inverse alike)
 *   y           Gauss-Jordan elimination of [M - sigma I | z] with partial
 *               pivoting, pivot rows broadcast through LDS
 *   phasors     datum N/2 or the last strongest date, floor 1e-3
 *               (tcoh.cl:696-724)
 *   gamma       |sum_{i<j} unit(T_ij) conj(z_i) z_j| / (N(N-1)/2)
 *               (tcoh.cl:726-744)
 */
#ifndef TCOH_REFINE_H
#define TCOH_REFINE_H

#include "eig.h"

/* LDS of one pixel. */
template <int N> struct RefineLds {
    double2 prow[N + 1]; /* pivot row (and right-hand side at N) */
    double inv[N * N];   /* |T|^-1, row-major */
    double prow_r[N];    /* real pivot row of the first exchange */
    int pk[N];           /* pivot row of each step */
    double2 z[N];        /* start vector, then the refined phasors */
    double red[2];
};

__device__ __forceinline__ double2 cmul64(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

/* conj(a) b */
__device__ __forceinline__ double2 cmulc64(double2 a, double2 b)
{
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

/* The unused row of largest |v| (ties to the lowest row) of the group:
   v is this lane's candidate (-1 when it has none). */
template <int G> __device__ __forceinline__ int pivot_row(double v, int r, int n)
{
    const double mx = gmaxd<G>(v);

    return gmini<G>(v == mx && v >= 0.0 ? r : n);
}

/* One EMI pixel's gamma (and phasors in sh.z) in FP64 from its T (packed
   strict upper triangle, float2) and the FP32 eigenvector zs (N float2).
   kappa: the pixel's |T| took the pseudo-inverse branch, so its inverse
   stands for the pseudo-inverse only within the condition bound (a pixel
   that passed the Cholesky test takes the inverse whatever its condition).
   Returns false (nothing written) when the bound fails or an exchange met a
   zero pivot: the FP32 result stands. */
template <int N, int G>
__device__ __forceinline__ bool refine_pixel(const float2 *__restrict__ tp,
                                             const float2 *__restrict__ zs, int lg,
                                             RefineLds<N> &sh, bool kappa, double &gamma)
{
    static_assert(G >= N, "one row per lane");
    const int r = lg;
    const bool own = r < N;
    double a[N];
    double fa = 0.0;

    /* |T|^-1 by exchange steps (as gj_inverse, one row per lane) */
#pragma unroll
    for (int c = 0; c < N; c++) {
        const float2 t = own ? load_t<N>(tp, r, c) : make_float2(0.f, 0.f);

        a[c] = own ? (double)hypotf(t.x, t.y) : 0.0;
        fa = fma(a[c], a[c], fa);
    }
    fa = gsumd<G>(fa);
    bool used = !own;
    int step = 0;

#pragma unroll
    for (int k = 0; k < N; k++) {
        const int p = pivot_row<G>(used ? -1.0 : fabs(a[k]), r, N);

        if (r == p) {
#pragma unroll
            for (int c = 0; c < N; c++)
                sh.prow_r[c] = a[c];
            used = true;
            step = k;
        }
        if (lg == 0)
            sh.pk[k] = p;
        wave_sync();
        const double rp = 1.0 / sh.prow_r[k];

        if (r == p) {
#pragma unroll
            for (int c = 0; c < N; c++)
                a[c] = c == k ? rp : -a[c] * rp;
        }
        else {
            const double f = a[k] * rp;

#pragma unroll
            for (int c = 0; c < N; c++)
                a[c] = c == k ? f : fma(-f, sh.prow_r[c], a[c]);
        }
        wave_sync();
        TCOH_STEP_FENCE();
    }
    double fi = 0.0;

    if (own) {
#pragma unroll
        for (int c = 0; c < N; c++) {
            fi = fma(a[c], a[c], fi);
            sh.inv[step * N + sh.pk[c]] = a[c];
        }
    }
    fi = gsumd<G>(fi);
    wave_sync();
    if (kappa && !(fa * fi <= PINV_IS_INV_KAPPA2))
        return false;

    /* row r of M - sigma I, sigma the Rayleigh quotient of the start */
    double2 m[N];
    const float2 zr = own ? zs[r] : make_float2(0.f, 0.f);

    if (own)
        sh.z[r] = make_double2(zr.x, zr.y);
    wave_sync();
    double2 mz = make_double2(0.0, 0.0);

#pragma unroll
    for (int c = 0; c < N; c++) {
        const float2 t = own ? load_t<N>(tp, r, c) : make_float2(0.f, 0.f);
        const double w = own ? sh.inv[r * N + c] : 0.0;

        m[c] = make_double2(w * (double)t.x, w * (double)t.y);
        const double2 mc = cmul64(m[c], sh.z[c]);

        mz.x += mc.x;
        mz.y += mc.y;
    }
    const double2 zd = make_double2(zr.x, zr.y);
    const double num = gsumd<G>(own ? zd.x * mz.x + zd.y * mz.y : 0.0);
    const double den = gsumd<G>(own ? zd.x * zd.x + zd.y * zd.y : 0.0);
    const double sigma = num / den;

    if (own)
        m[r].x -= sigma;
    wave_sync();

    /* (M - sigma I) y = z by elimination with partial pivoting: after step
       k the pivot row p_k alone holds column k; y_k = b_{p_k} / a_{p_k,k}. */
    double2 b = zd;
    used = !own;
    step = 0;
#pragma unroll
    for (int k = 0; k < N; k++) {
        const double mag = m[k].x * m[k].x + m[k].y * m[k].y;
        const int p = pivot_row<G>(used ? -1.0 : mag, r, N);

        if (r == p) {
#pragma unroll
            for (int c = 0; c < N; c++)
                if (c >= k)
                    sh.prow[c] = m[c];
            sh.prow[N] = b;
            used = true;
            step = k;
        }
        wave_sync();
        if (r != p) {
            const double2 pv = sh.prow[k];
            const double pm = pv.x * pv.x + pv.y * pv.y;
            /* f = m_k / pv */
            const double2 f = make_double2((m[k].x * pv.x + m[k].y * pv.y) / pm,
                                           (m[k].y * pv.x - m[k].x * pv.y) / pm);

#pragma unroll
            for (int c = 0; c < N; c++)
                if (c > k) {
                    const double2 q = cmul64(f, sh.prow[c]);

                    m[c].x -= q.x;
                    m[c].y -= q.y;
                }
            const double2 q = cmul64(f, sh.prow[N]);

            b.x -= q.x;
            b.y -= q.y;
            m[k] = make_double2(0.0, 0.0);
        }
        wave_sync();
        TCOH_STEP_FENCE();
    }
    double2 mk = make_double2(0.0, 0.0);

#pragma unroll
    for (int k = 0; k < N; k++)
        if (k == step)
            mk = m[k];
    const double mm = mk.x * mk.x + mk.y * mk.y;
    const double2 y = own ? make_double2((b.x * mk.x + b.y * mk.y) / mm,
                                         (b.y * mk.x - b.x * mk.y) / mm)
                          : make_double2(0.0, 0.0);
    const bool fin = !own || (isfinite(y.x) && isfinite(y.y));

    if (gmini<G>(fin ? 1 : 0) == 0)
        return false;

    /* phasors (tcoh.cl:696-724) from y, in date order */
    if (own)
        sh.z[step] = y;
    wave_sync();
    const double2 yc = own ? sh.z[r] : make_double2(0.0, 0.0);
    const double mag = sqrt(yc.x * yc.x + yc.y * yc.y);
    const double top = gmaxd<G>(own ? mag : 0.0);
    int ref = N / 2;

    {
        const double2 yr = sh.z[N / 2];

        if (sqrt(yr.x * yr.x + yr.y * yr.y) <= 1.0e-3 * top)
            ref = gmaxi<G>(own && mag == top ? r : -1);
    }
    const double2 zref = sh.z[ref];
    const double refm = sqrt(zref.x * zref.x + zref.y * zref.y);
    const double2 rot = make_double2(zref.x / refm, -zref.y / refm);
    const double2 v = cmul64(yc, rot);
    const double vm = sqrt(v.x * v.x + v.y * v.y);
    const double2 u = vm > 1.0e-3 * top ? make_double2(v.x / vm, v.y / vm)
                                        : make_double2(1.0, 0.0);

    wave_sync();
    if (own)
        sh.z[r] = u;
    wave_sync();

    /* gamma: lane r sums the pairs (r, j > r) */
    double2 s = make_double2(0.0, 0.0);

    if (own) {
#pragma unroll
        for (int j = 0; j < N; j++) {
            if (j <= r)
                continue;
            const float2 t = tp[upper_index(N, r, j)];
            const double tx = t.x, ty = t.y;
            const double tm = sqrt(tx * tx + ty * ty);
            const double2 un = tm > 0.0 ? make_double2(tx / tm, ty / tm)
                                        : make_double2(1.0, 0.0);
            const double2 w = cmulc64(u, sh.z[j]); /* conj(z_r) z_j */
            const double2 q = cmul64(un, w);

            s.x += q.x;
            s.y += q.y;
        }
    }
    s.x = gsumd<G>(s.x);
    s.y = gsumd<G>(s.y);
    gamma = sqrt(s.x * s.x + s.y * s.y) / (0.5 * N * (N - 1));
    return true;
}

#endif
