/* Linking stage of one pixel per group of G lanes: connectivity, EVD or
 * EMI phase linking, phasors and the temporal coherence (tcoh.cl:641-744),
 * from the pixel's coherence matrix T in the group's LDS (packed strict
 * upper triangle, unit diagonal implied).
 */
#ifndef TCOH_LINK_H
#define TCOH_LINK_H

#include "eig.h"

/* |T|^-1 of EMI by INV: the Cholesky of k_link, the Gauss-Jordan inverse
   of the first queue kernel, the Jacobi pseudo-inverse of the second. */
enum { INV_CHOLESKY = 0, INV_GAUSS_JORDAN = 1, INV_JACOBI = 2, INV_CHOLESKY_GLOBAL = 3 };

/* Column c of the pixel's T (lane c of a group of G lanes) from the packed
   strict upper triangle: T_rc at upper_index(N, r, c) = K(r) + c above
   the diagonal, the conjugate of upper_index(N, c, r) = Q(c) + r below
   it, both a lane base plus a compile-time offset once r is. Rows entirely
   above or below the columns of slot m take one read; rows among them read
   both and select. Reads whose values are discarded may leave the pixel's
   T but not the group's LDS: every index is >= 0 (Q(c) >= -1 for c < N,
   and r >= 1 where it is read) and at most G past the pixel's last entry,
   inside the scratch area that follows the group's T areas. */
template <int N, int G> struct TCol {
    const float2 *tp, *colp, *rowp;
    int c;

    __device__ __forceinline__ TCol(const float2 *t, int col)
        : tp(t), colp(t + col),
          /* (padding columns, c >= N, read from the start: their base
             would fall below the pixel's T for small N) */
          rowp(t + (col < N ? col * N - col * (col + 1) / 2 - col - 1 : 0)), c(col)
    {
    }
    /* T_rc, r < N, r != c for a column c < N (the diagonal gives 1) */
    __device__ __forceinline__ float2 get(int m, int r) const
    {
        const float2 id = make_float2(r == c ? 1.f : 0.f, 0.f);
        const int kr = r * N - r * (r + 1) / 2 - r - 1;

        if (r < m * G)
            return colp[kr]; /* every column of the slot is right of r */
        if (r > m * G + G - 1) {
            const float2 v = rowp[r]; /* every column left of r */

            return make_float2(v.x, -v.y);
        }
        if (r == 0) {
            /* only the first slot's lanes: c > 0 above the diagonal */
            const float2 v = tp[c > 0 ? c - 1 : 0];

            return c > 0 ? v : id;
        }
        const float2 up = colp[kr], lo = rowp[r];

        return r < c ? up : (r > c ? make_float2(lo.x, -lo.y) : id);
    }
};

/* gamma of the group's pixel, NaN when it is not estimated or (with the
   bias correction) its dates are not linked. EMI: a pixel whose |T| the
   INV step cannot invert (Cholesky pivot test, condition bound) is not
   linked here (deferred set, the result is meaningless) and goes to the
   next queue. */
template <int N, int G, int EMI, int INV>
__device__ __forceinline__ float link_group(const KParams &kp,
                                            const float2 *tp, bool est,
                                            int lg, EigLds<N> &sh,
                                            EmiLds<N> &em, bool &deferred,
                                            float2 *zraw = nullptr,
                                            bool *refine = nullptr)
{
    deferred = false;
    constexpr int CPL = (N + G - 1) / G;
    HMat<N, CPL> A;

    /* EVD keeps T in registers from here on; EMI reads it again after the
       inversion, so that it is not held live through it. */
    if (!EMI) {
#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const TCol<N, G> col(tp, lg + m * G);
            const bool use = est && lg + m * G < N;

#pragma unroll
            for (int r = 0; r < 4 * HMat<N, CPL>::NG; r++) {
                const float2 id = make_float2(r == lg + m * G ? 1.f : 0.f, 0.f);

                if (r < N) {
                    /* loaded in every lane, then selected: a load under
                       the lane condition became a branch around it */
                    const float2 t = col.get(m, r);

                    A.set(m, r, use ? t : id);
                }
                else
                    A.set(m, r, id);
            }
        }
    }

    /* Breadth-first search over the non-zero coherences (tcoh.cl:494-513,
       641-647), one bit per date. */
    if (kp.bias) {
        unsigned long long adj[CPL];
        unsigned long long reach = 1ull;
        const unsigned long long all = N == 64 ? ~0ull : ((1ull << N) - 1);

#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const int c = lg + m * G;

            adj[m] = 0ull;
#pragma unroll
            for (int r = 0; r < N; r++) {
                const float2 t = EMI ? (est && c < N ? load_t<N>(tp, r, c)
                                                     : make_float2(0.f, 0.f))
                                     : A.get(m, r);

                if (c < N && r != c && (t.x != 0.0f || t.y != 0.0f))
                    adj[m] |= 1ull << r;
            }
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
        hermitian_eigenvector<N, G, CPL, false>(A, sh, N - 1, lg, z);
    }
    else {
        float inv[CPL][N];
        bool ok = true;

        if constexpr (INV == INV_CHOLESKY || INV == INV_CHOLESKY_GLOBAL)
            ok = chol_inverse<N, G, CPL>(tp, est, em, lg, inv);
        else if constexpr (INV == INV_GAUSS_JORDAN)
            ok = gj_inverse<N, G, CPL>(tp, est, em, lg, inv);
        else
            jacobi_pinv<N, G, CPL>(tp, est, em, lg, inv);
        if (!ok) {
            deferred = true;
#pragma unroll
            for (int m = 0; m < CPL; m++)
#pragma unroll
                for (int r = 0; r < N; r++)
                    inv[m][r] = 0.0f;
        }
        /* M = |T|^-1 o T, reloaded so that T need not stay in registers
           through the inversion. */
#pragma unroll
        for (int m = 0; m < CPL; m++) {
            const int c = lg + m * G;

#pragma unroll
            for (int r = 0; r < 4 * HMat<N, CPL>::NG; r++) {
                float2 t = make_float2(r == c ? 1.f : 0.f, 0.f);

                if (r < N) {
                    /* T in LDS (k_link): the column loads above; the queue
                       kernels read T from global memory, one exact read
                       per entry */
                    if constexpr (INV == INV_CHOLESKY) {
                        if (est && c < N)
                            t = TCol<N, G>(tp, c).get(m, r);
                    }
                    else if (est && c < N)
                        t = load_t<N>(tp, r, c);
                    t = make_float2(t.x * inv[m][r], t.y * inv[m][r]);
                }
                A.set(m, r, t);
            }
        }
        float lam, norm;

        hermitian_eigenvector<N, G, CPL, true>(A, sh, 0, lg, z, &lam, &norm);
        /* pixels whose FP32 eigenvector may be off by more than REFINE_TOL
           are refined in FP64 (refine.h), from the eigenvector itself */
        if (refine)
            *refine = close_second_eigenvalue<N>(sh, lam, norm);
        if (zraw) {
#pragma unroll
            for (int m = 0; m < CPL; m++)
                if (lg + m * G < N)
                    zraw[lg + m * G] = z[m];
        }
    }

    /* Phasors relative to date N/2 or to the last strongest component
       (tcoh.cl:696-724). */
    float mags[CPL], lmax = 0.0f;

#pragma unroll
    for (int m = 0; m < CPL; m++) {
        const int c = lg + m * G;

        mags[m] = qsqrt<EMI>(fmaf(z[m].x, z[m].x, z[m].y * z[m].y));
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

        if (qsqrt<EMI>(fmaf(zr.x, zr.x, zr.y * zr.y)) <= PHASOR_FLOOR * lmax) {
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
    const float refm = EMI ? hypotf(zref.x, zref.y)
                           : sqrt_fast(fmaf(zref.x, zref.x, zref.y * zref.y));
    const float refr = qdiv<EMI>(zref.x, refm), refi = qdiv<EMI>(-zref.y, refm);
    const float floor_ = lmax * PHASOR_FLOOR;
    wave_sync();
#pragma unroll
    for (int m = 0; m < CPL; m++) {
        const int c = lg + m * G;
        const float re = z[m].x * refr - z[m].y * refi;
        const float im = z[m].x * refi + z[m].y * refr;
        const float mg = qsqrt<EMI>(fmaf(re, re, im * im));
        z[m] = mg > floor_ ? make_float2(qdiv<EMI>(re, mg), qdiv<EMI>(im, mg))
                           : make_float2(1.f, 0.f);
        if (c < N)
            sh.z[c] = z[m];
    }
    wave_sync();

    /* gamma = |sum_{i<j} unit(T_ij) conj(z_i) z_j| / (N(N-1)/2), unit(0)
       = 1 (tcoh.cl:726-744). Row r of the packed triangle holds its N-1-r
       entries T_{r,r+1..N-1} contiguously; lane lg takes the rows r1 = p
       and r2 = N - 2 - p of pairs p = lg, lg + G, ..., N entries per lane
       in all. Term t is entry t of row r1 for t < N-1-r1, else entry
       t-(N-1-r1) of row r2, whose column is t itself: each is a lane base
       plus a compile-time offset. The middle row of an even N pairs with
       itself and counts once (its second part weighs zero). This where T
       is in LDS (EVD, and EMI in k_link); the queue kernels, which read T
       from global memory, take the columns below. */
    float gr = 0.0f, gim = 0.0f;
    constexpr bool T_IN_LDS = !EMI || INV == INV_CHOLESKY;
    constexpr int NPR = N / 2; /* row pairs of rows 0 .. N-2 */

#pragma unroll
    for (int pp = 0; pp < (NPR + G - 1) / G; pp++) {
        const int r1 = lg + pp * G, r2 = N - 2 - r1, n1 = N - 1 - r1;

        if (!T_IN_LDS || !est || pp * G >= NPR || r1 >= NPR)
            continue;
        const float2 za = sh.z[r1];
        const float2 zb = r1 != r2 ? sh.z[r2] : make_float2(0.f, 0.f);
        const float2 *t1 = tp + upper_index(N, r1, r1 + 1);
        const float2 *t2 = tp + upper_index(N, r2, r2 + 1) - n1;
        const float2 *z1 = sh.z + r1 + 1;

#pragma unroll
        for (int t = 0; t < N; t++) {
            const bool first = t < n1;
            const float2 tv = (first ? t1 : t2)[t];
            const float2 zj = (first ? z1 : sh.z)[t];
            const float2 zi = first ? za : zb;
            const float m2 = fmaf(tv.x, tv.x, tv.y * tv.y);
            const float rm = __builtin_amdgcn_rsqf(m2);
            const float ur = m2 > 0.0f ? tv.x * rm : 1.0f;
            const float ui = m2 > 0.0f ? tv.y * rm : 0.0f;
            const float2 w = cmulc(zi, zj);

            gr = fmaf(ur, w.x, fmaf(-ui, w.y, gr));
            gim = fmaf(ur, w.y, fmaf(ui, w.x, gim));
        }
    }
    /* gamma by columns (the queue kernels): column j has j terms; lane lg
       takes the columns c1 = p and c2 = N - 1 - p of pairs p = lg, lg + G,
       ..., stepping the addresses, N - 1 terms a lane (the middle column of
       an odd N once). */
    constexpr int NPAIR = (N + 1) / 2;

#pragma unroll
    for (int pp = 0; pp < (NPAIR + G - 1) / G; pp++) {
        const int c1 = lg + pp * G, c2 = N - 1 - c1;

        if (!T_IN_LDS) {
            if (!est || pp * G >= NPAIR || c1 >= NPAIR)
                continue;
            const float2 z1 = sh.z[c1], z2 = sh.z[c2];
            /* T_{i,col} at upper_index(N, i, col): col - 1 at i = 0, then
               + (N - 2 - i) per row */
            int addr = c1 - 1, stp = N - 2, i = 0;

#pragma unroll 4
            for (int t = 0; t < N - 1; t++) {
                const bool first = t < c1;

                if (t == c1) {
                    addr = c2 - 1;
                    stp = N - 2;
                    i = 0;
                }
                const float2 zc = first ? z1 : z2;
                const float2 tv = tp[addr];
                const float m2 = fmaf(tv.x, tv.x, tv.y * tv.y);
                const float rm = __builtin_amdgcn_rsqf(m2);
                const float ur = m2 > 0.0f ? tv.x * rm : 1.0f;
                const float ui = m2 > 0.0f ? tv.y * rm : 0.0f;
                const float2 w = cmulc(sh.z[i], zc);
                const float wt = (first || c1 != c2) ? 1.0f : 0.0f;

                gr = fmaf(wt, fmaf(ur, w.x, -ui * w.y), gr);
                gim = fmaf(wt, fmaf(ur, w.y, ui * w.x), gim);
                addr += stp;
                stp -= 1;
                i += 1;
            }
        }
    }
    gr = gsum<G>(gr);
    gim = gsum<G>(gim);
    const float mod = EMI ? hypotf(gr, gim) : sqrt_fast(fmaf(gr, gr, gim * gim));

    return est ? mod / (0.5f * N * (N - 1)) : __builtin_nanf("");
}

#endif
