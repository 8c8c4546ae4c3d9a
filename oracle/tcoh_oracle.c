/* CPU oracle of the phase-linking hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker (or the timed CPU baseline);
 * the product path never calls it.
 *
 * A plain C restatement, written from the reference's algorithm:
 *   - SHP constants (stats.c:20-95) in FP64, and the kernel's view of them
 *     (ocl.c:222-229 prints them with %.9e and the OpenCL compiler reads
 *     the literal as a float);
 *   - the per-pixel kernel temporal_coherence (tcoh.cl:529-745) in FP32,
 *     operation for operation, with its helpers (tcoh.cl:34-527);
 *   - strip preparation (main.c:52-127);
 *   - orbit interpolation (orbit.c:324-346) and the flat-earth/topographic
 *     phase grid (geometry.c:26-291) in FP64.
 * Built with -ffp-contract=off so that no multiply-add is fused unless the
 * code says so. OpenCL built-ins map to C99: sqrt->sqrtf, hypot->hypotf,
 * rsqrt->1/sqrtf, copysign->copysignf, fmin/fmax->fminf/fmaxf. The OpenCL
 * log of the t-test is defined here as the float rounding of the double
 * logarithm (see oracle_logf); the reference leaves it to the OpenCL
 * runtime, so tlog decisions are parity-pinned to that definition only.
 */

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* Same layout as struct settings of local_proto.h:33-49. */
struct o_settings {
    int ndates;
    int win_az, win_rg;
    int test; /* 0 KS, 1 AD, 2 tlog */
    double alpha;
    int emi;
    int bias;
    int min_looks;
    int ks_max;
    double ad_norm, ad_sigma, ad_crit;
    double tlog_crit;
    int npairs;
    int pairs_all;
    const char *phase_reference;
    const char *orbit;
};

/* ------------------------------------------------------------------ */
/* SHP constants, stats.c:10-95                                        */
/* ------------------------------------------------------------------ */

/* Scholz & Stephens (1987) table 1, upper-tail points of T_1 for two
   samples (stats.c:13-15); the same values scipy.stats.anderson_ksamp
   interpolates. */
static const double ad_level_alpha[7] = {0.25,  0.10,  0.05, 0.025,
                                         0.01,  0.005, 0.001};
static const double ad_level_crit[7] = {0.325, 1.226, 1.961, 2.718,
                                        3.752, 4.592, 6.546};

/* Standard deviation of A2kN under H0 for samples of sizes na, nb:
   Scholz & Stephens eq. 4 (stats.c:20-42). */
double oracle_ad_sigma(int na, int nb)
{
    const int k = 2, tot = na + nb;
    const double hh = 1.0 / na + 1.0 / nb;
    double h = 0.0, g = 0.0, a, b, c, d, v;

    if (tot <= 3)
        return 1.0;
    for (int j = 1; j <= tot - 1; j++)
        h += 1.0 / j;
    for (int i = 1; i <= tot - 2; i++)
        for (int j = i + 1; j <= tot - 1; j++)
            g += 1.0 / ((double)(tot - i) * j);
    a = (4 * g - 6) * (k - 1) + (10 - 6 * g) * hh;
    b = (2 * g - 4) * k * k + 8 * h * k + (2 * g - 14 * h - 4) * hh - 8 * h +
        4 * g - 6;
    c = (6 * h + 2 * g - 2) * k * k + (4 * h - 4 * g + 6) * k +
        (2 * h - 6) * hh + 4 * h;
    d = (2 * h + 6) * k * k - 4 * h * k;
    v = (a * pow(tot, 3) + b * (double)tot * tot + c * tot + d) /
        ((tot - 1.0) * (tot - 2.0) * (tot - 3.0));
    return v > 0 ? sqrt(v) : 1.0;
}

/* P(|Z| > z) = alpha by bisection of erfc on [0, 40] (stats.c:45-59). */
double oracle_normal_quantile(double alpha)
{
    double lo = 0.0, hi = 40.0;

    for (int it = 0; it < 200; it++) {
        const double mid = 0.5 * (lo + hi);

        if (erfc(mid / sqrt(2.0)) > alpha)
            lo = mid;
        else
            hi = mid;
    }
    return 0.5 * (lo + hi);
}

/* AD critical value, linear in log(alpha) between table levels
   (stats.c:75-92). Returns NAN outside [0.001, 0.25]. */
double oracle_ad_critical(double alpha)
{
    int i;
    double f;

    if (alpha > ad_level_alpha[0] || alpha < ad_level_alpha[6])
        return NAN;
    for (i = 0; i < 6; i++)
        if (alpha >= ad_level_alpha[i + 1])
            break;
    if (i == 6)
        i = 5;
    f = (log(alpha) - log(ad_level_alpha[i])) /
        (log(ad_level_alpha[i + 1]) - log(ad_level_alpha[i]));
    return ad_level_crit[i] + f * (ad_level_crit[i + 1] - ad_level_crit[i]);
}

/* stats.c:61-95 */
void oracle_shp_constants(struct o_settings *s)
{
    const int n = s->ndates;

    s->ks_max = (int)floor(sqrt(-0.5 * log(s->alpha / 2.0)) * sqrt(2.0 / n) *
                           n);
    s->ad_norm = 1.0 / (2.0 * n) * (2.0 / n);
    s->ad_sigma = oracle_ad_sigma(n, n);
    s->ad_crit = s->test == 1 ? oracle_ad_critical(s->alpha) : 0.0;
    s->tlog_crit = oracle_normal_quantile(s->alpha);
}

/* The float the kernel sees for a double constant: ocl.c:222-229 prints it
   with "%.9e" and the OpenCL C compiler reads the literal as a float. */
float oracle_kernel_float(double x)
{
    char buf[64];

    snprintf(buf, sizeof(buf), "%.9e", x);
    return strtof(buf, NULL);
}

/* ------------------------------------------------------------------ */
/* Per-pixel kernel, tcoh.cl:26-745, FP32                              */
/* ------------------------------------------------------------------ */

#define O_MAX_SWEEPS 40        /* tcoh.cl:26 */
#define O_EIG_FLOOR 1.0e-6f    /* tcoh.cl:27 */
#define O_CHOL_FLOOR 1.0e-5f   /* tcoh.cl:30 */
#define O_PHASOR_FLOOR 1.0e-3f /* tcoh.cl:32 */

/* Kernel constants, as the compiled kernel holds them. */
struct kconst {
    int n, wa, wr, w, test, emi, bias, min_looks, ks_max, npairs, pairs_all;
    float ad_norm, ad_sigma, ad_crit, tlog_crit;
};

/* Work arrays of one pixel (the kernel's private memory, tcoh.cl:539-540,
   417-418). */
struct work {
    float *tr, *ti, *ar, *ai, *vr, *vi;
    float zr[64], zi[64];
    float d[64], e[64], taur[64], taui[64], u0[64], u1[64], u2[64], l[64];
    int piv[64];
};

/* OpenCL log(float) as defined by this oracle: the double logarithm
   rounded to float (correct rounding except when the double result lies
   within one double ulp of a float rounding boundary). */
float oracle_logf(float x) { return (float)log((double)x); }

/* tcoh.cl:34-54: merge walk, centre first on ties. */
static int ks_accept(const struct kconst *k, const float *a, const float *b)
{
    int ia = 0, ib = 0, dmax = 0;

    while (ia < k->n && ib < k->n) {
        int diff;

        if (a[ia] <= b[ib])
            ia++;
        else
            ib++;
        diff = abs(ia - ib);
        if (diff > dmax)
            dmax = diff;
    }
    return dmax <= k->ks_max;
}

/* tcoh.cl:56-75: Scholz-Stephens eq. 6 sum in j order. fused reports
   whether the decision would flip if the compiler contracted
   AD_NORM * s - 1 into one FMA. */
static float ad_statistic(const struct kconst *k, const float *a,
                          const float *b, float *fused)
{
    const int n = k->n;
    int ia = 0, ib = 0, m = 0;
    float s = 0.0f;

    for (int j = 1; j < 2 * n; j++) {
        float num;

        if (ia < n && (ib >= n || a[ia] <= b[ib])) {
            m++;
            ia++;
        }
        else
            ib++;
        num = (float)(2 * n * m - j * n);
        s += num * num / (float)(j * (2 * n - j));
    }
    if (fused)
        *fused = fmaf(k->ad_norm, s, -1.0f) / k->ad_sigma;
    return (k->ad_norm * s - 1.0f) / k->ad_sigma;
}

/* tcoh.cl:77-95 */
static int tlog_accept(const struct kconst *k, float mean_a, float var_a,
                       const float *b)
{
    const int n = k->n;
    float mean_b = 0.0f, var_b = 0.0f, se;

    for (int i = 0; i < n; i++)
        mean_b += oracle_logf(b[i]);
    mean_b /= n;
    for (int i = 0; i < n; i++) {
        const float dv = oracle_logf(b[i]) - mean_b;

        var_b += dv * dv;
    }
    var_b /= n - 1;
    se = sqrtf((var_a + var_b) / n);
    if (!(se > 0.0f))
        return 1;
    return fabsf(mean_a - mean_b) / se <= k->tlog_crit;
}

/* tcoh.cl:97-199: cyclic complex Jacobi; both triangles stored, matrix
   destroyed (eigenvalues on the diagonal), eigenvectors in the columns of
   vr + i vi. */
static void hermitian_eigen(int n, float *ar, float *ai, float *vr, float *vi)
{
    for (int p = 0; p < n * n; p++)
        vr[p] = vi[p] = 0.0f;
    for (int p = 0; p < n; p++)
        vr[p * n + p] = 1.0f;

    for (int sweep = 0; sweep < O_MAX_SWEEPS; sweep++) {
        float off = 0.0f, diag = 0.0f;

        for (int p = 0; p < n; p++) {
            diag += ar[p * n + p] * ar[p * n + p];
            for (int q = p + 1; q < n; q++)
                off += ar[p * n + q] * ar[p * n + q] +
                       ai[p * n + q] * ai[p * n + q];
        }
        if (off <= 1.0e-14f * diag)
            break;
        for (int p = 0; p < n - 1; p++) {
            for (int q = p + 1; q < n; q++) {
                const float xr = ar[p * n + q], xi = ai[p * n + q];
                const float r = hypotf(xr, xi);
                float cph, sph, theta, t, c, s, app, aqq;

                if (r == 0.0f)
                    continue;
                /* Phase of a_pq moved out of column/row q. */
                cph = xr / r;
                sph = xi / r;
                for (int k = 0; k < n; k++) {
                    float re = ar[k * n + q], im = ai[k * n + q];

                    ar[k * n + q] = re * cph + im * sph;
                    ai[k * n + q] = im * cph - re * sph;
                    re = ar[q * n + k];
                    im = ai[q * n + k];
                    ar[q * n + k] = re * cph - im * sph;
                    ai[q * n + k] = im * cph + re * sph;
                    re = vr[k * n + q];
                    im = vi[k * n + q];
                    vr[k * n + q] = re * cph + im * sph;
                    vi[k * n + q] = im * cph - re * sph;
                }
                /* Real rotation zeroing a_pq = r. */
                theta = (ar[q * n + q] - ar[p * n + p]) / (2.0f * r);
                if (fabsf(theta) > 1.0e18f)
                    t = 0.5f / theta;
                else
                    t = (theta >= 0.0f ? 1.0f : -1.0f) /
                        (fabsf(theta) + sqrtf(theta * theta + 1.0f));
                c = 1.0f / sqrtf(t * t + 1.0f);
                s = t * c;
                app = ar[p * n + p] - t * r;
                aqq = ar[q * n + q] + t * r;
                for (int k = 0; k < n; k++) {
                    float kpr, kpi, kqr, kqi;

                    if (k != p && k != q) {
                        kpr = ar[k * n + p];
                        kpi = ai[k * n + p];
                        kqr = ar[k * n + q];
                        kqi = ai[k * n + q];
                        ar[k * n + p] = c * kpr - s * kqr;
                        ai[k * n + p] = c * kpi - s * kqi;
                        ar[k * n + q] = s * kpr + c * kqr;
                        ai[k * n + q] = s * kpi + c * kqi;
                        ar[p * n + k] = ar[k * n + p];
                        ai[p * n + k] = -ai[k * n + p];
                        ar[q * n + k] = ar[k * n + q];
                        ai[q * n + k] = -ai[k * n + q];
                    }
                    kpr = vr[k * n + p];
                    kpi = vi[k * n + p];
                    kqr = vr[k * n + q];
                    kqi = vi[k * n + q];
                    vr[k * n + p] = c * kpr - s * kqr;
                    vi[k * n + p] = c * kpi - s * kqi;
                    vr[k * n + q] = s * kpr + c * kqr;
                    vi[k * n + q] = s * kpi + c * kqi;
                }
                ar[p * n + p] = app;
                ar[q * n + q] = aqq;
                ai[p * n + p] = ai[q * n + q] = 0.0f;
                ar[p * n + q] = ai[p * n + q] = 0.0f;
                ar[q * n + p] = ai[q * n + p] = 0.0f;
            }
        }
    }
}

/* tcoh.cl:201-293: Householder reduction to real tridiagonal (zhetd2,
   lower). Householder vectors are left below the subdiagonal. */
static void tridiagonalize(int n, float *ar, float *ai, struct work *w)
{
    float *wr = w->u0, *wi = w->u1;

    for (int k = 0; k < n - 1; k++) {
        const float alr = ar[(k + 1) * n + k], ali = ai[(k + 1) * n + k];
        float xnorm2 = 0.0f, beta, tr_ = 0.0f, ti_ = 0.0f;

        w->d[k] = ar[k * n + k];
        for (int i = k + 2; i < n; i++)
            xnorm2 += ar[i * n + k] * ar[i * n + k] +
                      ai[i * n + k] * ai[i * n + k];
        if (xnorm2 == 0.0f && ali == 0.0f)
            beta = alr;
        else {
            float dr, di, den, sr, si;

            beta = -copysignf(sqrtf(alr * alr + ali * ali + xnorm2), alr);
            tr_ = (beta - alr) / beta;
            ti_ = -ali / beta;
            dr = alr - beta;
            di = ali;
            den = dr * dr + di * di;
            sr = dr / den;
            si = -di / den;
            for (int i = k + 2; i < n; i++) {
                const float xr = ar[i * n + k], xi = ai[i * n + k];

                ar[i * n + k] = xr * sr - xi * si;
                ai[i * n + k] = xr * si + xi * sr;
            }
        }
        w->e[k] = beta;
        w->taur[k] = tr_;
        w->taui[k] = ti_;
        ar[(k + 1) * n + k] = 1.0f;
        ai[(k + 1) * n + k] = 0.0f;
        if (tr_ == 0.0f && ti_ == 0.0f)
            continue;
        {
            float pr = 0.0f, pi = 0.0f, hr, hi;

            for (int i = k + 1; i < n; i++) {
                float s_r = 0.0f, s_i = 0.0f;

                for (int j = k + 1; j < n; j++) {
                    const float a_r = ar[i * n + j], a_i = ai[i * n + j];
                    const float v_r = ar[j * n + k], v_i = ai[j * n + k];

                    s_r += a_r * v_r - a_i * v_i;
                    s_i += a_r * v_i + a_i * v_r;
                }
                wr[i] = tr_ * s_r - ti_ * s_i;
                wi[i] = tr_ * s_i + ti_ * s_r;
                pr += wr[i] * ar[i * n + k] + wi[i] * ai[i * n + k];
                pi += wr[i] * ai[i * n + k] - wi[i] * ar[i * n + k];
            }
            hr = -0.5f * (tr_ * pr - ti_ * pi);
            hi = -0.5f * (tr_ * pi + ti_ * pr);
            for (int i = k + 1; i < n; i++) {
                const float v_r = ar[i * n + k], v_i = ai[i * n + k];

                wr[i] += hr * v_r - hi * v_i;
                wi[i] += hr * v_i + hi * v_r;
            }
            for (int i = k + 1; i < n; i++) {
                const float vir = ar[i * n + k], vii = ai[i * n + k];

                for (int j = k + 1; j < n; j++) {
                    const float vjr = ar[j * n + k], vji = ai[j * n + k];

                    ar[i * n + j] -=
                        vir * wr[j] + vii * wi[j] + wr[i] * vjr + wi[i] * vji;
                    ai[i * n + j] -=
                        vii * wr[j] - vir * wi[j] + wi[i] * vjr - wr[i] * vji;
                }
            }
        }
    }
    w->d[n - 1] = ar[(n - 1) * n + n - 1];
}

/* tcoh.cl:295-308 */
static int sturm_count(int n, const float *d, const float *e, float x,
                       float pivmin)
{
    float q = d[0] - x;
    int count = q < 0.0f;

    for (int i = 1; i < n; i++) {
        if (fabsf(q) < pivmin)
            q = -pivmin;
        q = d[i] - x - e[i - 1] * e[i - 1] / q;
        count += q < 0.0f;
    }
    return count;
}

/* tcoh.cl:310-338 */
static float tridiagonal_eigenvalue(int n, const float *d, const float *e,
                                    int index, float *norm)
{
    float lo = d[0], hi = d[0], pivmin;

    for (int i = 0; i < n; i++) {
        const float r = (i > 0 ? fabsf(e[i - 1]) : 0.0f) +
                        (i < n - 1 ? fabsf(e[i]) : 0.0f);

        lo = fminf(lo, d[i] - r);
        hi = fmaxf(hi, d[i] + r);
    }
    *norm = fmaxf(fabsf(lo), fabsf(hi));
    pivmin = 1.0e-30f + 1.0e-7f * *norm * 1.0e-7f;
    for (int it = 0; it < 64; it++) {
        const float mid = 0.5f * (lo + hi);

        if (mid <= lo || mid >= hi)
            break;
        if (sturm_count(n, d, e, mid, pivmin) > index)
            hi = mid;
        else
            lo = mid;
    }
    return 0.5f * (lo + hi);
}

/* tcoh.cl:340-410: inverse iteration with a partially pivoted LU. */
static void tridiagonal_eigenvector(int n, struct work *w, float lambda,
                                    float norm, float *y)
{
    const float *d = w->d, *e = w->e;
    float *u0 = w->u0, *u1 = w->u1, *u2 = w->u2, *l = w->l;
    int *piv = w->piv;
    const float tiny = 1.0e-6f * fmaxf(norm, 1.0e-30f);
    float dd = d[0] - lambda, s1 = n > 1 ? e[0] : 0.0f;

    for (int k = 0; k < n - 1; k++) {
        const float sub = e[k], nd = d[k + 1] - lambda;
        const float ns = k + 1 < n - 1 ? e[k + 1] : 0.0f;

        if (fabsf(dd) >= fabsf(sub)) {
            if (fabsf(dd) < tiny)
                dd = dd < 0.0f ? -tiny : tiny;
            l[k] = sub / dd;
            piv[k] = 0;
            u0[k] = dd;
            u1[k] = s1;
            u2[k] = 0.0f;
            dd = nd - l[k] * s1;
            s1 = ns;
        }
        else {
            l[k] = dd / sub;
            piv[k] = 1;
            u0[k] = sub;
            u1[k] = nd;
            u2[k] = ns;
            dd = s1 - l[k] * nd;
            s1 = -l[k] * ns;
        }
    }
    if (fabsf(dd) < tiny)
        dd = dd < 0.0f ? -tiny : tiny;
    u0[n - 1] = dd;
    u1[n - 1] = u2[n - 1] = 0.0f;

    for (int k = 0; k < n; k++)
        y[k] = 1.0f;
    for (int it = 0; it < 3; it++) {
        float nrm = 0.0f;

        for (int k = 0; k < n - 1; k++) {
            if (piv[k]) {
                const float t = y[k];

                y[k] = y[k + 1];
                y[k + 1] = t;
            }
            y[k + 1] -= l[k] * y[k];
        }
        for (int k = n - 1; k >= 0; k--) {
            float v = y[k];

            if (k + 1 < n)
                v -= u1[k] * y[k + 1];
            if (k + 2 < n)
                v -= u2[k] * y[k + 2];
            y[k] = v / u0[k];
        }
        for (int k = 0; k < n; k++)
            nrm = fmaxf(nrm, fabsf(y[k]));
        for (int k = 0; k < n; k++)
            y[k] /= nrm;
    }
}

/* tcoh.cl:412-444 */
static void hermitian_eigenvector(int n, float *ar, float *ai, int index,
                                  struct work *w, float *zr, float *zi)
{
    float lambda, norm;

    tridiagonalize(n, ar, ai, w);
    lambda = tridiagonal_eigenvalue(n, w->d, w->e, index, &norm);
    tridiagonal_eigenvector(n, w, lambda, norm, zr);
    for (int k = 0; k < n; k++)
        zi[k] = 0.0f;
    for (int k = n - 2; k >= 0; k--) {
        float sr = 0.0f, si = 0.0f, hr, hi;

        if (w->taur[k] == 0.0f && w->taui[k] == 0.0f)
            continue;
        for (int i = k + 1; i < n; i++) {
            sr += ar[i * n + k] * zr[i] + ai[i * n + k] * zi[i];
            si += ar[i * n + k] * zi[i] - ai[i * n + k] * zr[i];
        }
        hr = w->taur[k] * sr - w->taui[k] * si;
        hi = w->taur[k] * si + w->taui[k] * sr;
        for (int i = k + 1; i < n; i++) {
            zr[i] -= ar[i * n + k] * hr - ai[i * n + k] * hi;
            zi[i] -= ar[i * n + k] * hi + ai[i * n + k] * hr;
        }
    }
}

/* tcoh.cl:446-492 */
static int cholesky_inverse(int n, const float *g, float *inv, float *wk)
{
    for (int j = 0; j < n; j++) {
        float s = g[j * n + j];

        for (int k = 0; k < j; k++)
            s -= wk[j * n + k] * wk[j * n + k];
        if (!(s > O_CHOL_FLOOR))
            return 0;
        wk[j * n + j] = sqrtf(s);
        for (int i = j + 1; i < n; i++) {
            float t = g[i * n + j];

            for (int k = 0; k < j; k++)
                t -= wk[i * n + k] * wk[j * n + k];
            wk[i * n + j] = t / wk[j * n + j];
        }
    }
    for (int j = 0; j < n; j++) {
        wk[j * n + j] = 1.0f / wk[j * n + j];
        for (int i = j + 1; i < n; i++) {
            float t = 0.0f;

            for (int k = j; k < i; k++)
                t -= wk[i * n + k] * wk[k * n + j];
            wk[i * n + j] = t / wk[i * n + i];
        }
    }
    for (int i = 0; i < n; i++)
        for (int j = i; j < n; j++) {
            float t = 0.0f;

            for (int k = j; k < n; k++)
                t += wk[k * n + i] * wk[k * n + j];
            inv[i * n + j] = inv[j * n + i] = t;
        }
    return 1;
}

/* tcoh.cl:494-513 */
static int connected(int n, const float *tr, const float *ti)
{
    int seen[64], stack[64], top = 0, reached = 1;

    for (int i = 0; i < n; i++)
        seen[i] = 0;
    seen[0] = 1;
    stack[top++] = 0;
    while (top > 0) {
        const int i = stack[--top];

        for (int j = 0; j < n; j++)
            if (!seen[j] && (tr[i * n + j] != 0.0f || ti[i * n + j] != 0.0f)) {
                seen[j] = 1;
                reached++;
                stack[top++] = j;
            }
    }
    return reached == n;
}

/* tcoh.cl:515-527 */
static void store_pairs(const struct kconst *k, float *pairs, size_t plane,
                        size_t pix, const float *tr, const float *ti)
{
    int p = 0;

    if (!k->npairs || !pairs)
        return;
    for (int i = 0; i < k->n - 1; i++)
        for (int j = i + 1; j < (k->pairs_all ? k->n : i + 2); j++, p++)
            pairs[p * plane + pix] =
                tr ? hypotf(tr[i * k->n + j], ti[i * k->n + j]) : NAN;
}

/* Statistics of one output pixel for the tests: shp mask bits of the
   window in dy-major order, and the number of AD decisions that an FMA
   contraction would flip. */
struct pixel_diag {
    uint32_t *mask;
    int words;
    long ad_flips;
};

/* tcoh.cl:529-745 for output pixel (x, y). */
static void pixel(const struct kconst *k, const float *slc, const float *amp,
                  const unsigned char *valid, int rows, int cols, int x,
                  int y, float *coh, int *count, float *pairs, struct work *w,
                  struct pixel_diag *dg)
{
    const int n = k->n;
    const size_t plane = (size_t)rows * cols, out = (size_t)y * cols + x;
    float *tr = w->tr, *ti = w->ti, *ar = w->ar, *ai = w->ai, *vr = w->vr,
          *vi = w->vi, *zr = w->zr, *zi = w->zi;
    float mean_a = 0.0f, var_a = 0.0f, sr, si, lmin;
    const float *a;
    const int centre = (y + k->wa / 2) * k->w + x + k->wr / 2;
    int looks = 0;

    if (dg && dg->mask)
        memset(dg->mask + out * dg->words, 0, dg->words * sizeof(uint32_t));
    if (!valid[centre]) {
        coh[out] = NAN;
        count[out] = -1;
        store_pairs(k, pairs, plane, out, NULL, NULL);
        return;
    }
    a = amp + (size_t)centre * n;
    if (k->test == 2) {
        for (int i = 0; i < n; i++)
            mean_a += oracle_logf(a[i]);
        mean_a /= n;
        for (int i = 0; i < n; i++) {
            const float dv = oracle_logf(a[i]) - mean_a;

            var_a += dv * dv;
        }
        var_a /= n - 1;
    }

    for (int i = 0; i < n * n; i++)
        tr[i] = ti[i] = 0.0f;
    for (int dy = 0; dy < k->wa; dy++) {
        for (int dx = 0; dx < k->wr; dx++) {
            const int cand = (y + dy) * k->w + x + dx;
            const float *b = amp + (size_t)cand * n;
            const float *s = slc + (size_t)cand * n * 2;
            int ok;

            if (!valid[cand])
                continue;
            if (k->test == 0)
                ok = ks_accept(k, a, b);
            else if (k->test == 1) {
                float fused;
                const float v = ad_statistic(k, a, b, &fused);

                ok = v <= k->ad_crit;
                if (dg && ok != (fused <= k->ad_crit))
                    dg->ad_flips++;
            }
            else
                ok = tlog_accept(k, mean_a, var_a, b);
            if (!ok)
                continue;
            if (dg && dg->mask) {
                const int bit = dy * k->wr + dx;

                dg->mask[out * dg->words + bit / 32] |= 1u << (bit % 32);
            }
            looks++;
            for (int i = 0; i < n; i++) {
                const float sxr = s[2 * i], sxi = s[2 * i + 1];

                tr[i * n + i] += sxr * sxr + sxi * sxi;
                for (int j = i + 1; j < n; j++) {
                    const float syr = s[2 * j], syi = s[2 * j + 1];

                    tr[i * n + j] += sxr * syr + sxi * syi;
                    ti[i * n + j] += sxi * syr - sxr * syi;
                }
            }
        }
    }
    count[out] = looks;
    if (looks < k->min_looks) {
        coh[out] = NAN;
        store_pairs(k, pairs, plane, out, NULL, NULL);
        return;
    }

    for (int i = 0; i < n; i++)
        zr[i] = sqrtf(tr[i * n + i]);
    for (int i = 0; i < n; i++) {
        for (int j = i + 1; j < n; j++) {
            const float den = zr[i] * zr[j];
            float re = 0.0f, im = 0.0f;

            if (den > 0.0f) {
                re = tr[i * n + j] / den;
                im = ti[i * n + j] / den;
                if (k->bias && looks > 1) {
                    const float m2 = re * re + im * im;
                    const float m2c = (looks * m2 - 1.0f) / (looks - 1.0f);
                    const float f =
                        (m2 > 0.0f && m2c > 0.0f) ? sqrtf(m2c / m2) : 0.0f;

                    re *= f;
                    im *= f;
                }
            }
            tr[i * n + j] = re;
            ti[i * n + j] = im;
            tr[j * n + i] = re;
            ti[j * n + i] = -im;
        }
        tr[i * n + i] = 1.0f;
        ti[i * n + i] = 0.0f;
    }
    store_pairs(k, pairs, plane, out, tr, ti);
    if (k->bias && !connected(n, tr, ti)) {
        coh[out] = NAN;
        return;
    }

    if (!k->emi) {
        memcpy(ar, tr, sizeof(float) * n * n);
        memcpy(ai, ti, sizeof(float) * n * n);
        hermitian_eigenvector(n, ar, ai, n - 1, w, zr, zi);
    }
    else {
        for (int i = 0; i < n * n; i++)
            vr[i] = hypotf(tr[i], ti[i]);
        if (!cholesky_inverse(n, vr, ar, vi)) {
            for (int i = 0; i < n * n; i++) {
                ar[i] = vr[i];
                ai[i] = 0.0f;
            }
            hermitian_eigen(n, ar, ai, vr, vi);
            lmin = 0.0f;
            for (int i = 0; i < n; i++) {
                zr[i] = ar[i * n + i];
                if (fabsf(zr[i]) > lmin)
                    lmin = fabsf(zr[i]);
            }
            lmin *= O_EIG_FLOOR;
            for (int i = 0; i < n; i++)
                zi[i] = fabsf(zr[i]) < lmin ? 0.0f : 1.0f / zr[i];
            for (int i = 0; i < n; i++)
                for (int j = i; j < n; j++) {
                    float g = 0.0f;

                    for (int q = 0; q < n; q++)
                        g += vr[i * n + q] * vr[j * n + q] * zi[q];
                    ar[i * n + j] = ar[j * n + i] = g;
                }
        }
        for (int i = 0; i < n * n; i++) {
            ai[i] = ar[i] * ti[i];
            ar[i] = ar[i] * tr[i];
        }
        hermitian_eigenvector(n, ar, ai, 0, w, zr, zi);
    }

    {
        int ref = n / 2;
        float refr, refi, refm;

        lmin = 0.0f;
        for (int i = 0; i < n; i++)
            lmin = fmaxf(lmin, hypotf(zr[i], zi[i]));
        if (hypotf(zr[ref], zi[ref]) <= O_PHASOR_FLOOR * lmin)
            for (int i = 0; i < n; i++)
                if (hypotf(zr[i], zi[i]) == lmin)
                    ref = i;
        refm = hypotf(zr[ref], zi[ref]);
        refr = zr[ref] / refm;
        refi = -zi[ref] / refm;
        lmin *= O_PHASOR_FLOOR;
        for (int i = 0; i < n; i++) {
            const float re = zr[i] * refr - zi[i] * refi;
            const float im = zr[i] * refi + zi[i] * refr;
            const float m = hypotf(re, im);

            zr[i] = m > lmin ? re / m : 1.0f;
            zi[i] = m > lmin ? im / m : 0.0f;
        }
    }

    sr = si = 0.0f;
    for (int i = 0; i < n; i++) {
        for (int j = i + 1; j < n; j++) {
            const float re = tr[i * n + j], im = ti[i * n + j];
            const float m = hypotf(re, im);
            const float ur = m > 0.0f ? re / m : 1.0f;
            const float ui = m > 0.0f ? im / m : 0.0f;
            const float wr = zr[i] * zr[j] + zi[i] * zi[j];
            const float wi = zr[i] * zi[j] - zi[i] * zr[j];

            sr += ur * wr - ui * wi;
            si += ur * wi + ui * wr;
        }
    }
    coh[out] = hypotf(sr, si) / (0.5f * n * (n - 1));
}

static void kconst_from(const struct o_settings *s, int padded_cols,
                        struct kconst *k)
{
    k->n = s->ndates;
    k->wa = s->win_az;
    k->wr = s->win_rg;
    k->w = padded_cols;
    k->test = s->test;
    k->emi = s->emi;
    k->bias = s->bias;
    k->min_looks = s->min_looks;
    k->ks_max = s->ks_max;
    k->npairs = s->npairs;
    k->pairs_all = s->pairs_all;
    k->ad_norm = oracle_kernel_float(s->ad_norm);
    k->ad_sigma = oracle_kernel_float(s->ad_sigma);
    k->ad_crit = oracle_kernel_float(s->ad_crit);
    k->tlog_crit = oracle_kernel_float(s->tlog_crit);
}

/* One strip through the FP32 kernel (ocl_run's contract, ocl.c:279-375).
   mask (optional, words = ceil(win_az*win_rg/32) per pixel) receives the
   accepted candidates; the return value counts AD decisions that FMA
   contraction would flip. Rows are spread over nthreads OpenMP threads
   (0: the OpenMP default). */
long oracle_run(const struct o_settings *s, int padded_cols, const float *slc,
                const float *amp, const unsigned char *valid, int padded_rows,
                int rows, int cols, float *coh, int *count, float *pairs,
                uint32_t *mask, int nthreads)
{
    struct kconst k;
    long flips = 0;
    const int words = (s->win_az * s->win_rg + 31) / 32;

    (void)padded_rows;
    kconst_from(s, padded_cols, &k);
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel reduction(+ : flips)
    {
        const size_t nn = (size_t)k.n * k.n;
        float *buf = malloc(6 * nn * sizeof(float));
        struct work *w = malloc(sizeof(*w));
        struct pixel_diag dg = {mask, words, 0};

        w->tr = buf;
        w->ti = buf + nn;
        w->ar = buf + 2 * nn;
        w->ai = buf + 3 * nn;
        w->vr = buf + 4 * nn;
        w->vi = buf + 5 * nn;
#pragma omp for schedule(dynamic, 1)
        for (int y = 0; y < rows; y++)
            for (int x = 0; x < cols; x++)
                pixel(&k, slc, amp, valid, rows, cols, x, y, coh, count,
                      pairs, w, &dg);
        flips += dg.ad_flips;
        free(w);
        free(buf);
    }
    return flips;
}

/* Single SHP decisions for the tests: tcoh.cl:34-95 on one pair. */
int oracle_ks_accept(int n, int ks_max, const float *a, const float *b)
{
    struct kconst k = {0};

    k.n = n;
    k.ks_max = ks_max;
    return ks_accept(&k, a, b);
}

float oracle_ad_statistic(int n, float ad_norm, float ad_sigma,
                          const float *a, const float *b)
{
    struct kconst k = {0};

    k.n = n;
    k.ad_norm = ad_norm;
    k.ad_sigma = ad_sigma;
    return ad_statistic(&k, a, b, NULL);
}

int oracle_tlog_accept(int n, float tlog_crit, const float *a, const float *b)
{
    struct kconst k = {0};
    float mean_a = 0.0f, var_a = 0.0f;

    k.n = n;
    k.tlog_crit = tlog_crit;
    for (int i = 0; i < n; i++)
        mean_a += oracle_logf(a[i]);
    mean_a /= n;
    for (int i = 0; i < n; i++) {
        const float dv = oracle_logf(a[i]) - mean_a;

        var_a += dv * dv;
    }
    var_a /= n - 1;
    return tlog_accept(&k, mean_a, var_a, b);
}

/* ------------------------------------------------------------------ */
/* Strip preparation, main.c:52-127                                    */
/* ------------------------------------------------------------------ */

/* i, q: float[n][nrows][ncols] with NaN nulls; phase: double[n][nrows]
   [ncols] or NULL; elev_null: u8[nrows][ncols] or NULL. */
void oracle_prepare_strip(int n, int nrows, int ncols, const float *iq_i,
                          const float *iq_q, const double *phase,
                          const unsigned char *elev_null, int row0,
                          int padded_rows, int halo_cols, float *slc,
                          float *amp, unsigned char *valid)
{
    const int pc = ncols + 2 * halo_cols;
    const size_t npad = (size_t)padded_rows * pc;
    const size_t plane = (size_t)nrows * ncols;

    memset(valid, 0, npad);
    memset(slc, 0, npad * n * 2 * sizeof(float));
    memset(amp, 0, npad * n * sizeof(float));
    for (int r = 0; r < padded_rows; r++) {
        const int row = row0 + r;
        const size_t base = (size_t)r * pc + halo_cols;

        if (row < 0 || row >= nrows)
            continue;
        memset(valid + base, 1, ncols);
        for (int k = 0; k < n; k++)
            for (int c = 0; c < ncols; c++) {
                const size_t pix = base + c, src = k * plane +
                                                   (size_t)row * ncols + c;
                const float vi = iq_i[src], vq = iq_q[src];
                float *s = slc + (pix * n + k) * 2;

                if (isnan(vi) || isnan(vq) || (vi == 0.0f && vq == 0.0f)) {
                    valid[pix] = 0;
                    continue;
                }
                s[0] = vi;
                s[1] = vq;
            }
        for (int c = 0; c < ncols; c++) {
            const size_t pix = base + c;
            float *a = amp + pix * n;

            if (elev_null && elev_null[(size_t)row * ncols + c])
                valid[pix] = 0;
            if (!valid[pix]) {
                memset(slc + pix * n * 2, 0, n * 2 * sizeof(float));
                continue;
            }
            for (int k = 0; k < n; k++) {
                const float *s = slc + (pix * n + k) * 2;
                const float v = sqrtf(s[0] * s[0] + s[1] * s[1]);
                int j = k;

                while (j > 0 && a[j - 1] > v) {
                    a[j] = a[j - 1];
                    j--;
                }
                a[j] = v;
            }
            if (phase) {
                for (int k = 0; k < n; k++) {
                    float *s = slc + (pix * n + k) * 2;
                    const double ph =
                        phase[k * plane + (size_t)row * ncols + c];
                    const double cs = cos(ph), sn = sin(ph);
                    const double re = s[0], im = s[1];

                    s[0] = (float)(re * cs - im * sn);
                    s[1] = (float)(re * sn + im * cs);
                }
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Orbit and geometry, orbit.c:324-346, geometry.c:20-291 (FP64)       */
/* ------------------------------------------------------------------ */

#define O_LAGRANGE 8 /* orbit.c LAGRANGE_POINTS */
#define O_C 299792458.0
#define O_WGS84_A 6378137.0
#define O_WGS84_B 6356752.314245
#define O_GRID_STEP 16

/* State vectors: t[n], pos[n][3], vel[n][3]. */
void oracle_orbit_state(int n, const double *t, const double *pos,
                        const double *vel, double tq, double *p, double *v)
{
    int first = 0;

    while (first + O_LAGRANGE < n && t[first + O_LAGRANGE / 2] < tq)
        first++;
    p[0] = p[1] = p[2] = v[0] = v[1] = v[2] = 0.0;
    for (int i = first; i < first + O_LAGRANGE; i++) {
        double wgt = 1.0;

        for (int j = first; j < first + O_LAGRANGE; j++)
            if (j != i)
                wgt *= (tq - t[j]) / (t[i] - t[j]);
        for (int c = 0; c < 3; c++) {
            p[c] += wgt * pos[3 * i + c];
            v[c] += wgt * vel[3 * i + c];
        }
    }
}

static double dot3(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* geometry.c:31-71. Returns NAN outside the orbit coverage. */
static double zero_doppler(int n, const double *t, const double *pos,
                           const double *vel, const double *p)
{
    double tz = t[0], best = -1.0;

    for (int i = 0; i < n; i++) {
        double dv[3], dist;

        for (int c = 0; c < 3; c++)
            dv[c] = p[c] - pos[3 * i + c];
        dist = dot3(dv, dv);
        if (best < 0 || dist < best) {
            best = dist;
            tz = t[i];
        }
    }
    for (int it = 0; it < 30; it++) {
        double s[3], v[3], dv[3], step;

        oracle_orbit_state(n, t, pos, vel, tz, s, v);
        for (int c = 0; c < 3; c++)
            dv[c] = p[c] - s[c];
        step = dot3(v, dv) / dot3(v, v);
        tz += step;
        if (fabs(step) < 1e-10)
            break;
    }
    if (tz < t[0] || tz > t[n - 1])
        return NAN;
    return tz;
}

/* geometry.c:74-95 */
static int solve3(double a[3][3], const double *b, double *x)
{
    const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                       a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                       a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);

    if (det == 0.0)
        return 0;
    for (int c = 0; c < 3; c++) {
        double m[3][3];

        memcpy(m, a, sizeof(m));
        for (int r = 0; r < 3; r++)
            m[r][c] = b[r];
        x[c] = (m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0])) /
               det;
    }
    return 1;
}

/* geometry.c:100-148. Returns 0 on failure. */
static int geolocate(int n, const double *t, const double *pos,
                     const double *vel, double tq, double range, double h,
                     int have_guess, double *p)
{
    const double a = O_WGS84_A + h, b = O_WGS84_B + h;
    double s[3], v[3];

    oracle_orbit_state(n, t, pos, vel, tq, s, v);
    if (!have_guess) {
        const double rs = sqrt(dot3(s, s));
        double right[3], nr, ground;

        right[0] = v[1] * s[2] - v[2] * s[1];
        right[1] = v[2] * s[0] - v[0] * s[2];
        right[2] = v[0] * s[1] - v[1] * s[0];
        nr = sqrt(dot3(right, right));
        ground = sqrt(fmax(range * range - (rs - a) * (rs - a), 0.0));
        for (int c = 0; c < 3; c++)
            p[c] = s[c] * a / rs + right[c] / nr * ground;
    }
    for (int it = 0; it < 50; it++) {
        double dv[3], jac[3][3], f[3], dx[3], dist;

        for (int c = 0; c < 3; c++)
            dv[c] = p[c] - s[c];
        dist = sqrt(dot3(dv, dv));
        f[0] = -dot3(v, dv);
        f[1] = -(dist - range);
        f[2] = -((p[0] * p[0] + p[1] * p[1]) / (a * a) + p[2] * p[2] / (b * b) -
                 1.0);
        for (int c = 0; c < 3; c++) {
            jac[0][c] = v[c];
            jac[1][c] = dv[c] / dist;
        }
        jac[2][0] = 2.0 * p[0] / (a * a);
        jac[2][1] = 2.0 * p[1] / (a * a);
        jac[2][2] = 2.0 * p[2] / (b * b);
        if (!solve3(jac, f, dx))
            return 0;
        for (int c = 0; c < 3; c++)
            p[c] += dx[c];
        if (sqrt(dot3(dx, dx)) < 1e-6)
            return 1;
    }
    return 0;
}

/* geometry.c:161-254 with the metadata already parsed. Orbits are packed:
   date k has nsv[k] state vectors starting at sv_off[k] in t/pos/vel.
   phase receives [grid_rows][grid_cols][nheights][n]; the grid shape and
   heights are returned. Returns 0 on success, 1 if geolocation failed, 2 if
   a point falls outside an orbit's coverage. */
int oracle_geometry_init(int n, int ref, const int *nsv, const int *sv_off,
                         const double *t, const double *pos, const double *vel,
                         double t_first, double dt, double tau0, double fs,
                         double lambda, double line0, double sample0,
                         int rows, int cols, double hmin, double hmax,
                         double *phase, int *grid_rows, int *grid_cols,
                         int *nheights, double *heights)
{
    const int gr_n = (rows - 1) / O_GRID_STEP + 2;
    const int gc_n = (cols - 1) / O_GRID_STEP + 2;
    int nh;
    double hs[3];
    int status = 0;

    if (hmax - hmin < 1.0) {
        nh = 1;
        hs[0] = 0.5 * (hmin + hmax);
    }
    else {
        nh = 3;
        hs[0] = hmin;
        hs[1] = 0.5 * (hmin + hmax);
        hs[2] = hmax;
    }
    *grid_rows = gr_n;
    *grid_cols = gc_n;
    *nheights = nh;
    for (int i = 0; i < nh; i++)
        heights[i] = hs[i];
    if (!phase)
        return 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : status)
    for (int gr = 0; gr < gr_n; gr++) {
        const double line = line0 + (double)gr * O_GRID_STEP;
        const double tl = t_first + line * dt;

        for (int gc = 0; gc < gc_n; gc++) {
            const double sample = sample0 + (double)gc * O_GRID_STEP;
            const double range = 0.5 * O_C * (tau0 + sample / fs);
            double p[3];

            for (int hl = 0; hl < nh; hl++) {
                double *ph = phase + (((size_t)gr * gc_n + gc) * nh + hl) * n;
                double r_ref = 0.0;

                if (!geolocate(nsv[ref], t + sv_off[ref],
                               pos + 3 * (size_t)sv_off[ref],
                               vel + 3 * (size_t)sv_off[ref], tl, range,
                               hs[hl], hl > 0, p)) {
                    status |= 1;
                    continue;
                }
                for (int k = 0; k < n; k++) {
                    const double *tk = t + sv_off[k];
                    const double *pk = pos + 3 * (size_t)sv_off[k];
                    const double *vk = vel + 3 * (size_t)sv_off[k];
                    const double tz = zero_doppler(nsv[k], tk, pk, vk, p);
                    double s[3], v[3], dv[3];

                    if (isnan(tz)) {
                        status |= 2;
                        ph[k] = NAN;
                        continue;
                    }
                    oracle_orbit_state(nsv[k], tk, pk, vk, tz, s, v);
                    for (int c = 0; c < 3; c++)
                        dv[c] = p[c] - s[c];
                    ph[k] = sqrt(dot3(dv, dv));
                    if (k == ref)
                        r_ref = ph[k];
                }
                for (int k = 0; k < n; k++)
                    ph[k] = 4.0 * M_PI / lambda * (ph[k] - r_ref);
            }
        }
    }
    return status;
}

/* geometry.c:257-291 */
void oracle_geometry_phases(int n, int grid_cols, int nheights,
                            const double *heights, const double *gphase,
                            int row, int col, double h, double *out)
{
    const int gr = row / O_GRID_STEP, gc = col / O_GRID_STEP;
    const double fr = (double)(row - gr * O_GRID_STEP) / O_GRID_STEP;
    const double fc = (double)(col - gc * O_GRID_STEP) / O_GRID_STEP;
    double wh[3] = {1.0, 0.0, 0.0};

    if (nheights == 3) {
        const double *x = heights;

        wh[0] = (h - x[1]) * (h - x[2]) / ((x[0] - x[1]) * (x[0] - x[2]));
        wh[1] = (h - x[0]) * (h - x[2]) / ((x[1] - x[0]) * (x[1] - x[2]));
        wh[2] = (h - x[0]) * (h - x[1]) / ((x[2] - x[0]) * (x[2] - x[1]));
    }
    for (int k = 0; k < n; k++)
        out[k] = 0.0;
    for (int hl = 0; hl < nheights; hl++) {
        const size_t stride = (size_t)nheights * n;
        const double *p00 =
            gphase + (((size_t)gr * grid_cols + gc) * nheights + hl) * n;
        const double *p01 = p00 + stride;
        const double *p10 = p00 + (size_t)grid_cols * stride;
        const double *p11 = p10 + stride;

        for (int k = 0; k < n; k++)
            out[k] += wh[hl] * ((1.0 - fr) * ((1.0 - fc) * p00[k] + fc * p01[k]) +
                                fr * ((1.0 - fc) * p10[k] + fc * p11[k]));
    }
}
