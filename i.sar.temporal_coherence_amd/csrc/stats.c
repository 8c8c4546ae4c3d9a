/* Decision thresholds of the SHP tests (stats.c:1-95 of the reference).
 *
 *   KS_MAX    = floor(c(alpha) sqrt(2/N) N), c(alpha) = sqrt(-ln(alpha/2)/2)
 *   AD_NORM   = (1/(2N)) (1/N + 1/N)   Scholz & Stephens (1987) eq. 6
 *   AD_SIGMA  = sd of A2 under H0, two samples of N (their eq. 4)
 *   AD_CRIT   = T_1 upper-tail point, linear in log(alpha) between the
 *               levels of their table 1 (alpha in [0.001, 0.25])
 *   TLOG_CRIT = two-sided standard normal quantile of alpha (Halley's
 *               method from Abramowitz and Stegun 26.2.23)
 */

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../include/tcoh_hip.h"

/* Scholz & Stephens (1987) table 1 for two samples (m = 1): significance
   levels and the interpolated T_1 critical points, as tabulated by the
   reference (stats.c:13-15) and scipy.stats.anderson_ksamp. */
static const double level_alpha[] = {0.25, 0.10, 0.05, 0.025,
                                     0.01, 0.005, 0.001};
static const double level_crit[] = {0.325, 1.226, 1.961, 2.718,
                                    3.752, 4.592, 6.546};
#define NLEVELS 7

static double harmonic(int m)
{
    double h = 0.0;

    for (int j = 1; j <= m; j++)
        h += 1.0 / j;
    return h;
}

/* Variance of A2kN for k = 2 samples of sizes a, b (eq. 4). */
static double anderson_sigma(int a, int b)
{
    const int k = 2, nt = a + b;
    const double hh = 1.0 / a + 1.0 / b, h = harmonic(nt - 1);
    double g = 0.0, ca, cb, cc, cd, var;

    if (nt <= 3)
        return 1.0;
    for (int i = 1; i <= nt - 2; i++)
        for (int j = i + 1; j <= nt - 1; j++)
            g += 1.0 / ((double)(nt - i) * j);
    ca = (4 * g - 6) * (k - 1) + (10 - 6 * g) * hh;
    cb = (2 * g - 4) * k * k + 8 * h * k + (2 * g - 14 * h - 4) * hh - 8 * h +
         4 * g - 6;
    cc = (6 * h + 2 * g - 2) * k * k + (4 * h - 4 * g + 6) * k +
         (2 * h - 6) * hh + 4 * h;
    cd = (2 * h + 6) * k * k - 4 * h * k;
    var = (ca * pow(nt, 3) + cb * (double)nt * nt + cc * nt + cd) /
          ((nt - 1.0) * (nt - 2.0) * (nt - 3.0));
    return var > 0 ? sqrt(var) : 1.0;
}

/* z with P(|Z| > z) = alpha for a standard normal Z: the upper-tail point
   of p = alpha / 2, started from the rational approximation 26.2.23 of
   Abramowitz and Stegun (Handbook of Mathematical Functions, 1964; error
   below 4.5e-4) and refined by Halley's method on Q(z) = erfc(z / sqrt 2) / 2
   - p, whose derivatives are -phi(z) and z phi(z): z += 2 Q / (2 phi - z Q).
   Cubic convergence: two or three steps reach double precision. */
static double two_sided_quantile(double alpha)
{
    const double p = 0.5 * alpha;
    const double t = sqrt(-2.0 * log(p < 0.5 ? p : 1.0 - p));
    double z = t - (2.515517 + t * (0.802853 + t * 0.010328)) /
                       (1.0 + t * (1.432788 + t * (0.189269 + t * 0.001308)));

    if (p > 0.5)
        z = -z;
    for (int it = 0; it < 20; it++) {
        const double q = 0.5 * erfc(z / sqrt(2.0)) - p;
        const double phi = exp(-0.5 * z * z) / sqrt(2.0 * M_PI);
        const double dz = 2.0 * q / (2.0 * phi - z * q);

        z += dz;
        if (!(fabs(dz) > 1e-15 * fabs(z)))
            break;
    }
    return z;
}

static double anderson_critical(double alpha)
{
    int i = 0;

    while (i < NLEVELS - 2 && alpha < level_alpha[i + 1])
        i++;
    return level_crit[i] + (log(alpha) - log(level_alpha[i])) /
                               (log(level_alpha[i + 1]) - log(level_alpha[i])) *
                               (level_crit[i + 1] - level_crit[i]);
}

void shp_constants(struct settings *s)
{
    const int n = s->ndates;

    s->ks_max =
        (int)floor(sqrt(-0.5 * log(s->alpha / 2.0)) * sqrt(2.0 / n) * n);
    s->ad_norm = 1.0 / (2.0 * n) * (2.0 / n);
    s->ad_sigma = anderson_sigma(n, n);
    s->ad_crit = 0.0;
    if (s->test == SHP_AD) {
        if (s->alpha > level_alpha[0] || s->alpha < level_alpha[NLEVELS - 1]) {
            fprintf(stderr,
                    "ERROR: SHP test ad: the significance level %g lies outside "
                    "the tabulated levels %g .. %g\n",
                    s->alpha, level_alpha[NLEVELS - 1], level_alpha[0]);
            exit(EXIT_FAILURE);
        }
        s->ad_crit = anderson_critical(s->alpha);
    }
    s->tlog_crit = two_sided_quantile(s->alpha);
}
