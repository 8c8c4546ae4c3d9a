/* Strip preparation (main.c:52-127 of the reference without the GRASS
 * row reads): validity, sorted amplitudes and the geometric phase
 * rotation of one padded strip, from planar complex dates.
 *
 * A pixel is valid when every date is non-null and non-zero (and it has a
 * height when elev_null is given); halo columns and rows outside the
 * region are invalid. Amplitudes are sqrtf(I^2 + Q^2), sorted increasingly
 * before the rotation; the rotation multiplies date k by exp(i phase_k)
 * in double and rounds to float. Rows are independent and run on all host
 * threads.
 */

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/tcoh_hip.h"

static void sort_small(float *a, int n)
{
    /* Insertion sort: n <= 64 and the result (ascending, ties kept) is
       all that matters. */
    for (int k = 1; k < n; k++) {
        const float v = a[k];
        int j = k;

        while (j > 0 && a[j - 1] > v) {
            a[j] = a[j - 1];
            j--;
        }
        a[j] = v;
    }
}

void tcoh_prepare_strip(int n, int nrows, int ncols, const float *iq_i,
                        const float *iq_q, const double *phase,
                        const unsigned char *elev_null, int row0,
                        int padded_rows, int halo_cols, float *slc,
                        float *amp, unsigned char *valid)
{
    const int pc = ncols + 2 * halo_cols;
    const size_t plane = (size_t)nrows * ncols;

#pragma omp parallel for schedule(static)
    for (int r = 0; r < padded_rows; r++) {
        const int row = row0 + r;
        const size_t base = (size_t)r * pc;

        memset(valid + base, 0, pc);
        memset(slc + base * n * 2, 0, (size_t)pc * n * 2 * sizeof(float));
        memset(amp + base * n, 0, (size_t)pc * n * sizeof(float));
        if (row < 0 || row >= nrows)
            continue;
        for (int c = 0; c < ncols; c++) {
            const size_t pix = base + halo_cols + c;
            const size_t src = (size_t)row * ncols + c;
            float *s = slc + pix * n * 2, *a = amp + pix * n;
            int ok = !(elev_null && elev_null[src]);

            for (int k = 0; ok && k < n; k++) {
                const float vi = iq_i[k * plane + src];
                const float vq = iq_q[k * plane + src];

                if (isnan(vi) || isnan(vq) || (vi == 0.0f && vq == 0.0f))
                    ok = 0;
                s[2 * k] = vi;
                s[2 * k + 1] = vq;
            }
            if (!ok) {
                memset(s, 0, (size_t)n * 2 * sizeof(float));
                continue;
            }
            valid[pix] = 1;
            for (int k = 0; k < n; k++)
                a[k] = sqrtf(s[2 * k] * s[2 * k] + s[2 * k + 1] * s[2 * k + 1]);
            sort_small(a, n);
            if (phase)
                for (int k = 0; k < n; k++) {
                    const double ph = phase[k * plane + src];
                    const double cs = cos(ph), sn = sin(ph);
                    const double re = s[2 * k], im = s[2 * k + 1];

                    s[2 * k] = (float)(re * cs - im * sn);
                    s[2 * k + 1] = (float)(re * sn + im * cs);
                }
        }
    }
}
