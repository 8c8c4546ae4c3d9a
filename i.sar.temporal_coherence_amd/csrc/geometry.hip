/* Flat-earth and topographic phase from the orbits, FP64 on the GPU
 * (geometry.c:1-296 and orbit.c:324-346 of the reference).
 *
 * The phase grid has a node every 16 lines and samples of the region
 * (plus one), and 1 or 3 height levels (hmin, mid, hmax when the heights
 * span at least 1 m). One thread per grid node:
 *   geolocate   ground point at each height seen from the reference orbit
 *               at the node's zero-Doppler time and slant range: Newton on
 *               (Doppler, range, ellipsoid) with a 3x3 Cramer solve,
 *               |dx| < 1e-6 m, at most 50 steps; the next height starts
 *               from the previous solution (geometry.c:100-148, 229-235)
 *   per date k  zero-Doppler time from the nearest state vector, Newton
 *               t += v.(p-s)/|v|^2, |dt| < 1e-10 s, at most 30 steps
 *               (geometry.c:31-71); slant range R_k
 *   phase_k     4 pi / lambda (R_k - R_ref)            (geometry.c:248-249)
 * Positions and velocities come from 8-point Lagrange interpolation of the
 * state vectors (orbit.c:324-346). Per-pixel interpolation stays on the
 * host (geometry_phases, geometry.c:256-291).
 */

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/tcoh_hip.h"
#include <hip/hip_runtime.h>

#define LAGRANGE_POINTS 8
#define SPEED_OF_LIGHT 299792458.0
#define WGS84_A 6378137.0
#define WGS84_B 6356752.314245
#define GRID_STEP 16

struct OrbitView {
    int n;
    const double *t, *pos, *vel; /* pos/vel [n][3] */
};

__host__ __device__ static void lagrange_state(const OrbitView &o, double t,
                                               double *p, double *v)
{
    int first = 0;

    while (first + LAGRANGE_POINTS < o.n &&
           o.t[first + LAGRANGE_POINTS / 2] < t)
        first++;
    for (int c = 0; c < 3; c++)
        p[c] = v[c] = 0.0;
    for (int i = first; i < first + LAGRANGE_POINTS; i++) {
        double w = 1.0;

        for (int j = first; j < first + LAGRANGE_POINTS; j++)
            if (j != i)
                w *= (t - o.t[j]) / (o.t[i] - o.t[j]);
        for (int c = 0; c < 3; c++) {
            p[c] += w * o.pos[3 * i + c];
            v[c] += w * o.vel[3 * i + c];
        }
    }
}

__device__ static double dot3(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ static double det3(const double m[3][3])
{
    return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
           m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}

/* Zero-Doppler time of p on orbit o; NAN outside its coverage. */
__device__ static double zero_doppler_time(const OrbitView &o, const double *p)
{
    double t = o.t[0], best = -1.0;

    for (int i = 0; i < o.n; i++) {
        const double d[3] = {p[0] - o.pos[3 * i], p[1] - o.pos[3 * i + 1],
                             p[2] - o.pos[3 * i + 2]};
        const double dist = dot3(d, d);

        if (best < 0 || dist < best) {
            best = dist;
            t = o.t[i];
        }
    }
    for (int it = 0; it < 30; it++) {
        double s[3], v[3], d[3];

        lagrange_state(o, t, s, v);
        for (int c = 0; c < 3; c++)
            d[c] = p[c] - s[c];
        const double dt = dot3(v, d) / dot3(v, v);

        t += dt;
        if (fabs(dt) < 1e-10)
            break;
    }
    return (t < o.t[0] || t > o.t[o.n - 1]) ? NAN : t;
}

/* Returns false when the Newton solve fails. */
__device__ static bool geolocate_point(const OrbitView &o, double t,
                                       double range, double h, bool guess,
                                       double *p)
{
    const double a = WGS84_A + h, b = WGS84_B + h;
    double s[3], v[3];

    lagrange_state(o, t, s, v);
    if (!guess) {
        const double rs = sqrt(dot3(s, s));
        const double right[3] = {v[1] * s[2] - v[2] * s[1],
                                 v[2] * s[0] - v[0] * s[2],
                                 v[0] * s[1] - v[1] * s[0]};
        const double nr = sqrt(dot3(right, right));
        const double ground = sqrt(fmax(range * range - (rs - a) * (rs - a), 0.0));

        for (int c = 0; c < 3; c++)
            p[c] = s[c] * a / rs + right[c] / nr * ground;
    }
    for (int it = 0; it < 50; it++) {
        double d[3], jac[3][3], f[3], dx[3];

        for (int c = 0; c < 3; c++)
            d[c] = p[c] - s[c];
        const double dist = sqrt(dot3(d, d));

        f[0] = -dot3(v, d);
        f[1] = -(dist - range);
        f[2] = -((p[0] * p[0] + p[1] * p[1]) / (a * a) + p[2] * p[2] / (b * b) -
                 1.0);
        for (int c = 0; c < 3; c++) {
            jac[0][c] = v[c];
            jac[1][c] = d[c] / dist;
        }
        jac[2][0] = 2.0 * p[0] / (a * a);
        jac[2][1] = 2.0 * p[1] / (a * a);
        jac[2][2] = 2.0 * p[2] / (b * b);
        const double det = det3(jac);

        if (det == 0.0)
            return false;
        for (int c = 0; c < 3; c++) {
            double m[3][3];

            for (int r = 0; r < 3; r++)
                for (int q = 0; q < 3; q++)
                    m[r][q] = q == c ? f[r] : jac[r][q];
            dx[c] = det3(m) / det;
        }
        for (int c = 0; c < 3; c++)
            p[c] += dx[c];
        if (sqrt(dot3(dx, dx)) < 1e-6)
            return true;
    }
    return false;
}

struct GeoArgs {
    int n, ref, grid_rows, grid_cols, nheights;
    double heights[3];
    double t_first, dt, tau0, fs, lambda, line0, sample0;
    const int *nsv, *off;
    const double *t, *pos, *vel;
};

__global__ __launch_bounds__(64) void k_geometry(GeoArgs ga,
                                                double *__restrict__ phase,
                                                int *__restrict__ status)
{
    const int node = blockIdx.x * 64 + threadIdx.x;

    if (node >= ga.grid_rows * ga.grid_cols)
        return;
    const int gr = node / ga.grid_cols, gc = node % ga.grid_cols;
    const double line = ga.line0 + (double)gr * GRID_STEP;
    const double t = ga.t_first + line * ga.dt;
    const double sample = ga.sample0 + (double)gc * GRID_STEP;
    const double range = 0.5 * SPEED_OF_LIGHT * (ga.tau0 + sample / ga.fs);
    const OrbitView ro = {ga.nsv[ga.ref], ga.t + ga.off[ga.ref],
                          ga.pos + 3 * (size_t)ga.off[ga.ref],
                          ga.vel + 3 * (size_t)ga.off[ga.ref]};
    double p[3];

    for (int hl = 0; hl < ga.nheights; hl++) {
        double *ph = phase + ((size_t)node * ga.nheights + hl) * ga.n;
        double r_ref = 0.0;

        if (!geolocate_point(ro, t, range, ga.heights[hl], hl > 0, p)) {
            atomicOr(status, 1);
            return;
        }
        for (int k = 0; k < ga.n; k++) {
            const OrbitView o = {ga.nsv[k], ga.t + ga.off[k],
                                 ga.pos + 3 * (size_t)ga.off[k],
                                 ga.vel + 3 * (size_t)ga.off[k]};
            const double tz = zero_doppler_time(o, p);
            double s[3], v[3], d[3];

            if (isnan(tz)) {
                atomicOr(status, 2);
                return;
            }
            lagrange_state(o, tz, s, v);
            for (int c = 0; c < 3; c++)
                d[c] = p[c] - s[c];
            ph[k] = sqrt(dot3(d, d));
            if (k == ga.ref)
                r_ref = ph[k];
        }
        for (int k = 0; k < ga.n; k++)
            ph[k] = 4.0 * M_PI / ga.lambda * (ph[k] - r_ref);
    }
}

[[noreturn]] static void geo_fatal(const char *msg)
{
    fprintf(stderr, "ERROR: %s\n", msg);
    exit(EXIT_FAILURE);
}

static void geo_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        char buf[256];

        snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s", (int)e,
                 hipGetErrorString(e), what);
        geo_fatal(buf);
    }
}

extern "C" void orbit_state(const struct orbit *o, double t, double *pos,
                            double *vel)
{
    const OrbitView v = {o->n, o->t, &o->pos[0][0], &o->vel[0][0]};

    lagrange_state(v, t, pos, vel);
}

extern "C" void tcoh_geometry_init(struct geometry *g,
                                   const struct tcoh_scene *sc, int n, int ref,
                                   const struct orbit *orbits, double hmin,
                                   double hmax, double h0)
{
    GeoArgs ga;
    int total = 0, status = 0;

    if (ref < 0 || ref >= n)
        geo_fatal("The reference date is not one of the inputs");
    g->ndates = n;
    g->step = GRID_STEP;
    g->grid_rows = (sc->rows - 1) / GRID_STEP + 2;
    g->grid_cols = (sc->cols - 1) / GRID_STEP + 2;
    g->h0 = h0;
    if (hmax - hmin < 1.0) {
        g->nheights = 1;
        g->heights[0] = 0.5 * (hmin + hmax);
    }
    else {
        g->nheights = 3;
        g->heights[0] = hmin;
        g->heights[1] = 0.5 * (hmin + hmax);
        g->heights[2] = hmax;
    }
    for (int k = 0; k < n; k++) {
        if (orbits[k].n < LAGRANGE_POINTS)
            geo_fatal("An orbit has fewer than 8 state vectors");
        total += orbits[k].n;
    }

    const size_t nodes = (size_t)g->grid_rows * g->grid_cols;
    const size_t nphase = nodes * g->nheights * n;
    int *hoff = (int *)malloc(sizeof(int) * 2 * n);
    double *ht = (double *)malloc(sizeof(double) * total * 7);
    int *d_ns, *d_status;
    double *d_t, *d_phase;

    for (int k = 0, o = 0; k < n; o += orbits[k].n, k++) {
        hoff[k] = orbits[k].n;
        hoff[n + k] = o;
        memcpy(ht + o, orbits[k].t, sizeof(double) * orbits[k].n);
        memcpy(ht + total + 3 * o, orbits[k].pos, sizeof(double) * 3 * orbits[k].n);
        memcpy(ht + 4 * total + 3 * o, orbits[k].vel,
               sizeof(double) * 3 * orbits[k].n);
    }
    geo_check(hipMalloc(&d_ns, sizeof(int) * 2 * n), "hipMalloc");
    geo_check(hipMalloc(&d_status, sizeof(int)), "hipMalloc");
    geo_check(hipMalloc(&d_t, sizeof(double) * total * 7), "hipMalloc");
    geo_check(hipMalloc(&d_phase, sizeof(double) * nphase), "hipMalloc");
    geo_check(hipMemcpy(d_ns, hoff, sizeof(int) * 2 * n, hipMemcpyHostToDevice),
              "hipMemcpy");
    geo_check(hipMemcpy(d_t, ht, sizeof(double) * total * 7,
                        hipMemcpyHostToDevice),
              "hipMemcpy");
    geo_check(hipMemcpy(d_status, &status, sizeof(int), hipMemcpyHostToDevice),
              "hipMemcpy");
    ga.n = n;
    ga.ref = ref;
    ga.grid_rows = g->grid_rows;
    ga.grid_cols = g->grid_cols;
    ga.nheights = g->nheights;
    for (int i = 0; i < 3; i++)
        ga.heights[i] = g->heights[i];
    ga.t_first = sc->t_first;
    ga.dt = sc->dt;
    ga.tau0 = sc->tau0;
    ga.fs = sc->fs;
    ga.lambda = sc->lambda;
    ga.line0 = sc->line0;
    ga.sample0 = sc->sample0;
    ga.nsv = d_ns;
    ga.off = d_ns + n;
    ga.t = d_t;
    ga.pos = d_t + total;
    ga.vel = d_t + 4 * total;
    hipLaunchKernelGGL(k_geometry, dim3((unsigned)((nodes + 63) / 64)), dim3(64),
                       0, 0, ga, d_phase, d_status);
    geo_check(hipGetLastError(), "geometry kernel");
    g->phase = (double *)malloc(sizeof(double) * nphase);
    geo_check(hipMemcpy(g->phase, d_phase, sizeof(double) * nphase,
                        hipMemcpyDeviceToHost),
              "hipMemcpy");
    geo_check(hipMemcpy(&status, d_status, sizeof(int), hipMemcpyDeviceToHost),
              "hipMemcpy");
    (void)hipFree(d_ns);
    (void)hipFree(d_status);
    (void)hipFree(d_t);
    (void)hipFree(d_phase);
    free(hoff);
    free(ht);
    if (status & 1)
        geo_fatal("Geolocation did not converge (singular geometry)");
    if (status & 2)
        geo_fatal("A ground point is seen outside of an orbit's coverage: is "
                  "the stack coregistered on option reference?");
}

extern "C" void geometry_phases(const struct geometry *g, int row, int col,
                                double h, double *phase)
{
    const int n = g->ndates;
    const int gr = row / g->step, gc = col / g->step;
    const double fr = (double)(row - gr * g->step) / g->step;
    const double fc = (double)(col - gc * g->step) / g->step;
    double wh[3] = {1.0, 0.0, 0.0};

    if (g->nheights == 3) {
        const double *x = g->heights;

        wh[0] = (h - x[1]) * (h - x[2]) / ((x[0] - x[1]) * (x[0] - x[2]));
        wh[1] = (h - x[0]) * (h - x[2]) / ((x[1] - x[0]) * (x[1] - x[2]));
        wh[2] = (h - x[0]) * (h - x[1]) / ((x[2] - x[0]) * (x[2] - x[1]));
    }
    for (int k = 0; k < n; k++)
        phase[k] = 0.0;
    for (int hl = 0; hl < g->nheights; hl++) {
        const size_t stride = (size_t)g->nheights * n;
        const double *p00 =
            g->phase + (((size_t)gr * g->grid_cols + gc) * g->nheights + hl) * n;
        const double *p01 = p00 + stride;
        const double *p10 = p00 + (size_t)g->grid_cols * stride;
        const double *p11 = p10 + stride;

        for (int k = 0; k < n; k++)
            phase[k] += wh[hl] * ((1.0 - fr) * ((1.0 - fc) * p00[k] + fc * p01[k]) +
                                  fr * ((1.0 - fc) * p10[k] + fc * p11[k]));
    }
}

extern "C" void geometry_free(struct geometry *g)
{
    free(g->phase);
    g->phase = NULL;
}
