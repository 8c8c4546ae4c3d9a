/* Flat-earth and topographic phase from the orbits, FP64 on the GPU: the
 * phase grid of the reference's geometry_init (geometry.c:161-254), its
 * per-pixel interpolation (geometry.c:256-291) and the orbit state
 * interpolation (orbit.c:324-346), each computed here its own way.
 *
 * Grid: a node every 16 lines and samples of the region (and one past
 * it), at 1 height level (the middle of [hmin, hmax] when that spans
 * under 1 m) or 3 (hmin, the middle, hmax), the phase of every date
 * relative to the reference date: phi_k = 4 pi / lambda (R_k - R_ref).
 *
 * One wavefront per grid node (a block of 64 threads):
 *  1. ground points (lanes < levels): the node's line time on the
 *     reference orbit gives s, v; its sample gives the slant range R. The
 *     point lies on the circle of radius R around s in the plane normal to
 *     v (zero Doppler), right of the track; one parameter, the look angle
 *     theta from the nadir direction, is solved by Newton's method on the
 *     ellipsoid raised by h, (x^2 + y^2) / (a + h)^2 + z^2 / (b + h)^2 = 1,
 *     from the look angle at which a sphere of the local geocentric radius
 *     plus h is seen. The points go to LDS.
 *  2. ranges (lane k = date k, for every level): the zero-Doppler time of
 *     the point on date k's orbit, v(t).(p - s(t)) = 0, bracketed between
 *     the two state vectors where the Doppler changes sign and refined by
 *     the secant method on the interpolated states; R_k = |p - s(t)|.
 *  3. phases from the reference date's range (through LDS).
 * Orbit states at any t: the Lagrange polynomial through the 8 state
 * vectors around t (the window of orbit.c:324-346), evaluated in the
 * barycentric form, velocities interpolated as data like positions.
 */

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/tcoh_hip.h"
#include "geo_interp.h"
#include <hip/hip_runtime.h>

#define WINDOW_POINTS 8
#define LIGHT_SPEED 299792458.0
#define ELLIPSOID_A 6378137.0
#define ELLIPSOID_B 6356752.314245
#define NODE_SPACING 16

/* One date's state vectors: times t[n], positions and velocities [n][3]. */
struct StateVectors {
    int n;
    const double *t, *pos, *vel;
};

/* First of the WINDOW_POINTS state vectors around t: half of them before t
   where the orbit allows (orbit.c:324-346 takes the same window). */
__host__ __device__ static int window_start(const StateVectors &o, double t)
{
    int lo = 0, hi = o.n; /* number of state vectors before t, by bisection */

    while (lo < hi) {
        const int mid = (lo + hi) / 2;

        if (o.t[mid] < t)
            lo = mid + 1;
        else
            hi = mid;
    }
    const int first = lo - WINDOW_POINTS / 2;

    return first < 0 ? 0 : (first > o.n - WINDOW_POINTS ? o.n - WINDOW_POINTS : first);
}

/* Position and velocity at t: the window's interpolating polynomial in the
   barycentric (second) form, y_m + sum_i c_i (y_i - y_m) / sum_i c_i with
   c_i = w_i / (t - t_i), w_i = 1 / prod_{j != i} (t_i - t_j) and y_m the
   window's middle state vector: the differences are tens of kilometres,
   not thousands, so their rounding is a hundred times smaller. With base,
   the position comes back as base - p, without rounding p itself to the
   metre-scale grid of its thousands of kilometres (the ranges below are
   differences of that size). */
__host__ __device__ static void interpolate_state(const StateVectors &o, double t,
                                                  double p[3], double v[3],
                                                  const double *base = nullptr)
{
    const int f = window_start(o, t);
    double c[WINDOW_POINTS], sum = 0.0;

    for (int i = 0; i < WINDOW_POINTS; i++) {
        const double dt = t - o.t[f + i];

        if (dt == 0.0) { /* on a state vector */
            for (int d = 0; d < 3; d++) {
                p[d] = base ? base[d] - o.pos[3 * (f + i) + d] : o.pos[3 * (f + i) + d];
                v[d] = o.vel[3 * (f + i) + d];
            }
            return;
        }
        double w = 1.0;

        for (int j = 0; j < WINDOW_POINTS; j++)
            if (j != i)
                w *= o.t[f + i] - o.t[f + j];
        c[i] = 1.0 / (w * dt);
        sum += c[i];
    }
    const int m = f + WINDOW_POINTS / 2;

    for (int d = 0; d < 3; d++) {
        const double pm = o.pos[3 * m + d], vm = o.vel[3 * m + d];
        double ps = 0.0, vs = 0.0;

        for (int i = 0; i < WINDOW_POINTS; i++) {
            ps += c[i] * (o.pos[3 * (f + i) + d] - pm);
            vs += c[i] * (o.vel[3 * (f + i) + d] - vm);
        }
        p[d] = base ? (base[d] - pm) - ps / sum : pm + ps / sum;
        v[d] = vm + vs / sum;
    }
}

__device__ static double dot(const double a[3], const double b[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ static void unit(double a[3])
{
    const double r = 1.0 / sqrt(dot(a, a));

    for (int d = 0; d < 3; d++)
        a[d] *= r;
}

/* The point at slant range R, zero Doppler and height h (right-looking)
   seen from state (s, v); false without convergence. */
__device__ static bool ground_point(const double s[3], const double v[3], double range,
                                    double h, double p[3])
{
    const double A = ELLIPSOID_A + h, B = ELLIPSOID_B + h;
    const double vn = sqrt(dot(v, v));
    const double sv = dot(s, v) / vn;
    /* e1: towards nadir within the zero-Doppler plane; e2: right of track */
    double e1[3], e2[3];

    for (int d = 0; d < 3; d++)
        e1[d] = sv * v[d] / vn - s[d];
    unit(e1);
    e2[0] = v[1] * s[2] - v[2] * s[1];
    e2[1] = v[2] * s[0] - v[0] * s[2];
    e2[2] = v[0] * s[1] - v[1] * s[0];
    unit(e2);
    /* start: the look angle of a sphere of the local geocentric radius */
    const double rs = sqrt(dot(s, s));
    const double sl = s[2] / rs, cl2 = 1.0 - sl * sl;
    const double rho = A * B / sqrt(B * B * cl2 + A * A * sl * sl);
    double ct = (rs * rs + range * range - rho * rho) / (2.0 * rs * range);
    double theta = acos(fmin(1.0, fmax(-1.0, ct)));

    for (int it = 0; it < 50; it++) {
        const double c = cos(theta), sn = sin(theta);
        double q[3], dq[3];

        for (int d = 0; d < 3; d++) {
            q[d] = s[d] + range * (c * e1[d] + sn * e2[d]);
            dq[d] = range * (c * e2[d] - sn * e1[d]);
        }
        const double e = (q[0] * q[0] + q[1] * q[1]) / (A * A) + q[2] * q[2] / (B * B) - 1.0;
        const double de =
            2.0 * ((q[0] * dq[0] + q[1] * dq[1]) / (A * A) + q[2] * dq[2] / (B * B));

        if (!(de != 0.0))
            return false;
        const double step = e / de;

        theta -= step;
        /* converged to 0.1 um along the circle: the step after this one
           (quadratic convergence) would be below the rounding of theta */
        if (fabs(step) * range < 1e-7) {
            const double c2 = cos(theta), s2 = sin(theta);

            for (int d = 0; d < 3; d++)
                p[d] = s[d] + range * (c2 * e1[d] + s2 * e2[d]);
            return true;
        }
    }
    return false;
}

/* Doppler term v(t).(p - s(t)) of point p on orbit o at t; d = p - s(t). */
__device__ static double doppler(const StateVectors &o, const double p[3], double t,
                                 double d[3])
{
    double v[3];

    interpolate_state(o, t, d, v, p);
    return dot(v, d);
}

/* Slant range from orbit o to p at its zero-Doppler time; NAN when the
   orbit's state vectors do not span that time. */
__device__ static double zero_doppler_range(const StateVectors &o, const double p[3])
{
    /* bracket: consecutive state vectors with the Doppler term going from
       positive (p ahead) to non-positive */
    int i = -1;
    double gi = 0.0, gn = 0.0;

    for (int k = 0; k + 1 < o.n && i < 0; k++) {
        const double *s0 = o.pos + 3 * k, *v0 = o.vel + 3 * k;
        const double *s1 = s0 + 3, *v1 = v0 + 3;
        const double g0 = v0[0] * (p[0] - s0[0]) + v0[1] * (p[1] - s0[1]) +
                          v0[2] * (p[2] - s0[2]);
        const double g1 = v1[0] * (p[0] - s1[0]) + v1[1] * (p[1] - s1[1]) +
                          v1[2] * (p[2] - s1[2]);

        if (g0 > 0.0 && g1 <= 0.0) {
            i = k;
            gi = g0;
            gn = g1;
        }
    }
    if (i < 0)
        return NAN;
    double s[3];
    double ta = o.t[i] + (o.t[i + 1] - o.t[i]) * gi / (gi - gn);
    double tb = ta + 1e-3;
    double ga = doppler(o, p, ta, s), gb = doppler(o, p, tb, s);

    for (int it = 0; it < 40 && gb != 0.0 && gb != ga; it++) {
        const double tc = tb - gb * (tb - ta) / (gb - ga);

        ta = tb;
        ga = gb;
        tb = tc;
        gb = doppler(o, p, tb, s);
        if (fabs(tb - ta) < 1e-9)
            break;
    }
    if (!(tb >= o.t[0] && tb <= o.t[o.n - 1]))
        return NAN;
    /* s: p - s(tb), from the last evaluation */
    return sqrt(dot(s, s));
}

struct GridArgs {
    int n, ref, grid_rows, grid_cols, levels;
    double heights[3];
    double t_first, dt, tau0, fs, lambda, line0, sample0;
    const int *count, *first; /* per date: state vectors and their offset */
    const double *t, *pos, *vel;
};

__device__ static StateVectors date_orbit(const GridArgs &ga, int k)
{
    const size_t f = (size_t)ga.first[k];

    return {ga.count[k], ga.t + f, ga.pos + 3 * f, ga.vel + 3 * f};
}

/* status bits: 1 a ground point failed, 2 a date's orbit does not span the
   zero-Doppler time of a node */
__global__ __launch_bounds__(64) void k_geometry(GridArgs ga, double *__restrict__ phase,
                                                int *__restrict__ status)
{
    __shared__ double pts[3][3];
    __shared__ double rng[3][64];
    __shared__ int bad;
    const int node = blockIdx.x, lane = threadIdx.x;
    const int gr = node / ga.grid_cols, gc = node - (node / ga.grid_cols) * ga.grid_cols;

    if (lane == 0)
        bad = 0;
    __syncthreads();
    if (lane < ga.levels) {
        const double line = ga.line0 + (double)(gr * NODE_SPACING);
        const double sample = ga.sample0 + (double)(gc * NODE_SPACING);
        const double range = 0.5 * LIGHT_SPEED * (ga.tau0 + sample / ga.fs);
        double s[3], v[3], p[3];

        interpolate_state(date_orbit(ga, ga.ref), ga.t_first + line * ga.dt, s, v);
        if (ground_point(s, v, range, ga.heights[lane], p)) {
            for (int d = 0; d < 3; d++)
                pts[lane][d] = p[d];
        }
        else {
            atomicOr(&bad, 1);
        }
    }
    __syncthreads();
    if (bad) {
        if (lane == 0)
            atomicOr(status, bad);
        return;
    }
    if (lane < ga.n) {
        const StateVectors o = date_orbit(ga, lane);

        for (int l = 0; l < ga.levels; l++) {
            const double r = zero_doppler_range(o, pts[l]);

            rng[l][lane] = r;
            if (isnan(r))
                atomicOr(&bad, 2);
        }
    }
    __syncthreads();
    if (bad) {
        if (lane == 0)
            atomicOr(status, bad);
        return;
    }
    if (lane < ga.n) {
        const double scale = 4.0 * M_PI / ga.lambda;

        for (int l = 0; l < ga.levels; l++)
            phase[((size_t)node * ga.levels + l) * ga.n + lane] =
                lane == ga.ref ? 0.0 : scale * (rng[l][lane] - rng[l][ga.ref]);
    }
}

extern "C" void tcoh_report_error(const char *msg);
extern "C" void tcoh_clear_error(void);

/* Report a failure as every entry point does (message and exit, or, in
   error mode 1, kept for tcoh_last_error and the caller returns). */
static void geometry_error(const char *fmt, ...)
{
    char buf[512];
    va_list ap;

    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    tcoh_report_error(buf);
}

static bool hip_ok(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        geometry_error("Phase grid: HIP error %d (%s) in %s", (int)e,
                       hipGetErrorString(e), what);
    return e == hipSuccess;
}

extern "C" void orbit_state(const struct orbit *o, double t, double *pos, double *vel)
{
    const StateVectors sv = {o->n, o->t, &o->pos[0][0], &o->vel[0][0]};

    interpolate_state(sv, t, pos, vel);
}

extern "C" void tcoh_geometry_init(struct geometry *g, const struct tcoh_scene *sc, int n,
                                   int ref, const struct orbit *orbits, double hmin,
                                   double hmax, double h0)
{
    tcoh_clear_error();
    g->phase = NULL;
    if (n < 1 || n > 64 || ref < 0 || ref >= n) {
        geometry_error("Phase grid: reference date %d of %d dates (1 to 64 dates)", ref, n);
        return;
    }
    size_t total = 0;

    for (int k = 0; k < n; k++) {
        if (orbits[k].n < WINDOW_POINTS) {
            geometry_error("Phase grid: date %d has %d orbit state vectors, at least %d "
                           "are needed", k + 1, orbits[k].n, WINDOW_POINTS);
            return;
        }
        total += (size_t)orbits[k].n;
    }
    g->ndates = n;
    g->step = NODE_SPACING;
    g->grid_rows = (sc->rows - 1) / NODE_SPACING + 2;
    g->grid_cols = (sc->cols - 1) / NODE_SPACING + 2;
    g->h0 = h0;
    g->nheights = hmax - hmin < 1.0 ? 1 : 3;
    if (g->nheights == 1) {
        g->heights[0] = 0.5 * (hmin + hmax);
    }
    else {
        g->heights[0] = hmin;
        g->heights[1] = 0.5 * (hmin + hmax);
        g->heights[2] = hmax;
    }

    /* every date's state vectors in one device buffer: counts and offsets
       (ints, then the status word), and times, positions, velocities */
    const size_t nodes = (size_t)g->grid_rows * g->grid_cols;
    const size_t nphase = nodes * g->nheights * n;
    int *idx = (int *)malloc(sizeof(int) * (2 * (size_t)n + 1));
    double *sv = (double *)malloc(sizeof(double) * 7 * total);
    double *out = (double *)malloc(sizeof(double) * nphase);
    size_t at = 0;

    for (int k = 0; k < n; k++) {
        const size_t c = (size_t)orbits[k].n;

        idx[k] = orbits[k].n;
        idx[n + k] = (int)at;
        memcpy(sv + at, orbits[k].t, sizeof(double) * c);
        memcpy(sv + total + 3 * at, orbits[k].pos, sizeof(double) * 3 * c);
        memcpy(sv + 4 * total + 3 * at, orbits[k].vel, sizeof(double) * 3 * c);
        at += c;
    }
    idx[2 * n] = 0;
    int *d_idx = nullptr;
    double *d_sv = nullptr, *d_phase = nullptr;
    int status = -1;
    bool ok = hip_ok(hipMalloc(&d_idx, sizeof(int) * (2 * (size_t)n + 1)), "hipMalloc") &&
              hip_ok(hipMalloc(&d_sv, sizeof(double) * 7 * total), "hipMalloc") &&
              hip_ok(hipMalloc(&d_phase, sizeof(double) * nphase), "hipMalloc") &&
              hip_ok(hipMemcpy(d_idx, idx, sizeof(int) * (2 * (size_t)n + 1),
                               hipMemcpyHostToDevice), "hipMemcpy") &&
              hip_ok(hipMemcpy(d_sv, sv, sizeof(double) * 7 * total,
                               hipMemcpyHostToDevice), "hipMemcpy");

    if (ok) {
        GridArgs ga;

        ga.n = n;
        ga.ref = ref;
        ga.grid_rows = g->grid_rows;
        ga.grid_cols = g->grid_cols;
        ga.levels = g->nheights;
        for (int l = 0; l < 3; l++)
            ga.heights[l] = g->heights[l];
        ga.t_first = sc->t_first;
        ga.dt = sc->dt;
        ga.tau0 = sc->tau0;
        ga.fs = sc->fs;
        ga.lambda = sc->lambda;
        ga.line0 = sc->line0;
        ga.sample0 = sc->sample0;
        ga.count = d_idx;
        ga.first = d_idx + n;
        ga.t = d_sv;
        ga.pos = d_sv + total;
        ga.vel = d_sv + 4 * total;
        hipLaunchKernelGGL(k_geometry, dim3((unsigned)nodes), dim3(64), 0, 0, ga, d_phase,
                           d_idx + 2 * n);
        ok = hip_ok(hipGetLastError(), "the phase grid kernel") &&
             hip_ok(hipMemcpy(out, d_phase, sizeof(double) * nphase, hipMemcpyDeviceToHost),
                    "hipMemcpy") &&
             hip_ok(hipMemcpy(&status, d_idx + 2 * n, sizeof(int), hipMemcpyDeviceToHost),
                    "hipMemcpy");
    }
    (void)hipFree(d_idx);
    (void)hipFree(d_sv);
    (void)hipFree(d_phase);
    free(idx);
    free(sv);
    if (ok && status == 0) {
        g->phase = out;
        return;
    }
    free(out);
    if (!ok)
        return;
    if (status & 1)
        geometry_error("Phase grid: no ground point at the slant range of a grid node "
                       "(the look-angle iteration did not converge)");
    else
        geometry_error("Phase grid: a grid node's zero-Doppler time falls outside a "
                       "date's orbit state vectors; are the dates coregistered on the "
                       "reference date?");
}

extern "C" void geometry_phases(const struct geometry *g, int row, int col, double h,
                                double *phase)
{
    const GridCell cell = grid_cell(row, col, g->step, g->grid_cols);

    for (int k = 0; k < g->ndates; k++)
        phase[k] = cell_phase(g->phase, cell, g->grid_cols, g->nheights, g->ndates,
                              g->heights, h, k);
}

extern "C" void geometry_free(struct geometry *g)
{
    free(g->phase);
    g->phase = NULL;
}
