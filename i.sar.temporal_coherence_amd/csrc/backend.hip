/* HIP backend of the phase-linking module: the ocl_* entry points of
 * local_proto.h:104-112 (ocl.c:105-391 in the reference) and the
 * extensions declared in include/tcoh_hip.h.
 *
 * One handle = one device, one in-order HIP stream and the settings fixed
 * at open (ocl.c:191-254). A strip runs as
 *   prep      amplitudes to date-planar order (+ log statistics for tlog)
 *   shp       SHP mask and count per output pixel
 *   [split]   hi/lo 16-bit copy of the samples (mixed precision only)
 *   link      per band of rows, fused: coherence matrix T in LDS, pair
 *             coherences, phase linking and temporal coherence
 * Buffers only grow (ocl.c:264-277). The reference splits its launches
 * into ~0.5 s chunks against GPU watchdogs (ocl.c:329-363); here bands of
 * about a million pixels (the tuned band size, tiles.h) bound every launch
 * instead.
 */

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <stdexcept>
#include <string>
#include <algorithm>
#include <mutex>
#include <thread>
#include <vector>

#include "common.h"
#include "tiles.h"
#include "../../include/tcoh_hip.h"

namespace {

int g_error_mode = 0;
thread_local std::string g_last_error;

struct tcoh_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

[[noreturn]] void fail(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;

    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw tcoh_error(buf);
}

void check(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        fail("HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
}

/* Report an error as the reference does (G_fatal_error: message, exit)
   or, in error mode 1, keep it for tcoh_last_error(). */
void report(const char *msg)
{
    g_last_error = msg;
    if (g_error_mode == 0) {
        fprintf(stderr, "ERROR: %s\n", msg);
        fflush(stderr);
        exit(EXIT_FAILURE);
    }
}

template <class F> auto guarded(F f, decltype(f()) on_error) -> decltype(f())
{
    g_last_error.clear();
    try {
        return f();
    }
    catch (const std::exception &e) {
        report(e.what());
        return on_error;
    }
}

template <class F> void guarded_void(F f)
{
    g_last_error.clear();
    try {
        f();
    }
    catch (const std::exception &e) {
        report(e.what());
    }
}

/* The float a kernel compiled with "-DX=%.9ef" holds (ocl.c:222-229). */
float kernel_float(double x)
{
    char buf[64];

    snprintf(buf, sizeof(buf), "%.9e", x);
    return strtof(buf, nullptr);
}

struct DevBuf {
    void *p = nullptr;
    size_t size = 0;

    void ensure(size_t bytes, const char *what)
    {
        if (p && size >= bytes)
            return;
        if (p)
            check(hipFree(p), "hipFree");
        p = nullptr;
        size = 0;
        check(hipMalloc(&p, bytes ? bytes : 1), what);
        size = bytes;
    }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        size = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};


} // namespace

/* Stage timing: one event pair around every kernel launch of a run. */
struct StageTimer {
    std::vector<hipEvent_t> ev;
    std::vector<int> stage;
    size_t used = 0;

    void begin() { used = 0; stage.clear(); }
    hipEvent_t next()
    {
        if (used == ev.size()) {
            hipEvent_t e;

            check(hipEventCreate(&e), "hipEventCreate");
            ev.push_back(e);
        }
        return ev[used++];
    }
    void mark(int st, hipStream_t s, bool start)
    {
        check(hipEventRecord(next(), s), "hipEventRecord");
        if (start)
            stage.push_back(st);
    }
    void release()
    {
        for (hipEvent_t e : ev)
            (void)hipEventDestroy(e);
        ev.clear();
    }
};

/* Scratch of the allocation-free entry point (tcoh_run_device_noalloc):
   a file-scope device array, of which every GPU the library runs on holds
   its own copy for the life of the process. Host code reaches a device's
   copy only through the address hipGetSymbolAddress gives with that device
   current (noalloc_base): the array's name is a host-side stand-in. */
#define TCOH_NOALLOC_BYTES ((size_t)256 << 20)
#define TCOH_NOALLOC_ALIGN 256
#define TCOH_MAX_DEVICES 64
static __device__ __align__(256) unsigned char scratch[TCOH_NOALLOC_BYTES];

namespace {
std::mutex g_noalloc_mu;
void *g_noalloc_base[TCOH_MAX_DEVICES];
} // namespace

/* The device address of the current device's scratch array (looked up
   once per device, at the device's first ocl_open). */
static void *noalloc_base(int device)
{
    std::lock_guard<std::mutex> lock(g_noalloc_mu);

    if (device < 0 || device >= TCOH_MAX_DEVICES)
        fail("Device %d beyond the %d the library tracks", device, TCOH_MAX_DEVICES);
    if (!g_noalloc_base[device]) {
        void *p = nullptr;

        check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_noalloc_base[device] = p;
    }
    return g_noalloc_base[device];
}

/* What a captured strip depends on: a graph is replayed only for the same
   buffers, shape, stream and arithmetic. */
struct GraphKey {
    const void *slc, *amp, *valid, *coh, *count, *pairs, *phase;
    int padded_rows, rows, cols, precision;
    hipStream_t stream;
    int tiles[3];

    bool operator==(const GraphKey &k) const
    {
        return slc == k.slc && amp == k.amp && valid == k.valid && coh == k.coh &&
               count == k.count && pairs == k.pairs && phase == k.phase &&
               padded_rows == k.padded_rows && rows == k.rows && cols == k.cols &&
               precision == k.precision && stream == k.stream &&
               !memcmp(tiles, k.tiles, sizeof(tiles));
    }
};

struct GraphCache {
    bool valid = false;
    GraphKey key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    void release()
    {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        valid = false;
    }
};

struct ocl {
    int device;
    hipStream_t stream;
    struct settings s;
    int padded_cols;
    int precision;
    /* launch shapes: SHP and linking stripe widths, pixels per linking
       launch (tiles.h), and whether tcoh_set_tiles fixed them */
    int tiles[3];
    bool tiles_set;
    int timing;
    double stage_ms[4];
    int stage_launches[4];
    StageTimer timer;
    /* ocl_run's transfer buffers */
    DevBuf slc, amp, valid, coh, count, pairs, phase;
    /* scratch */
    DevBuf ampT, lstat, mask, adtab, hl, ex;
    /* EMI band state (EmiQueue): T, eigenvectors and inversion kind of the
       band's pixels, the two queues and their counts */
    DevBuf qt, qz, qref, qpix, qn, qnext;
    /* the phase grid of tcoh_prepare_strip_device / tcoh_run_region */
    DevBuf grid;
    /* tcoh_run_region: double-buffered date planes and outputs on the
       device, pinned host staging, one strip's prepared buffers, the copy
       stream and the events ordering the two streams */
    DevBuf rin[2], rhgt[2], rout[2], rslc, ramp, rvalid;
    void *hin[2] = {nullptr, nullptr}, *hout[2] = {nullptr, nullptr};
    size_t hin_size[2] = {0, 0}, hout_size[2] = {0, 0};
    hipStream_t copy = nullptr;
    hipEvent_t ev_h2d[2], ev_prep[2], ev_comp[2], ev_d2h[2];
    bool region_ready = false;
    /* graph mode (tcoh_set_graph) and the captured strip */
    bool graph = false;
    struct GraphCache *gcache = nullptr;
    /* this device's copy of the allocation-free entry point's scratch */
    void *noalloc = nullptr;
};

static int env_precision()
{
    const char *e = getenv("TCOH_PRECISION");

    if (!e || !*e)
        return TCOH_MIXED_BF16;
    if (!strcmp(e, "plain"))
        return TCOH_PLAIN;
    if (!strcmp(e, "f16"))
        return TCOH_MIXED_F16;
    return TCOH_MIXED_BF16;
}

/* The tuned shapes of a handle's estimator, form and depth (tiles.h):
   the nearest tuned depth with the same lanes per pixel, else the nearest
   one, else the built-in shapes. */
static void tile_lookup(struct ocl *o)
{
    const int n = o->s.ndates, emi = o->s.emi ? 1 : 0;
    const int mixed = o->precision != TCOH_PLAIN && !emi;
    auto bucket = [](int d) { return d <= 16 ? 0 : (d <= 32 ? 1 : 2); };
    const TileEntry *best = nullptr;
    long best_d = 0;

    for (const TileEntry &e : TILE_TABLE) {
        if (e.n <= 0 || e.emi != emi || e.mixed != mixed)
            continue;
        const long d = (bucket(e.n) != bucket(n) ? 1000L : 0L) + std::abs(e.n - n);

        if (!best || d < best_d) {
            best = &e;
            best_d = d;
        }
    }
    o->tiles[0] = best ? best->shp_stripe : TILE_STRIPE_DEFAULT;
    o->tiles[1] = best ? best->link_stripe : TILE_STRIPE_DEFAULT;
    o->tiles[2] = best && best->band_px > 0 ? best->band_px : TILE_BAND_PX;
}

/* ocl_open's max_alloc: main.c cuts strips so that the samples (8N
   bytes a padded pixel) and the pair maps (4P bytes an output pixel) each
   stay within max_alloc (main.c:731-742). A strip of S pixels needs about
   S * F bytes on the device in all, F the per-pixel footprint of every
   buffer (inputs, outputs, SHP masks, date-planar amplitudes and, for the
   split forms, which a handle may switch to after ocl_open, the 16-bit
   samples); EMI adds its queues, sized by the linking band. max_alloc =
   usable * max(8N, 4P) / F therefore keeps such a strip within 90 % of the
   free memory. The SHP kernel indexes the date planes with 32-bit offsets,
   which caps a strip at INT32_MAX / bucket pixels. */
static size_t strip_alloc_limit(const struct settings &s, int band_px, size_t free_bytes)
{
    const double n = s.ndates, nb = shp_bucket(s.ndates), nt = n * (n - 1) / 2;
    const double words = (s.win_az * s.win_rg + 31) / 32;
    double f = 12 * n + 1 + 4 * nb + 4 + 4 + 4.0 * s.npairs + 4 * words;
    double fixed = 0;

    if (s.test == SHP_TLOG)
        f += 8;
    if (s.emi)
        fixed = (double)band_px * (8 * nt + 8 * n + 12);
    else
        f += 8 * nb + 2;
    const double usable = 0.9 * (double)free_bytes - fixed;
    const double big = std::max(8 * n, 4.0 * s.npairs);
    const double lim = usable > 0 ? usable * big / f : 0.0;
    const double cap = (double)(INT32_MAX / shp_bucket(s.ndates)) * 8 * n;

    return (size_t)std::min(lim, cap);
}

static bool eig_supported(int n);
static hipError_t launch_link(const KParams &kp, const float2 *slc,
                              const uint2 *hl, const short *ex, const uint32_t *mask,
                              const int *count, float *coh, float *pairs,
                              int row0, int nrows, int precision,
                              const EmiQueue &q, hipStream_t s);
static hipError_t launch_emi_queue(const KParams &kp, const EmiQueue &q, int pbase,
                                   int cap, float *coh, hipStream_t s);

extern "C" void ocl_list_devices(void)
{
    guarded_void([] {
        int nd = 0;

        check(hipGetDeviceCount(&nd), "hipGetDeviceCount");
        if (nd == 0)
            fail("No HIP device found");
        fprintf(stdout, "platform|device|type|name|version|compute_units|"
                        "global_memory_mb\n");
        for (int d = 0; d < nd; d++) {
            hipDeviceProp_t pr;
            int rt = 0;

            check(hipGetDeviceProperties(&pr, d), "hipGetDeviceProperties");
            (void)hipRuntimeGetVersion(&rt);
            fprintf(stdout, "0|%d|GPU|%s (%s)|HIP %d.%d|%d|%lu\n", d, pr.name,
                    pr.gcnArchName, rt / 10000000, (rt / 100000) % 100,
                    pr.multiProcessorCount,
                    (unsigned long)(pr.totalGlobalMem / (1024 * 1024)));
        }
    });
}

extern "C" struct ocl *ocl_open(int platform, int device,
                                const struct settings *s, int padded_cols,
                                size_t *max_alloc)
{
    return guarded(
        [&]() -> struct ocl * {
            int nd = 0;

            check(hipGetDeviceCount(&nd), "hipGetDeviceCount");
            if (platform >= 0 && platform != 0)
                fail("Platform %d not found (1 platform, see -l)", platform);
            if (platform < 0 && device >= 0)
                fail("Option device requires option platform");
            if (nd == 0)
                fail("No HIP device found");
            if (device < 0)
                device = 0;
            if (device >= nd)
                fail("Device %d not found on platform 0 (%d devices, see -l)",
                     device, nd);
            if (s->ndates < 3 || s->ndates > MAX_DATES)
                fail("The stack must have 3 to %d dates, not %d", MAX_DATES,
                     s->ndates);
            if (!eig_supported(s->ndates))
                fail("No kernel built for %d dates", s->ndates);
            if (s->win_az < 1 || s->win_rg < 1 || s->win_az % 2 == 0 ||
                s->win_rg % 2 == 0)
                fail("Window sizes must be odd, got %d,%d", s->win_az,
                     s->win_rg);
            if (s->win_az * s->win_rg > TCOH_MAX_WINDOW)
                fail("Windows of more than %d pixels are not supported",
                     TCOH_MAX_WINDOW);
            if (padded_cols < s->win_rg)
                fail("Padded strip width %d below the window width %d",
                     padded_cols, s->win_rg);

            struct ocl *o = new struct ocl();
            hipDeviceProp_t pr;

            o->gcache = new GraphCache();
            size_t fr = 0, tot = 0;

            o->device = device;
            o->s = *s;
            o->padded_cols = padded_cols;
            o->precision = env_precision();
            o->tiles_set = false;
            tile_lookup(o);
            o->timing = 0;
            check(hipSetDevice(device), "hipSetDevice");
            o->noalloc = noalloc_base(device);
            check(hipStreamCreateWithFlags(&o->stream, hipStreamNonBlocking),
                  "hipStreamCreate");
            check(hipGetDeviceProperties(&pr, device), "hipGetDeviceProperties");
            check(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
            if (max_alloc)
                *max_alloc = strip_alloc_limit(*s, o->tiles[2], fr);
            return o;
        },
        nullptr);
}

extern "C" void tcoh_set_precision(struct ocl *o, int precision)
{
    guarded_void([&] {
        if (precision < TCOH_PLAIN || precision > TCOH_MIXED_BF16)
            fail("Unknown precision %d", precision);
        o->precision = precision;
        if (!o->tiles_set)
            tile_lookup(o);
    });
}

extern "C" void tcoh_set_timing(struct ocl *o, int on) { o->timing = on; }

extern "C" void tcoh_set_tiles(struct ocl *o, const int tiles[3])
{
    guarded_void([&] {
        if (!tiles) {
            o->tiles_set = false;
            tile_lookup(o);
            return;
        }
        if (tiles[0] < 0 || tiles[1] < 0 || tiles[2] < 0)
            fail("Launch shapes must not be negative (%d, %d, %d)", tiles[0], tiles[1],
                 tiles[2]);
        o->tiles[0] = tiles[0];
        o->tiles[1] = tiles[1];
        o->tiles[2] = tiles[2] > 0 ? tiles[2] : TILE_BAND_PX;
        o->tiles_set = true;
    });
}

extern "C" void tcoh_get_tiles(const struct ocl *o, int tiles[3])
{
    for (int k = 0; k < 3; k++)
        tiles[k] = o->tiles[k];
}

extern "C" void tcoh_stage_times(const struct ocl *o, double ms[4])
{
    for (int i = 0; i < 4; i++)
        ms[i] = o->stage_ms[i];
}

extern "C" void tcoh_stage_launches(const struct ocl *o, int launches[4])
{
    for (int i = 0; i < 4; i++)
        launches[i] = o->stage_launches[i];
}

extern "C" void tcoh_set_error_mode(int mode) { g_error_mode = mode; }

extern "C" const char *tcoh_last_error(void) { return g_last_error.c_str(); }

/* For the host-only C sources of the library (orbit ingestion): report a
   failure under the error mode, and clear the message at an entry point. */
extern "C" void tcoh_report_error(const char *msg) { report(msg); }

extern "C" void tcoh_clear_error(void) { g_last_error.clear(); }

/* Device scratch of one strip: where each buffer the kernels use besides
   the caller's inputs and outputs lives. */
struct StripScratch {
    float *ampT = nullptr;
    float2 *lstat = nullptr;
    uint32_t *mask = nullptr;
    float *adtab = nullptr;
    uint32_t *hl = nullptr;
    short *ex = nullptr;
    EmiQueue q = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
};

/* Sizes of the strip's scratch buffers (bytes; 0: not used), in the order
   of StripScratch. */
enum { SC_AMPT, SC_LSTAT, SC_MASK, SC_ADTAB, SC_HL, SC_EX, SC_QT, SC_QZ, SC_QREF, SC_QPIX,
       SC_QNEXT, SC_QN, SC_COUNT };
struct ScratchPlan {
    size_t bytes[SC_COUNT];
    int chunk, cap; /* rows per link launch, pixels per band */
    bool mixed, bf16;

    size_t total(size_t align) const
    {
        size_t t = 0;

        for (size_t b : bytes)
            t += (b + align - 1) / align * align;
        return t;
    }
};

/* Pixels of one linking group: 64 / (lanes per pixel), the lanes being 8
   up to 16 dates, 16 up to 32 and 64 beyond (k_link.hip's LinkShape). */
static int link_group_pixels(int n)
{
    return n <= 16 ? 8 : (n <= 32 ? 4 : 1);
}

/* The strip's settings as the kernels take them, after the strip checks. */
static KParams strip_params(const struct ocl *o, int padded_rows, int rows, int cols,
                            const void *pairs, void *phase)
{
    const struct settings &s = o->s;
    KParams kp;

    if (cols + s.win_rg - 1 != o->padded_cols)
        fail("Strip of %d columns does not match the padded width %d", cols,
             o->padded_cols);
    if (padded_rows < rows + s.win_az - 1)
        fail("Padded strip of %d rows too small for %d rows", padded_rows,
             rows);
    if ((size_t)padded_rows * o->padded_cols > (size_t)INT32_MAX)
        fail("Strip of %d x %d pixels is too large", padded_rows,
             o->padded_cols);
    if (s.npairs && !pairs)
        fail("Pair coherences requested without an output buffer");
    kp.n = s.ndates;
    kp.wa = s.win_az;
    kp.wr = s.win_rg;
    kp.w = o->padded_cols;
    kp.rows = rows;
    kp.cols = cols;
    kp.npad = padded_rows * o->padded_cols;
    kp.test = (int)s.test;
    kp.emi = s.emi;
    kp.bias = s.bias;
    kp.min_looks = s.min_looks;
    kp.ks_max = s.ks_max;
    kp.npairs = s.npairs;
    kp.pairs_all = s.pairs_all;
    kp.words = (s.win_az * s.win_rg + 31) / 32;
    kp.ad_norm = kernel_float(s.ad_norm);
    kp.ad_sigma = kernel_float(s.ad_sigma);
    kp.ad_crit = kernel_float(s.ad_crit);
    kp.tlog_crit = kernel_float(s.tlog_crit);
    kp.phase = (float *)phase;
    /* stripes per row -> stripe width in tiles (SHP tiles: 64 pixels;
       linking tiles: a group's pixels) */
    kp.shp_stripe = o->tiles[0] > 0 ? ((cols + 63) / 64 + o->tiles[0] - 1) / o->tiles[0] : 0;
    {
        const int ppw = link_group_pixels(s.ndates), nbx = (cols + ppw - 1) / ppw;

        kp.link_stripe = o->tiles[1] > 0 ? (nbx + o->tiles[1] - 1) / o->tiles[1] : 0;
    }
    if ((size_t)kp.npad * shp_bucket(kp.n) > (size_t)INT32_MAX)
        fail("Strip of %d pixels too large for %d dates (at most %d pixels; "
             "use strips within ocl_open's max_alloc)",
             kp.npad, kp.n, INT32_MAX / shp_bucket(kp.n));
    return kp;
}

static ScratchPlan scratch_plan(const KParams &kp, int precision, int band_px)
{
    ScratchPlan pl;
    const size_t npad = (size_t)kp.npad, nout = (size_t)kp.rows * kp.cols;
    const size_t n = kp.n, nb = shp_bucket(kp.n), nt = n * (n - 1) / 2;

    for (size_t &b : pl.bytes)
        b = 0;
    /* EMI takes the plain form with FP64 sums whatever the setting */
    pl.mixed = precision != TCOH_PLAIN && !kp.emi;
    pl.bf16 = precision == TCOH_MIXED_BF16;
    pl.chunk = std::max(1, band_px / kp.cols);
    if (pl.mixed) {
        /* a launch's window rows of split samples within LIST_SPAN_MAX */
        const size_t row_bytes = (size_t)kp.w * nb * 8;
        const long fit = (long)(LIST_SPAN_MAX / row_bytes) - (kp.wa - 1);

        if (fit < 1)
            fail("Strip rows of %d padded pixels at %d dates are too wide for the split "
                 "forms with a %d-line window", kp.w, kp.n, kp.wa);
        pl.chunk = (int)std::min<long>(pl.chunk, fit);
    }
    pl.cap = std::min(pl.chunk, kp.rows) * kp.cols;
    pl.bytes[SC_AMPT] = npad * nb * sizeof(float);
    pl.bytes[SC_MASK] = nout * kp.words * sizeof(uint32_t);
    if (kp.test == 2)
        pl.bytes[SC_LSTAT] = npad * sizeof(float2);
    if (kp.test == 1)
        pl.bytes[SC_ADTAB] = sizeof(float) * 2 * n * (n + 1);
    if (pl.mixed) {
        /* samples as (hi, lo) 16-bit pairs, one zero pixel appended for the
           padding of candidate lists; FP16 with each pixel's exponent */
        pl.bytes[SC_HL] = (npad + 1) * nb * 2 * sizeof(uint32_t);
        if (!pl.bf16)
            pl.bytes[SC_EX] = (npad + 1) * sizeof(short);
    }
    if (kp.emi) {
        const size_t cap = (size_t)pl.cap;

        pl.bytes[SC_QT] = cap * nt * sizeof(float2);
        pl.bytes[SC_QZ] = cap * n * sizeof(float2);
        pl.bytes[SC_QREF] = cap * sizeof(int);
        pl.bytes[SC_QPIX] = cap * sizeof(int);
        pl.bytes[SC_QNEXT] = cap * sizeof(int);
        pl.bytes[SC_QN] = 3 * sizeof(int);
    }
    return pl;
}

static StripScratch scratch_pointers(void *const p[SC_COUNT])
{
    StripScratch sc;

    sc.ampT = (float *)p[SC_AMPT];
    sc.lstat = (float2 *)p[SC_LSTAT];
    sc.mask = (uint32_t *)p[SC_MASK];
    sc.adtab = (float *)p[SC_ADTAB];
    sc.hl = (uint32_t *)p[SC_HL];
    sc.ex = (short *)p[SC_EX];
    sc.q = {(float2 *)p[SC_QT], (float2 *)p[SC_QZ], (int *)p[SC_QREF],
            (int *)p[SC_QPIX], (int *)p[SC_QN], (int *)p[SC_QNEXT]};
    return sc;
}

/* The handle's own scratch, grown to the plan (buffers only grow). */
static StripScratch handle_scratch(struct ocl *o, const ScratchPlan &pl)
{
    DevBuf *bufs[SC_COUNT] = {&o->ampT, &o->lstat, &o->mask, &o->adtab, &o->hl, &o->ex,
                              &o->qt,   &o->qz,    &o->qref, &o->qpix,  &o->qnext, &o->qn};
    void *p[SC_COUNT];

    for (int k = 0; k < SC_COUNT; k++) {
        p[k] = nullptr;
        if (pl.bytes[k]) {
            bufs[k]->ensure(pl.bytes[k], "hipMalloc(strip scratch)");
            p[k] = bufs[k]->p;
        }
    }
    return scratch_pointers(p);
}

/* One strip's launches on st into the given scratch. Nothing here
   allocates, copies from the host or waits; timing (events around the
   stages) only when asked. */
static void enqueue_kernels(struct ocl *o, const KParams &kp, const ScratchPlan &pl,
                            const StripScratch &sc, const void *slc, const void *amp,
                            const void *valid, void *coh, void *count, void *pairs,
                            hipStream_t st, bool timing)
{
    StageTimer &tm = o->timer;
    auto mark = [&](int stage, bool start) {
        if (timing)
            tm.mark(stage, st, start);
    };
    const int rows = kp.rows, cols = kp.cols;

    if (timing)
        tm.begin();
    mark(0, true);
    if (kp.test == 1)
        check(launch_adtab(kp.n, sc.adtab, st), "AD table kernel");
    check(launch_prep(kp, (const float *)amp, sc.ampT, (const unsigned char *)valid,
                      sc.lstat, st),
          "prep kernel");
    mark(0, false);
    mark(1, true);
    check(launch_shp(kp, sc.ampT, (const unsigned char *)valid, sc.lstat, sc.adtab,
                     sc.mask, (int *)count, st),
          "shp kernel");
    mark(1, false);
    if (pl.mixed) {
        /* counted with the preparation stage */
        mark(0, true);
        check(launch_split(kp, (const float2 *)slc, sc.hl, pl.bf16 ? nullptr : sc.ex,
                           pl.bf16, st),
              "split kernel");
        mark(0, false);
    }

    /* Fused covariance + linking, in launches of about a million output
       pixels (each well under the GPU watchdog, ocl.c:22-23). EMI defers
       the pixels whose |T| fails the Cholesky test to queue 1 and links
       them in more launches (Gauss-Jordan, the Jacobi for what that
       leaves, the FP64 refinement); those kernels read the queue counts on
       the device, so nothing waits on the host. */
    for (int r0 = 0; r0 < rows; r0 += pl.chunk) {
        const int nr = std::min(pl.chunk, rows - r0);

        mark(2, true);
        if (kp.emi)
            check(launch_zero_words((uint32_t *)sc.q.n, 3, st), "queue count kernel");
        check(launch_link(kp, (const float2 *)slc, (const uint2 *)sc.hl, sc.ex, sc.mask,
                          (const int *)count, (float *)coh, (float *)pairs, r0, nr,
                          o->precision, sc.q, st),
              "link kernel");
        if (kp.emi)
            check(launch_emi_queue(kp, sc.q, r0 * cols, nr * cols, (float *)coh, st),
                  "EMI queue kernels");
        mark(2, false);
    }
    if (timing) {
        for (int i = 0; i < 4; i++) {
            o->stage_ms[i] = 0.0;
            o->stage_launches[i] = 0;
        }
        check(hipEventSynchronize(tm.ev[tm.used - 1]), "hipEventSynchronize");
        for (size_t k = 0; k < tm.stage.size(); k++) {
            float ms;

            check(hipEventElapsedTime(&ms, tm.ev[2 * k], tm.ev[2 * k + 1]),
                  "hipEventElapsedTime");
            o->stage_ms[tm.stage[k]] += ms;
            o->stage_launches[tm.stage[k]]++;
        }
    }
}

/* One strip's launches on st, in the handle's scratch (allocated here
   when it has to grow). With capture set nothing is timed: the stream is
   being captured into a graph (the eager run of the same call has
   allocated everything), or the strip is one of a batch. */
static void enqueue_strip(struct ocl *o, const void *slc, const void *amp,
                          const void *valid, int padded_rows, int rows, int cols,
                          void *coh, void *count, void *pairs, void *phase, hipStream_t st,
                          bool capture)
{
    if (rows <= 0 || cols <= 0)
        return;
    const KParams kp = strip_params(o, padded_rows, rows, cols, pairs, phase);

    check(hipSetDevice(o->device), "hipSetDevice");
    const ScratchPlan pl = scratch_plan(kp, o->precision, o->tiles[2]);
    const StripScratch sc = handle_scratch(o, pl);

    enqueue_kernels(o, kp, pl, sc, slc, amp, valid, coh, count, pairs, st,
                    o->timing != 0 && !capture);
}

/* A strip on the device, eagerly or (graph mode, timing off) as a HIP
   graph: the first call with a new key runs eagerly, then captures the
   same launches into a graph that later calls with that key replay with
   one hipGraphLaunch. Every launch of a strip is a kernel (the queue
   counters are cleared by one), and none waits on the host. */
static void run_device(struct ocl *o, const void *slc, const void *amp,
                       const void *valid, int padded_rows, int rows, int cols,
                       void *coh, void *count, void *pairs, void *phase, hipStream_t st)
{
    GraphCache &gc = *o->gcache;
    const GraphKey key = {slc,  amp,  valid,        coh, count,
                          pairs, phase, padded_rows, rows, cols,
                          o->precision, st, {o->tiles[0], o->tiles[1], o->tiles[2]}};

    if (!o->graph || o->timing) {
        enqueue_strip(o, slc, amp, valid, padded_rows, rows, cols, coh, count, pairs, phase,
                      st, false);
        return;
    }
    if (gc.valid && gc.key == key) {
        check(hipSetDevice(o->device), "hipSetDevice");
        check(hipGraphLaunch(gc.exec, st), "hipGraphLaunch");
        return;
    }
    enqueue_strip(o, slc, amp, valid, padded_rows, rows, cols, coh, count, pairs, phase, st,
                  false);
    gc.release();
    check(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed),
          "hipStreamBeginCapture");
    try {
        enqueue_strip(o, slc, amp, valid, padded_rows, rows, cols, coh, count, pairs, phase,
                      st, true);
    }
    catch (...) {
        hipGraph_t g = nullptr;

        (void)hipStreamEndCapture(st, &g);
        if (g)
            (void)hipGraphDestroy(g);
        throw;
    }
    check(hipStreamEndCapture(st, &gc.graph), "hipStreamEndCapture");
    check(hipGraphInstantiate(&gc.exec, gc.graph, nullptr, nullptr, 0),
          "hipGraphInstantiate");
    gc.key = key;
    gc.valid = true;
}

extern "C" void tcoh_set_graph(struct ocl *o, int on)
{
    guarded_void([&] {
        o->graph = on != 0;
        if (!o->graph)
            o->gcache->release();
    });
}

extern "C" void tcoh_run_device(struct ocl *o, const void *slc,
                                const void *amp, const void *valid,
                                int padded_rows, int rows, int cols, void *coh,
                                void *count, void *pairs, void *stream)
{
    guarded_void([&] {
        run_device(o, slc, amp, valid, padded_rows, rows, cols, coh, count,
                   pairs, nullptr, stream ? (hipStream_t)stream : o->stream);
    });
}

extern "C" void tcoh_run_device_phases(struct ocl *o, const void *slc,
                                       const void *amp, const void *valid,
                                       int padded_rows, int rows, int cols,
                                       void *coh, void *count, void *pairs,
                                       void *phases, void *stream)
{
    guarded_void([&] {
        if (!phases)
            fail("Linked phases requested without an output buffer");
        run_device(o, slc, amp, valid, padded_rows, rows, cols, coh, count,
                   pairs, phases, stream ? (hipStream_t)stream : o->stream);
    });
}

extern "C" void tcoh_run_device_noalloc(struct ocl *o, const void *slc,
                                        const void *amp, const void *valid,
                                        int padded_rows, int rows, int cols,
                                        void *coh, void *count, void *pairs,
                                        void *stream)
{
    guarded_void([&] {
        if (rows <= 0 || cols <= 0)
            return;
        const KParams kp = strip_params(o, padded_rows, rows, cols, pairs, nullptr);
        const ScratchPlan pl = scratch_plan(kp, o->precision, o->tiles[2]);
        const size_t need = pl.total(TCOH_NOALLOC_ALIGN);

        if (need > TCOH_NOALLOC_BYTES)
            fail("Strip of %d x %d output pixels needs %zu bytes of scratch; the "
                 "allocation-free entry point has %zu (smaller strips, or "
                 "tcoh_run_device)", rows, cols, need, TCOH_NOALLOC_BYTES);
        check(hipSetDevice(o->device), "hipSetDevice");
        unsigned char *base = (unsigned char *)o->noalloc;
        void *p[SC_COUNT];
        size_t at = 0;

        for (int k = 0; k < SC_COUNT; k++) {
            p[k] = pl.bytes[k] ? base + at : nullptr;
            at += (pl.bytes[k] + TCOH_NOALLOC_ALIGN - 1) / TCOH_NOALLOC_ALIGN *
                  TCOH_NOALLOC_ALIGN;
        }
        enqueue_kernels(o, kp, pl, scratch_pointers(p), slc, amp, valid, coh, count,
                        pairs, stream ? (hipStream_t)stream : o->stream, false);
    });
}

/* Batches: the strips of several handles (stacks of any depth, settings
   and form) on one device, issued as one unit of work on a caller's
   stream. The strips run concurrently on the batch's own streams, which
   fork from and join back into the caller's stream through events, in a
   longest-first list schedule of their estimated work. In graph mode the
   whole batch (every stage of every strip, the fork and the join) is
   captured once and replayed as one graph. */
struct tcoh_batch {
    int device = -1;
    int nstreams = 0;
    std::vector<hipStream_t> streams;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    std::vector<int> order, lane; /* the last schedule */
    bool graph = false;
    /* the captured batch and what it depends on */
    bool valid = false;
    std::vector<tcoh_batch_item> key;
    std::vector<int> key_state; /* per item: precision and launch shapes */
    hipStream_t key_stream = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t exec = nullptr;

    void release_graph()
    {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (g)
            (void)hipGraphDestroy(g);
        exec = nullptr;
        g = nullptr;
        valid = false;
    }
};

extern "C" double tcoh_strip_cost(const struct settings *s, int rows, int cols)
{
    const double n = s->ndates, w = (double)s->win_az * s->win_rg;

    return (double)rows * cols * (w * n * n + n * n * n);
}

extern "C" void tcoh_batch_schedule(int n_items, const double *cost, int streams,
                                    int *order, int *lane)
{
    std::vector<int> idx(n_items > 0 ? n_items : 0);
    std::vector<double> load(streams > 0 ? streams : 1, 0.0);

    for (int i = 0; i < n_items; i++)
        idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int a, int b) { return cost[a] > cost[b]; });
    for (int k = 0; k < n_items; k++) {
        const int i = idx[k];
        const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());

        order[k] = i;
        lane[i] = l;
        load[l] += cost[i];
    }
}

extern "C" struct tcoh_batch *tcoh_batch_create(int streams)
{
    return guarded(
        [&]() -> struct tcoh_batch * {
            if (streams < 0 || streams > 64)
                fail("A batch takes 1 to 64 streams, not %d", streams);
            struct tcoh_batch *b = new struct tcoh_batch();

            b->nstreams = streams ? streams : 4;
            return b;
        },
        nullptr);
}

extern "C" void tcoh_batch_set_graph(struct tcoh_batch *b, int on)
{
    guarded_void([&] {
        b->graph = on != 0;
        if (!b->graph)
            b->release_graph();
    });
}

extern "C" void tcoh_batch_plan(const struct tcoh_batch *b, int *order, int *lane)
{
    for (size_t k = 0; k < b->order.size(); k++) {
        order[k] = b->order[k];
        lane[k] = b->lane[k];
    }
}

/* Fork, every strip on its lane's stream in schedule order, join. */
static void batch_enqueue(struct tcoh_batch *b, const struct tcoh_batch_item *it,
                          int n, hipStream_t st)
{
    std::vector<char> used(b->nstreams, 0);

    for (int i = 0; i < n; i++)
        used[b->lane[i]] = 1;
    check(hipEventRecord(b->fork, st), "hipEventRecord(batch fork)");
    for (int l = 0; l < b->nstreams; l++)
        if (used[l])
            check(hipStreamWaitEvent(b->streams[l], b->fork, 0),
                  "hipStreamWaitEvent(batch fork)");
    for (int k = 0; k < n; k++) {
        const struct tcoh_batch_item &x = it[b->order[k]];

        enqueue_strip(x.o, x.slc, x.amp, x.valid, x.padded_rows, x.rows, x.cols, x.coh,
                      x.count, x.pairs, nullptr, b->streams[b->lane[b->order[k]]], true);
    }
    for (int l = 0; l < b->nstreams; l++)
        if (used[l]) {
            check(hipEventRecord(b->join[l], b->streams[l]), "hipEventRecord(batch join)");
            check(hipStreamWaitEvent(st, b->join[l], 0), "hipStreamWaitEvent(batch join)");
        }
}

extern "C" void tcoh_batch_run(struct tcoh_batch *b, const struct tcoh_batch_item *items,
                               int n_items, void *stream)
{
    guarded_void([&] {
        if (n_items <= 0)
            return;
        if (!items)
            fail("Batch of %d strips without items", n_items);
        const int device = items[0].o ? items[0].o->device : -1;
        std::vector<double> cost(n_items);
        std::vector<int> state;

        for (int i = 0; i < n_items; i++) {
            const struct ocl *o = items[i].o;

            if (!o)
                fail("Batch item %d has no handle", i);
            if (o->device != device)
                fail("Batch item %d is on device %d, item 0 on device %d", i, o->device,
                     device);
            for (int j = 0; j < i; j++)
                if (items[j].o == o)
                    fail("Batch items %d and %d share a handle (each strip needs its "
                         "own handle's scratch)", j, i);
            cost[i] = tcoh_strip_cost(&o->s, items[i].rows, items[i].cols);
            state.insert(state.end(), {o->precision, o->tiles[0], o->tiles[1], o->tiles[2]});
        }
        check(hipSetDevice(device), "hipSetDevice");
        if (b->device < 0) {
            b->device = device;
            b->streams.resize(b->nstreams);
            b->join.resize(b->nstreams);
            check(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming),
                  "hipEventCreate(batch)");
            for (int l = 0; l < b->nstreams; l++) {
                check(hipStreamCreateWithFlags(&b->streams[l], hipStreamNonBlocking),
                      "hipStreamCreate(batch)");
                check(hipEventCreateWithFlags(&b->join[l], hipEventDisableTiming),
                      "hipEventCreate(batch)");
            }
        }
        else if (b->device != device)
            fail("Batch created on device %d given strips on device %d", b->device, device);
        b->order.assign(n_items, 0);
        b->lane.assign(n_items, 0);
        tcoh_batch_schedule(n_items, cost.data(), b->nstreams, b->order.data(),
                            b->lane.data());
        hipStream_t st = stream ? (hipStream_t)stream : items[0].o->stream;

        if (!b->graph) {
            batch_enqueue(b, items, n_items, st);
            return;
        }
        if (b->valid && b->key_stream == st && (int)b->key.size() == n_items &&
            !memcmp(b->key.data(), items, n_items * sizeof(*items)) && b->key_state == state) {
            check(hipGraphLaunch(b->exec, st), "hipGraphLaunch(batch)");
            return;
        }
        /* eagerly first (the handles' scratch grows here), then captured */
        batch_enqueue(b, items, n_items, st);
        b->release_graph();
        check(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed),
              "hipStreamBeginCapture(batch)");
        try {
            batch_enqueue(b, items, n_items, st);
        }
        catch (...) {
            hipGraph_t g = nullptr;

            (void)hipStreamEndCapture(st, &g);
            if (g)
                (void)hipGraphDestroy(g);
            throw;
        }
        check(hipStreamEndCapture(st, &b->g), "hipStreamEndCapture(batch)");
        check(hipGraphInstantiate(&b->exec, b->g, nullptr, nullptr, 0),
              "hipGraphInstantiate(batch)");
        b->key.assign(items, items + n_items);
        b->key_state = state;
        b->key_stream = st;
        b->valid = true;
    });
}

extern "C" void tcoh_batch_destroy(struct tcoh_batch *b)
{
    if (!b)
        return;
    guarded_void([&] {
        if (b->device >= 0) {
            check(hipSetDevice(b->device), "hipSetDevice");
            for (hipStream_t s : b->streams)
                (void)hipStreamSynchronize(s);
        }
        b->release_graph();
        for (hipStream_t s : b->streams)
            (void)hipStreamDestroy(s);
        for (hipEvent_t e : b->join)
            (void)hipEventDestroy(e);
        if (b->fork)
            (void)hipEventDestroy(b->fork);
    });
    delete b;
}

/* ocl_run's transfers around run_device, with the optional phases. */
static void host_run(struct ocl *o, const float *slc, const float *amp,
                     const unsigned char *valid, int padded_rows, int rows,
                     int cols, float *coh, int *count, float *pairs, float *phases)
{
    {
        const size_t npad = (size_t)padded_rows * o->padded_cols;
        const size_t nout = (size_t)rows * cols;
        const size_t n = o->s.ndates;
        hipStream_t st = o->stream;

        check(hipSetDevice(o->device), "hipSetDevice");
        o->slc.ensure(npad * n * 2 * sizeof(float), "hipMalloc(slc)");
        o->amp.ensure(npad * n * sizeof(float), "hipMalloc(amp)");
        o->valid.ensure(npad, "hipMalloc(valid)");
        o->coh.ensure(nout * sizeof(float), "hipMalloc(coh)");
        o->count.ensure(nout * sizeof(int), "hipMalloc(count)");
        o->pairs.ensure((o->s.npairs ? o->s.npairs * nout : 1) * sizeof(float),
                        "hipMalloc(pairs)");
        check(hipMemcpyAsync(o->slc.p, slc, npad * n * 2 * sizeof(float),
                             hipMemcpyHostToDevice, st),
              "hipMemcpy(slc)");
        check(hipMemcpyAsync(o->amp.p, amp, npad * n * sizeof(float),
                             hipMemcpyHostToDevice, st),
              "hipMemcpy(amp)");
        check(hipMemcpyAsync(o->valid.p, valid, npad, hipMemcpyHostToDevice,
                             st),
              "hipMemcpy(valid)");
        if (phases)
            o->phase.ensure(n * nout * sizeof(float), "hipMalloc(phases)");
        run_device(o, o->slc.p, o->amp.p, o->valid.p, padded_rows, rows, cols,
                   o->coh.p, o->count.p, o->pairs.p, phases ? o->phase.p : nullptr, st);
        check(hipMemcpyAsync(coh, o->coh.p, nout * sizeof(float),
                             hipMemcpyDeviceToHost, st),
              "hipMemcpy(coh)");
        if (o->s.npairs)
            check(hipMemcpyAsync(pairs, o->pairs.p,
                                 o->s.npairs * nout * sizeof(float),
                                 hipMemcpyDeviceToHost, st),
                  "hipMemcpy(pairs)");
        check(hipMemcpyAsync(count, o->count.p, nout * sizeof(int),
                             hipMemcpyDeviceToHost, st),
              "hipMemcpy(count)");
        if (phases)
            check(hipMemcpyAsync(phases, o->phase.p, n * nout * sizeof(float),
                                 hipMemcpyDeviceToHost, st),
                  "hipMemcpy(phases)");
        check(hipStreamSynchronize(st), "hipStreamSynchronize");
    }
}

extern "C" void ocl_run(struct ocl *o, const float *slc, const float *amp,
                        const unsigned char *valid, int padded_rows, int rows,
                        int cols, float *coh, int *count, float *pairs)
{
    guarded_void([&] {
        host_run(o, slc, amp, valid, padded_rows, rows, cols, coh, count, pairs, nullptr);
    });
}

extern "C" void tcoh_run_phases(struct ocl *o, const float *slc, const float *amp,
                                const unsigned char *valid, int padded_rows, int rows,
                                int cols, float *coh, int *count, float *pairs,
                                float *phases)
{
    guarded_void([&] {
        if (!phases)
            fail("Linked phases requested without an output buffer");
        host_run(o, slc, amp, valid, padded_rows, rows, cols, coh, count, pairs, phases);
    });
}

/* The phase grid of g on the device, checked against the region
   (geometry_phases reads nodes gr and gr + 1, gc and gc + 1 of every region
   cell, geometry.c:264-284). */
static const double *upload_grid(struct ocl *o, const struct geometry *g,
                                 int nrows, int ncols, hipStream_t st)
{
    const int n = o->s.ndates;

    if (g->ndates != n)
        fail("Phase grid of %d dates for a stack of %d", g->ndates, n);
    if (!g->phase || g->step < 1 || (g->nheights != 1 && g->nheights != 3))
        fail("Invalid phase grid");
    if ((nrows - 1) / g->step + 1 >= g->grid_rows ||
        (ncols - 1) / g->step + 1 >= g->grid_cols)
        fail("Phase grid of %d x %d nodes does not cover a %d x %d region",
             g->grid_rows, g->grid_cols, nrows, ncols);
    const size_t bytes =
        (size_t)g->grid_rows * g->grid_cols * g->nheights * n * sizeof(double);

    o->grid.ensure(bytes, "hipMalloc(phase grid)");
    check(hipMemcpyAsync(o->grid.p, g->phase, bytes, hipMemcpyHostToDevice, st),
          "hipMemcpy(phase grid)");
    return o->grid.as<double>();
}

/* Arguments of k_strip for planes of plane_rows rows holding region rows
   plane_row0 .. plane_row0 + plane_rows - 1; row0 is the strip's first
   padded row in plane rows. */
static StripArgs strip_args(const struct ocl *o, int plane_rows, int ncols,
                            const void *i, const void *q,
                            const struct geometry *g, const double *dgrid,
                            const void *heights, int row0, int padded_rows,
                            int halo_cols, int plane_row0, void *slc, void *amp,
                            void *valid)
{
    StripArgs sa;

    sa.i = (const float *)i;
    sa.q = (const float *)q;
    sa.heights = (const double *)heights;
    sa.grid = g ? dgrid : nullptr;
    sa.slc = (float2 *)slc;
    sa.amp = (float *)amp;
    sa.valid = (unsigned char *)valid;
    sa.n = o->s.ndates;
    sa.nrows = plane_rows;
    sa.ncols = ncols;
    sa.row0 = row0;
    sa.prow = padded_rows;
    sa.halo = halo_cols;
    sa.pc = ncols + 2 * halo_cols;
    sa.plane_row0 = plane_row0;
    sa.grid_cols = g ? g->grid_cols : 0;
    sa.step = g ? g->step : 1;
    sa.nh = g ? g->nheights : 1;
    for (int k = 0; k < 3; k++)
        sa.hx[k] = g ? g->heights[k] : 0.0;
    sa.h0 = g ? g->h0 : 0.0;
    return sa;
}

extern "C" void tcoh_prepare_strip_device(struct ocl *o, int nrows, int ncols,
                                          const void *i, const void *q,
                                          const struct geometry *g,
                                          const void *heights, int row0,
                                          int padded_rows, int halo_cols,
                                          void *slc, void *amp, void *valid,
                                          void *stream)
{
    guarded_void([&] {
        const int pc = ncols + 2 * halo_cols;
        hipStream_t st = stream ? (hipStream_t)stream : o->stream;

        if (nrows <= 0 || ncols <= 0 || padded_rows < 0 || halo_cols < 0)
            fail("Invalid strip: %d x %d region, %d padded rows, %d halo "
                 "columns", nrows, ncols, padded_rows, halo_cols);
        if (pc != o->padded_cols)
            fail("Strip of %d columns and halo %d does not match the padded "
                 "width %d", ncols, halo_cols, o->padded_cols);
        if ((size_t)padded_rows * pc > (size_t)INT32_MAX)
            fail("Strip of %d x %d pixels is too large", padded_rows, pc);
        if (!i || !q || !slc || !amp || !valid)
            fail("Strip preparation without date planes or output buffers");
        check(hipSetDevice(o->device), "hipSetDevice");
        const double *dgrid = g ? upload_grid(o, g, nrows, ncols, st) : nullptr;
        const StripArgs sa = strip_args(o, nrows, ncols, i, q, g, dgrid, heights,
                                        row0, padded_rows, halo_cols, 0, slc,
                                        amp, valid);

        check(launch_strip(sa, st), "strip kernel");
        if (g)
            /* the grid buffer is reused by the next call */
            check(hipStreamSynchronize(st), "hipStreamSynchronize");
    });
}

/* Parts 0 .. n_parts-1 of a host copy on up to 16 threads (at most the
   hardware's): the pinned staging copies of tcoh_run_region run while the
   GPU computes, and one thread alone would not keep up with the device. */
template <class F> static void host_parallel(int n_parts, F part)
{
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int nt = std::min({16, hw, n_parts});

    if (nt <= 1) {
        for (int k = 0; k < n_parts; k++)
            part(k);
        return;
    }
    std::vector<std::thread> th;

    for (int t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            for (int k = t; k < n_parts; k += nt)
                part(k);
        });
    for (auto &x : th)
        x.join();
}

static void ensure_pinned(void **p, size_t *size, size_t bytes)
{
    if (*p && *size >= bytes)
        return;
    if (*p)
        check(hipHostFree(*p), "hipHostFree");
    *p = nullptr;
    *size = 0;
    check(hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault),
          "hipHostMalloc");
    *size = bytes;
}

extern "C" void tcoh_run_region(struct ocl *o, int nrows, int ncols,
                                const float *i, const float *q,
                                const struct geometry *g, const double *heights,
                                float *coh, int *count, float *pairs,
                                int strip_rows)
{
    guarded_void([&] {
        const struct settings &s = o->s;
        const int n = s.ndates, ha = s.win_az / 2, hr = s.win_rg / 2;
        const int np = s.npairs;
        hipStream_t cs = o->stream;

        if (nrows <= 0 || ncols <= 0)
            fail("Invalid region of %d x %d pixels", nrows, ncols);
        if (ncols + 2 * hr != o->padded_cols)
            fail("Region of %d columns does not match the padded width %d",
                 ncols, o->padded_cols);
        if (!i || !q || !coh || !count || (np && !pairs))
            fail("Region run without date planes or output buffers");
        check(hipSetDevice(o->device), "hipSetDevice");
        int R = strip_rows > 0 ? strip_rows : std::max(16, (1 << 20) / ncols);

        R = std::min(R, nrows);
        if (!o->region_ready) {
            check(hipStreamCreateWithFlags(&o->copy, hipStreamNonBlocking),
                  "hipStreamCreate");
            for (int b = 0; b < 2; b++) {
                hipEvent_t *evs[] = {&o->ev_h2d[b], &o->ev_prep[b], &o->ev_comp[b],
                                     &o->ev_d2h[b]};
                for (hipEvent_t *e : evs) {
                    check(hipEventCreateWithFlags(e, hipEventDisableTiming),
                          "hipEventCreate");
                    check(hipEventRecord(*e, cs), "hipEventRecord");
                }
            }
            o->region_ready = true;
        }
        const int hmax = R + 2 * ha;            /* staged plane rows */
        const size_t plane_in = (size_t)hmax * ncols;
        const size_t in_bytes = 2 * (size_t)n * plane_in * sizeof(float);
        const size_t hgt_bytes = heights ? plane_in * sizeof(double) : 0;
        const size_t out_px = (size_t)R * ncols;
        const size_t out_bytes = (2 + (size_t)np) * out_px * sizeof(float);
        const size_t pr_max = (size_t)R + s.win_az - 1;
        const size_t npad = pr_max * o->padded_cols;
        const double *dgrid = g ? upload_grid(o, g, nrows, ncols, cs) : nullptr;

        for (int b = 0; b < 2; b++) {
            /* buffers of an earlier call may still be in use */
            check(hipEventSynchronize(o->ev_d2h[b]), "hipEventSynchronize");
            check(hipEventSynchronize(o->ev_h2d[b]), "hipEventSynchronize");
            o->rin[b].ensure(in_bytes, "hipMalloc(region planes)");
            if (heights)
                o->rhgt[b].ensure(hgt_bytes, "hipMalloc(region heights)");
            o->rout[b].ensure(out_bytes, "hipMalloc(region outputs)");
            ensure_pinned(&o->hin[b], &o->hin_size[b], in_bytes + hgt_bytes);
            ensure_pinned(&o->hout[b], &o->hout_size[b], out_bytes);
        }
        check(hipStreamSynchronize(cs), "hipStreamSynchronize");
        o->rslc.ensure(npad * n * 2 * sizeof(float), "hipMalloc(strip slc)");
        o->ramp.ensure(npad * n * sizeof(float), "hipMalloc(strip amp)");
        o->rvalid.ensure(npad, "hipMalloc(strip valid)");

        const int nstrips = (nrows + R - 1) / R;
        const size_t plane = (size_t)nrows * ncols;

        /* outputs of strip t from pinned staging to the caller's arrays */
        auto finish = [&](int t) {
            const int b = t % 2, r0 = t * R, rs = std::min(R, nrows - r0);
            const size_t px = (size_t)rs * ncols;
            const float *h = (const float *)o->hout[b];

            check(hipEventSynchronize(o->ev_d2h[b]), "hipEventSynchronize");
            host_parallel(2 + np, [&](int k) {
                const float *src = h + (size_t)k * px;

                if (k == 0)
                    memcpy(coh + (size_t)r0 * ncols, src, px * sizeof(float));
                else if (k == 1)
                    memcpy(count + (size_t)r0 * ncols, src, px * sizeof(int));
                else
                    memcpy(pairs + (size_t)(k - 2) * plane + (size_t)r0 * ncols,
                           src, px * sizeof(float));
            });
        };

        for (int t = 0; t < nstrips; t++) {
            const int b = t % 2;
            const int r0 = t * R, rs = std::min(R, nrows - r0);
            const int a = std::max(0, r0 - ha), e = std::min(nrows, r0 + rs + ha);
            const int hrows = e - a, prow = rs + s.win_az - 1;
            const size_t pin = (size_t)hrows * ncols;
            float *hbuf = (float *)o->hin[b];

            /* 1. stage the strip's rows of every date (and heights) */
            check(hipEventSynchronize(o->ev_h2d[b]), "hipEventSynchronize");
            host_parallel(2 * n + (heights ? 1 : 0), [&](int k) {
                if (k == 2 * n) {
                    memcpy(hbuf + 2 * (size_t)n * pin, heights + (size_t)a * ncols,
                           pin * sizeof(double));
                    return;
                }
                const float *src = (k < n ? i : q) + (size_t)(k % n) * plane +
                                   (size_t)a * ncols;

                memcpy(hbuf + (size_t)k * pin, src, pin * sizeof(float));
            });
            /* 2. to the device once strip t - 2 has read the buffer */
            check(hipStreamWaitEvent(o->copy, o->ev_prep[b], 0), "hipStreamWaitEvent");
            check(hipMemcpyAsync(o->rin[b].p, hbuf, 2 * (size_t)n * pin * sizeof(float),
                                 hipMemcpyHostToDevice, o->copy),
                  "hipMemcpy(region planes)");
            if (heights)
                check(hipMemcpyAsync(o->rhgt[b].p, hbuf + 2 * (size_t)n * pin,
                                     pin * sizeof(double), hipMemcpyHostToDevice,
                                     o->copy),
                      "hipMemcpy(region heights)");
            check(hipEventRecord(o->ev_h2d[b], o->copy), "hipEventRecord");
            /* 3. prepare and estimate on the compute stream, once the planes
               are in and strip t - 2's outputs are out */
            check(hipStreamWaitEvent(cs, o->ev_h2d[b], 0), "hipStreamWaitEvent");
            check(hipStreamWaitEvent(cs, o->ev_d2h[b], 0), "hipStreamWaitEvent");
            const float *di = o->rin[b].as<float>();
            const StripArgs sa = strip_args(
                o, hrows, ncols, di, di + (size_t)n * pin, g, dgrid,
                heights ? o->rhgt[b].p : nullptr, (r0 - ha) - a, prow, hr, a,
                o->rslc.p, o->ramp.p, o->rvalid.p);

            check(launch_strip(sa, cs), "strip kernel");
            check(hipEventRecord(o->ev_prep[b], cs), "hipEventRecord");
            float *dout = o->rout[b].as<float>();
            const size_t px = (size_t)rs * ncols;

            enqueue_strip(o, o->rslc.p, o->ramp.p, o->rvalid.p, prow, rs, ncols, dout,
                          dout + px, np ? dout + 2 * px : nullptr, nullptr, cs, false);
            check(hipEventRecord(o->ev_comp[b], cs), "hipEventRecord");
            /* 4. outputs back on the copy stream */
            check(hipStreamWaitEvent(o->copy, o->ev_comp[b], 0), "hipStreamWaitEvent");
            check(hipMemcpyAsync(o->hout[b], dout, (2 + (size_t)np) * px * sizeof(float),
                                 hipMemcpyDeviceToHost, o->copy),
                  "hipMemcpy(region outputs)");
            check(hipEventRecord(o->ev_d2h[b], o->copy), "hipEventRecord");
            /* 5. the previous strip's outputs to the caller */
            if (t > 0)
                finish(t - 1);
        }
        finish(nstrips - 1);
        check(hipStreamSynchronize(o->copy), "hipStreamSynchronize");
        check(hipStreamSynchronize(cs), "hipStreamSynchronize");
    });
}

extern "C" void ocl_close(struct ocl *o)
{
    if (!o)
        return;
    guarded_void([&] {
        (void)hipSetDevice(o->device);
        (void)hipStreamSynchronize(o->stream);
        DevBuf *bufs[] = {&o->slc,  &o->amp,   &o->valid, &o->coh,
                          &o->count, &o->pairs, &o->phase, &o->ampT,  &o->lstat,
                          &o->mask, &o->adtab, &o->hl,
                          &o->ex, &o->qt, &o->qz, &o->qref, &o->qpix, &o->qn, &o->qnext,
                          &o->grid, &o->rin[0], &o->rin[1],
                          &o->rhgt[0], &o->rhgt[1], &o->rout[0], &o->rout[1],
                          &o->rslc, &o->ramp, &o->rvalid};
        if (o->copy)
            (void)hipStreamSynchronize(o->copy);
        for (DevBuf *b : bufs)
            b->release();
        for (int b = 0; b < 2; b++) {
            if (o->hin[b])
                (void)hipHostFree(o->hin[b]);
            if (o->hout[b])
                (void)hipHostFree(o->hout[b]);
        }
        if (o->region_ready) {
            for (int b = 0; b < 2; b++) {
                (void)hipEventDestroy(o->ev_h2d[b]);
                (void)hipEventDestroy(o->ev_prep[b]);
                (void)hipEventDestroy(o->ev_comp[b]);
                (void)hipEventDestroy(o->ev_d2h[b]);
            }
            (void)hipStreamDestroy(o->copy);
        }
        o->timer.release();
        o->gcache->release();
        delete o->gcache;
        (void)hipStreamDestroy(o->stream);
        delete o;
    });
}

/* Stack depths with a compiled eigen kernel: TCOH_EIG_FOR_EACH(X) expands
   to X(3) X(4) ... for the depths the build compiled. */
#define TCOH_EIG_CASE(NN)                                                    \
    case NN:                                                                 \
        return launch_link_n<NN>(kp, slc, hl, ex, mask, count, coh, pairs,   \
                                 row0, nrows, precision, q, s);

static hipError_t launch_link(const KParams &kp, const float2 *slc,
                              const uint2 *hl, const short *ex, const uint32_t *mask,
                              const int *count, float *coh, float *pairs,
                              int row0, int nrows, int precision,
                              const EmiQueue &q, hipStream_t s)
{
    switch (kp.n) {
        TCOH_EIG_FOR_EACH(TCOH_EIG_CASE)
    default:
        return hipErrorInvalidValue;
    }
}

#define TCOH_QUEUE_CASE(NN)                                                  \
    case NN:                                                                 \
        return launch_emi_queue_n<NN>(kp, q, pbase, cap, coh, s);

static hipError_t launch_emi_queue(const KParams &kp, const EmiQueue &q, int pbase,
                                   int cap, float *coh, hipStream_t s)
{
    switch (kp.n) {
        TCOH_EIG_FOR_EACH(TCOH_QUEUE_CASE)
    default:
        return hipErrorInvalidValue;
    }
}

#define TCOH_EIG_ITEM(NN) NN,
static bool eig_supported(int n)
{
    static const int list[] = {TCOH_EIG_FOR_EACH(TCOH_EIG_ITEM)};

    for (int v : list)
        if (v == n)
            return true;
    return false;
}
