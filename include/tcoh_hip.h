/* C ABI of the MI355X (gfx950) phase-linking backend.
 *
 * Drop-in for the OpenCL backend of i.sar.temporal_coherence: the types and
 * the four ocl_* entry points keep the layout, names and meaning of the
 * reference's local_proto.h, so that main.c links against libtcoh_hip.so
 * instead of ocl.c unchanged. Everything else below is an extension (marked
 * as such) that the reference has no counterpart for.
 *
 * Errors: like the reference (G_fatal_error, ocl.c:44-48), every failure
 * prints a message and ends the process, unless tcoh_set_error_mode(1) was
 * called, in which case the failing entry point returns and tcoh_last_error()
 * holds the message (used by bindings that must not exit).
 */
#ifndef TCOH_HIP_H
#define TCOH_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A C file that already has the reference's types (it includes
   local_proto.h) defines TCOH_HAVE_REFERENCE_TYPES before this header and
   takes only the entry points; the layouts are the same. */
#ifndef TCOH_HAVE_REFERENCE_TYPES

/* local_proto.h:9-10: largest supported stack depth. */
#define MAX_DATES 64

/* local_proto.h:12 */
enum shp_test { SHP_KS = 0, SHP_AD = 1, SHP_TLOG = 2 };

/* local_proto.h:33-49, field for field. */
struct settings {
    int ndates;
    int win_az, win_rg;
    enum shp_test test;
    double alpha;
    int emi;
    int bias;
    int min_looks;
    int ks_max;
    double ad_norm, ad_sigma, ad_crit;
    double tlog_crit;
    int npairs;    /* pair coherence maps to write, 0 for none */
    int pairs_all; /* all pairs i < j instead of consecutive ones */
    const char *phase_reference; /* none, ellipsoid or elevation */
    const char *orbit;           /* orbit source, NULL without phase */
};

/* local_proto.h:15-18, field for field. */
struct utc {
    int year, month, day, hour, minute;
    double second;
};

/* local_proto.h:52-57, field for field. */
struct orbit {
    int n;
    double *t;
    double (*pos)[3], (*vel)[3];
    char source[32];
};

/* local_proto.h:61-68, field for field. */
struct geometry {
    int ndates;
    int grid_rows, grid_cols, step;
    int nheights;
    double heights[3];
    double *phase; /* [node][height][date] */
    double h0;     /* height used without elevation map */
};

#endif /* TCOH_HAVE_REFERENCE_TYPES */

/* Visible at file scope either way, for the prototypes below. */
struct settings;
struct utc;
struct orbit;
struct geometry;

/* stats.c:61-95: KS_MAX, AD_NORM, AD_SIGMA, AD_CRIT, TLOG_CRIT from
   s->ndates, s->alpha and s->test. */
void shp_constants(struct settings *s);

/* orbit.c:324-346: position and velocity at t by 8-point Lagrange
   interpolation of the state vectors. */
void orbit_state(const struct orbit *o, double t, double *pos, double *vel);

/* ocl.c:105-136: print platform|device|type|name|version|compute_units|
   global_memory_mb, one line per HIP device. */
void ocl_list_devices(void);

/* ocl.c:191-254: select the device (platform -1 or 0 is the HIP runtime;
   device -1 is the first GPU), bind the settings (fixed for the life of the
   handle) and the padded strip width, and report the largest allocation the
   strip sizing of main.c:731-742 may use. */
struct ocl;
struct ocl *ocl_open(int platform, int device, const struct settings *s,
                     int padded_cols, size_t *max_alloc);

/* ocl.c:279-375: estimate one strip. slc is float[padded_rows][padded_cols]
   [N][2], amp float[..][N] sorted increasingly, valid u8[..]; outputs are
   row-major rows x cols: coh (NaN where not estimated), count (-1 for an
   invalid centre) and, when s->npairs > 0, pairs planar [npairs][rows][cols]
   (pairs may be NULL otherwise). Synchronous. */
void ocl_run(struct ocl *o, const float *slc, const float *amp,
             const unsigned char *valid, int padded_rows, int rows, int cols,
             float *coh, int *count, float *pairs);

/* ocl.c:377-391 */
void ocl_close(struct ocl *o);

/* ---- Flat-earth and topographic phase (geometry.c, FP64 on the GPU) ---- */

/* The numbers geometry_init reads from the reference date's metadata
   (geometry.c:166-197): first line time (s since 1970), azimuth time
   interval, slant range time of the first sample, range sampling rate,
   wavelength, and the line/sample of the first region cell; plus the region
   size. Extension: the reference parses them from description.json. */
struct tcoh_scene {
    double t_first, dt, tau0, fs, lambda;
    double line0, sample0;
    int rows, cols;
};

/* geometry.c:161-254 with the metadata lookups replaced by scene: the phase
   grid [node][height][date] on 16-pixel nodes, computed on the GPU in FP64
   (geolocation Newton on the reference orbit, zero-Doppler Newton per date).
   Orbits are host structs; n dates, ref the reference date. */
void tcoh_geometry_init(struct geometry *g, const struct tcoh_scene *scene,
                        int n, int ref, const struct orbit *orbits,
                        double hmin, double hmax, double h0);

/* geometry.c:256-291 */
void geometry_phases(const struct geometry *g, int row, int col, double h,
                     double *phase);

/* geometry.c:293-296 */
void geometry_free(struct geometry *g);

/* ---- Orbit ingestion (host; orbit.c, jsonget.c, epoch.c) ----
   The reference reads a date's metadata through struct epoch, a GRASS
   type; these take the date's description.json text (meta) and its name
   (for messages) instead. Returns: 0 or 1 as stated on success, -1 after a
   failure that error mode 1 turned into a return. */

/* epoch.c:28-32: seconds since 1970-01-01T00:00:00 UTC. */
double tcoh_utc_seconds(const struct utc *t);

/* epoch.c:34-52 */
void tcoh_utc_from_seconds(double sec, struct utc *t);

/* orbit.c:47-56: "YYYY-MM-DDThh:mm:ss[.ffffff]" to seconds; 1 when read,
   0 otherwise. */
int tcoh_parse_utc(const char *text, double *sec);

/* jsonget.c:171-190: value at a dotted path (a numeric component indexes an
   array) as a new string, to be released with tcoh_free; NULL if the path
   is absent, null, an object or an array. */
char *tcoh_json_get(const char *text, const char *path);

/* jsonget.c:192-205: the number at path, NAN if absent or not a number. */
double tcoh_json_get_number(const char *text, const char *path);

/* jsonget.c:207-225: elements of the array at path, -1 if not an array. */
int tcoh_json_array_length(const char *text, const char *path);

void tcoh_free(void *p);

/* orbit.c:59-71 with epoch.c:70-85's fallback: swath.start_time and
   stop_time, else the first and last line times. 0 on success. */
int tcoh_acquisition_span(const char *meta, const char *name, double *start,
                          double *stop);

/* orbit.c:81-111: swath.orbit_state_vectors of the metadata; fails unless
   8 or more vectors cover the acquisition. 0 on success; o is released
   with tcoh_orbit_free. */
int tcoh_orbit_from_annotation(const char *meta, const char *name,
                               struct orbit *o);

/* orbit.c:175-203: the <OSV> records of an EOF document with UTC time in
   [t0, t1]; source names the orbit in messages. 0 on success. */
int tcoh_orbit_from_eof(const char *text, double t0, double t1,
                        const char *source, struct orbit *o);

/* orbit.c:114-151: the .EOF file of dir whose name starts with prefix and
   whose validity _V<start>_<stop> covers [t0, t1], the greatest name when
   several do (the latest production). 1 when found (name in out). */
int tcoh_orbit_pick_file(const char *dir, const char *prefix, double t0,
                         double t1, char *out, size_t out_size);

/* orbit.c:205-301 for a local cache_dir: the <mission>_OPER_AUX_<kind>_OPOD_
   file (kind POEORB or RESORB) covering the acquisition with 120 s of
   margin, its state vectors within that margin. 1 when read, 0 when no file
   covers it and required is 0. The reference's download from the ESA
   mirror is not done (no network access is assumed). */
int tcoh_orbit_from_file(const char *meta, const char *name, const char *kind,
                         const char *cache_dir, int required, struct orbit *o);

/* orbit.c:304-318: precise, else restituted, else annotation. 0 on
   success. */
int tcoh_orbit_best(const char *meta, const char *name, const char *cache_dir,
                    struct orbit *o);

/* orbit.c:348-353 */
void tcoh_orbit_free(struct orbit *o);

/* ---- The module's data either side of the estimate (host; main.c) ----
   Raw files hold what GRASS stores for a cell, uncompressed, in native byte
   order: float32 with null as the all-ones word, int32 with null as
   INT32_MIN. Extensions: the reference reads and writes GRASS rasters. */

/* main.c:602-645: the names <prefix>_<a>_<b> of the pair maps, in the pair
   order of ocl_run's output, into names (npairs strings of stride bytes).
   a and b are the dates' YYYYMMDD, with Thhmmss when two consecutive dates
   share a day, or their 1-based ranks when any date has no time (starts
   NULL, or has_time[i] 0; has_time NULL means every date has one). *dated
   (if not NULL) tells which. Returns the number of pairs. */
int tcoh_pair_names(const char *prefix, int n, int pairs_all,
                    const struct utc *starts, const int *has_time,
                    char *names, size_t stride, int *dated);

/* The first bytes of a raw file mapped read-only (for instance the I or Q
   planes [N][rows][cols] float32 that tcoh_run_region reads, as the rows
   read_strip takes from <date>_i / <date>_q, main.c:70-72). 0 on success. */
int tcoh_raw_map(const char *path, size_t bytes, const void **addr);
void tcoh_raw_unmap(const void *addr, size_t bytes);

/* main.c:796-824: the output maps <dir>/<name>.raw (+ .hdr) of a region of
   rows x cols: coherence (output), SHP count (shp_count, or NULL for none)
   and npairs pair maps. put_strip appends the rows x cols outputs of one
   ocl_run (count < 0 and NaN become null); close writes the headers and
   gives main.c:921-922's summary: estimated pixels and their mean. */
struct tcoh_raw_out;
struct tcoh_raw_out *tcoh_raw_create(const char *dir, const char *output,
                                     const char *shp_count,
                                     const char *const *pair_names, int npairs,
                                     int rows, int cols);
int tcoh_raw_put_strip(struct tcoh_raw_out *w, int rows, const float *coh,
                       const int *count, const float *pairs);
int tcoh_raw_close(struct tcoh_raw_out *w, long *estimated, double *mean);

/* geometry.c:165-191: the scene numbers geometry_init reads from the
   reference date's metadata, and its check that every date has the same
   wavelength. line0/sample0 (the first region cell in the reference image)
   and rows/cols come from the caller's region. 0 on success. */
int tcoh_scene_from_metadata(const char *const *metas,
                             const char *const *names, int n, int ref,
                             double line0, double sample0, int rows, int cols,
                             struct tcoh_scene *sc);

/* ---- Extensions (no counterpart in the reference) ---- */

/* Arithmetic of the covariance stage. PLAIN: FP32 FMA on the vector ALUs,
   as the reference kernel. MIXED_F16 / MIXED_BF16: hi/lo split 16-bit
   operands, three products on the matrix cores with FP32 accumulation
   (FP16: each pixel under its own power-of-two scale, each window rescaled
   exactly to its largest candidate, so results depend on the window only).
   The EMI estimator ignores the setting: its covariance always sums in
   FP64 on the vector ALUs, since |T|^-1 o T amplifies the rounding of FP32
   sums beyond 1e-4 in gamma where |T| is near singular. */
enum tcoh_precision {
    TCOH_PLAIN = 0,
    TCOH_MIXED_F16 = 1,
    TCOH_MIXED_BF16 = 2
};

/* Select the covariance arithmetic of a handle (default: the environment
   variable TCOH_PRECISION = plain | f16 | bf16, else bf16, the split form
   with no scaling of the samples). */
void tcoh_set_precision(struct ocl *o, int precision);

/* main.c:52-127 (read_strip) without the GRASS row reads: build the padded
   strip from planar complex dates. i, q are float[n][nrows][ncols] (NaN is
   a null); rows row0 .. row0+padded_rows-1 of the region (outside: invalid),
   halo_cols invalid columns on each side. With phase (double
   [n][nrows][ncols], may be NULL) date k is multiplied by exp(i phase_k)
   after the amplitudes are taken; elev_null (u8[nrows][ncols], may be NULL)
   marks pixels without a height. Outputs as ocl_run's inputs. */
void tcoh_prepare_strip(int n, int nrows, int ncols, const float *i,
                        const float *q, const double *phase,
                        const unsigned char *elev_null, int row0,
                        int padded_rows, int halo_cols, float *slc,
                        float *amp, unsigned char *valid);

/* read_strip (main.c:52-127) on the device, for date planes already in
   device memory: the same strip as tcoh_prepare_strip (validity and
   amplitudes bit for bit; a rotated sample may differ by one float ulp,
   from the device's FP64 sin/cos) into device buffers slc, amp, valid with
   ocl_run's layouts, for the handle's stack depth and padded width
   (ncols + 2 halo_cols). i, q: float[n][nrows][ncols] (NaN is a null).
   g (host struct, may be NULL): the flat-earth and topographic phase of
   geometry_phases (geometry.c:256-291) is removed from every date, at the
   height heights[row][col] (double[nrows][ncols] in device memory, NaN a
   null that makes the pixel invalid) or, with heights NULL, at g->h0.
   Enqueued on stream (NULL: the handle's stream); returns without waiting,
   except after copying g's grid to the device. Extension: the reference
   prepares strips on the host. */
void tcoh_prepare_strip_device(struct ocl *o, int nrows, int ncols,
                               const void *i, const void *q,
                               const struct geometry *g, const void *heights,
                               int row0, int padded_rows, int halo_cols,
                               void *slc, void *amp, void *valid, void *stream);

/* The whole region from host date planes, streamed strip by strip: the
   strip loop of main.c:724-825 with ocl_run (ocl.c:289-374) pipelined. Each
   strip's rows (plus the window halo) of every date are staged in pinned
   memory and copied on a second stream while the previous strip computes;
   the strip is prepared on the device (tcoh_prepare_strip_device, so g and
   heights as there, heights on the host here), estimated, and its maps
   copied back while the next strip computes. i, q: float[n][nrows][ncols];
   outputs row-major nrows x ncols, pairs planar [npairs][nrows][ncols]
   (may be NULL without pairs). ncols + win_rg - 1 must be the handle's
   padded width. strip_rows: output rows per strip, 0 for about a million
   pixels. The maps equal a single-strip run's (strips are seamless).
   Synchronous. Extension: the reference streams nothing. */
void tcoh_run_region(struct ocl *o, int nrows, int ncols, const float *i,
                     const float *q, const struct geometry *g,
                     const double *heights, float *coh, int *count,
                     float *pairs, int strip_rows);

/* ocl_run on device-resident buffers (hipMalloc'd, same layouts), enqueued
   on stream (a hipStream_t; NULL is the handle's own stream). Returns
   without synchronising: no host wait anywhere in the strip, EMI included.
   Every call on a handle uses that handle's scratch buffers, so calls on
   one handle must run in order (one stream, or streams the caller orders
   with events); handles are independent. */
void tcoh_run_device(struct ocl *o, const void *slc, const void *amp,
                     const void *valid, int padded_rows, int rows, int cols,
                     void *coh, void *count, void *pairs, void *stream);

/* tcoh_run_device plus the linked phase of every date: phases (device,
   float[N][rows][cols], planar like pairs) receives arg(z_k) in radians of
   the unit phasors that gamma is fitted with (tcoh.cl:696-724: relative to
   date N/2 or the last strongest date), NaN where coh is NaN. Extension:
   the reference keeps the phasors inside the kernel. */
void tcoh_run_device_phases(struct ocl *o, const void *slc, const void *amp,
                            const void *valid, int padded_rows, int rows,
                            int cols, void *coh, void *count, void *pairs,
                            void *phases, void *stream);

/* ocl_run plus the linked phases (host buffer as above). Synchronous. */
void tcoh_run_phases(struct ocl *o, const float *slc, const float *amp,
                     const unsigned char *valid, int padded_rows, int rows,
                     int cols, float *coh, int *count, float *pairs,
                     float *phases);

/* tcoh_run_device without any allocation: for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving
   loops, harnesses that forbid hipMalloc). The strip's scratch (planar
   amplitudes, SHP masks, split samples, EMI queues, the AD table) lives in
   a fixed array of 256 MiB that every GPU the library runs on holds for the
   process's life; strips needing more are rejected with an error (about
   600 k output pixels at N=30 in the split forms). Calls on one GPU share
   that array: calls that overlap on one GPU race. Not timed. No
   counterpart in the reference. */
void tcoh_run_device_noalloc(struct ocl *o, const void *slc, const void *amp,
                             const void *valid, int padded_rows, int rows,
                             int cols, void *coh, void *count, void *pairs,
                             void *stream);

/* Graph mode (default off): tcoh_run_device and tcoh_run_device_phases
   (and ocl_run) capture a strip's launches into a HIP graph on the first
   call with given buffers, shape, stream and precision, and replay it with
   one hipGraphLaunch on later calls with the same ones (a call with others
   captures anew). Not while timing is on. The stream must be capturable
   (not the legacy null stream); calls on one handle must not overlap.
   Extension. */
void tcoh_set_graph(struct ocl *o, int on);

/* Launch shapes of a handle: tiles[0], tiles[1] the number of XCD stripes
   a row is cut into by the SHP and linking launches (0 = 8, one stripe per
   XCD; 1 = the whole row, every XCD a band of rows), tiles[2] the output
   pixels of one linking launch (0 = the default, about a million). ocl_open takes
   them from the tuned table compiled into the library (tiles.h, by stack
   depth, estimator and form); tcoh_set_tiles overrides it (NULL goes back
   to the table). Only the speed depends on them, not the maps. Extension:
   the tuner sweeps them through these calls. */
void tcoh_set_tiles(struct ocl *o, const int tiles[3]);
void tcoh_get_tiles(const struct ocl *o, int tiles[3]);

/* ---- Batches of stacks ----
   Several stacks (a handle each: own depth, settings and precision) on
   one device, run as one unit of work: each item is one strip with
   tcoh_run_device's buffers. tcoh_batch_run forks from stream (NULL: the
   first item's handle stream) onto the batch's own streams, enqueues every
   strip in a longest-first list schedule of tcoh_strip_cost over them, and
   joins back into stream; it returns without synchronising. Items must be
   on one device and have distinct handles; the maps equal tcoh_run_device's
   for each item. In graph mode (tcoh_batch_set_graph) the first call with
   given items and stream runs eagerly and captures the whole batch into
   one HIP graph, which later calls with the same items, stream, precisions
   and launch shapes replay with one hipGraphLaunch. Calls on one batch, and
   on the handles it runs, must not overlap. Extension: the reference runs
   one stack per module invocation (main.c:724-825). */
struct tcoh_batch_item {
    struct ocl *o;
    const void *slc, *amp, *valid;
    int padded_rows, rows, cols;
    void *coh, *count, *pairs;
};
struct tcoh_batch;

/* streams: concurrent lanes of the batch (0: 4, the HIP runtime's hardware
   queues per process). */
struct tcoh_batch *tcoh_batch_create(int streams);
void tcoh_batch_set_graph(struct tcoh_batch *b, int on);
void tcoh_batch_run(struct tcoh_batch *b, const struct tcoh_batch_item *items,
                    int n_items, void *stream);
/* The last run's schedule: order[k] the item issued k-th, lane[i] the
   stream of item i. */
void tcoh_batch_plan(const struct tcoh_batch *b, int *order, int *lane);
void tcoh_batch_destroy(struct tcoh_batch *b);

/* Estimated work of a strip for the schedule: output pixels x (W N^2 +
   N^3), W the window's candidates (the bound of the SHP count L that the
   covariance's L N^2 takes) and N^3 the eigen solve. */
double tcoh_strip_cost(const struct settings *s, int rows, int cols);
/* Longest processing time first: items by cost, largest first (ties in
   item order), each to the least loaded of `streams` lanes (ties to the
   lowest); order[k] the item issued k-th, lane[i] the lane of item i. */
void tcoh_batch_schedule(int n_items, const double *cost, int streams,
                         int *order, int *lane);

/* Device time of the last run's stages in milliseconds: [0] preparation
   (planar amplitudes, and the hi/lo split in the mixed forms), [1] SHP
   selection, [2] the fused covariance + linking kernel, [3] reserved (0);
   summed over the stage's kernel launches, measured with HIP events on the
   run's stream when timing is on; and the number of launches of each. */
void tcoh_set_timing(struct ocl *o, int on);
void tcoh_stage_times(const struct ocl *o, double ms[4]);
void tcoh_stage_launches(const struct ocl *o, int launches[4]);

/* 0 (default): fatal errors end the process. 1: they return. */
void tcoh_set_error_mode(int mode);
const char *tcoh_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
