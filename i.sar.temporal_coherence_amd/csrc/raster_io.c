/* The module's data either side of the estimate, without GRASS: pair-map
 * names (main.c:602-645), raw band-sequential date planes read in place
 * (the rows read_strip takes from <date>_i / <date>_q, main.c:70-72), and
 * the output maps written with the module's null conventions and summary
 * (main.c:796-824, 921-922).
 *
 * Raw files hold what GRASS stores for a cell, uncompressed: float32 maps
 * with null as the all-ones word (Rast_set_f_null_value), int32 maps with
 * null as INT32_MIN (Rast_set_c_null_value); native byte order. Each output
 * <name>.raw gets an ENVI-style <name>.hdr naming its size, type and null.
 */
#include <fcntl.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "../../include/tcoh_hip.h"

void tcoh_report_error(const char *msg);
void tcoh_clear_error(void);

static int fail(const char *what, const char *path)
{
    char buf[2400];

    snprintf(buf, sizeof(buf), "%s <%s>", what, path);
    tcoh_report_error(buf);
    return -1;
}

/* ---- pair-map names ---- */

static void tag_of(const struct utc *t, int with_time, char *buf, size_t len)
{
    if (with_time)
        snprintf(buf, len, "%04d%02d%02dT%02d%02d%02d", t->year, t->month,
                 t->day, t->hour, t->minute, (int)t->second);
    else
        snprintf(buf, len, "%04d%02d%02d", t->year, t->month, t->day);
}

int tcoh_pair_names(const char *prefix, int n, int pairs_all,
                    const struct utc *starts, const int *has_time,
                    char *names, size_t stride, int *dated_out)
{
    int dated = starts != NULL, with_time = 0, p = 0;

    tcoh_clear_error();
    /* every date must have a time; two consecutive dates on one day ask
       for the time of day in the tags */
    for (int i = 0; i < n && dated; i++) {
        dated = !has_time || has_time[i];
        if (dated && i > 0 && starts[i].year == starts[i - 1].year &&
            starts[i].month == starts[i - 1].month &&
            starts[i].day == starts[i - 1].day)
            with_time = 1;
    }
    for (int i = 0; i < n - 1; i++)
        for (int j = i + 1; j < (pairs_all ? n : i + 2); j++, p++) {
            char a[32], b[32];
            int len;

            if (dated) {
                tag_of(&starts[i], with_time, a, sizeof(a));
                tag_of(&starts[j], with_time, b, sizeof(b));
            }
            else {
                snprintf(a, sizeof(a), "%d", i + 1);
                snprintf(b, sizeof(b), "%d", j + 1);
            }
            len = snprintf(names + (size_t)p * stride, stride, "%s_%s_%s",
                           prefix, a, b);
            if (len < 0 || (size_t)len >= stride)
                return fail("Pair map name too long for", prefix);
        }
    if (dated_out)
        *dated_out = dated;
    return p;
}

/* ---- raw date planes, read in place ---- */

int tcoh_raw_map(const char *path, size_t bytes, const void **addr)
{
    struct stat st;
    void *m;
    int fd;

    tcoh_clear_error();
    *addr = NULL;
    if ((fd = open(path, O_RDONLY)) < 0)
        return fail("Unable to open", path);
    if (fstat(fd, &st) != 0 || (size_t)st.st_size < bytes) {
        close(fd);
        return fail("Raw file shorter than its declared size:", path);
    }
    m = bytes ? mmap(NULL, bytes, PROT_READ, MAP_SHARED, fd, 0) : NULL;
    close(fd);
    if (m == MAP_FAILED)
        return fail("Unable to map", path);
    if (m)
        madvise(m, bytes, MADV_SEQUENTIAL);
    *addr = m;
    return 0;
}

void tcoh_raw_unmap(const void *addr, size_t bytes)
{
    if (addr && bytes)
        munmap((void *)addr, bytes);
}

/* ---- output maps ---- */

struct tcoh_raw_out {
    int rows, cols, npairs, written;
    FILE *coh, *count, **pairs;
    char **paths; /* header stems: coh, count (or NULL), pairs */
    double sum;
    long estimated;
};

static FILE *open_out(const char *dir, const char *name, char **stem)
{
    const size_t len = strlen(dir) + strlen(name) + 2;
    char *path = malloc(len + 4);
    FILE *fp;

    if (!path) {
        fail("Out of memory for", name);
        return NULL;
    }
    snprintf(path, len + 4, "%s/%s.raw", dir, name);
    fp = fopen(path, "wb");
    if (!fp) {
        fail("Unable to create", path);
        free(path);
        return NULL;
    }
    path[len - 1] = '\0'; /* the stem, without ".raw" */
    *stem = path;
    return fp;
}

/* Close every file (0 when all closed cleanly: buffered data reached the
   file system) and free w. */
static int release(struct tcoh_raw_out *w)
{
    int bad = 0;

    if (w->coh)
        bad |= fclose(w->coh) != 0;
    if (w->count)
        bad |= fclose(w->count) != 0;
    for (int p = 0; p < w->npairs && w->pairs; p++)
        if (w->pairs[p])
            bad |= fclose(w->pairs[p]) != 0;
    for (int k = 0; k < w->npairs + 2 && w->paths; k++)
        free(w->paths[k]);
    free(w->pairs);
    free(w->paths);
    free(w);
    return bad;
}

struct tcoh_raw_out *tcoh_raw_create(const char *dir, const char *output,
                                     const char *shp_count,
                                     const char *const *pair_names, int npairs,
                                     int rows, int cols)
{
    struct tcoh_raw_out *w = calloc(1, sizeof(*w));

    tcoh_clear_error();
    if (!w) {
        fail("Out of memory for", output);
        return NULL;
    }
    w->rows = rows;
    w->cols = cols;
    w->npairs = npairs;
    w->pairs = calloc(npairs > 0 ? npairs : 1, sizeof(FILE *));
    w->paths = calloc(npairs + 2, sizeof(char *));
    if (!w->pairs || !w->paths) {
        fail("Out of memory for", output);
        release(w);
        return NULL;
    }
    if (!(w->coh = open_out(dir, output, &w->paths[0])) ||
        (shp_count && !(w->count = open_out(dir, shp_count, &w->paths[1])))) {
        release(w);
        return NULL;
    }
    for (int p = 0; p < npairs; p++)
        if (!(w->pairs[p] = open_out(dir, pair_names[p], &w->paths[2 + p]))) {
            release(w);
            return NULL;
        }
    return w;
}

/* One strip of rows x cols outputs of ocl_run, appended to the maps:
   NaN -> float null, count < 0 -> int null; the estimated pixels are
   summed in the module's order for the summary. */
int tcoh_raw_put_strip(struct tcoh_raw_out *w, int rows, const float *coh,
                       const int *count, const float *pairs)
{
    const uint32_t fnull = 0xffffffffu;
    const int32_t cnull = INT32_MIN;
    size_t plane;
    uint32_t *fb;
    int32_t *cb;
    int ok;

    tcoh_clear_error();
    if (!w || rows < 0)
        return fail("Invalid map handle or row count for", "tcoh_raw_put_strip");
    plane = (size_t)rows * w->cols;
    fb = malloc((size_t)w->cols * sizeof(uint32_t));
    cb = malloc((size_t)w->cols * sizeof(int32_t));
    ok = fb && cb;
    if (!ok)
        fail("Out of memory for a row of", w->paths[0]);
    else if (w->written + rows > w->rows)
        ok = 0, fail("More rows than the map holds:", w->paths[0]);
    for (int r = 0; r < rows && ok; r++) {
        const float *cr = coh + (size_t)r * w->cols;

        for (int c = 0; c < w->cols; c++) {
            if (isnan(cr[c]))
                fb[c] = fnull;
            else {
                memcpy(&fb[c], &cr[c], 4);
                w->sum += cr[c];
                w->estimated++;
            }
        }
        ok = fwrite(fb, 4, (size_t)w->cols, w->coh) == (size_t)w->cols;
        if (ok && w->count) {
            const int *kr = count + (size_t)r * w->cols;

            for (int c = 0; c < w->cols; c++)
                cb[c] = kr[c] < 0 ? cnull : kr[c];
            ok = fwrite(cb, 4, (size_t)w->cols, w->count) == (size_t)w->cols;
        }
        for (int p = 0; p < w->npairs && ok; p++) {
            const float *pr = pairs + p * plane + (size_t)r * w->cols;

            for (int c = 0; c < w->cols; c++) {
                if (isnan(pr[c]))
                    fb[c] = fnull;
                else
                    memcpy(&fb[c], &pr[c], 4);
            }
            ok = fwrite(fb, 4, (size_t)w->cols, w->pairs[p]) == (size_t)w->cols;
        }
        if (!ok)
            fail("Unable to write the maps of", w->paths[0]);
    }
    free(fb);
    free(cb);
    if (!ok)
        return -1;
    w->written += rows;
    return 0;
}

static int header(const char *stem, int rows, int cols, int type,
                  const char *null)
{
    char path[1200];
    FILE *fp;

    snprintf(path, sizeof(path), "%s.hdr", stem);
    if (!(fp = fopen(path, "w")))
        return fail("Unable to create", path);
    fprintf(fp,
            "ENVI\nsamples = %d\nlines = %d\nbands = 1\nheader offset = 0\n"
            "file type = ENVI Standard\ndata type = %d\ninterleave = bsq\n"
            "byte order = %d\ndata ignore value = %s\n",
            cols, rows, type, (int)(*(const unsigned char *)&(int){1} == 0),
            null);
    return fclose(fp) == 0 ? 0 : fail("Unable to write", path);
}

/* Close the maps and write their headers; the summary of main.c:921-922
   (estimated pixels and their mean gamma, NaN if none). */
int tcoh_raw_close(struct tcoh_raw_out *w, long *estimated, double *mean)
{
    int rc = 0;

    tcoh_clear_error();
    if (!w)
        return fail("Invalid map handle for", "tcoh_raw_close");
    if (w->written != w->rows)
        rc = fail("Fewer rows written than the map holds:", w->paths[0]);
    if (!rc)
        rc = header(w->paths[0], w->rows, w->cols, 4, "nan");
    if (!rc && w->count)
        rc = header(w->paths[1], w->rows, w->cols, 3, "-2147483648");
    for (int p = 0; p < w->npairs && !rc; p++)
        rc = header(w->paths[2 + p], w->rows, w->cols, 4, "nan");
    if (estimated)
        *estimated = w->estimated;
    if (mean)
        *mean = w->estimated ? w->sum / w->estimated : NAN;
    {
        char stem[1200];

        snprintf(stem, sizeof(stem), "%s", w->paths[0]);
        if (release(w) && !rc)
            rc = fail("Unable to write the maps of", stem);
    }
    return rc;
}
