/* Orbit ingestion on the host: the inputs of the geometry stage read from a
 * date's metadata (the description.json text that r.in.s1slc writes) and
 * from Sentinel-1 orbit files (POEORB / RESORB EOF) in a local directory.
 *
 *   UTC calendar      epoch.c:28-52   (days from the civil date, proleptic
 *                                      Gregorian, and back)
 *   JSON lookup       jsonget.c:99-225 (value at a dotted path, numeric
 *                                      components index arrays)
 *   state vectors     orbit.c:47-111   (annotation), 114-151 (file name
 *                                      validity, choice of file), 175-203
 *                                      (EOF parsing), 73-79 (coverage)
 *   orbit sources     orbit.c:205-318  (local files only: the download from
 *                                      the ESA mirror needs a network)
 *   scene numbers     geometry.c:150-191 (what geometry_init reads)
 *
 * Failures are reported like the rest of the library (tcoh_report_error:
 * message and exit, or, in error mode 1, a return value and
 * tcoh_last_error()).
 */
#include <dirent.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/tcoh_hip.h"

/* Defined with the library's error state (backend.hip). */
void tcoh_report_error(const char *msg);
void tcoh_clear_error(void);

/* Margin of state vectors kept around an acquisition, seconds
   (orbit.c:21), and the Lagrange order the interpolation needs
   (orbit.c:23). */
#define ORBIT_MARGIN 120.0
#define ORBIT_POINTS 8

static int report(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;

    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    tcoh_report_error(buf);
    return -1;
}

/* ---- UTC ---- */

/* Days from 1970-01-01 to y-m-d, proleptic Gregorian: March-based years
   split into 400-year eras of 146097 days (the civil-from-days method
   published by H. Hinnant). */
static long civil_to_days(long y, long m, long d)
{
    const long ym = m <= 2 ? y - 1 : y;
    const long era = (ym >= 0 ? ym : ym - 399) / 400;
    const long year_of_era = ym - era * 400;
    const long mar = m > 2 ? m - 3 : m + 9; /* months since March */
    const long day_of_year = (153 * mar + 2) / 5 + d - 1;
    const long day_of_era = year_of_era * 365 + year_of_era / 4 -
                            year_of_era / 100 + day_of_year;

    return era * 146097 + day_of_era - 719468;
}

double tcoh_utc_seconds(const struct utc *t)
{
    return (double)civil_to_days(t->year, t->month, t->day) * 86400.0 +
           t->hour * 3600.0 + t->minute * 60.0 + t->second;
}

void tcoh_utc_from_seconds(double sec, struct utc *t)
{
    long days = (long)floor(sec / 86400.0);
    const double rest = sec - days * 86400.0;
    long era, day_of_era, year_of_era, day_of_year, mp;

    days += 719468; /* from 0000-03-01 */
    era = (days >= 0 ? days : days - 146096) / 146097;
    day_of_era = days - era * 146097;
    year_of_era = (day_of_era - day_of_era / 1460 + day_of_era / 36524 -
                   day_of_era / 146096) / 365;
    day_of_year = day_of_era - (365 * year_of_era + year_of_era / 4 -
                                year_of_era / 100);
    mp = (5 * day_of_year + 2) / 153;
    t->day = (int)(day_of_year - (153 * mp + 2) / 5 + 1);
    t->month = (int)(mp < 10 ? mp + 3 : mp - 9);
    t->year = (int)(year_of_era + era * 400 + (t->month <= 2));
    t->hour = (int)(rest / 3600.0);
    t->minute = (int)((rest - t->hour * 3600.0) / 60.0);
    t->second = rest - t->hour * 3600.0 - t->minute * 60.0;
}

/* ISO time such as 2023-01-12T14:25:09.020000 (the fields the reference
   scans with "%d-%d-%dT%d:%d:%lf", orbit.c:47-56). */
int tcoh_parse_utc(const char *text, double *sec)
{
    struct utc t;

    if (!text || sscanf(text, "%d-%d-%dT%d:%d:%lf", &t.year, &t.month, &t.day,
                        &t.hour, &t.minute, &t.second) != 6)
        return 0;
    *sec = tcoh_utc_seconds(&t);
    return 1;
}

/* ---- JSON: value at a dotted path ---- */

static const char *ws(const char *p)
{
    while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')
        p++;
    return p;
}

/* End of the string whose opening quote is at p (after the closing quote),
   NULL if it does not end. */
static const char *string_end(const char *p)
{
    for (++p; *p && *p != '"'; p++)
        if (*p == '\\' && !*++p)
            return NULL;
    return *p ? p + 1 : NULL;
}

/* End of the value at p (leading blanks skipped), NULL if malformed. A
   scalar runs to the next delimiter; containers are matched by nesting. */
static const char *value_end(const char *p)
{
    int depth = 0;

    p = ws(p);
    if (*p != '{' && *p != '[' && *p != '"') {
        if (!*p || *p == '}' || *p == ']')
            return NULL;
        while (*p && !strchr(",}] \t\r\n", *p))
            p++;
        return p;
    }
    while (*p) {
        if (*p == '"') {
            if (!(p = string_end(p)))
                return NULL;
            if (depth == 0)
                return p;
            continue;
        }
        if (*p == '{' || *p == '[')
            depth++;
        else if ((*p == '}' || *p == ']') && --depth == 0)
            return p + 1;
        p++;
    }
    return NULL;
}

/* Start of the value at path in text, NULL if absent. */
static const char *json_find(const char *text, const char *path)
{
    const char *p = ws(text);

    while (*path) {
        const char *dot = strchr(path, '.');
        const size_t len = dot ? (size_t)(dot - path) : strlen(path);
        const char *comp = path;

        path += len + (dot != NULL);
        if (*p == '[') {
            char num[32], *tail;
            long index, i;

            /* an empty component reads as index 0, as strtol has it */
            if (len >= sizeof(num))
                return NULL;
            memcpy(num, comp, len);
            num[len] = '\0';
            index = strtol(num, &tail, 10);
            if (*tail || index < 0)
                return NULL;
            p = ws(p + 1);
            for (i = 0; i < index; i++) {
                if (*p == ']' || !(p = value_end(p)))
                    return NULL;
                p = ws(p);
                if (*p != ',')
                    return NULL;
                p = ws(p + 1);
            }
            if (*p == ']')
                return NULL;
            continue;
        }
        if (*p != '{')
            return NULL;
        for (p = ws(p + 1);; ) {
            const char *kend;
            int same;

            if (*p != '"' || !(kend = string_end(p)))
                return NULL;
            /* keys compare as written, escapes included */
            same = (size_t)(kend - p - 2) == len && !strncmp(p + 1, comp, len);
            p = ws(kend);
            if (*p != ':')
                return NULL;
            p = ws(p + 1);
            if (same)
                break;
            if (!(p = value_end(p)))
                return NULL;
            p = ws(p);
            if (*p == ',')
                p = ws(p + 1);
        }
    }
    return p;
}

/* The value at path as a newly allocated string (free with tcoh_free): the
   contents of a string with \n and \t decoded and any other escaped
   character kept as itself, or the text of a number, true or false. NULL
   for an absent path, null, an object or an array. */
char *tcoh_json_get(const char *text, const char *path)
{
    const char *p = text ? json_find(text, path) : NULL, *end;
    char *out, *o;

    if (!p || *p == '{' || *p == '[' || !strncmp(p, "null", 4))
        return NULL;
    if (*p == '"') {
        if (!(end = string_end(p)))
            return NULL;
        out = o = malloc((size_t)(end - p));
        if (!out)
            return NULL;
        for (p++; p < end - 1; p++) {
            char ch = *p;

            if (ch == '\\') {
                ch = *++p;
                ch = ch == 'n' ? '\n' : ch == 't' ? '\t' : ch;
            }
            *o++ = ch;
        }
        *o = '\0';
        return out;
    }
    end = value_end(p);
    if (!end || end == p)
        return NULL;
    out = malloc((size_t)(end - p) + 1);
    if (!out)
        return NULL;
    memcpy(out, p, (size_t)(end - p));
    out[end - p] = '\0';
    return out;
}

/* The number at path, NAN when absent or not wholly a number. */
double tcoh_json_get_number(const char *text, const char *path)
{
    char *s = tcoh_json_get(text, path), *tail;
    double x = NAN;

    if (s) {
        x = strtod(s, &tail);
        if (tail == s || *tail)
            x = NAN;
        free(s);
    }
    return x;
}

/* Elements of the array at path, -1 when it is not an array. */
int tcoh_json_array_length(const char *text, const char *path)
{
    const char *p = text ? json_find(text, path) : NULL;
    int n = 0;

    if (!p || *p != '[')
        return -1;
    for (p = ws(p + 1); *p && *p != ']'; n++) {
        if (!(p = value_end(p)))
            return -1;
        p = ws(p);
        if (*p == ',')
            p = ws(p + 1);
    }
    return n;
}

void tcoh_free(void *p) { free(p); }

/* ---- state vectors ---- */

static int orbit_reserve(struct orbit *o, int n, const char *source)
{
    const size_t m = n > 0 ? (size_t)n : 1;

    o->n = 0;
    o->t = malloc(m * sizeof(double));
    o->pos = malloc(m * sizeof(*o->pos));
    o->vel = malloc(m * sizeof(*o->vel));
    snprintf(o->source, sizeof(o->source), "%s", source);
    if (!o->t || !o->pos || !o->vel) {
        tcoh_orbit_free(o);
        return report("Out of memory for %d state vectors", n);
    }
    return 0;
}

static void orbit_push(struct orbit *o, double t, const double v[6])
{
    o->t[o->n] = t;
    for (int k = 0; k < 3; k++) {
        o->pos[o->n][k] = v[k];
        o->vel[o->n][k] = v[3 + k];
    }
    o->n++;
}

void tcoh_orbit_free(struct orbit *o)
{
    free(o->t);
    free(o->pos);
    free(o->vel);
    o->t = NULL;
    o->pos = NULL;
    o->vel = NULL;
    o->n = 0;
}

/* Acquisition start and stop: swath.start_time / stop_time, else the first
   and last line times (the epoch's own time span, epoch.c:70-85). */
static int span_of(const char *meta, const char *name, double *start,
                   double *stop)
{
    static const char *keys[2][2] = {
        {"swath.start_time", "swath.stop_time"},
        {"raster_geometry.first_line_time", "raster_geometry.last_line_time"}};
    int ok = 0;

    for (int s = 0; s < 2 && !ok; s++) {
        char *a = tcoh_json_get(meta, keys[s][0]);
        char *b = tcoh_json_get(meta, keys[s][1]);

        if (s == 0)
            ok = tcoh_parse_utc(a, start) && tcoh_parse_utc(b, stop);
        else if ((ok = tcoh_parse_utc(a, start)) && !tcoh_parse_utc(b, stop))
            *stop = *start; /* no last line time: a single instant */
        free(a);
        free(b);
    }
    return ok ? 0 : report("No acquisition time for <%s>", name);
}

int tcoh_acquisition_span(const char *meta, const char *name, double *start,
                          double *stop)
{
    tcoh_clear_error();
    return span_of(meta, name, start, stop);
}

static int orbit_covers(struct orbit *o, const char *name, double start,
                        double stop)
{
    if (o->n < ORBIT_POINTS || o->t[0] > start || o->t[o->n - 1] < stop) {
        tcoh_orbit_free(o);
        return report("The %s orbit of <%s> does not cover its acquisition "
                      "with %d state vectors",
                      o->source, name, ORBIT_POINTS);
    }
    return 0;
}

int tcoh_orbit_from_annotation(const char *meta, const char *name,
                               struct orbit *o)
{
    static const char *base = "swath.orbit_state_vectors";
    const int n = meta ? tcoh_json_array_length(meta, base) : -1;
    double start, stop;

    tcoh_clear_error();
    if (n <= 0)
        return report("No orbit state vectors in the metadata of <%s>", name);
    if (span_of(meta, name, &start, &stop) || orbit_reserve(o, n, "annotation"))
        return -1;
    for (int i = 0; i < n; i++) {
        char path[160], *time;
        double t = 0.0, v[6];
        int ok;

        snprintf(path, sizeof(path), "%s.%d.time", base, i);
        time = tcoh_json_get(meta, path);
        ok = tcoh_parse_utc(time, &t);
        free(time);
        for (int k = 0; k < 6 && ok; k++) {
            snprintf(path, sizeof(path), "%s.%d.%s.%d", base, i,
                     k < 3 ? "position" : "velocity", k % 3);
            v[k] = tcoh_json_get_number(meta, path);
            ok = !isnan(v[k]);
        }
        if (!ok) {
            tcoh_orbit_free(o);
            return report("Invalid orbit state vector %d of <%s>", i, name);
        }
        orbit_push(o, t, v);
    }
    return orbit_covers(o, name, start, stop);
}

/* First occurrence of needle in [p, end), NULL if none. */
static const char *find_before(const char *p, const char *end,
                               const char *needle)
{
    const char *q = strstr(p, needle);

    return q && q < end ? q : NULL;
}

/* The <OSV> records of an EOF document whose UTC time lies in [t0, t1]. A
   record without a UTC time is skipped; one with a missing or unreadable
   coordinate is an error. */
int tcoh_orbit_from_eof(const char *text, double t0, double t1,
                        const char *source, struct orbit *o)
{
    static const char *tags[6] = {"<X", "<Y", "<Z", "<VX", "<VY", "<VZ"};
    int n = 0;

    tcoh_clear_error();
    for (const char *p = strstr(text, "<OSV>"); p; p = strstr(p + 5, "<OSV>"))
        n++;
    if (orbit_reserve(o, n, source))
        return -1;
    for (const char *p = strstr(text, "<OSV>"); p; p = strstr(p + 5, "<OSV>")) {
        const char *end = strstr(p, "</OSV>"), *u;
        double t, v[6];

        if (!end || !(u = find_before(p, end, "<UTC>UTC=")) ||
            !tcoh_parse_utc(u + 9, &t) || t < t0 || t > t1)
            continue;
        for (int k = 0; k < 6; k++) {
            /* the tag may carry attributes (<X unit="m">): the value
               follows the tag's '>' */
            const char *q = find_before(p, end, tags[k]);

            q = q ? find_before(q, end, ">") : NULL;
            if (!q || sscanf(q + 1, "%lf", &v[k]) != 1) {
                tcoh_orbit_free(o);
                return report("Invalid state vector in orbit file");
            }
        }
        orbit_push(o, t, v);
    }
    return 0;
}

/* Validity _VYYYYMMDDThhmmss_YYYYMMDDThhmmss of an orbit file name. */
static int name_validity(const char *name, double *v0, double *v1)
{
    const char *p = strstr(name, "_V");
    struct utc a, b;

    if (!p || sscanf(p, "_V%4d%2d%2dT%2d%2d%lf_%4d%2d%2dT%2d%2d%lf", &a.year,
                     &a.month, &a.day, &a.hour, &a.minute, &a.second, &b.year,
                     &b.month, &b.day, &b.hour, &b.minute, &b.second) != 12)
        return 0;
    *v0 = tcoh_utc_seconds(&a);
    *v1 = tcoh_utc_seconds(&b);
    return 1;
}

/* The .EOF file of dir named <prefix>... whose validity covers [t0, t1];
   among several, the greatest name (the latest production time). Written to
   out; 1 when found. */
int tcoh_orbit_pick_file(const char *dir, const char *prefix, double t0,
                         double t1, char *out, size_t out_size)
{
    DIR *d = opendir(dir);
    struct dirent *de;
    int found = 0;

    if (!d)
        return 0;
    while ((de = readdir(d))) {
        const char *nm = de->d_name;
        const size_t len = strlen(nm);
        double v0, v1;

        if (strncmp(nm, prefix, strlen(prefix)) || len < 4 ||
            strcmp(nm + len - 4, ".EOF") || !name_validity(nm, &v0, &v1) ||
            v0 > t0 || v1 < t1)
            continue;
        if (len + 1 > out_size)
            continue;
        if (!found || strcmp(nm, out) > 0) {
            memcpy(out, nm, len + 1);
            found = 1;
        }
    }
    closedir(d);
    return found;
}

static char *read_file(const char *path)
{
    FILE *fp = fopen(path, "rb");
    char *text = NULL;
    long size;

    if (!fp)
        return NULL;
    if (fseek(fp, 0, SEEK_END) == 0 && (size = ftell(fp)) >= 0 &&
        fseek(fp, 0, SEEK_SET) == 0 && (text = malloc((size_t)size + 1))) {
        if (fread(text, 1, (size_t)size, fp) == (size_t)size)
            text[size] = '\0';
        else {
            free(text);
            text = NULL;
        }
    }
    fclose(fp);
    return text;
}

int tcoh_orbit_from_file(const char *meta, const char *name, const char *kind,
                         const char *cache_dir, int required, struct orbit *o)
{
    char *mission, prefix[128], file[1024], path[2048], *text;
    double start, stop;
    int rc;

    tcoh_clear_error();
    mission = tcoh_json_get(meta, "swath.mission");
    if (!mission)
        return report("No mission in the metadata of <%s>", name);
    if (span_of(meta, name, &start, &stop)) {
        free(mission);
        return -1;
    }
    snprintf(prefix, sizeof(prefix), "%s_OPER_AUX_%s_OPOD_", mission, kind);
    if (!tcoh_orbit_pick_file(cache_dir, prefix, start - ORBIT_MARGIN,
                              stop + ORBIT_MARGIN, file, sizeof(file))) {
        rc = required ? report("No %s orbit file of %s covers <%s> in <%s>",
                               kind, mission, name, cache_dir)
                      : 0;
        free(mission);
        return rc;
    }
    free(mission);
    snprintf(path, sizeof(path), "%s/%s", cache_dir, file);
    if (!(text = read_file(path)))
        return report("Unable to read <%s>", path);
    rc = tcoh_orbit_from_eof(text, start - ORBIT_MARGIN, stop + ORBIT_MARGIN,
                             kind, o);
    free(text);
    if (rc)
        return -1;
    return orbit_covers(o, name, start, stop) ? -1 : 1;
}

int tcoh_orbit_best(const char *meta, const char *name, const char *cache_dir,
                    struct orbit *o)
{
    static const char *kinds[2] = {"POEORB", "RESORB"};

    tcoh_clear_error();
    for (int k = 0; k < 2 && cache_dir; k++) {
        const int rc = tcoh_orbit_from_file(meta, name, kinds[k], cache_dir, 0, o);

        if (rc)
            return rc < 0 ? -1 : 0;
    }
    return tcoh_orbit_from_annotation(meta, name, o);
}

/* ---- scene numbers for the geometry stage ---- */

static int number_of(const char *meta, const char *name, const char *path,
                     double *x)
{
    *x = meta ? tcoh_json_get_number(meta, path) : NAN;
    return isnan(*x) ? report("No <%s> in the metadata of <%s>, needed for the "
                              "flat-earth and topographic phase",
                              path, name)
                     : 0;
}

int tcoh_scene_from_metadata(const char *const *metas,
                             const char *const *names, int n, int ref,
                             double line0, double sample0, int rows, int cols,
                             struct tcoh_scene *sc)
{
    const char *meta, *name;
    char *first;
    int ok;

    tcoh_clear_error();
    if (ref < 0 || ref >= n)
        return report("The reference date is not one of the inputs");
    meta = metas[ref];
    name = names[ref];
    /* the numbers are looked up before the first line time, so a document
       missing both reports the number (geometry.c:165-180) */
    if (number_of(meta, name, "raster_geometry.azimuth_time_interval", &sc->dt) ||
        number_of(meta, name, "raster_geometry.slant_range_time_first_sample",
                  &sc->tau0) ||
        number_of(meta, name, "swath.range_sampling_rate", &sc->fs) ||
        number_of(meta, name, "swath.wavelength", &sc->lambda))
        return -1;
    first = meta ? tcoh_json_get(meta, "raster_geometry.first_line_time") : NULL;
    ok = tcoh_parse_utc(first, &sc->t_first);
    free(first);
    if (!ok)
        return report("No raster_geometry.first_line_time in the metadata of "
                      "<%s>",
                      name);
    for (int k = 0; k < n; k++) {
        double lk;

        if (number_of(metas[k], names[k], "swath.wavelength", &lk))
            return -1;
        if (fabs(lk - sc->lambda) > 1e-9 * sc->lambda)
            return report("Inputs <%s> and <%s> have different radar "
                          "wavelengths",
                          names[k], name);
    }
    sc->line0 = line0;
    sc->sample0 = sample0;
    sc->rows = rows;
    sc->cols = cols;
    return 0;
}
