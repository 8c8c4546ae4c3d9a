/* Drop-in for the module's geometry.c: the same three entry points
 * (local_proto.h:96-102) over libtcoh_hip.so, so that main.c, which calls
 * geometry_init once and geometry_phases per pixel (main.c:99-124, 710),
 * links unchanged. Compiled in the module, against GRASS's headers and the
 * module's local_proto.h, in place of geometry.c.
 *
 * geometry_init (geometry.c:161-254) takes the reference date's line and
 * sample of the region's first cell from the rasters' headers, as the
 * module does; the metadata reads and the wavelength check are
 * tcoh_scene_from_metadata's, and the phase grid is computed on the GPU by
 * tcoh_geometry_init (FP64). geometry_phases and geometry_free are the
 * library's own (same grid layout [node][height][date]).
 */
#include <grass/gis.h>

#include "local_proto.h"

#define TCOH_HAVE_REFERENCE_TYPES
#include "tcoh_hip.h"

void geometry_init(struct geometry *g, const struct epoch *epochs, int n,
                   int ref, const struct orbit *orbits,
                   const struct Cell_head *region, double hmin, double hmax,
                   double h0)
{
    const struct epoch *e = &epochs[ref];
    const char *metas[MAX_DATES], *names[MAX_DATES];
    struct tcoh_scene sc;
    /* the reference image spans 0..cols east and 0..rows north */
    const double line0 = (e->head.north - region->north) / region->ns_res;
    const double sample0 = (region->west - e->head.west) / region->ew_res;
    int k;

    for (k = 0; k < n; k++) {
        metas[k] = epochs[k].meta;
        names[k] = epochs[k].basename;
    }
    /* both end the process with the reference's messages on failure */
    tcoh_scene_from_metadata(metas, names, n, ref, line0, sample0, region->rows,
                             region->cols, &sc);
    tcoh_geometry_init(g, &sc, n, ref, orbits, hmin, hmax, h0);
}
