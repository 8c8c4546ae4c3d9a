/* Launch shapes of the strip kernels, as the tuner (tune.py, run on an
 * MI355X) chose them: for an estimator, a covariance form and a stack
 * depth, the number of XCD stripes a row is cut into by the SHP and the
 * linking launches (0 takes 8, one stripe per XCD; 1 the whole row, every
 * XCD a band of rows) and the output pixels of one linking launch. A handle takes the entry of its estimator and form at
 * the nearest tuned depth with the same lanes per pixel, else the nearest
 * depth, else the built-in shapes (TILE_STRIPE_DEFAULT, TILE_BAND_PX).
 *
 * Written by `python i.sar.temporal_coherence_amd/tune.py --write`; the
 * table's last line is an end marker (n = 0).
 */
#ifndef TCOH_TILES_H
#define TCOH_TILES_H

#define TILE_STRIPE_DEFAULT 0
#define TILE_BAND_PX (1 << 20)

struct TileEntry {
    int emi, mixed, n;
    int shp_stripe, link_stripe, band_px;
};

static const TileEntry TILE_TABLE[] = {
    /* emi mixed  n  shp  link  band */
    {0, 0, 0, 0, 0, 0},
};

#endif
