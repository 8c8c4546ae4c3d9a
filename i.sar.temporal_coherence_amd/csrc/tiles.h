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
    {0, 1, 8, 2, 1, 1048576}, /* evd f16 2048x3072: 22.15 -> 21.03 ms */
    {0, 1, 12, 0, 1, 1048576}, /* evd f16 2048x2048: 16.55 -> 15.16 ms */
    {0, 1, 20, 2, 1, 1048576}, /* evd f16 1536x2048: 28.25 -> 27.28 ms */
    {0, 1, 30, 1, 0, 1048576}, /* evd f16 2048x2048: 37.83 -> 37.12 ms */
    {0, 1, 45, 0, 1, 1048576}, /* evd f16 768x768: 52.58 -> 49.63 ms */
    {0, 1, 60, 0, 1, 1048576}, /* evd f16 512x512: 37.09 -> 33.84 ms */
    {0, 0, 8, 2, 0, 1048576}, /* evd plain 2048x3072: 22.03 -> 21.60 ms */
    {0, 0, 12, 0, 0, 1048576}, /* evd plain 2048x2048: 20.49 -> 20.49 ms */
    {0, 0, 20, 0, 0, 1048576}, /* evd plain 1536x2048: 49.01 -> 49.01 ms */
    {0, 0, 30, 0, 0, 1048576}, /* evd plain 2048x2048: 102.37 -> 102.37 ms */
    {0, 0, 45, 0, 1, 1048576}, /* evd plain 768x768: 96.85 -> 79.88 ms */
    {0, 0, 60, 0, 1, 1048576}, /* evd plain 512x512: 61.24 -> 46.44 ms */
    {1, 0, 30, 0, 0, 1048576}, /* emi plain 1024x1024: 93.17 -> 93.17 ms */
    {0, 0, 0, 0, 0, 0},
};

#endif
