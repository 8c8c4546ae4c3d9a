"""Strip preparation on the device (tcoh_prepare_strip_device) against the
oracle's restatement of read_strip (main.c:52-127), through the C ABI.

Validity and amplitudes must match bit for bit, and so must the samples
without a geometric phase. With one, the samples are rotated by the FP64
phase of the grid (geometry.c:256-291, here evaluated in another form, so
within a few ulp of the oracle's: 1e-11 rad at phases of 1e4 rad) through
the device's FP64 sin/cos instead of the host's: a rotated sample may then
differ from the oracle's by the float rounding of its components, which is
small against the sample's magnitude though not always against a component
near zero; asserted: within 2.5e-7 of the magnitude, and nearly all
identical.
"""

import numpy as np
import pytest

import synth
from test_gpu_parity import M2C_BAND
from test_oracle import TOLERANCE

pytestmark = pytest.mark.gpu

WINDOW = (9, 5)


def planes(tcoh, n, rows, cols, seed):
    """I/Q date planes with nulls and zero samples planted."""
    stack = synth.ds_stack(n, 0.6, rows=rows, cols=cols, seed=seed)
    i = np.ascontiguousarray(stack.real, np.float32)
    q = np.ascontiguousarray(stack.imag, np.float32)
    rng = np.random.default_rng(seed + 100)
    for _ in range(12):
        k, r, c = rng.integers(n), rng.integers(rows), rng.integers(cols)
        (i if rng.integers(2) else q)[k, r, c] = np.nan
    for _ in range(6):
        k, r, c = rng.integers(n), rng.integers(rows), rng.integers(cols)
        i[k, r, c] = q[k, r, c] = 0.0
    i[1, 2, 3] = 0.0  # one zero component alone keeps the pixel valid
    return i, q


def prepare_on_device(tcoh, i, q, row0, prow, halo, geometry=None, heights=None):
    D = tcoh.api.DeviceBuffer
    n, rows, cols = i.shape
    pc = cols + 2 * halo
    s = tcoh.make_settings(n, window=(WINDOW[0], 2 * halo + 1), min_shp=3)
    di, dq = D.from_array(i), D.from_array(q)
    dh = None if heights is None else D.from_array(np.asarray(heights, np.float64))
    # outputs start as garbage (0x5a bytes): every byte must be written
    slc = D(prow * pc * n * 8, fill=0x5A)
    amp = D(prow * pc * n * 4, fill=0x5A)
    valid = D(prow * pc, fill=0x5A)
    with tcoh.Backend(s, pc, platform=0, device=0) as be:
        be.prepare_strip_device(di.ptr, dq.ptr, rows, cols, row0, prow, halo, slc.ptr,
                                amp.ptr, valid.ptr, geometry=geometry,
                                heights_ptr=None if dh is None else dh.ptr)
        out = (slc.to_array(np.float32), amp.to_array(np.float32),
               valid.to_array(np.uint8))
    return out


def meets_oracle(cref, fp64, n, opts, i, q, maps, precision, ph=None, en=None):
    """A region's maps (coh, count, pairs (npairs, px)) against the oracle
    on the host-prepared whole region (cref.prepare_strip, with the
    geometric phase ph removed where given): SHP counts equal to the FP32
    restatement's; gamma and the pair coherences within TOLERANCE of the
    FP64 oracle with the same NaN pattern. With the bias correction the
    pairs must be all pairs: their magnitudes give the oracle this run's
    decisions inside the form's m2c band."""
    _, rows, cols = i.shape
    ha, hr = opts["window"][0] // 2, opts["window"][1] // 2
    prow, pc = rows + 2 * ha, cols + 2 * hr
    slc, amp, valid = cref.prepare_strip(i, q, -ha, prow, hr, ph, en)
    so = cref.settings(n, **opts)
    _, ocount, _, mask, _ = cref.run(so, pc, slc, amp, valid, prow, rows, cols)
    coh, count, pairs = maps
    assert np.array_equal(count, ocount), "SHP counts differ from the reference kernel"
    adopt = None
    if opts.get("bias"):
        assert opts.get("pairs") == "all"
        adopt = pairs.T.astype(np.float64)
    c64, p64, _, _ = fp64.run(slc, ocount, mask, cref.settings_dict(so), pc, rows, cols,
                              band=M2C_BAND[precision], adopt=adopt)
    assert np.array_equal(np.isnan(coh), np.isnan(c64))
    err = np.nanmax(np.abs(coh - c64))
    assert err <= TOLERANCE, "max |d gamma| %.3g" % err
    if pairs is not None and np.isfinite(p64).any():
        assert np.array_equal(np.isnan(pairs.T), np.isnan(p64))
        perr = np.nanmax(np.abs(pairs.T - p64))
        assert perr <= TOLERANCE, "max |d pair| %.3g" % perr
    return int(np.sum(np.isfinite(coh)))


def ulps(a, b):
    ai = a.view(np.int32).astype(np.int64)
    bi = b.view(np.int32).astype(np.int64)
    ai = np.where(ai < 0, -(ai & 0x7FFFFFFF), ai)
    bi = np.where(bi < 0, -(bi & 0x7FFFFFFF), bi)
    return np.abs(ai - bi)


@pytest.mark.parametrize("n", [6, 30, 45])
def test_strip_without_geometry_is_bit_exact(tcoh, cref, n):
    rows, cols = 37, 45
    ha, hr = WINDOW[0] // 2, WINDOW[1] // 2
    i, q = planes(tcoh, n, rows, cols, seed=n)
    # the whole region with its halo, a strip inside it, a strip past its end
    for row0, prow in ((-ha, rows + 2 * ha), (7, 16), (rows - 5, 12)):
        got = prepare_on_device(tcoh, i, q, row0, prow, hr)
        want = cref.prepare_strip(i, q, row0, prow, hr)
        for g, w, name in zip(got, want, ("slc", "amp", "valid")):
            assert np.array_equal(g.view(np.uint8), w.view(np.uint8)), (name, row0)
        assert want[2].sum() > 0


@pytest.mark.parametrize("n,elevation", [(6, True), (30, True), (30, False)])
def test_strip_with_geometric_phase(tcoh, cref, n, elevation):
    rows, cols = 37, 45
    ha, hr = WINDOW[0] // 2, WINDOW[1] // 2
    i, q = planes(tcoh, n, rows, cols, seed=3 * n)
    orbits, scene, ref = synth.geometry_case(n=n, rows=rows, cols=cols, seed=4)
    grid, hs, status = cref.geometry_grid(orbits, scene, ref, 0.0, 500.0)
    assert status == 0 and grid.shape[2] == 3
    rng = np.random.default_rng(5)
    heights = rng.uniform(-20.0, 520.0, (rows, cols)) if elevation else None
    if elevation:
        heights[rng.integers(rows, size=9), rng.integers(cols, size=9)] = np.nan
    h0 = 123.5
    g = tcoh.api.make_geometry(grid, hs, h0=h0)
    row0, prow = -ha, rows + 2 * ha
    got = prepare_on_device(tcoh, i, q, row0, prow, hr, geometry=g, heights=heights)
    ph = cref.pixel_phases(grid, hs, rows, cols, elevation=heights, h0=h0)
    en = None if heights is None else np.isnan(heights).astype(np.uint8)
    want = cref.prepare_strip(i, q, row0, prow, hr, ph, en)
    assert np.array_equal(got[2], want[2])
    assert np.array_equal(got[1].view(np.uint32), want[1].view(np.uint32))
    gc = got[0].reshape(-1, 2).astype(np.float64)
    wc = want[0].reshape(-1, 2).astype(np.float64)
    mag = np.hypot(wc[:, 0], wc[:, 1])
    rel = np.hypot(*(gc - wc).T) / np.where(mag > 0, mag, 1.0)
    assert rel.max() <= 2.5e-7, rel.max()
    assert np.mean(rel == 0) > 0.99
    # the rotation is real: the rotated samples differ from the raw ones
    raw = cref.prepare_strip(i, q, row0, prow, hr)
    assert not np.array_equal(raw[0], want[0])


def test_device_strip_feeds_the_estimate(tcoh, cref):
    """tcoh_run_device on the device-prepared strip gives the maps of
    ocl_run on the host-prepared strip, bit for bit."""
    D = tcoh.api.DeviceBuffer
    n, rows, cols = 8, 40, 33
    ha, hr = WINDOW[0] // 2, WINDOW[1] // 2
    i, q = planes(tcoh, n, rows, cols, seed=21)
    prow, pc = rows + 2 * ha, cols + 2 * hr
    s = tcoh.make_settings(n, window=WINDOW, min_shp=9, pairs="consecutive",
                           bias=True)
    slc_h, amp_h, valid_h = cref.prepare_strip(i, q, -ha, prow, hr)
    with tcoh.Backend(s, pc, platform=0, device=0) as be:
        coh_h, count_h, pairs_h = be.run(slc_h, amp_h, valid_h, prow, rows, cols)
        di, dq = D.from_array(i), D.from_array(q)
        slc, amp, valid = D(prow * pc * n * 8), D(prow * pc * n * 4), D(prow * pc)
        coh, count = D(rows * cols * 4), D(rows * cols * 4)
        pairs = D(s.npairs * rows * cols * 4)
        be.prepare_strip_device(di.ptr, dq.ptr, rows, cols, -ha, prow, hr, slc.ptr,
                                amp.ptr, valid.ptr)
        be.run_device(slc.ptr, amp.ptr, valid.ptr, prow, rows, cols, coh.ptr,
                      count.ptr, pairs.ptr)
        coh_d = coh.to_array(np.float32)
        count_d = count.to_array(np.int32)
        pairs_d = pairs.to_array(np.float32, (s.npairs, rows * cols))
    assert np.array_equal(count_d, count_h)
    assert np.array_equal(coh_d.view(np.uint32), coh_h.view(np.uint32))
    assert np.array_equal(pairs_d.view(np.uint32), pairs_h.view(np.uint32))


def test_strip_errors_are_reported(tcoh):
    n, rows, cols = 6, 20, 20
    s = tcoh.make_settings(n, window=WINDOW, min_shp=3)
    buf = tcoh.api.DeviceBuffer(16384)
    p = buf.ptr
    with tcoh.Backend(s, cols + 4, platform=0, device=0) as be:
        with pytest.raises(RuntimeError, match="padded width"):
            be.prepare_strip_device(p, p, rows, cols + 1, 0, 1, 2, p, p, p)
        grid = np.zeros((2, 2, 1, n))  # covers 16 x 16 cells, not 20 x 20
        g = tcoh.api.make_geometry(grid, [0.0])
        with pytest.raises(RuntimeError, match="does not cover"):
            be.prepare_strip_device(p, p, rows, cols, 0, 1, 2, p, p, p, geometry=g)


def same_maps(got, want):
    for g, w in zip(got, want):
        if w is None:
            assert g is None
            continue
        assert np.array_equal(np.asarray(g).view(np.uint32), np.asarray(w).view(np.uint32))


@pytest.mark.parametrize("precision", ["plain", "f16", "bf16"])
def test_region_stream_equals_one_strip(tcoh, cref, fp64, precision):
    """tcoh_run_region (strips staged, prepared and estimated in a
    pipeline) gives the maps of one ocl_run over the whole host-prepared
    region, bit for bit, for several strip heights (reference
    test:118-135: strips are seamless), and they meet the oracle."""
    n, rows, cols = 8, 70, 33
    ha, hr = WINDOW[0] // 2, WINDOW[1] // 2
    i, q = planes(tcoh, n, rows, cols, seed=31)
    s = tcoh.make_settings(n, window=WINDOW, min_shp=9, pairs="all", bias=True)
    prow, pc = rows + 2 * ha, cols + 2 * hr
    slc_h, amp_h, valid_h = cref.prepare_strip(i, q, -ha, prow, hr)
    with tcoh.Backend(s, pc, platform=0, device=0, precision=precision) as be:
        want = be.run(slc_h, amp_h, valid_h, prow, rows, cols)
        for strip_rows in (16, 23, 0):
            got = be.run_region(i, q, strip_rows=strip_rows)
            same_maps(got, want)
    opts = dict(window=WINDOW, min_shp=9, pairs="all", bias=True)
    assert meets_oracle(cref, fp64, n, opts, i, q, got, precision) > 100


def test_region_stream_with_geometric_phase(tcoh, cref, fp64):
    """With the flat-earth / topographic phase removed at elevation heights:
    the streamed region equals the whole region prepared on the device
    (the same strip kernel over the region's full planes) and estimated in
    one strip, and meets the oracle on the region prepared on the host with
    the oracle's phases (the rotations agree to a float rounding)."""
    n, rows, cols = 6, 58, 40
    ha, hr = WINDOW[0] // 2, WINDOW[1] // 2
    i, q = planes(tcoh, n, rows, cols, seed=41)
    orbits, scene, ref = synth.geometry_case(n=n, rows=rows, cols=cols, seed=6)
    grid, hs, status = cref.geometry_grid(orbits, scene, ref, 0.0, 400.0)
    assert status == 0
    heights = np.random.default_rng(7).uniform(0.0, 400.0, (rows, cols))
    heights[10, 11] = np.nan
    g = tcoh.api.make_geometry(grid, hs, h0=50.0)
    s = tcoh.make_settings(n, window=WINDOW, min_shp=9, pairs="consecutive")
    prow, pc = rows + 2 * ha, cols + 2 * hr
    D = tcoh.api.DeviceBuffer
    with tcoh.Backend(s, pc, platform=0, device=0) as be:
        di, dq, dh = D.from_array(i), D.from_array(q), D.from_array(heights)
        slc, amp, valid = D(prow * pc * n * 8), D(prow * pc * n * 4), D(prow * pc)
        coh, count = D(rows * cols * 4), D(rows * cols * 4)
        pairs = D(s.npairs * rows * cols * 4)
        be.prepare_strip_device(di.ptr, dq.ptr, rows, cols, -ha, prow, hr, slc.ptr,
                                amp.ptr, valid.ptr, geometry=g, heights_ptr=dh.ptr)
        be.run_device(slc.ptr, amp.ptr, valid.ptr, prow, rows, cols, coh.ptr,
                      count.ptr, pairs.ptr)
        want = (coh.to_array(np.float32), count.to_array(np.int32),
                pairs.to_array(np.float32, (s.npairs, rows * cols)))
        for strip_rows in (13, 0):
            got = be.run_region(i, q, geometry=g, heights=heights, strip_rows=strip_rows)
            same_maps(got, want)
    assert want[1][10 * cols + 11] == -1
    ph = cref.pixel_phases(grid, hs, rows, cols, elevation=heights, h0=50.0)
    en = np.isnan(heights).astype(np.uint8)
    opts = dict(window=WINDOW, min_shp=9, pairs="consecutive")
    # (the handle's default form: split BF16)
    assert meets_oracle(cref, fp64, n, opts, i, q, got, "bf16", ph=ph, en=en) > 100


def test_raw_region_writes_the_module_maps(tcoh, cref, fp64, tmp_path):
    """The module's loop without GRASS: raw I/Q date planes mapped in place,
    streamed through the GPU, written as maps with GRASS's null words, the
    reference's pair-map names and main.c:921-922's summary; the values are
    those of the in-memory region run, which meets the oracle."""
    import datetime as dt

    api = tcoh.api
    n, rows, cols = 6, 45, 38
    i, q = planes(tcoh, n, rows, cols, seed=51)
    i.tofile(tmp_path / "stack_i.raw")
    q.tofile(tmp_path / "stack_q.raw")
    s = tcoh.make_settings(n, window=WINDOW, min_shp=9, pairs="consecutive")
    pc = cols + 2 * (WINDOW[1] // 2)
    t0 = dt.datetime(2023, 1, 12, 14, 25, 9)
    starts = [t0 + dt.timedelta(days=12 * k) for k in range(n)]
    with tcoh.Backend(s, pc, platform=0, device=0) as be:
        coh, count, pairs = be.run_region(i, q, strip_rows=17)
        est, mean, names = api.run_raw_region(
            be, tmp_path / "stack_i.raw", tmp_path / "stack_q.raw", rows, cols,
            tmp_path, "coh", shp_count="shp", pairs_prefix="p", starts=starts,
            strip_rows=17)
    assert names[0] == "p_20230112_20230124" and len(names) == n - 1
    opts = dict(window=WINDOW, min_shp=9, pairs="consecutive")
    assert meets_oracle(cref, fp64, n, opts, i, q, (coh, count, pairs), "bf16") > 100
    nan = np.isnan(coh)
    f = np.fromfile(tmp_path / "coh.raw", np.uint32)
    assert np.all(f[nan] == 0xFFFFFFFF) and np.array_equal(f[~nan], coh.view(np.uint32)[~nan])
    c = np.fromfile(tmp_path / "shp.raw", np.int32)
    assert np.array_equal(c, np.where(count < 0, np.iinfo(np.int32).min, count))
    for p, name in enumerate(names):
        g = np.fromfile(tmp_path / (name + ".raw"), np.float32)
        assert np.array_equal(np.isnan(g), np.isnan(pairs[p]))
        assert np.array_equal(g[~np.isnan(g)], pairs[p][~np.isnan(pairs[p])])
    assert est == int(np.sum(~nan)) and est > 100
    assert mean == pytest.approx(float(np.mean(coh[~nan], dtype=np.float64)), rel=1e-12)


@pytest.mark.parametrize("precision", ["plain", "f16", "bf16"])
def test_depth60_with_geometric_phase(tcoh, cref, fp64, precision):
    """The full-geometry workload's path on a crop: N=60 dates carrying
    exp(-i phi_geo) of a 0-500 m DEM (phi_geo from the library's FP64 grid
    on circular orbits), removed again by the device strip preparation
    (k_strip at 64 date slots) at the DEM heights, then estimated by the
    N=60 kernels (KS, EVD, bias, consecutive pairs). Against the oracle's
    strip (read_strip with geometry_phases at the same heights), its SHP
    counts (bit for bit) and the FP64 gamma and pairs (1e-4)."""
    from test_gpu_parity import M2C_BAND
    D = tcoh.api.DeviceBuffer
    n, rows, cols = 60, 34, 30
    window = (21, 7)
    ha, hr = window[0] // 2, window[1] // 2
    orbits, scene, ref = synth.geometry_case(n=n, rows=rows, cols=cols, ref=0, seed=7)
    grid, hs = tcoh.api.geometry_grid(orbits, scene, ref, 0.0, 500.0)
    g = tcoh.api.make_geometry(grid, hs)
    yy, xx = np.mgrid[0:rows, 0:cols]
    dem = 250.0 + 240.0 * np.sin(yy / 7.0) * np.cos(xx / 5.0)
    dem[5, 6] = np.nan
    phi = cref.pixel_phases(grid, hs, rows, cols, elevation=dem)
    reg, params = synth.config_voronoi(n=n, rows=rows, cols=cols, seed=3,
                                            gamma0=(0.3, 0.9))
    obs = synth.scene_stack(n, rows, cols, reg, params, seed=3) * np.exp(-1j * phi)
    i = np.ascontiguousarray(obs.real, np.float32)
    q = np.ascontiguousarray(obs.imag, np.float32)
    prow, pc = rows + 2 * ha, cols + 2 * hr
    want = cref.prepare_strip(i, q, -ha, prow, hr, phase=phi,
                              elev_null=np.isnan(dem).astype(np.uint8))
    s = tcoh.make_settings(n, window=window, min_shp=20, bias=True, pairs="consecutive")
    s_all = tcoh.make_settings(n, window=window, min_shp=20, bias=True, pairs="all")
    di, dq, dh = D.from_array(i), D.from_array(q), D.from_array(dem)
    slc, amp, valid = D(prow * pc * n * 8), D(prow * pc * n * 4), D(prow * pc)
    npx = rows * cols
    coh, count, pairs = D(npx * 4), D(npx * 4), D((n - 1) * npx * 4)
    pall = D(n * (n - 1) // 2 * npx * 4)
    with tcoh.Backend(s, pc, platform=0, device=0, precision=precision) as be, \
            tcoh.Backend(s_all, pc, platform=0, device=0, precision=precision) as ball:
        be.prepare_strip_device(di.ptr, dq.ptr, rows, cols, -ha, prow, hr, slc.ptr,
                                amp.ptr, valid.ptr, geometry=g, heights_ptr=dh.ptr)
        be.run_device(slc.ptr, amp.ptr, valid.ptr, prow, rows, cols, coh.ptr, count.ptr,
                      pairs.ptr)
        got_coh = coh.to_array(np.float32)
        got_count = count.to_array(np.int32)
        got_pairs = pairs.to_array(np.float32, (n - 1, npx))
        ball.run_device(slc.ptr, amp.ptr, valid.ptr, prow, rows, cols, coh.ptr, count.ptr,
                        pall.ptr)
        all_pairs = pall.to_array(np.float32, (n * (n - 1) // 2, npx))
        assert np.array_equal(coh.to_array(np.float32).view(np.uint32),
                              got_coh.view(np.uint32))
        got_valid = valid.to_array(np.uint8)
    assert np.array_equal(got_valid, want[2])
    so = cref.settings(n, window=window, min_shp=20, bias=True, pairs="consecutive")
    _c, ocount, _p, mask, _f = cref.run(so, pc, want[0], want[1], want[2], prow, rows, cols)
    assert np.array_equal(got_count, ocount)
    assert (ocount >= so.min_looks).sum() > 500
    c64, p64, _near, _m = fp64.run(want[0], ocount, mask, cref.settings_dict(so), pc, rows,
                                   cols, band=M2C_BAND[precision],
                                   adopt=all_pairs.T.astype(np.float64))
    assert np.array_equal(np.isnan(got_coh), np.isnan(c64))
    err = np.nanmax(np.abs(got_coh - c64))
    assert err <= 1e-4, err
    assert np.nanmax(np.abs(got_pairs.T - p64)) <= 1e-4
    assert np.isnan(got_coh.reshape(rows, cols)[5, 6])
