"""The module's data either side of the estimate, without GRASS (host C in
libtcoh_hip.so): pair-map names (main.c:602-645), raw date planes mapped in
place, output maps with the module's nulls and summary (main.c:796-824,
921-922). CPU only.
"""

import datetime as dt
import struct

import numpy as np
import pytest

# The reference tests' dates: T0 and one date every 12 days
# (tests/conftest.py:23, test:304-306).
T0 = dt.datetime(2023, 1, 12, 14, 25, 9)


@pytest.fixture()
def api(tcoh):
    tcoh.api.lib().tcoh_set_error_mode(1)
    return tcoh.api


def ref_name(prefix, i, j):
    """The name the reference's tests expect for pair (i, j) of dates
    T0 + 12 k days (test:304-306)."""
    day = [(T0 + dt.timedelta(days=12 * k)).strftime("%Y%m%d") for k in (i, j)]
    return "{}_{}_{}".format(prefix, *day)


@pytest.mark.parametrize("mode", ["consecutive", "all"])
def test_pair_names_match_the_reference_tests(api, mode):
    n = 5
    starts = [T0 + dt.timedelta(days=12 * k) for k in range(n)]
    names, dated = api.pair_names("p", n, mode == "all", starts)
    pairs = ([(i, j) for i in range(n - 1) for j in range(i + 1, n)] if mode == "all"
             else [(i, i + 1) for i in range(n - 1)])
    assert dated and names == [ref_name("p", i, j) for i, j in pairs]
    assert names[0] == "p_20230112_20230124"


def test_pair_names_with_time_and_ranks(api):
    # two consecutive dates on one day: every tag carries the time of day
    starts = [dt.datetime(2023, 1, 12, 5, 43, 21, 900000),
              dt.datetime(2023, 1, 12, 17, 2, 3),
              dt.datetime(2023, 1, 24, 5, 43, 22)]
    names, dated = api.pair_names("c", 3, True, starts)
    assert dated and names == ["c_20230112T054321_20230112T170203",
                               "c_20230112T054321_20230124T054322",
                               "c_20230112T170203_20230124T054322"]
    # the same day but not consecutive: dates only
    names, _ = api.pair_names("c", 3, False, [starts[0], starts[2], starts[1]])
    assert names == ["c_20230112_20230124", "c_20230124_20230112"]
    # a date without a time: 1-based ranks for every pair
    names, dated = api.pair_names("r", 4, False, starts + [starts[2]],
                                  has_time=[1, 1, 0, 1])
    assert not dated and names == ["r_1_2", "r_2_3", "r_3_4"]
    names, dated = api.pair_names("r", 3, True, None)
    assert not dated and names == ["r_1_2", "r_1_3", "r_2_3"]


def test_raw_maps_hold_grass_nulls_and_the_summary(api, tmp_path):
    rng = np.random.default_rng(0)
    rows, cols, npairs = 7, 5, 2
    coh = rng.random((rows, cols)).astype(np.float32)
    coh[rng.random((rows, cols)) < 0.3] = np.nan
    count = rng.integers(-1, 200, (rows, cols)).astype(np.int32)
    pairs = rng.random((npairs, rows, cols)).astype(np.float32)
    pairs[0, 2, 3] = np.nan
    names = ["p_20230112_20230124", "p_20230124_20230205"]
    maps = api.RawMaps(tmp_path, "coh", rows, cols, shp_count="shp", pair_maps=names)
    for r0, r1 in ((0, 3), (3, 4), (4, 7)):  # three strips in row order
        maps.put_strip(coh[r0:r1], count[r0:r1],
                       pairs[:, r0:r1].reshape(npairs, -1))
    est, mean = maps.close()
    # the summary: estimated pixels and their mean, summed in row order
    total, k = 0.0, 0
    for v in coh.reshape(-1):
        if not np.isnan(v):
            total += float(v)
            k += 1
    assert est == k and mean == total / k
    # GRASS's null words: all ones for float, INT32_MIN for int
    f = np.fromfile(tmp_path / "coh.raw", np.uint32).reshape(rows, cols)
    assert np.all(f[np.isnan(coh)] == 0xFFFFFFFF)
    assert np.array_equal(f[~np.isnan(coh)], coh.view(np.uint32)[~np.isnan(coh)])
    c = np.fromfile(tmp_path / "shp.raw", np.int32).reshape(rows, cols)
    assert np.array_equal(c, np.where(count < 0, np.iinfo(np.int32).min, count))
    for p, name in enumerate(names):
        g = np.fromfile(tmp_path / (name + ".raw"), np.uint32).reshape(rows, cols)
        want = np.where(np.isnan(pairs[p]), np.uint32(0xFFFFFFFF), pairs[p].view(np.uint32))
        assert np.array_equal(g, want)
    hdr = (tmp_path / "shp.hdr").read_text().splitlines()
    assert hdr[0] == "ENVI" and "samples = 5" in hdr and "lines = 7" in hdr
    assert "data type = 3" in hdr and "data ignore value = -2147483648" in hdr
    assert "data type = 4" in (tmp_path / "coh.hdr").read_text().splitlines()
    order = "byte order = %d" % (struct.pack("=i", 1)[0] == 0)
    assert order in hdr


def test_raw_maps_errors(api, tmp_path):
    maps = api.RawMaps(tmp_path, "coh", 2, 3)
    with pytest.raises(RuntimeError, match="More rows than the map holds"):
        maps.put_strip(np.zeros((3, 3)), np.zeros((3, 3)))
    maps.put_strip(np.zeros((1, 3)), np.zeros((1, 3)))
    with pytest.raises(RuntimeError, match="Fewer rows written"):
        maps.close()
    with pytest.raises(RuntimeError, match="Unable to create"):
        api.RawMaps(tmp_path / "absent", "coh", 2, 3)
    # an empty region: nothing estimated, mean NaN
    maps = api.RawMaps(tmp_path, "nothing", 1, 2)
    maps.put_strip(np.full((1, 2), np.nan), np.full((1, 2), -1))
    est, mean = maps.close()
    assert est == 0 and np.isnan(mean)


def test_raw_map_reads_in_place(api, tmp_path):
    import ctypes as C

    a = np.arange(3 * 4 * 5, dtype=np.float32).reshape(3, 4, 5)
    a.tofile(tmp_path / "i.raw")
    L = api.lib()
    addr = C.c_void_p()
    assert L.tcoh_raw_map(str(tmp_path / "i.raw").encode(), a.nbytes, C.byref(addr)) == 0
    got = np.ctypeslib.as_array(C.cast(addr, C.POINTER(C.c_float)), shape=(a.size,))
    assert np.array_equal(got.reshape(a.shape), a)
    L.tcoh_raw_unmap(addr, a.nbytes)
    assert L.tcoh_raw_map(str(tmp_path / "i.raw").encode(), a.nbytes + 4, C.byref(addr)) == -1
    assert b"shorter than its declared size" in L.tcoh_last_error()
    assert L.tcoh_raw_map(str(tmp_path / "none.raw").encode(), 4, C.byref(addr)) == -1
    assert b"Unable to open" in L.tcoh_last_error()


def test_raw_maps_reject_a_missing_handle(api):
    L = api.lib()
    assert L.tcoh_raw_put_strip(None, 1, None, None, None) == -1
    assert b"Invalid map handle" in L.tcoh_last_error()
    assert L.tcoh_raw_close(None, None, None) == -1
