"""Orbit ingestion (host C in libtcoh_hip.so): UTC calendar (epoch.c:28-52),
JSON path lookup (jsonget.c:99-225), annotation and EOF state vectors and
the choice of orbit file (orbit.c:47-318), scene numbers (geometry.c:165-191).

Each function is checked against an independent implementation of what it
must compute: Python's datetime for the calendar, the json module for the
lookups, the decimal text that was written for the parsed numbers, and the
file-selection and fallback rules of orbit.c on constructed directories.
None of these needs a GPU.
"""

import datetime as dt
import json
import math
import os

import numpy as np
import pytest

import synth

EPOCH = dt.datetime(1970, 1, 1)
T2023 = (dt.datetime(2023, 1, 12, 5, 43, 21, 250000) - EPOCH).total_seconds()


@pytest.fixture()
def api(tcoh):
    tcoh.api.lib().tcoh_set_error_mode(1)
    return tcoh.api


# ---- UTC ---------------------------------------------------------------


def test_utc_seconds_match_datetime(api):
    """epoch.c:28-32 against datetime over 1600-2400, leap days included
    (2000-02-29 yes, 1900-02-29 and 2100-02-29 do not exist)."""
    rng = np.random.default_rng(0)
    days = [dt.datetime(2000, 2, 29), dt.datetime(1900, 3, 1), dt.datetime(2100, 2, 28),
            dt.datetime(1970, 1, 1), dt.datetime(1969, 12, 31, 23, 59, 59),
            dt.datetime(1600, 1, 1), dt.datetime(2399, 12, 31, 23, 59, 59)]
    for _ in range(400):
        days.append(dt.datetime(1600, 1, 1) + dt.timedelta(
            seconds=int(rng.integers(0, 800 * 365 * 86400))))
    for d in days:
        want = (d - EPOCH).total_seconds()
        assert api.utc_seconds(d.year, d.month, d.day, d.hour, d.minute,
                               float(d.second)) == want, d


def test_utc_from_seconds_round_trip(api):
    """epoch.c:34-52: the broken-down fields of datetime, whole seconds
    exact, and seconds back to the same value."""
    rng = np.random.default_rng(1)
    for sec in list(rng.integers(-200 * 365 * 86400, 400 * 365 * 86400, 300)) + [0, -1, 951782400]:
        d = EPOCH + dt.timedelta(seconds=int(sec))
        got = api.utc_from_seconds(float(sec))
        assert got[:5] == (d.year, d.month, d.day, d.hour, d.minute), (sec, got)
        assert got[5] == d.second
        assert api.utc_seconds(*got) == float(sec)


def test_parse_utc(api):
    assert api.parse_utc("2023-01-12T05:43:21.250000") == T2023
    assert api.parse_utc("2023-01-12T05:43:21") == T2023 - 0.25
    assert api.parse_utc("2023-01-12 05:43:21") is None
    assert api.parse_utc("garbage") is None


# ---- JSON --------------------------------------------------------------

DOC = {
    "swath": {
        "mission": "S1B",
        "missionx": "not this one",
        "polarization": "VV",
        "wavelength": 0.05546576,
        "count": 12,
        "neg": -3.5e-7,
        "flag": True,
        "off": False,
        "nothing": None,
        "empty": [],
        "grid": [[1, 2], [3, [4, 5, {"deep": "yes"}]], []],
        "orbit_state_vectors": [
            {"time": "2023-01-12T05:42:21.250000", "position": [1.5, -2.25, 3e6]},
            {"time": "2023-01-12T05:42:31.250000", "position": [4.0, 5.0, 6.0]},
        ],
    },
    "raster_geometry": {"first_line_time": "2023-01-12T05:43:21.250000",
                        "azimuth_time_interval": 0.002055556},
    "quoted": "say \"hi\" / back\\slash",
    "s": "tab\there\nnewline",
}


def _walk(obj, path=()):
    """(dotted path, value) for every node of a JSON document."""
    yield ".".join(path), obj
    if isinstance(obj, dict):
        for k, v in obj.items():
            yield from _walk(v, path + (k,))
    elif isinstance(obj, list):
        for i, v in enumerate(obj):
            yield from _walk(v, path + (str(i),))


@pytest.mark.parametrize("indent", [None, 2, "\t"])
def test_json_get_matches_the_json_module(api, indent):
    """Every node of a document, compact and indented: strings by content,
    numbers and booleans by their text, containers and null as absent."""
    text = json.dumps(DOC, indent=indent)
    n = 0
    for path, v in _walk(DOC):
        if not path:
            continue
        got = api.json_get(text, path)
        if isinstance(v, (dict, list)) or v is None:
            assert got is None, path
        elif isinstance(v, bool):
            assert got == ("true" if v else "false"), path
        elif isinstance(v, str):
            assert got == v, path
        else:
            assert got == json.dumps(v), path
            assert api.json_get_number(text, path) == v, path
        if isinstance(v, list):
            assert api.json_array_length(text, path) == len(v), path
        else:
            assert api.json_array_length(text, path) == -1, path
        n += 1
    assert n > 30


def test_json_absent_paths_and_numbers(api):
    text = json.dumps(DOC)
    for path in ["swath.missio", "swath.mission.x", "nope", "swath.grid.3",
                 "swath.grid.-1", "swath.grid.x", "swath.empty.0",
                 "swath.orbit_state_vectors.2.time", "raster_geometry.x"]:
        assert api.json_get(text, path) is None, path
        assert math.isnan(api.json_get_number(text, path)), path
    # a string that is wholly a number reads as one; partly, as NaN
    assert api.json_get_number('{"a": "12.5"}', "a") == 12.5
    assert math.isnan(api.json_get_number('{"a": "12.5 m"}', "a"))
    assert math.isnan(api.json_get_number(text, "swath.mission"))
    # a key is compared as written: the first of two equal keys wins
    assert api.json_get('{"a": 1, "a": 2}', "a") == "1"


def test_json_escapes_follow_the_reference(api):
    """jsonget.c:74-95: \\n and \\t are decoded, any other escaped character
    is kept as itself (so \\r reads as r and \\u00e9 as u00e9)."""
    text = r'{"a": "x\ny\tz\"q\\w\/e\r\u00e9"}'
    assert api.json_get(text, "a") == 'x\ny\tz"q\\w/er' + "u00e9"


# ---- state vectors -----------------------------------------------------


def _case(tcoh, half_span=60.0, n=3):
    """Orbits of a 3-date stack shifted to January 2023, the reference
    date's scene and every date's first line time."""
    orbits, scene, ref = synth.geometry_case(n=n, rows=200, cols=300, seed=3)
    t0s = [o[0][len(o[0]) // 2] for o in orbits]
    base = [o[1][0, 1] for o in orbits]
    shift = T2023 - t0s[ref]
    t0s = [t + shift for t in t0s]
    orbits = synth.circular_orbits(t0s, base, half_span=half_span)
    scene = dict(scene, t_first=t0s[ref])
    return orbits, scene, ref, t0s


def test_annotation_state_vectors_are_read_exactly(tcoh, api):
    orbits, scene, ref, t0s = _case(tcoh)
    for k, o in enumerate(orbits):
        meta = synth.s1_metadata(o, scene, t0s[k])
        t, pos, vel, source = api.orbit_from_annotation(meta, f"s{k}")
        assert source == "annotation"
        # times: the microsecond text of each state vector
        want_t = [api.parse_utc(synth.iso_utc(x)) for x in o[0]]
        assert np.array_equal(t, want_t)
        assert np.max(np.abs(t - o[0])) < 1e-6
        # numbers: JSON carries them at full precision
        assert np.array_equal(pos, o[1]) and np.array_equal(vel, o[2])
        start, stop = api.acquisition_span(meta)
        assert start == api.parse_utc(synth.iso_utc(t0s[k]))
        assert stop - start == pytest.approx((scene["rows"] - 1) * scene["dt"], abs=2e-6)


def test_eof_state_vectors_are_read_exactly(tcoh, api):
    """The EOF numbers are the 6-decimal text written; records outside
    [t0, t1] are skipped."""
    orbits, scene, ref, t0s = _case(tcoh, half_span=600.0)
    o = orbits[0]
    text = synth.eof_text(o)
    t, pos, vel, source = api.orbit_from_eof(text, -1e30, 1e30, "POEORB")
    assert source == "POEORB" and len(t) == len(o[0]) == 121
    assert np.array_equal(pos, np.array([[float(f"{x:.6f}") for x in r] for r in o[1]]))
    assert np.array_equal(vel, np.array([[float(f"{x:.6f}") for x in r] for r in o[2]]))
    t, pos, vel, _ = api.orbit_from_eof(text, t0s[0] - 120.0, t0s[0] + 125.0)
    assert len(t) == 25 and t[0] >= t0s[0] - 120.0 and t[-1] <= t0s[0] + 125.0
    # a record without a UTC time is skipped, one without a coordinate fails
    no_utc = text.replace("<UTC>UTC=", "<UTC>", 1)
    assert len(api.orbit_from_eof(no_utc, -1e30, 1e30)[0]) == 120
    broken = text.replace('<VZ unit="m/s">', "<VQ>", 1)
    with pytest.raises(RuntimeError, match="Invalid state vector"):
        api.orbit_from_eof(broken, -1e30, 1e30)


def test_eof_and_annotation_give_the_same_orbit(tcoh, api, cref):
    """Both sources of one orbit interpolate (orbit.c:324-346) to the same
    state within the EOF's 1 um rounding, and give the same phase grid
    (geometry.c:161-254, the FP64 restatement) to 1e-3 rad (4 pi / lambda x
    2 um = 4.5e-4 rad at C-band); the reference's own check of the two
    sources compares coherence maps to 1e-6."""
    import ctypes as C

    orbits_ann, scene, ref, t0s = _case(tcoh)
    orbits_eof = synth.circular_orbits(t0s, [o[1][0, 1] for o in orbits_ann],
                                            half_span=600.0)
    got_ann, got_eof = [], []
    for k in range(len(t0s)):
        meta = synth.s1_metadata(orbits_ann[k], scene, t0s[k])
        start, stop = api.acquisition_span(meta)
        a = api.orbit_from_annotation(meta)
        e = api.orbit_from_eof(synth.eof_text(orbits_eof[k]), start - 120, stop + 120)
        got_ann.append(a[:3])
        got_eof.append(e[:3])
        for tq in np.linspace(start, stop, 7):
            pa, va, pe, ve = (np.zeros(3) for _ in range(4))
            for (t, p, v), P, V in ((a[:3], pa, va), (e[:3], pe, ve)):
                o = api.Orbit()
                keep = [np.ascontiguousarray(x, np.float64) for x in (t, p, v)]
                o.n = len(t)
                o.t = keep[0].ctypes.data_as(C.POINTER(C.c_double))
                o.pos = keep[1].ctypes.data_as(C.POINTER(C.c_double))
                o.vel = keep[2].ctypes.data_as(C.POINTER(C.c_double))
                api.lib().orbit_state(C.byref(o), tq, P.ctypes.data, V.ctypes.data)
            assert np.max(np.abs(pa - pe)) < 2e-6
            assert np.max(np.abs(va - ve)) < 2e-6
    ga, _, sa = cref.geometry_grid(got_ann, scene, ref, 0.0, 400.0)
    ge, _, se = cref.geometry_grid(got_eof, scene, ref, 0.0, 400.0)
    assert sa == 0 and se == 0
    assert np.max(np.abs(ga - ge)) < 1e-3


def test_annotation_coverage_and_errors(tcoh, api):
    orbits, scene, ref, t0s = _case(tcoh)
    o = orbits[0]
    meta = synth.s1_metadata(o, scene, t0s[0])
    # fewer than 8 vectors, or not covering the acquisition
    short = tuple(x[:7] for x in o)
    with pytest.raises(RuntimeError, match="does not cover"):
        api.orbit_from_annotation(synth.s1_metadata(short, scene, t0s[0]))
    late = tuple(x[7:] for x in o)
    with pytest.raises(RuntimeError, match="does not cover"):
        api.orbit_from_annotation(synth.s1_metadata(late, scene, t0s[0]))
    d = json.loads(meta)
    d["swath"]["orbit_state_vectors"][3]["velocity"][1] = "fast"
    with pytest.raises(RuntimeError, match="Invalid orbit state vector 3"):
        api.orbit_from_annotation(json.dumps(d))
    d = json.loads(meta)
    d["swath"]["orbit_state_vectors"] = []
    with pytest.raises(RuntimeError, match="No orbit state vectors"):
        api.orbit_from_annotation(json.dumps(d))
    # no swath times: the first and last line times; then a single instant
    d = json.loads(meta)
    del d["swath"]["start_time"]
    start, stop = api.acquisition_span(json.dumps(d))
    assert start == api.parse_utc(d["raster_geometry"]["first_line_time"])
    assert stop == api.parse_utc(d["raster_geometry"]["last_line_time"])
    del d["raster_geometry"]["last_line_time"]
    assert api.acquisition_span(json.dumps(d)) == (start, start)
    del d["raster_geometry"]["first_line_time"]
    with pytest.raises(RuntimeError, match="No acquisition time for <late>"):
        api.acquisition_span(json.dumps(d), "late")
    # a success after a failure leaves no message behind
    api.orbit_from_annotation(meta)
    assert not api.lib().tcoh_last_error()


# ---- orbit files ---------------------------------------------------------


def test_orbit_file_choice(tcoh, api, tmp_path):
    """orbit.c:114-151, 205-318 on a local directory: prefix and .EOF
    suffix, validity covering the acquisition with 120 s of margin, the
    latest production among several; precise before restituted before
    annotation."""
    syn = synth
    orbits, scene, ref, t0s = _case(tcoh, half_span=600.0)
    o = orbits[0]
    ann = tuple(x[54:67] for x in o)  # 13 vectors, t0 +- 60 s
    meta = syn.s1_metadata(ann, scene, t0s[0])
    start, stop = api.acquisition_span(meta)
    eof = syn.eof_text(o)
    cover = (t0s[0] - 599, t0s[0] + 599)
    files = {
        syn.eof_name(*cover, produced="20230201T120000"): eof,
        syn.eof_name(*cover, produced="20230202T120000"): eof,      # latest
        syn.eof_name(*cover, produced="20230203T120000") + ".zip": eof,  # zipped
        syn.eof_name(*cover, produced="20230204T120000", mission="S1B"): eof,
        syn.eof_name(start - 100, stop + 500, produced="20230205T120000"): eof,  # margin short
        syn.eof_name(*cover, kind="RESORB", produced="20230112T090000"): eof,
    }
    for name, text in files.items():
        (tmp_path / name).write_text(text)
    prefix = "S1A_OPER_AUX_POEORB_OPOD_"
    assert api.orbit_pick_file(tmp_path, prefix, start - 120, stop + 120) == \
        syn.eof_name(*cover, produced="20230202T120000")
    assert api.orbit_pick_file(tmp_path, prefix, start - 600, stop + 120) is None
    assert api.orbit_pick_file(tmp_path / "absent", prefix, start, stop) is None

    t, pos, vel, source = api.orbit_best(meta, tmp_path)
    assert source == "POEORB"
    assert len(t) == 25 and t[0] >= start - 120 and t[-1] <= stop + 120
    for name in list(files):
        if "_POEORB_" in name and "S1A" in name:
            os.remove(tmp_path / name)
    assert api.orbit_best(meta, tmp_path)[3] == "RESORB"
    assert api.orbit_from_file(meta, "POEORB", tmp_path) is None
    with pytest.raises(RuntimeError, match="No POEORB orbit file of S1A covers <s0>"):
        api.orbit_from_file(meta, "POEORB", tmp_path, required=True, name="s0")
    os.remove(tmp_path / syn.eof_name(*cover, kind="RESORB", produced="20230112T090000"))
    t, pos, vel, source = api.orbit_best(meta, tmp_path)
    assert source == "annotation" and np.array_equal(pos, ann[1])
    assert api.orbit_best(meta, None)[3] == "annotation"
    d = json.loads(meta)
    del d["swath"]["mission"]
    with pytest.raises(RuntimeError, match="No mission"):
        api.orbit_best(json.dumps(d), tmp_path)


def test_scene_from_metadata(tcoh, api):
    orbits, scene, ref, t0s = _case(tcoh)
    metas = [synth.s1_metadata(o, scene, t0s[k]) for k, o in enumerate(orbits)]
    sc = api.scene_from_metadata(metas, ref, 12.0, 34.0, 200, 300)
    assert sc["t_first"] == api.parse_utc(synth.iso_utc(t0s[ref]))
    for key in ("dt", "tau0", "fs", "lambda"):
        assert sc[key] == scene[key], key
    assert (sc["line0"], sc["sample0"], sc["rows"], sc["cols"]) == (12.0, 34.0, 200, 300)
    d = json.loads(metas[0])
    d["swath"]["wavelength"] *= 1.001
    bad = [json.dumps(d)] + metas[1:]
    with pytest.raises(RuntimeError, match="different radar wavelengths"):
        api.scene_from_metadata(bad, ref, 0.0, 0.0, 200, 300)
    d = json.loads(metas[ref])
    del d["raster_geometry"]["azimuth_time_interval"]
    del d["raster_geometry"]["first_line_time"]
    bad = list(metas)
    bad[ref] = json.dumps(d)
    with pytest.raises(RuntimeError, match="azimuth_time_interval"):
        api.scene_from_metadata(bad, ref, 0.0, 0.0, 200, 300)
    d = json.loads(metas[ref])
    del d["raster_geometry"]["first_line_time"]
    bad[ref] = json.dumps(d)
    with pytest.raises(RuntimeError, match="first_line_time"):
        api.scene_from_metadata(bad, ref, 0.0, 0.0, 200, 300)
    with pytest.raises(RuntimeError, match="not one of the inputs"):
        api.scene_from_metadata(metas, 5, 0.0, 0.0, 200, 300)
