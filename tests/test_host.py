"""The C ABI library and its host code, without a GPU."""

import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import synth
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "tcoh_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    # "type name(" at a declaration; "double (*pos)[3]" is a field, not one
    found = re.findall(r"\b(?:void|int|double|float|char|ocl|tcoh_[a-z_]+)\s*\**\s*"
                       r"([a-z_][a-z0-9_]*)\s*\(", text)
    return sorted(set(found))


def test_library_exports_every_declared_symbol(tcoh):
    lib = tcoh.api.lib()
    names = declared_functions()
    assert "ocl_run" in names and "ocl_open" in names
    for name in names:
        assert hasattr(lib, name), name


C_CALLER = r"""
#include <stdio.h>
#include "tcoh_hip.h"

int main(void)
{
    struct settings s = {0};

    s.ndates = 30;
    s.alpha = 0.05;
    s.test = SHP_KS;
    shp_constants(&s);
    printf("%d %.9e\n", s.ks_max, s.tlog_crit);
    return 0;
}
"""

# A caller that has the reference's types from its own header.
C_REFERENCE_TYPES = r"""
struct settings;
struct orbit;
struct geometry;
#define TCOH_HAVE_REFERENCE_TYPES
#include "tcoh_hip.h"
void caller(struct ocl *o) { ocl_close(o); }
"""


def test_c_caller_links_against_the_library(tcoh, tmp_path):
    """A C program built against include/tcoh_hip.h links and calls the
    library as main.c would (no GPU call)."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no gcc")
    lib = os.path.dirname(tcoh.api.LIB_PATH)
    src, exe = tmp_path / "caller.c", tmp_path / "caller"
    src.write_text(C_CALLER)
    subprocess.run([gcc, "-std=c11", "-Wall", "-Werror", str(src), "-I",
                    os.path.join(ROOT, "include"), "-L", lib, "-ltcoh_hip",
                    "-Wl,-rpath," + lib, "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    assert out.split() == ["10", "1.959963985e+00"]
    ref = tmp_path / "reftypes.c"
    ref.write_text(C_REFERENCE_TYPES)
    subprocess.run([gcc, "-std=c11", "-Wall", "-Werror", "-fsyntax-only", str(ref),
                    "-I", os.path.join(ROOT, "include")], check=True)


# Headers of the shape the module compiles module/geometry.c against: the
# fields of GRASS's struct Cell_head and of local_proto.h's struct epoch
# that geometry_init reads (a compile fixture, not the real headers).
GIS_H = r"""
struct Cell_head { int rows, cols; double ew_res, ns_res, north, south, east, west; };
"""
LOCAL_PROTO_H = r"""
#include "tcoh_hip.h"
#define MAX_DATES 64
struct epoch { char basename[256]; struct Cell_head head; char *meta; };
void geometry_init(struct geometry *g, const struct epoch *epochs, int n, int ref,
                   const struct orbit *orbits, const struct Cell_head *region,
                   double hmin, double hmax, double h0);
"""


def test_module_geometry_adapter_links_against_the_library(tcoh, tmp_path):
    """module/geometry.c, the drop-in for the module's geometry.c, compiles
    against headers of the module's shape with geometry_init's reference
    signature (local_proto.h:96-99) and links against the library, which
    supplies what it calls and geometry_phases / geometry_free."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no gcc")
    (tmp_path / "grass").mkdir()
    (tmp_path / "grass" / "gis.h").write_text(GIS_H)
    (tmp_path / "local_proto.h").write_text(LOCAL_PROTO_H)
    lib = os.path.dirname(tcoh.api.LIB_PATH)
    obj = tmp_path / "geometry.o"
    subprocess.run([gcc, "-std=c11", "-Wall", "-Werror", "-c",
                    os.path.join(ROOT, "i.sar.temporal_coherence_amd", "module", "geometry.c"),
                    "-I", str(tmp_path), "-I", os.path.join(ROOT, "include"), "-o", str(obj)],
                   check=True)
    so = tmp_path / "libmodule.so"
    subprocess.run([gcc, "-shared", str(obj), "-L", lib, "-ltcoh_hip", "-Wl,-rpath," + lib,
                    "-Wl,--no-undefined", "-o", str(so)], check=True)
    syms = subprocess.run(["nm", "-D", str(so)], check=True, capture_output=True,
                          text=True).stdout
    assert re.search(r" T geometry_init$", syms, re.M)
    for name in ("tcoh_scene_from_metadata", "tcoh_geometry_init"):
        assert re.search(r" U %s$" % name, syms, re.M), name
    for name in ("geometry_phases", "geometry_free"):
        assert hasattr(tcoh.api.lib(), name)


def test_settings_layout_matches_reference_struct(tcoh, cref):
    """struct settings of local_proto.h:33-49 in both bindings."""
    a, b = tcoh.api.Settings, cref.OSettings
    assert C.sizeof(a) == C.sizeof(b)
    for (fa, _), (fb, _) in zip(a._fields_, b._fields_):
        assert getattr(a, fa).offset == getattr(b, fb).offset


@pytest.mark.parametrize("n", [3, 6, 10, 30, 45, 64])
@pytest.mark.parametrize("test", ["ks", "ad", "tlog"])
def test_shp_constants_match_oracle(tcoh, cref, n, test):
    for alpha in (0.01, 0.05, 0.2):
        s = tcoh.make_settings(n, window=(21, 7), shp_test=test, alpha=alpha, min_shp=3)
        o = cref.settings(n, shp_test=test, alpha=alpha, min_shp=3)
        for f in ("ks_max", "ad_norm", "ad_sigma", "ad_crit"):
            assert getattr(s, f) == getattr(o, f), f
        # the normal quantile by two different methods (Halley's iteration
        # here, the oracle's bisection): equal to an ulp, and equal as the
        # kernel constant they become ("%.9e" then float, ocl.c:222-229)
        a, b = s.tlog_crit, o.tlog_crit
        assert abs(a - b) <= 4e-16 * b
        assert np.float32(float("%.9e" % a)) == np.float32(float("%.9e" % b))


def test_settings_derivation(tcoh):
    """main.c:585-608."""
    s = tcoh.make_settings(8, window=(5, 3), min_shp=4, pairs="all")
    assert s.min_looks == 8 and s.npairs == 28 and s.pairs_all == 1
    s = tcoh.make_settings(8, min_shp=30, pairs="consecutive")
    assert s.min_looks == 30 and s.npairs == 7
    with pytest.raises(ValueError):
        tcoh.make_settings(8, window=(4, 3))
    with pytest.raises(ValueError):
        tcoh.make_settings(40, window=(5, 5))
    assert tcoh.pair_list(4, "all")[:3] == [(0, 1), (0, 2), (0, 3)]


def test_prepare_strip_matches_oracle(tcoh, cref):
    """read_strip (main.c:52-127): validity, sorted sqrtf amplitudes and the
    double-precision rotation, bit for bit."""
    stack = synth.ds_stack(6, 0.6, rows=20, cols=17, seed=11)
    stack[1, 3:5, 2:6] = np.nan
    stack[4, 7, 7] = 0
    rng = np.random.default_rng(0)
    phase = rng.uniform(-10, 10, stack.shape)
    elev_null = np.zeros(stack.shape[1:], np.uint8)
    elev_null[12, 3] = 1
    i = np.ascontiguousarray(stack.real, np.float32)
    q = np.ascontiguousarray(stack.imag, np.float32)
    for row0, prow, halo, ph, en in ((-3, 26, 2, None, None), (5, 9, 3, phase, elev_null)):
        got = tcoh.prepare_strip(i, q, row0, prow, halo, ph, en)
        want = cref.prepare_strip(i, q, row0, prow, halo, ph, en)
        for g, w in zip(got, want):
            assert np.array_equal(g, w)


def test_orbit_state_and_geometry_phases_match_oracle(tcoh, cref):
    api = tcoh.api
    n = 13
    t = np.arange(n) * 10.0 + 1000.0
    pos = np.stack([7e6 * np.cos(t / 900.0), 1e5 + 0 * t, 7e6 * np.sin(t / 900.0)], 1)
    vel = np.gradient(pos, axis=0) / 10.0
    o = api.Orbit()
    o.n = n
    tt = np.ascontiguousarray(t)
    pp = np.ascontiguousarray(pos)
    vv = np.ascontiguousarray(vel)
    o.t = tt.ctypes.data_as(C.POINTER(C.c_double))
    o.pos = pp.ctypes.data_as(C.POINTER(C.c_double))
    o.vel = vv.ctypes.data_as(C.POINTER(C.c_double))
    for tq in (1000.0, 1033.3, 1061.0, 1120.0):
        p1, v1 = np.zeros(3), np.zeros(3)
        p2, v2 = np.zeros(3), np.zeros(3)
        api.lib().orbit_state(C.byref(o), tq, p1.ctypes.data, v1.ctypes.data)
        cref.lib().oracle_orbit_state(n, tt, pp.reshape(-1), vv.reshape(-1), tq, p2, v2)
        # the same interpolating polynomial (same 8-vector window), evaluated
        # in another form: equal to rounding (7e6 m x 1e-15 is 7 nm)
        assert np.allclose(p1, p2, rtol=1e-15, atol=1e-8)
        assert np.allclose(v1, v2, rtol=1e-15, atol=1e-11)
    # geometry_phases on a hand-made grid
    g = api.Geometry()
    g.ndates, g.grid_rows, g.grid_cols, g.step, g.nheights = 3, 4, 5, 16, 3
    for k, h in enumerate((0.0, 250.0, 500.0)):
        g.heights[k] = h
    grid = np.ascontiguousarray(rng_grid := np.random.default_rng(1).normal(
        size=4 * 5 * 3 * 3))
    g.phase = grid.ctypes.data_as(C.POINTER(C.c_double))
    hs = np.array([0.0, 250.0, 500.0])
    for row, col, h in ((0, 0, 0.0), (17, 40, 123.0), (47, 63, 480.0)):
        a, b = np.zeros(3), np.zeros(3)
        api.lib().geometry_phases(C.byref(g), row, col, h, a.ctypes.data)
        cref.lib().oracle_geometry_phases(3, 5, 3, hs, rng_grid, row, col, h, b)
        assert np.allclose(a, b, rtol=1e-12, atol=1e-12)


def _analytic_node_phases(orbits_exact, scene, ref, gr, gc, h):
    """Phases of grid node (gr, gc) at height h with the exact orbit states:
    geolocation on the reference orbit at the node's azimuth time (Doppler,
    range, ellipsoid), zero-Doppler time of every date, 4 pi / lambda times
    the range difference to the reference date."""
    a_e, b_e = 6378137.0 + h, 6356752.314245 + h
    line = scene["line0"] + 16.0 * gr
    sample = scene["sample0"] + 16.0 * gc
    t = scene["t_first"] + line * scene["dt"]
    rng = 0.5 * 299792458.0 * (scene["tau0"] + sample / scene["fs"])
    s, v = orbits_exact[ref](t)
    rs = np.linalg.norm(s)
    p = s * (a_e / rs)
    p[1] += math.sqrt(max(rng ** 2 - (rs - a_e) ** 2, 0.0))
    for _ in range(100):
        d = p - s
        dist = np.linalg.norm(d)
        f = np.array([v @ d, dist - rng,
                      (p[0] ** 2 + p[1] ** 2) / a_e ** 2 + p[2] ** 2 / b_e ** 2 - 1.0])
        jac = np.array([v, d / dist, [2 * p[0] / a_e ** 2, 2 * p[1] / a_e ** 2,
                                      2 * p[2] / b_e ** 2]])
        step = np.linalg.solve(jac, -f)
        p = p + step
        if np.max(np.abs(step)) < 1e-9:
            break
    ranges = []
    for state in orbits_exact:
        tz = t
        for _ in range(50):
            s, v = state(tz)
            ds = 1e-3
            s2, v2 = state(tz + ds)
            f = v @ (p - s)
            df = (v2 @ (p - s2) - f) / ds
            tz -= f / df
            if abs(f / df) < 1e-12:
                break
        s, _ = state(tz)
        ranges.append(np.linalg.norm(p - s))
    ranges = np.array(ranges)
    return 4.0 * math.pi / scene["lambda"] * (ranges - ranges[ref])


def test_geometry_oracle_matches_exact_orbits(tcoh, cref):
    """The FP64 geometry restatement (geometry.c:31-254, orbit.c:324-346)
    against the same range-Doppler solution with exact circular-orbit
    states (the model of tests/geometry_model.py): the two differ only by
    the 8-point Lagrange interpolation of 10 s state vectors and the Newton
    stopping rules, together well below a micrometre of range on a circular
    orbit (4 pi / lambda x 1 um = 2.3e-4 rad at C-band); asserted: 1e-5 rad."""
    orbits, scene, ref = synth.geometry_case(n=4, rows=100, cols=200, seed=5)
    radius, omega = 6378137.0 + 700e3, 7500.0 / (6378137.0 + 700e3)

    def exact(t0, b):
        def state(t):
            th = math.radians(25.0) + omega * (t - t0)
            return (np.array([radius * math.cos(th), b, radius * math.sin(th)]),
                    np.array([-radius * omega * math.sin(th), 0.0,
                              radius * omega * math.cos(th)]))
        return state

    t0s = [o[0][len(o[0]) // 2] for o in orbits]
    bs = [o[1][0, 1] for o in orbits]
    states = [exact(t0, b) for t0, b in zip(t0s, bs)]
    ph, hs, status = cref.geometry_grid(orbits, scene, ref, 0.0, 400.0)
    assert status == 0 and ph.shape == (100 // 16 + 2, 200 // 16 + 2, 3, 4)
    worst = 0.0
    for gr, gc in [(0, 0), (3, 7), (ph.shape[0] - 1, ph.shape[1] - 1)]:
        for hl, h in enumerate(hs):
            want = _analytic_node_phases(states, scene, ref, gr, gc, h)
            worst = max(worst, np.max(np.abs(ph[gr, gc, hl] - want)))
    assert worst < 1e-5, worst
    # the phases vary with the baseline: a real geometric signal, not zeros
    assert np.ptp(ph[0, 0, 0]) > 100.0


def test_pixel_phases_match_the_library(tcoh, cref):
    """The oracle's per-cell phase planes (the checker of the device strip
    preparation) against the library's host geometry_phases
    (geometry.c:256-291), bit for bit, at elevation heights and at h0."""
    rng = np.random.default_rng(8)
    n, rows, cols = 4, 19, 35
    grid = rng.normal(0.0, 30.0, ((rows - 1) // 16 + 2, (cols - 1) // 16 + 2, 3, n))
    hs = np.array([-10.0, 240.0, 490.0])
    g = tcoh.api.make_geometry(grid, hs, h0=77.0)
    elev = rng.uniform(0.0, 500.0, (rows, cols))
    elev[3, 4] = np.nan
    for e in (elev, None):
        got = cref.pixel_phases(grid, hs, rows, cols, elevation=e, h0=77.0)
        for r, c in ((0, 0), (5, 17), (18, 34), (3, 4)):
            h = 77.0 if e is None else e[r, c]
            want = np.zeros(n)
            if not np.isnan(h):
                tcoh.api.lib().geometry_phases(C.byref(g), r, c, h, want.ctypes.data)
            assert np.allclose(got[:, r, c], want, rtol=1e-12, atol=1e-12)
            assert np.isnan(h) == (not want.any())


def test_geometry_phases_reproduce_tensor_polynomials(tcoh, cref):
    """Bilinear interpolation in (row, col) times quadratic interpolation in
    height (geometry.c:256-291) reproduces exactly, up to rounding, any
    field f(r, c, h) = (a + b r + c c + d r c) (e + f h + g h^2) sampled on
    the 16-pixel nodes at the three heights, at every pixel and height:
    checked for the oracle's restatement and the library's own code. With
    one height level the field is the level's value."""
    rng = np.random.default_rng(9)
    n, rows, cols, step = 3, 40, 50, 16
    gr, gcn = (rows - 1) // step + 2, (cols - 1) // step + 2
    hs = np.array([12.0, 260.0, 505.0])
    co = rng.normal(size=(n, 7))

    def f(r, c, h):
        a, b, cc, d, e, ff, g = co.T
        return (a + b * r + cc * c + d * r * c) * (e + ff * h + g * h * h)

    grid = np.zeros((gr, gcn, 3, n))
    for i in range(gr):
        for j in range(gcn):
            for lvl in range(3):
                grid[i, j, lvl] = f(i * step, j * step, hs[lvl])
    flat = np.ascontiguousarray(grid.reshape(-1))
    g = tcoh.api.make_geometry(grid, hs)
    for r, c, h in ((0, 0, 12.0), (17, 33, 300.0), (39, 49, -40.0), (31, 2, 505.0),
                    (8, 24, 777.0)):
        want = f(r, c, h)
        a, b = np.zeros(n), np.zeros(n)
        tcoh.api.lib().geometry_phases(C.byref(g), r, c, h, a.ctypes.data)
        cref.lib().oracle_geometry_phases(n, gcn, 3, hs, flat, r, c, h, b)
        scale = np.max(np.abs(grid))
        assert np.allclose(a, want, rtol=0, atol=1e-12 * scale)
        assert np.allclose(b, want, rtol=0, atol=1e-12 * scale)
    one = np.ascontiguousarray(grid[:, :, :1])
    g1 = tcoh.api.make_geometry(one, hs[:1])
    a = np.zeros(n)
    tcoh.api.lib().geometry_phases(C.byref(g1), 17, 33, 999.0, a.ctypes.data)
    u, v = 1 / 16, 1 / 16
    want = ((1 - u) * (1 - v) * one[1, 2, 0] + (1 - u) * v * one[1, 3, 0] +
            u * (1 - v) * one[2, 2, 0] + u * v * one[2, 3, 0])
    assert np.allclose(a, want, rtol=1e-14)


def test_linking_device_functions_are_inlined():
    """Every device function of the linking kernels' headers is forced
    inline. These functions take references into the group's LDS and the
    lane's private arrays; an out-of-line EMI inversion routine once
    received them as flat pointers and its accesses faulted the GPU
    (memory aperture violation in the N=30 EMI linking kernel), so no
    function there may become a call again."""
    csrc = os.path.join(ROOT, "i.sar.temporal_coherence_amd", "csrc")
    bad = []
    for name in ("common.h", "cov.h", "eig.h", "link.h", "refine.h", "k_link.hip"):
        text = re.sub(r"/\*.*?\*/", "", open(os.path.join(csrc, name)).read(), flags=re.S)
        # a declaration: template header (optional) then qualifiers up to the name's "("
        for m in re.finditer(r"((?:__host__\s+)?__device__[^;{()]*?)\b([A-Za-z_]\w*)\s*\(",
                             text):
            quals = m.group(1)
            if "__global__" in quals or "__shared__" in quals:
                continue
            if "__forceinline__" not in quals:
                bad.append("%s: %s" % (name, m.group(2)))
    assert not bad, bad
