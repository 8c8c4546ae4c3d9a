"""Parity of the HIP path (through the C ABI) with the oracle, on a GPU.

SHP counts must equal the FP32 restatement of the reference kernel bit
for bit; gamma and the pair coherences must be within TOLERANCE of the
FP64 oracle, with the same NaN pattern (reference test:53-62, 309-315).
"""

import numpy as np
import pytest

import synth
from conftest import strip_of
from test_oracle import CASES, TOLERANCE

pytestmark = pytest.mark.gpu

PRECISIONS = ["plain", "f16", "bf16"]

# Bias-correction threshold band of each form: a bound on the error its
# rounding makes in m2c = (L m2 - 1) / (L - 1) (tcoh.cl:620-633), inside
# which the zero / non-zero decision on T_ij is a matter of rounding. The
# BF16 split keeps ~2^-18 of each sample, so |dT| <~ 3e-5 and |dm2c| <~
# 2 |T| |dT| L / (L - 1) ~ 1e-5 where |T|^2 ~ 1/L; FP16 keeps 2^-22 (~7e-7);
# plain FP32 sums stay below 1e-6; EMI sums in FP64 (1e-10 is far above
# their rounding). (With the splits emulated in double on the N=30 crop
# below: 1.7e-6 BF16, 1.9e-7 FP16.)
M2C_BAND = {"plain": 1e-6, "f16": 1e-6, "bf16": 1e-5, "emi": 1e-10}


def tol_for(precision):
    return TOLERANCE


def settings_for(tcoh, n, opts):
    return tcoh.make_settings(
        n, window=opts.get("window", (21, 7)), shp_test=opts.get("shp_test", "ks"),
        alpha=opts.get("alpha", 0.05), min_shp=opts.get("min_shp", 20),
        estimator=opts.get("estimator", "evd"), bias=opts.get("bias", False),
        pairs=opts.get("pairs", "consecutive"))


def all_pair_magnitudes(tcoh, n, opts, precision, slc, amp, valid, prow, pcols, rows,
                        cols, coh, count):
    """|T_ij| of every pair i < j of the same run (pairs=all): gamma and the
    SHP counts must not depend on the pair mode, bit for bit."""
    s_all = settings_for(tcoh, n, dict(opts, pairs="all"))
    with tcoh.Backend(s_all, pcols, precision=precision) as be:
        coh_a, count_a, pairs_a = be.run(slc, amp, valid, prow, rows, cols)
    assert np.array_equal(count_a, count)
    assert np.array_equal(coh_a.view(np.uint32), coh.view(np.uint32))
    return pairs_a.T.astype(np.float64)


def check_against_oracles(tcoh, cref, fp64, stack, opts, precision, tol=TOLERANCE):
    """SHP counts equal to the FP32 restatement's; gamma and the pair
    coherences within tol of the FP64 oracle at every pixel, with the same
    NaN pattern. With the bias correction the oracle takes this run's
    zero / non-zero decision where |m2c| is within the form's rounding of
    the threshold (M2C_BAND), from the all-pairs magnitudes of the run."""
    n, rows, cols = stack.shape
    s = settings_for(tcoh, n, opts)
    window = (s.win_az, s.win_rg)
    slc, amp, valid, prow, pcols = strip_of(stack, window)
    so = cref.settings(n, window=window, shp_test=opts.get("shp_test", "ks"),
                       alpha=opts.get("alpha", 0.05), min_shp=opts.get("min_shp", 20),
                       estimator=opts.get("estimator", "evd"),
                       bias=opts.get("bias", False), pairs=opts.get("pairs", "consecutive"))
    _ocoh, ocount, _opairs, mask, _ = cref.run(so, pcols, slc, amp, valid, prow, rows, cols)
    with tcoh.Backend(s, pcols, precision=precision) as be:
        coh, count, pairs = be.run(slc, amp, valid, prow, rows, cols)
    assert np.array_equal(count, ocount), "SHP counts differ from the reference kernel"
    band = M2C_BAND["emi" if s.emi else precision]
    adopt = None
    if s.bias:
        adopt = all_pair_magnitudes(tcoh, n, opts, precision, slc, amp, valid, prow, pcols,
                                    rows, cols, coh, count)
    c64, p64, _near, _m2c = fp64.run(slc, ocount, mask, cref.settings_dict(so), pcols,
                                     rows, cols, band=band, adopt=adopt)
    assert np.array_equal(np.isnan(coh), np.isnan(c64))
    both = np.isfinite(coh) & np.isfinite(c64)
    err = np.where(both, np.abs(coh - c64), 0.0)
    worst = int(np.argmax(err))
    assert err[worst] <= tol, "max |d gamma| %.3g at pixel %d" % (err[worst], worst)
    assert np.array_equal(np.isnan(pairs.T), np.isnan(p64))
    if np.isfinite(p64).any():
        perr = np.nanmax(np.abs(pairs.T - p64))
        assert perr <= tol, "max |d pair| %.3g" % perr
    return coh, count, pairs


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("case", CASES, ids=[str(i) for i in range(len(CASES))])
def test_reference_cases(tcoh, cref, fp64, case, precision):
    stack = synth.ds_stack(case["n"], case["g"], seed=case["seed"])
    check_against_oracles(tcoh, cref, fp64, stack, case["opts"], precision,
                          tol_for(precision))


# Shape edges: the group-size boundaries of the linking stage (N = 3, 4, 16
# with 8 lanes a pixel; 17, 31, 32 with 16), a 41x11 window (15 mask
# words) and a 33x31 window next to the 1024-pixel limit, across the SHP
# tests, both estimators, bias correction and both pair modes.
EDGE_CASES = [
    dict(n=3, g=0.8, seed=11, opts=dict(window=(5, 3), min_shp=3)),
    dict(n=4, g=0.6, seed=12, opts=dict(window=(3, 3), shp_test="tlog", estimator="emi",
                                        min_shp=4, bias=True)),
    dict(n=16, g=0.7, seed=13, opts=dict(window=(9, 5), shp_test="ad", min_shp=16,
                                         pairs="all")),
    dict(n=17, g=0.7, seed=14, opts=dict(window=(9, 5), estimator="emi", min_shp=17,
                                         bias=True)),
    dict(n=31, g=0.8, seed=15, opts=dict(shp_test="tlog", bias=True, pairs="consecutive")),
    dict(n=32, g=0.85, seed=16, opts=dict(estimator="emi")),
    dict(n=12, g=0.7, seed=17, opts=dict(window=(41, 11), bias=True)),
    dict(n=5, g=0.7, seed=18, opts=dict(window=(33, 31), shp_test="ad", pairs="all")),
]


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("case", EDGE_CASES,
                         ids=["n%d_%dx%d" % (c["n"], *c["opts"].get("window", (21, 7)))
                              for c in EDGE_CASES])
def test_shape_edges(tcoh, cref, fp64, case, precision):
    stack = synth.ds_stack(case["n"], case["g"], seed=case["seed"])
    coh, count, _ = check_against_oracles(tcoh, cref, fp64, stack, case["opts"],
                                          precision, tol_for(precision))
    assert np.isfinite(coh).sum() > 200


@pytest.mark.parametrize("precision", PRECISIONS)
def test_nulls_and_zero_samples(tcoh, cref, fp64, precision):
    stack = synth.ds_stack(5, 0.8, seed=6)
    stack[2, 10:13, 4:9] = np.nan
    stack[4, 20, 20] = 0
    coh, count, _ = check_against_oracles(tcoh, cref, fp64, stack, dict(min_shp=5),
                                          precision, tol_for(precision))
    coh, count = coh.reshape(30, 24), count.reshape(30, 24)
    assert np.isnan(coh[10:13, 4:9]).all() and (count[10:13, 4:9] == -1).all()
    assert np.isnan(coh[20, 20]) and count[20, 20] == -1


@pytest.mark.parametrize("precision", PRECISIONS)
def test_high_dynamic_range(tcoh, cref, fp64, precision):
    """Point scatterers 60 dB above the distributed background (amplitude
    x1000 at every date) next to a block 60 dB below it: every form holds
    the 1e-4 bar, the FP16 split through its per-pixel scale and per-window
    rescale (a bright pixel costs the windows without it no precision)."""
    stack = synth.ds_stack(12, 0.6, rows=40, cols=48, seed=8)
    stack[:, 12, 15] *= 1000.0
    stack[:, 25, 30] *= 1000.0
    stack[:, 30:33, 5:9] *= 1e-3
    stack[:, 8:11, 40] *= 1e-3
    check_against_oracles(tcoh, cref, fp64, stack,
                          dict(window=(9, 5), min_shp=5, bias=True, pairs="all"),
                          precision, tol_for(precision))


@pytest.mark.parametrize("precision", PRECISIONS)
def test_strips_are_seamless(tcoh, precision):
    """reference test:118-135: a tiny memory budget gives several strips and
    the same maps as one strip."""
    stack = synth.ds_stack(8, 0.6, rows=30, cols=400, seed=4)
    one = tcoh.temporal_coherence(stack, window=(7, 5), min_shp=10, precision=precision)
    many = tcoh.temporal_coherence(stack, window=(7, 5), min_shp=10, memory=1,
                                   precision=precision)
    s = tcoh.make_settings(8, window=(7, 5), min_shp=10)
    assert tcoh.api.strip_rows_for(s, 400, 1, 0, 30) < 30
    assert np.array_equal(one[1], many[1])
    assert np.array_equal(one[0], many[0], equal_nan=True)


def test_coherent_and_incoherent(tcoh):
    for est in ("evd", "emi"):
        means = {}
        for g in (0.95, 0.0):
            coh, _, _ = tcoh.temporal_coherence(synth.ds_stack(10, g, seed=5),
                                                estimator=est)
            means[g] = np.nanmean(coh)
        assert means[0.95] > 0.9 and means[0.0] < 0.5


@pytest.mark.parametrize("precision", PRECISIONS)
def test_small_scene(tcoh, cref, fp64, precision):
    """N=10 dates, 256 x 256, four coherence quadrants, KS, EVD, min_shp 10."""
    stack = synth.config_small(seed=0)
    check_against_oracles(tcoh, cref, fp64, stack, dict(min_shp=10), precision,
                          tol_for(precision))


@pytest.mark.parametrize("precision", PRECISIONS)
def test_depth30_scene_crop(tcoh, cref, fp64, precision):
    """N=30, Voronoi decorrelation scene crop (96 x 96), KS, EVD, min_shp 30,
    bias correction, consecutive pairs."""
    reg, params = synth.config_voronoi(n=30, rows=96, cols=96, seed=1)
    stack = synth.scene_stack(30, 96, 96, reg, params, seed=1)
    check_against_oracles(tcoh, cref, fp64, stack,
                          dict(min_shp=30, bias=True, pairs="consecutive"), precision,
                          tol_for(precision))


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("estimator", ["evd", "emi"])
@pytest.mark.parametrize("n", [33, 45, 64])
def test_deep_stacks(tcoh, cref, fp64, n, estimator, precision):
    """Stacks deeper than 32 dates (one pixel per wavefront, 64 date slots;
    N=45 and N=60-64 are the depths of the batched and full-geometry
    workloads): Voronoi scene crop, KS, bias, consecutive pairs."""
    reg, params = synth.config_voronoi(n=n, rows=36, cols=30, seed=n,
                                            gamma0=(0.3, 0.9))
    stack = synth.scene_stack(n, 36, 30, reg, params, seed=n)
    check_against_oracles(tcoh, cref, fp64, stack,
                          dict(min_shp=n, estimator=estimator, bias=True,
                               pairs="consecutive"), precision, tol_for(precision))


def emi_fallback_scene(tcoh):
    """N=30 crop of the AD+EMI workload (gamma0 ~ U(0.6, 0.99)): |T| fails
    the Cholesky floor at many pixels, which then take the Jacobi
    pseudo-inverse (tcoh.cl:657-687)."""
    reg, params = synth.config_voronoi(n=30, rows=40, cols=40, seed=0,
                                            gamma0=(0.6, 0.99))
    return synth.scene_stack(30, 40, 40, reg, params, seed=0)


@pytest.mark.parametrize("precision", ["plain", "f16"])
def test_emi_fallback_scene(tcoh, cref, fp64, precision):
    check_against_oracles(tcoh, cref, fp64, emi_fallback_scene(tcoh),
                          dict(shp_test="ad", min_shp=20, estimator="emi",
                               bias=True, pairs="all"), precision, tol_for(precision))


def test_errors_are_reported(tcoh):
    s = tcoh.make_settings(6, window=(5, 3), min_shp=6)
    with pytest.raises(RuntimeError, match="platform"):
        tcoh.Backend(s, 20, platform=-1, device=0)
    with pytest.raises(RuntimeError, match="Platform 3"):
        tcoh.Backend(s, 20, platform=3)
    with tcoh.Backend(s, 20) as be:
        with pytest.raises(RuntimeError, match="padded width"):
            be.run(np.zeros(30 * 22 * 12, np.float32), np.zeros(30 * 22 * 6, np.float32),
                   np.zeros(30 * 22, np.uint8), 30, 20, 20)


def test_geometry_grid_matches_oracle(tcoh, cref):
    """tcoh_geometry_init (FP64 on the GPU) against the FP64 C restatement
    of geometry.c:161-254: the same Newton iterations in the same
    precision, so the grids agree to far below the stopping rules' effect
    (geolocation steps < 1e-6 m: 4 pi / lambda x 1e-6 m = 2.3e-4 rad);
    asserted 1e-6 rad, on three heights and a reference date in the
    middle."""
    orbits, scene, ref = synth.geometry_case(n=6, rows=400, cols=700, seed=2)
    got, hs = tcoh.api.geometry_grid(orbits, scene, ref, 10.0, 900.0)
    want, ohs, status = cref.geometry_grid(orbits, scene, ref, 10.0, 900.0)
    assert status == 0 and got.shape == want.shape and np.array_equal(hs, ohs)
    assert np.all(got[..., ref] == 0.0)
    err = np.max(np.abs(got - want))
    assert err <= 1e-6, "max |d phase| %.3g rad" % err
    # under 1 m of relief: one height, the middle (geometry.c:204-213)
    flat, fhs = tcoh.api.geometry_grid(orbits, scene, ref, 100.0, 100.5)
    assert flat.shape[2] == 1 and fhs[0] == 100.25


def test_geometry_from_ingested_orbits(tcoh, cref, tmp_path):
    """The geometry inputs read as main.c would have them: state vectors
    from POEORB files in a local directory (orbit.c:205-318) and from the
    annotation (orbit.c:81-111), scene numbers from the reference date's
    metadata (geometry.c:165-191). The GPU grid on the file orbits equals the
    FP64 restatement on the same orbits to 1e-6 rad, and the two orbit
    sources agree to 1e-3 rad (the EOF's 1 um rounding is 4.5e-4 rad at
    C-band)."""
    api, syn = tcoh.api, synth
    api.lib().tcoh_set_error_mode(1)
    orbits, scene, ref = syn.geometry_case(n=4, rows=300, cols=500, seed=4)
    t0s = [o[0][len(o[0]) // 2] for o in orbits]
    base = [o[1][0, 1] for o in orbits]
    wide = syn.circular_orbits(t0s, base, half_span=600.0)
    metas = []
    for k in range(len(t0s)):
        metas.append(syn.s1_metadata(orbits[k], scene, t0s[k]))
        (tmp_path / syn.eof_name(t0s[k] - 599, t0s[k] + 599)).write_text(
            syn.eof_text(wide[k]))
    sc = api.scene_from_metadata(metas, ref, 0.0, 0.0, scene["rows"], scene["cols"])
    eof = [api.orbit_best(m, tmp_path) for m in metas]
    assert all(o[3] == "POEORB" for o in eof)
    eof = [o[:3] for o in eof]
    ann = [api.orbit_from_annotation(m)[:3] for m in metas]
    got, hs = api.geometry_grid(eof, sc, ref, 0.0, 600.0)
    want, ohs, status = cref.geometry_grid(eof, sc, ref, 0.0, 600.0)
    assert status == 0 and got.shape == want.shape and np.array_equal(hs, ohs)
    err = np.max(np.abs(got - want))
    assert err <= 1e-6, "max |d phase| %.3g rad" % err
    from_ann, _ = api.geometry_grid(ann, sc, ref, 0.0, 600.0)
    err = np.max(np.abs(got - from_ann))
    assert err <= 1e-3, "EOF vs annotation: max |d phase| %.3g rad" % err


def wrapped_phase_errors(ph, z64, zmag):
    """|arg(exp(i ph) conj(z64))| per date and linked pixel, leaving out the
    components whose FP64 eigenvector magnitude is within a factor 2 of the
    phasor floor (1e-3 of the largest, tcoh.cl:715-722), where taking phase
    0 or not is a matter of rounding. ph: (N, P) radians; z64, zmag: (P, N)."""
    both = np.isfinite(ph.T) & np.isfinite(z64.real)
    clear = np.isfinite(zmag) & ((zmag > 2e-3) | (zmag < 0.5e-3))
    d = np.abs(np.angle(np.exp(1j * np.where(both, ph.T, 0.0)) * np.conj(
        np.where(both, z64, 1.0))))
    return d[both & clear]


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("scene", ["n30_crop", "n10_quadrants", "emi"])
def test_linked_phases(tcoh, cref, fp64, scene, precision):
    """The linked phases (tcoh_run_phases) against the FP64 oracle's unit
    phasors: the per-date wrapped phase error, max and RMS."""
    if scene == "n30_crop":
        reg, params = synth.config_voronoi(n=30, rows=64, cols=64, seed=1)
        stack = synth.scene_stack(30, 64, 64, reg, params, seed=1)
        opts = dict(min_shp=30, bias=True, pairs="all")
    elif scene == "n10_quadrants":
        stack = synth.config_small(seed=0)[:, 96:160, 96:160]
        opts = dict(min_shp=10)
    else:
        stack = emi_fallback_scene(tcoh)
        opts = dict(shp_test="ad", min_shp=20, estimator="emi", bias=True, pairs="all")
    n, rows, cols = stack.shape
    s = settings_for(tcoh, n, opts)
    slc, amp, valid, prow, pcols = strip_of(stack, (s.win_az, s.win_rg))
    so = cref.settings(n, window=(s.win_az, s.win_rg), shp_test=opts.get("shp_test", "ks"),
                       min_shp=opts["min_shp"], estimator=opts.get("estimator", "evd"),
                       bias=opts.get("bias", False), pairs=opts.get("pairs"))
    _c, ocount, _p, mask, _ = cref.run(so, pcols, slc, amp, valid, prow, rows, cols)
    with tcoh.Backend(s, pcols, precision=precision) as be:
        coh, count, pairs, ph = be.run(slc, amp, valid, prow, rows, cols, phases=True)
        coh2, _, _ = be.run(slc, amp, valid, prow, rows, cols)
    assert np.array_equal(coh.view(np.uint32), coh2.view(np.uint32))
    assert np.array_equal(count, ocount)
    assert np.array_equal(np.isnan(ph), np.broadcast_to(np.isnan(coh), ph.shape))
    z64 = np.empty((rows * cols, n), complex)
    zmag = np.empty((rows * cols, n))
    band = M2C_BAND["emi" if s.emi else precision]
    c64, _p64, _n, _m = fp64.run(slc, ocount, mask, cref.settings_dict(so), pcols, rows,
                                 cols, band=band,
                                 adopt=pairs.T.astype(np.float64) if s.bias else None,
                                 z_out=z64, zmag_out=zmag)
    assert np.array_equal(np.isnan(coh), np.isnan(c64))
    d = wrapped_phase_errors(ph, z64, zmag)
    assert d.size > 1000
    print("linked phases %s %s: max %.3g rad, rms %.3g rad over %d" %
          (scene, precision, d.max(), np.sqrt(np.mean(d ** 2)), d.size))
    assert np.sqrt(np.mean(d ** 2)) <= 1e-4 and d.max() <= 1e-2


def test_geometry_grid_matches_exact_orbits(tcoh):
    """The GPU phase grid against the range-Doppler solution with the exact
    circular-orbit states (tests/test_host.py's model): they differ only by
    the 8-point interpolation of 10 s state vectors and the iterations'
    stopping rules, both far below a micrometre of range (2.3e-4 rad at
    C-band); asserted 1e-5 rad."""
    import math
    from test_host import _analytic_node_phases
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
    states = [exact(t0, o[1][0, 1]) for t0, o in zip(t0s, orbits)]
    ph, hs = tcoh.api.geometry_grid(orbits, scene, ref, 0.0, 400.0)
    worst = 0.0
    for gr, gc in [(0, 0), (3, 7), (ph.shape[0] - 1, ph.shape[1] - 1)]:
        for hl, h in enumerate(hs):
            want = _analytic_node_phases(states, scene, ref, gr, gc, h)
            worst = max(worst, np.max(np.abs(ph[gr, gc, hl] - want)))
    assert worst < 1e-5, worst
