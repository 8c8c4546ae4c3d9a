"""Pins of the CPU oracle (no GPU): known answers, scipy, closed forms,
brute force, and the FP32 restatement against the FP64 oracle."""

import math

import numpy as np
import pytest

import synth
from conftest import strip_of

TOLERANCE = 1e-4  # largest |d gamma|, |d pair| vs FP64 (reference test:27)

# Appendix values computed independently from stats.c's closed forms.
KAT = {
    8: (5, 1.562500000e-02, 6.812341293e-01),
    10: (6, 1.000000000e-02, 6.962061929e-01),
    12: (6, 6.944444444e-03, 7.063865824e-01),
    20: (8, 2.500000000e-03, 7.273664994e-01),
    30: (10, 1.111111111e-03, 7.382626231e-01),
    45: (12, 4.938271605e-04, 7.457303009e-01),
    60: (14, 2.777777778e-04, 7.495401734e-01),
}


@pytest.mark.parametrize("n", sorted(KAT))
def test_shp_constants_kat(cref, n):
    s = cref.settings(n, shp_test="ad")
    ks, norm, sigma = KAT[n]
    assert s.ks_max == ks
    assert "%.9e" % s.ad_norm == "%.9e" % norm
    assert "%.9e" % s.ad_sigma == "%.9e" % sigma
    assert s.ad_crit == pytest.approx(1.961, abs=1e-12)
    assert "%.9e" % s.tlog_crit == "1.959963985e+00"


def test_constants_match_scipy(cref):
    stats = pytest.importorskip("scipy.stats")
    L = cref.lib()
    for alpha in (0.001, 0.01, 0.05, 0.1, 0.25):
        assert L.oracle_normal_quantile(alpha) == pytest.approx(
            stats.norm.isf(alpha / 2), rel=1e-12)
    # Table points of Scholz & Stephens as scipy interpolates them.
    assert L.oracle_ad_critical(0.05) == pytest.approx(1.961)
    assert L.oracle_ad_critical(0.01) == pytest.approx(3.752)
    assert math.isnan(L.oracle_ad_critical(0.3))
    for n in (4, 12, 30):
        crit = math.sqrt(-0.5 * math.log(0.025)) * math.sqrt(2.0 / n)
        assert cref.settings(n).ks_max == math.floor(crit * n)


@pytest.mark.filterwarnings("ignore:p-value capped")
@pytest.mark.filterwarnings("ignore:p-value floored")
def test_shp_statistics_match_scipy(cref):
    """reference test:30-46, on the FP32 restatement."""
    stats = pytest.importorskip("scipy.stats")
    L = cref.lib()
    rng = np.random.default_rng(1)
    n = 12
    s = cref.settings(n, shp_test="ad")
    norm, sig = L.oracle_kernel_float(s.ad_norm), L.oracle_kernel_float(s.ad_sigma)
    for _ in range(200):
        a = np.sort(rng.rayleigh(size=n)).astype(np.float32)
        b = np.sort(rng.rayleigh(scale=1.5, size=n)).astype(np.float32)
        expected = stats.anderson_ksamp([a.astype(np.float64), b.astype(np.float64)],
                                        midrank=False).statistic
        got = L.oracle_ad_statistic(n, norm, sig, a, b)
        assert got == pytest.approx(expected, rel=2e-5, abs=2e-5)
        d = stats.ks_2samp(a, b).statistic
        assert L.oracle_ks_accept(n, s.ks_max, a, b) == int(round(d * n) <= s.ks_max)


def ks_closed_form(a, b, k):
    """The GPU kernel's formulation: for i >= K, b[i-K] < a[i], a[i-K] <= b[i]."""
    if k < 1:
        return 0
    for i in range(k, len(a)):
        if not (b[i - k] < a[i] and a[i - k] <= b[i]):
            return 0
    return 1


def test_ks_closed_form_equals_merge_walk(cref):
    L = cref.lib()
    rng = np.random.default_rng(7)
    for trial in range(4000):
        n = int(rng.integers(3, 65))
        k = int(rng.integers(0, n + 1))
        if trial % 3 == 0:  # heavy ties
            a = np.sort(rng.integers(0, 4, n)).astype(np.float32)
            b = np.sort(rng.integers(0, 4, n)).astype(np.float32)
        else:
            a = np.sort(rng.rayleigh(size=n)).astype(np.float32)
            b = np.sort(rng.rayleigh(scale=rng.uniform(0.6, 1.6), size=n)).astype(np.float32)
        assert L.oracle_ks_accept(n, k, a, b) == ks_closed_form(a, b, k)
    # exhaustive: every pair of sorted 0/1/2 samples of length 4
    vals = [np.array(v, np.float32) for v in np.ndindex(3, 3, 3, 3)]
    vals = [np.sort(v) for v in vals]
    for a in vals:
        for b in vals:
            for k in range(0, 5):
                assert L.oracle_ks_accept(4, k, a, b) == ks_closed_form(a, b, k)


def test_tlog_matches_double_welch(cref):
    """Welch t on the sorted log amplitudes (tcoh.cl:77-95) vs float64."""
    L = cref.lib()
    rng = np.random.default_rng(3)
    crit = cref.settings(10, alpha=0.01).tlog_crit
    cf = L.oracle_kernel_float(crit)
    agree = 0
    for _ in range(500):
        a = np.sort(rng.rayleigh(size=10)).astype(np.float32)
        b = np.sort(rng.rayleigh(scale=rng.uniform(0.7, 1.5), size=10)).astype(np.float32)
        la, lb = np.log(a.astype(np.float64)), np.log(b.astype(np.float64))
        se = math.sqrt((la.var(ddof=1) + lb.var(ddof=1)) / 10)
        t = abs(la.mean() - lb.mean()) / se
        if abs(t - crit) < 1e-4:
            continue
        assert L.oracle_tlog_accept(10, cf, a, b) == int(t <= crit)
        agree += 1
    assert agree > 450


def test_logf_is_rounded_double_log(cref):
    L = cref.lib()
    for x in (1.0, 2.5, 1e-20, 3.0e30, 7.0):
        assert L.oracle_logf(x) == np.float32(math.log(np.float32(x)))


# ---- FP64 oracle pins -------------------------------------------------------

def test_fp64_gamma_is_one_for_consistent_phases(fp64):
    """T = Gamma o u u^H, Gamma real positive definite, |u_i| = 1: the EVD
    eigenvector is D_u times Gamma's (positive, Perron) top eigenvector, and
    by Fiedler's inequality Gamma^-1 o Gamma >= I with eigenvector 1, so
    EMI returns u; both give gamma = 1 and the phases of u."""
    rng = np.random.default_rng(0)
    n = 7
    t_days = 12.0 * np.arange(n)
    gam = 0.6 * np.exp(-np.abs(t_days[:, None] - t_days[None, :]) / 48.0) + 0.3
    np.fill_diagonal(gam, 1.0)
    u = np.exp(1j * rng.uniform(-np.pi, np.pi, (5, n)))
    t = gam[None] * u[:, :, None] * u[:, None, :].conj()
    want = u * np.conj(u[:, n // 2])[:, None]
    for emi in (False, True):
        z = fp64.estimate(t, emi=emi)
        assert np.allclose(fp64.gamma(t, z), 1.0, atol=1e-9)
        assert np.allclose(z, want, atol=1e-9)


def test_fp64_phasor_conventions(fp64):
    # datum = date N/2; a negligible datum falls back to the LAST largest
    u = np.array([[1 + 0j, 1j, 2.0, 1e-6, -1j]])
    z = fp64.phasors(u)
    assert np.allclose(z[0, 2], 1.0)
    assert np.allclose(z[0, 1], 1j)
    u = np.array([[2 + 0j, 1j, 1e-9, 2j, 0.5]])
    z = fp64.phasors(u)           # |u_2| tiny -> datum = last of |u| == 2 (index 3)
    assert np.allclose(z[0, 3], 1.0)
    assert np.allclose(z[0, 0], -1j)
    assert z[0, 2] == 1.0         # negligible component -> phasor 1


def test_fp64_gamma_global_phase_invariance(fp64):
    rng = np.random.default_rng(2)
    n, p = 6, 50
    x = rng.standard_normal((p, 40, n)) + 1j * rng.standard_normal((p, 40, n))
    c = np.einsum("pki,pkj->pij", x, x.conj())
    t = fp64.coherence(c, np.full(p, 40), bias=False)
    z = fp64.estimate(t, emi=False)
    g1 = fp64.gamma(t, z)
    g2 = fp64.gamma(t, fp64.phasors(z * np.exp(0.7j)))
    assert np.allclose(g1, g2, atol=1e-12)
    # EVD eigenvector against power iteration on T + 2I
    v = np.ones((p, n), complex)
    for _ in range(3000):
        v = np.einsum("pij,pj->pi", t + 2 * np.eye(n), v)
        v /= np.linalg.norm(v, axis=1, keepdims=True)
    assert np.allclose(fp64.phasors(v), z, atol=1e-6)


def test_fp64_coherence_and_bias_closed_form(fp64):
    c = np.array([[[4.0, 2 + 2j], [2 - 2j, 9.0]]])
    t = fp64.coherence(c, np.array([10]), bias=False)
    assert np.allclose(t[0, 0, 1], (2 + 2j) / 6.0)
    tb = fp64.coherence(c, np.array([10]), bias=True)
    m2 = abs((2 + 2j) / 6.0) ** 2
    m2c = (10 * m2 - 1) / 9
    assert np.allclose(tb[0, 0, 1], (2 + 2j) / 6.0 * math.sqrt(m2c / m2))
    # small coherence with few looks: zeroed by the correction
    c2 = np.array([[[1.0, 0.1], [0.1, 1.0]]])
    assert fp64.coherence(c2, np.array([5]), bias=True)[0, 0, 1] == 0


def test_fp64_connected(fp64):
    t = np.eye(4, dtype=complex)[None].repeat(2, 0)
    t[0, 0, 1] = t[0, 1, 0] = 0.3
    t[0, 1, 2] = t[0, 2, 1] = 0.2j
    t[0, 3, 2] = t[0, 2, 3] = 0.1
    t[1, 0, 1] = t[1, 1, 0] = 0.3
    t[1, 2, 3] = t[1, 3, 2] = 0.5
    assert list(fp64.connected(t)) == [True, False]


def test_fp64_emi_inverse_branches(fp64):
    g = np.array([[[1.0, 0.5], [0.5, 1.0]], [[1.0, 1.0], [1.0, 1.0]]]) + 0j
    inv, ok = fp64.abs_inverse(g)
    assert list(ok) == [True, False]
    assert np.allclose(inv[0], np.linalg.inv(g[0].real))
    # singular: pseudo-inverse of [[1,1],[1,1]] = [[.25,.25],[.25,.25]]
    assert np.allclose(inv[1], 0.25)


# ---- FP32 restatement of the kernel against the FP64 oracle -----------------

CASES = [
    dict(n=6, g=0.7, seed=0, opts={}),
    dict(n=6, g=0.7, seed=0, opts=dict(shp_test="ad", estimator="emi", min_shp=9, window=(9, 5))),
    dict(n=6, g=0.7, seed=0, opts=dict(shp_test="tlog", alpha=0.01, min_shp=9, window=(7, 5))),
    dict(n=6, g=0.3, seed=3, opts=dict(bias=True, estimator="evd", window=(5, 3), min_shp=6)),
    dict(n=6, g=0.3, seed=3, opts=dict(bias=True, estimator="emi", window=(5, 3), min_shp=6)),
    dict(n=5, g=0.7, seed=20, opts=dict(min_shp=5, pairs="all")),
    dict(n=10, g=0.95, seed=5, opts={}),
]


def run_oracles(cref, fp64, tcoh_synth, case):
    n, opts = case["n"], dict(case["opts"])
    stack = tcoh_synth.ds_stack(n, case["g"], seed=case["seed"])
    window = opts.get("window", (21, 7))
    s = cref.settings(n, window=window, shp_test=opts.get("shp_test", "ks"),
                      alpha=opts.get("alpha", 0.05), min_shp=opts.get("min_shp", 20),
                      estimator=opts.get("estimator", "evd"), bias=opts.get("bias", False),
                      pairs=opts.get("pairs", "consecutive"))
    slc, amp, valid, prow, pcols = strip_of(stack, window)
    rows, cols = stack.shape[1:]
    coh, count, pairs, mask, flips = cref.run(s, pcols, slc, amp, valid, prow, rows, cols)
    c64, p64 = fp64.run(slc, count, mask, cref.settings_dict(s), pcols, rows, cols)
    return s, coh, count, pairs, c64, p64, flips


@pytest.mark.parametrize("case", CASES, ids=[str(i) for i in range(len(CASES))])
def test_fp32_restatement_within_tolerance_of_fp64(cref, fp64, tcoh, case):
    s, coh, count, pairs, c64, p64, flips = run_oracles(cref, fp64, synth, case)
    assert flips == 0
    assert np.array_equal(np.isnan(coh), np.isnan(c64))
    assert np.nanmax(np.abs(coh - c64)) <= TOLERANCE
    assert np.array_equal(np.isnan(pairs.T), np.isnan(p64))
    assert np.nanmax(np.abs(pairs.T - p64)) <= TOLERANCE
    assert count.max() <= s.win_az * s.win_rg


def test_bias_unlinked_dates_give_nan_with_finite_pairs(cref, fp64, tcoh):
    """reference test:357-378: with -b, unlinked dates -> NaN gamma, pairs
    still defined."""
    s, coh, count, pairs, c64, p64, _ = run_oracles(cref, fp64, synth, CASES[3])
    unlinked = np.isnan(coh) & (count >= s.min_looks)
    assert unlinked.any()
    assert np.isfinite(pairs[0][unlinked]).all()


def test_bias_pairs_are_sqrt_m2c(cref, fp64, tcoh):
    """Debiased |T_ij| = sqrt(max(m2c, 0)) (tcoh.cl:620-633): the closed form
    the parity tests use to bound pair errors near the threshold."""
    case = CASES[3]
    n, opts = case["n"], case["opts"]
    stack = synth.ds_stack(n, case["g"], seed=case["seed"])
    s = cref.settings(n, window=opts["window"], bias=True, min_shp=opts["min_shp"],
                      pairs="all")
    slc, amp, valid, prow, pcols = strip_of(stack, opts["window"])
    rows, cols = stack.shape[1:]
    _c, count, _p, mask, _f = cref.run(s, pcols, slc, amp, valid, prow, rows, cols)
    _c64, p64, near, m2c = fp64.run(slc, count, mask, cref.settings_dict(s), pcols,
                                    rows, cols, band=1e-3)
    fin = np.isfinite(p64)
    assert fin.any() and (m2c[fin] <= 0).any() and (m2c[fin] > 0).any()
    assert np.allclose(p64[fin], np.sqrt(np.maximum(m2c[fin], 0)), rtol=0, atol=1e-12)
    assert np.array_equal(near, np.sum(np.abs(np.nan_to_num(m2c, nan=1.0)) < 1e-3, axis=1))


def test_gamma_step_at_a_zeroed_coherence(fp64):
    """A coherence that the bias correction zeroes counts as unit(0) = 1 in
    gamma (tcoh.cl:726-744): moving one entry from a tiny value to zero
    changes gamma by at most 2 / (N(N-1)/2), the allowance of the parity
    tests at the threshold."""
    rng = np.random.default_rng(8)
    n = 12
    z = np.exp(1j * rng.uniform(-np.pi, np.pi, n))
    t = 0.6 * np.outer(z, z.conj()) + 0.4 * np.eye(n)
    worst = 0.0
    for i, j in [(0, 1), (3, 7), (10, 11)]:
        for ph in np.linspace(-np.pi, np.pi, 9):
            a, b = t.copy(), t.copy()
            a[i, j] = 1e-9 * np.exp(1j * ph)
            b[i, j] = 0.0
            a[j, i], b[j, i] = np.conj(a[i, j]), 0.0
            g = fp64.gamma(np.stack([a, b]), np.stack([z, z]))
            worst = max(worst, abs(g[0] - g[1]))
    assert 1.0 / (n * (n - 1) / 2) < worst <= 2.0 / (n * (n - 1) / 2) + 1e-12


def test_coherent_and_incoherent(cref, fp64, tcoh):
    """reference test:138-151."""
    means = {}
    for g in (0.95, 0.0):
        stack = synth.ds_stack(10, g, seed=5)
        s = cref.settings(10)
        slc, amp, valid, prow, pcols = strip_of(stack, (21, 7))
        coh, *_ = cref.run(s, pcols, slc, amp, valid, prow, 30, 24)
        means[g] = np.nanmean(coh)
    assert means[0.95] > 0.9 and means[0.0] < 0.5


def test_nulls_and_zero_samples(cref, tcoh):
    """reference test:154-169."""
    stack = synth.ds_stack(5, 0.8, seed=6)
    stack[2, 10:13, 4:9] = np.nan
    stack[4, 20, 20] = 0
    s = cref.settings(5, min_shp=5)
    slc, amp, valid, prow, pcols = strip_of(stack, (21, 7))
    coh, count, *_ = cref.run(s, pcols, slc, amp, valid, prow, 30, 24)
    coh, count = coh.reshape(30, 24), count.reshape(30, 24)
    assert np.isnan(coh[10:13, 4:9]).all() and (count[10:13, 4:9] == -1).all()
    assert np.isnan(coh[20, 20]) and count[20, 20] == -1
    assert np.isfinite(coh[25, 5])


def test_emi_fallback_scene_reaches_the_jacobi_path(cref, fp64, tcoh):
    """The GPU test scene of the EMI fallback really holds pixels whose |T|
    fails the Cholesky floor (so the Jacobi pseudo-inverse is exercised),
    and others that pass it."""
    reg, params = synth.config_voronoi(n=30, rows=40, cols=40, seed=0,
                                            gamma0=(0.6, 0.99))
    stack = synth.scene_stack(30, 40, 40, reg, params, seed=0)
    slc, amp, valid, prow, pcols = strip_of(stack, (21, 7))
    s = cref.settings(30, shp_test="ad", min_shp=20, estimator="emi", bias=True,
                      pairs="all")
    _c, count, _p, mask, _f = cref.run(s, pcols, slc, amp, valid, prow, 40, 40)
    est = np.flatnonzero(count >= s.min_looks)
    y, x = est // 40, est % 40
    v = np.asarray(slc, np.float32).reshape(-1, 30, 2)
    sc = v[..., 0].astype(np.float64) + 1j * v[..., 1]
    c = fp64.covariance(sc, fp64.mask_bits(mask[est], 147), y * pcols + x,
                        fp64.window_offsets(21, 7, pcols))
    t = fp64.coherence(c, count[est], True)
    with np.errstate(all="ignore"):
        _piv, ok = fp64.cholesky_pivots(np.abs(t))
    assert 0.05 < np.mean(~ok) < 0.6, np.mean(~ok)


def test_fp64_split_forms(fp64):
    """The mixed forms' covariance in the FP64 oracle: round to nearest even
    (BF16 ties to even by bits, FP16 by numpy's IEEE conversion); FP16
    scales each pixel so that its largest component lies in [2^13, 2^14);
    hi + lo within the form's precision of the (scaled) sample; and the
    form's C equal to the sum of the hi.hi, hi.lo and lo.hi products of the
    real Gram matrices (C_re = Re Re^T + Im Im^T, C_im = Im Re^T - Re Im^T)
    of the candidates rescaled to the window's largest exponent."""
    # 1 + 2^-8 is halfway between BF16 neighbours 1 and 1 + 2^-7: to even
    x = np.array([1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, -2.5, 0.0], np.float32)
    assert list(fp64.round_bf16(x)) == [1.0, 1 + 2.0 ** -6, -2.5, 0.0]
    rng = np.random.default_rng(3)
    n, k = 7, 40
    xs = (rng.normal(size=(k, n, 2)) * 300).astype(np.float32)
    # one pixel 60 dB (x1000) brighter, one 2^-20 dimmer: other exponents
    xs[5] *= 1000.0
    xs[9] *= 2.0 ** -20
    for form, prec in (("f16", 2.0 ** -21), ("bf16", 2.0 ** -15)):
        hi, lo, e = fp64.split_parts(xs, form)
        scale = np.ldexp(1.0, 14 - e) if form == "f16" else np.ones(k)
        xsc = xs * scale[:, None, None]
        if form == "f16":
            top = np.abs(hi).reshape(k, -1).max(axis=1)
            assert np.all((top >= 2.0 ** 13) & (top <= 2.0 ** 14))
        assert np.all(np.abs(hi + lo - xsc) <= prec * np.abs(xsc) + 1e-30)
        acc = np.ones((1, k), bool)
        c = fp64.form_covariance((hi, lo, e), acc, np.array([0]), np.arange(k), form)[0]
        if form == "f16":
            # v_pk_mul_f16 by 2^-shift: FP16 results
            f = np.ldexp(1.0, -(e.max() - e))[:, None, None]
            h16 = (hi * f).astype(np.float16).astype(np.float64)
            l16 = (lo * f).astype(np.float16).astype(np.float64)
            rh, ih, rl, il = h16[..., 0], h16[..., 1], l16[..., 0], l16[..., 1]
        else:
            rh, ih, rl, il = hi[..., 0], hi[..., 1], lo[..., 0], lo[..., 1]

        def gram(a_h, a_l, b_h, b_l):
            return a_h.T @ b_h + a_h.T @ b_l + a_l.T @ b_h

        cre = gram(rh, rl, rh, rl) + gram(ih, il, ih, il)
        cim = gram(ih, il, rh, rl) - gram(rh, rl, ih, il)
        tol = 1e-9 * np.abs(cre).max()
        assert np.allclose(c.real, cre, rtol=1e-12, atol=tol)
        assert np.allclose(c.imag, cim, rtol=1e-12, atol=tol)
        # and the form's T close to the exact one despite the 60 dB spread
        x64 = xs[..., 0].astype(np.float64) + 1j * xs[..., 1]
        c64 = fp64.covariance(x64, acc, np.array([0]), np.arange(k))[0]
        t64 = fp64.coherence(c64[None], np.array([k]), False)[0]
        tf = fp64.coherence(c[None], np.array([k]), False)[0]
        assert np.max(np.abs(tf - t64)) < (1e-6 if form == "f16" else 1e-4)


def test_fp16_split_depends_on_the_window_only(fp64):
    """Under the FP16 form a window's covariance does not depend on samples
    outside it: a pixel elsewhere of any finite magnitude (here 1e30 and
    1e-30) leaves its T bit for bit unchanged, as in the reference's FP32
    arithmetic."""
    rng = np.random.default_rng(4)
    xs = (rng.normal(size=(30, 6, 2)) * 50).astype(np.float32)
    acc = np.ones((1, 20), bool)
    base = fp64.form_covariance(fp64.split_parts(xs, "f16"), acc, np.array([0]),
                                np.arange(20), "f16")
    for v in (1e30, 1e-30):
        ys = xs.copy()
        ys[25] = v
        other = fp64.form_covariance(fp64.split_parts(ys, "f16"), acc, np.array([0]),
                                     np.arange(20), "f16")
        assert np.array_equal(base, other)


def test_prepare_strip_against_numpy(cref):
    """oracle_prepare_strip (main.c:52-127) against a numpy statement of the
    same contract: rows outside the region and the halo invalid; a pixel
    valid when every date is non-null and not (0, 0) and its height is not
    null; amplitudes float32 sqrt(I^2 + Q^2), each operation rounded to
    float32, sorted increasingly, taken before the rotation; samples times
    exp(+i phase) in double, then rounded to float32; invalid pixels all
    zero."""
    rng = np.random.default_rng(12)
    n, nrows, ncols, halo = 6, 11, 13, 3
    i = rng.normal(0, 50, (n, nrows, ncols)).astype(np.float32)
    q = rng.normal(0, 50, (n, nrows, ncols)).astype(np.float32)
    i[2, 4, 5] = np.nan
    i[1, 6, 7] = q[1, 6, 7] = 0.0
    q[3, 2, 2] = np.nan
    phase = rng.uniform(-40.0, 40.0, (n, nrows, ncols))
    en = np.zeros((nrows, ncols), np.uint8)
    en[8, 1] = 1
    row0, prow = -2, 15
    pc = ncols + 2 * halo
    for ph, el in ((None, None), (phase, en)):
        slc, amp, valid = cref.prepare_strip(i, q, row0, prow, halo, ph, el)
        slc = slc.reshape(prow, pc, n, 2)
        amp = amp.reshape(prow, pc, n)
        valid = valid.reshape(prow, pc)
        want_v = np.zeros((prow, pc), np.uint8)
        want_s = np.zeros((prow, pc, n, 2), np.float32)
        want_a = np.zeros((prow, pc, n), np.float32)
        for r in range(prow):
            row = row0 + r
            if not 0 <= row < nrows:
                continue
            vi, vq = i[:, row, :], q[:, row, :]  # (n, ncols)
            ok = ~(np.isnan(vi) | np.isnan(vq) | ((vi == 0) & (vq == 0)))
            ok = ok.all(axis=0)
            if el is not None:
                ok &= el[row] == 0
            a = np.sqrt((vi * vi + vq * vq).astype(np.float32)).astype(np.float32)
            a = np.sort(a, axis=0)
            re, im = vi.astype(np.float64), vq.astype(np.float64)
            if ph is not None:
                cs, sn = np.cos(ph[:, row, :]), np.sin(ph[:, row, :])
                re, im = re * cs - im * sn, re * sn + im * cs
            for c in np.flatnonzero(ok):
                want_v[r, halo + c] = 1
                want_a[r, halo + c] = a[:, c]
                want_s[r, halo + c, :, 0] = re[:, c].astype(np.float32)
                want_s[r, halo + c, :, 1] = im[:, c].astype(np.float32)
        assert np.array_equal(valid, want_v)
        assert valid.sum() > 100 and valid.sum() < nrows * ncols
        assert np.array_equal(amp, want_a)
        # numpy's and the C library's double sin/cos may differ in the last
        # bit; after rounding to float32 that shows at most as one float ulp
        d = np.abs(slc.view(np.int32).astype(np.int64) - want_s.view(np.int32))
        assert d.max() <= 1 and np.mean(d == 0) > 0.999
