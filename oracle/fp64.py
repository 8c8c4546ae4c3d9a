"""FP64 oracle of the phase-linking estimate -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's accuracy report import
this module, as the checker; the product path never does.

Given the SHP decisions of each output pixel (taken from the FP32 C
restatement of the reference kernel, whose KS/AD/tlog decisions are exact
integer or IEEE-float logic on the float inputs), everything downstream is
computed in double precision from the float32 samples:

  covariance and coherence matrix with optional bias correction
      tcoh.cl:567-636 (reference.py:124-142 divides by looks; T is the same)
  pair coherences |T_ij|                  tcoh.cl:515-527, 637-639
  connectivity of the non-zero coherences tcoh.cl:494-513, 641-647
  EVD: eigenvector of the largest eigenvalue of T      tcoh.cl:649-656
  EMI: smallest eigenvalue of |T|^-1 o T; |T|^-1 by Cholesky when every
       pivot exceeds 1e-5, else the eigen pseudo-inverse with relative
       floor 1e-6 (tcoh.cl:657-694; reference.py:150-157 always uses the
       pseudo-inverse with absolute floor 1e-9, which agrees whenever the
       Cholesky succeeds)
  phasors relative to date N/2, or to the last strongest component when
       date N/2 is below 1e-3 of it; components below 1e-3 take phase 0
       (tcoh.cl:696-724)
  gamma = |sum_{i<j} unit(T_ij) conj(z_i) z_j| / (N(N-1)/2), unit(0)=1
       (tcoh.cl:726-744; reference.py:172-184 in angle form)

Eigenvectors come from numpy.linalg.eigh (LAPACK), the same library the
reference's own oracle uses (reference.py:148-157).

With form="f16" or "bf16", the covariance is the one the library's mixed
forms define (no reference counterpart): each sample split as hi + lo
16-bit values (round to nearest even), every product taken as hi.hi +
hi.lo + lo.hi, so C_form = C(hi + lo) - lo lo^H. FP16 first scales each
pixel by the power of two that puts its largest component magnitude in
[2^13, 2^14), then, per output pixel, multiplies every candidate's hi and
lo by 2^-(its exponent below the window's largest accepted one) in FP16
(exact unless the product is subnormal); BF16 takes no scale. This is the
arithmetic of the form, not of the reference: its distance from the exact
C is the form's own error.

With adopt (bias correction only), the oracle takes the tested side's
decision where the bias threshold is within rounding: for every entry
whose |m2c| lies below band, |T_ij| is the tested side's (adopt: its
|T_ij| for all pairs i < j, as a pairs=all run gives them; 0 where it
zeroed the entry) and the phase is the FP64 one. Elsewhere everything is
FP64.
"""

import numpy as np

CHOL_FLOOR = 1.0e-5
EIG_FLOOR = 1.0e-6
PHASOR_FLOOR = 1.0e-3


def window_offsets(win_az, win_rg, padded_cols):
    """Candidate offsets from the window origin, dy-major (tcoh.cl:572-576)."""
    dy, dx = np.mgrid[0:win_az, 0:win_rg]
    return (dy * padded_cols + dx).ravel()


def mask_bits(mask, nbits):
    """Unpack uint32 mask words (pixels, words) into bools (pixels, nbits)."""
    b = np.unpackbits(mask.view(np.uint8).reshape(mask.shape[0], -1),
                      axis=1, bitorder="little")
    return b[:, :nbits].astype(bool)


def covariance(slc, accepted, origin, offsets):
    """Hermitian C = sum_k s_k s_k^H over accepted candidates, complex128.

    slc: complex (npad, N); accepted: bool (P, K); origin: window origin
    pixel index (P,); offsets (K,).
    """
    idx = origin[:, None] + offsets[None, :]
    x = slc[idx] * accepted[..., None]  # (P, K, N)
    return np.einsum("pki,pkj->pij", x, x.conj(), optimize=True)


def round_bf16(x):
    """float32 -> nearest BF16 (ties to even), as float32 (finite x)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def split16(xs, form):
    """(hi, lo) float64 of float32 values xs under the form's rounding."""
    xs = np.asarray(xs, np.float32)
    if form == "bf16":
        hi = round_bf16(xs)
        lo = round_bf16((xs - hi).astype(np.float32))
    else:
        hi = xs.astype(np.float16).astype(np.float32)
        lo = (xs - hi).astype(np.float32).astype(np.float16).astype(np.float32)
    return hi.astype(np.float64), lo.astype(np.float64)


def split_exponents(x):
    """Per pixel (leading axis) the FP16 form's scale exponent: that of the
    largest finite component magnitude (frexp), at least -113; None for a
    pixel without a non-zero one."""
    x = np.asarray(x, np.float32).reshape(x.shape[0], -1)
    m = np.where(np.isfinite(x), np.abs(x), 0).max(axis=1)
    e = np.maximum(np.frexp(m)[1], -113)
    return np.where(m > 0, e, -32768)


def split_parts(slc32, form):
    """Per-pixel split of float32 samples (P, N, 2): (hi, lo) float64 of
    the scaled values and the scale exponents (FP16: each pixel times
    2^(14 - e)); BF16 unscaled, exponents 0."""
    x = np.asarray(slc32, np.float32)
    if form == "bf16":
        hi, lo = split16(x, form)
        return hi, lo, np.zeros(x.shape[0], np.int64)
    e = split_exponents(x)
    sc = np.ldexp(np.float32(1), np.where(e == -32768, 0, 14 - e)).astype(np.float32)
    hi, lo = split16(x * sc.reshape((-1,) + (1,) * (x.ndim - 1)), form)
    return hi, lo, e


def form_covariance(parts, accepted, origin, offsets, form):
    """C of a mixed form (P, N, N) complex128 from split_parts output: per
    output pixel, FP16 rescales each accepted candidate's hi and lo by
    2^-(window max exponent - its exponent) with FP16 rounding, then C =
    sum (hi + lo)(hi + lo)^H - lo lo^H (in units of the window's common
    scale, to which T is invariant)."""
    hi, lo, e = parts
    idx = origin[:, None] + offsets[None, :]
    h, l = hi[idx], lo[idx]  # (P, K, N, 2)
    if form != "bf16":
        ek = np.where(accepted, e[idx], np.iinfo(np.int64).min)
        emax = ek.max(axis=1, keepdims=True)
        sh = np.minimum(emax - np.where(accepted, e[idx], emax), 31)
        f = np.where(sh <= 24, np.ldexp(1.0, -sh), 0.0)[..., None, None]
        h = (h * f).astype(np.float16).astype(np.float64)
        l = (l * f).astype(np.float16).astype(np.float64)
    acc = accepted[..., None]
    full = (h[..., 0] + 1j * h[..., 1] + l[..., 0] + 1j * l[..., 1]) * acc
    low = (l[..., 0] + 1j * l[..., 1]) * acc
    return (np.einsum("pki,pkj->pij", full, full.conj(), optimize=True) -
            np.einsum("pki,pkj->pij", low, low.conj(), optimize=True))


def coherence(c, looks, bias, band=None, adopt=None):
    """tcoh.cl:610-636 in double: normalised, optionally debiased T. With
    band and adopt (P, N(N-1)/2): the tested side's |T_ij| where |m2c| <
    band (see the module notes)."""
    n = c.shape[-1]
    d = np.sqrt(np.maximum(np.real(np.einsum("pii->pi", c)), 0.0))
    den = d[:, :, None] * d[:, None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(den > 0, c / np.where(den > 0, den, 1.0), 0.0)
        if bias:
            m2 = np.abs(t) ** 2
            lk = looks[:, None, None].astype(np.float64)
            m2c = (lk * m2 - 1.0) / (lk - 1.0)
            f = np.where((m2 > 0) & (m2c > 0) & (lk > 1),
                         np.sqrt(np.where(m2 > 0, m2c / np.where(m2 > 0, m2, 1), 0)),
                         np.where(lk > 1, 0.0, 1.0))
            if adopt is not None:
                iu, ju = np.triu_indices(n, 1)
                want = np.zeros_like(m2)
                want[:, iu, ju] = adopt
                near = (np.abs(m2c) < band) & np.triu(np.ones((n, n), bool), 1)[None]
                f = np.where(near, want / np.sqrt(np.where(m2 > 0, m2, 1.0)), f)
                f = np.where(near & ~(m2 > 0), 0.0, f)
            t = t * f
    iu = np.arange(n)
    t[:, iu, iu] = 1.0
    # Hermitian from the upper triangle, as the kernel builds it.
    up = np.triu(np.ones((n, n), bool), 1)
    t = np.where(up, t, np.conj(np.swapaxes(t, 1, 2)))
    t[:, iu, iu] = 1.0
    return t


def bias_m2c(c, looks):
    """Debiased squared magnitudes m2c = (L m2 - 1) / (L - 1) (tcoh.cl:620-633).

    The bias correction zeroes T_ij where m2c <= 0 and otherwise gives
    |T_ij| = sqrt(m2c); gamma counts a zero T_ij as unit(0) = 1. An entry
    whose m2c lies within a form's rounding of zero is therefore a threshold
    decision on which gamma may differ by up to 2 / (N(N-1)/2) between
    correct implementations, and sqrt(m2c) is steep there.
    """
    d = np.sqrt(np.maximum(np.real(np.einsum("pii->pi", c)), 0.0))
    den = d[:, :, None] * d[:, None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        m2 = np.where(den > 0, np.abs(c) ** 2 / np.where(den > 0, den, 1.0) ** 2, 0.0)
    lk = looks[:, None, None].astype(np.float64)
    return (lk * m2 - 1.0) / np.maximum(lk - 1.0, 1.0)


def connected(t):
    """Whether the non-zero coherences link all dates (tcoh.cl:494-513)."""
    n = t.shape[-1]
    adj = (t != 0).astype(np.float64)
    reach = np.zeros(t.shape[:2])
    reach[:, 0] = 1.0
    for _ in range(n):
        new = np.minimum(reach + np.einsum("pi,pij->pj", reach, adj), 1.0)
        if np.array_equal(new, reach):
            break
        reach = new
    return reach.sum(axis=1) == n


def cholesky_pivots(g):
    """Pivots s_j = g_jj - sum_k L_jk^2 of the Cholesky of g (P, N, N)."""
    p, n, _ = g.shape
    low = np.zeros_like(g)
    piv = np.zeros((p, n))
    ok = np.ones(p, bool)
    for j in range(n):
        s = g[:, j, j] - np.sum(low[:, j, :j] ** 2, axis=1)
        piv[:, j] = s
        ok &= s > CHOL_FLOOR
        root = np.sqrt(np.where(s > 0, s, 1.0))
        low[:, j, j] = root
        if j + 1 < n:
            low[:, j + 1:, j] = (g[:, j + 1:, j] -
                                 np.einsum("pik,pk->pi", low[:, j + 1:, :j],
                                           low[:, j, :j])) / root[:, None]
    return piv, ok


def abs_inverse(t, loading=0.0):
    """|T|^-1 as the EMI estimator inverts it (tcoh.cl:657-687)."""
    g = np.abs(t)
    if loading:
        g = g + loading * np.eye(g.shape[-1])
    _piv, ok = cholesky_pivots(g)
    inv = np.empty_like(g)
    if ok.any():
        inv[ok] = np.linalg.inv(g[ok])
    bad = ~ok
    if bad.any():
        w, v = np.linalg.eigh(g[bad])
        floor = EIG_FLOOR * np.max(np.abs(w), axis=1, keepdims=True)
        winv = np.where(np.abs(w) < floor, 0.0,
                        1.0 / np.where(w == 0, 1.0, w))
        inv[bad] = np.einsum("pik,pk,pjk->pij", v, winv, v)
    return inv, ok


def phasors(u):
    """Unit phasors relative to the datum (tcoh.cl:696-724), complex (P, N)."""
    n = u.shape[1]
    mag = np.abs(u)
    top = mag.max(axis=1)
    ref = np.full(u.shape[0], n // 2)
    weak = mag[:, n // 2] <= PHASOR_FLOOR * top
    if weak.any():
        last = n - 1 - np.argmax(mag[:, ::-1] == top[:, None], axis=1)
        ref = np.where(weak, last, ref)
    zref = u[np.arange(u.shape[0]), ref]
    z = u * (np.conj(zref) / np.abs(zref))[:, None]
    m = np.abs(z)
    keep = m > PHASOR_FLOOR * top[:, None]
    return np.where(keep, z / np.where(keep, m, 1.0), 1.0 + 0.0j)


def gamma(t, z):
    """Pepe-Lanari fit (tcoh.cl:726-744)."""
    n = t.shape[-1]
    m = np.abs(t)
    unit = np.where(m > 0, t / np.where(m > 0, m, 1.0), 1.0 + 0.0j)
    w = np.conj(z)[:, :, None] * z[:, None, :]
    up = np.triu(np.ones((n, n), bool), 1)
    s = np.sum(np.where(up, unit * w, 0.0), axis=(1, 2))
    return np.abs(s) / (0.5 * n * (n - 1))


def estimate(t, emi, loading=0.0, mag_out=None):
    """Linked phases of each T (P, N, N): unit phasors (P, N). mag_out (P, N)
    optional: receives |u_k| / max |u| of the eigenvector (the phasor
    floor's test, tcoh.cl:696-724)."""
    if not emi:
        _w, v = np.linalg.eigh(t)
        u = v[:, :, -1]
    else:
        inv, _ok = abs_inverse(t, loading)
        _w, v = np.linalg.eigh(inv * t)
        u = v[:, :, 0]
    if mag_out is not None:
        m = np.abs(u)
        mag_out[...] = m / m.max(axis=1, keepdims=True)
    return phasors(u)


def run(slc, count, mask, settings, padded_cols, rows, cols, pixels=None,
        chunk=4096, emi_loading=0.0, band=None, form=None, adopt=None, z_out=None,
        zmag_out=None):
    """FP64 coherence and pair coherences of the strip's output pixels.

    slc: float32 (npad, N, 2) or complex; count: int (rows*cols,) from the
    SHP decisions; mask: uint32 (rows*cols, words) accepted candidates.
    settings: dict with ndates, win_az, win_rg, min_looks, bias, emi,
    npairs, pairs_all. pixels: flat output indices to evaluate (default
    all). Returns (coh (len(pixels),), pairs (len(pixels), npairs)) with NaN
    where not estimated, like the kernel. With band (bias correction only)
    also returns near (len(pixels),): the number of pairs i < j with
    |m2c| < band, and m2c (len(pixels), npairs) of the output pairs (NaN
    where not estimated). adopt (len(pixels), N(N-1)/2): the tested side's
    |T_ij| of every pair i < j, taken where |m2c| < band (module notes).
    z_out (len(pixels), N) complex, optional: receives the unit phasors of
    the linked pixels (NaN elsewhere); zmag_out (len(pixels), N) their
    eigenvector magnitudes relative to the largest.
    """
    n = settings["ndates"]
    wa, wr = settings["win_az"], settings["win_rg"]
    parts = None
    if form in ("f16", "bf16"):
        parts = split_parts(np.asarray(slc, np.float32).reshape(-1, n, 2), form)
    elif slc.dtype != np.complex128:
        slc = np.asarray(slc, np.float32).reshape(-1, n, 2)
        slc = slc[..., 0].astype(np.float64) + 1j * slc[..., 1].astype(np.float64)
    if pixels is None:
        pixels = np.arange(rows * cols)
    pixels = np.asarray(pixels)
    offs = window_offsets(wa, wr, padded_cols)
    npairs = settings.get("npairs", 0)
    if npairs:
        if settings.get("pairs_all"):
            pi, pj = np.triu_indices(n, 1)
        else:
            pi, pj = np.arange(n - 1), np.arange(1, n)
    coh = np.full(len(pixels), np.nan)
    pairs = np.full((len(pixels), npairs), np.nan)
    near = np.zeros(len(pixels), np.int64)
    pm2c = np.full((len(pixels), npairs), np.nan)
    if z_out is not None:
        z_out[...] = np.nan
    if zmag_out is not None:
        zmag_out[...] = np.nan
    use_adopt = adopt is not None and settings["bias"]
    if use_adopt and band is None:
        raise ValueError("adopt needs band")
    for lo in range(0, len(pixels), chunk):
        sel = pixels[lo:lo + chunk]
        cnt = count[sel]
        est = cnt >= settings["min_looks"]
        if not est.any():
            continue
        p = sel[est]
        y, x = p // cols, p % cols
        origin = y * padded_cols + x
        acc = mask_bits(mask[p], wa * wr)
        if parts is not None:
            c = form_covariance(parts, acc, origin, offs, form)
        else:
            c = covariance(slc, acc, origin, offs)
        out = np.flatnonzero(est) + lo
        t = coherence(c, cnt[est], settings["bias"], band,
                      adopt[out] if use_adopt else None)
        if band is not None and settings["bias"]:
            m2c = bias_m2c(c, cnt[est])
            up = np.triu(np.ones((n, n), bool), 1)
            near[out] = np.sum((np.abs(m2c) < band) & up, axis=(1, 2))
            if npairs:
                pm2c[out] = m2c[:, pi, pj]
        if npairs:
            pairs[out] = np.abs(t[:, pi, pj])
        ok = connected(t) if settings["bias"] else np.ones(len(p), bool)
        if ok.any():
            mg = np.empty((int(ok.sum()), n)) if zmag_out is not None else None
            z = estimate(t[ok], settings["emi"], emi_loading, mg)
            coh[out[ok]] = gamma(t[ok], z)
            if z_out is not None:
                z_out[out[ok]] = z
            if zmag_out is not None:
                zmag_out[out[ok]] = mg
    if band is not None:
        return coh, pairs, near, pm2c
    return coh, pairs
