"""Synthetic coregistered SLC stacks for tests and benchmarks.

ds_stack      the distributed-scatterer model of the reference's test
              fixtures (tests/conftest.py:28-44): s_k = sqrt(g) z0 e^{j phi_k}
              + sqrt(1-g) z_k, phi_k = 0.7k + 0.05k(x + 0.5y), scaled by 300
              and rounded (integer DNs, hence amplitude ties). Same random
              draw order, so a seed gives the reference's stack.
scene_stack   decorrelating scenes: dates t_i = 12 i days, per region
              Gamma_ij = (g0 - ginf) exp(-|t_i - t_j| / tau) + ginf (i != j),
              Gamma_ii = 1, Sigma = Gamma o exp(j(phi_i - phi_j)),
              phi_i = i v, v ~ U(-pi/4, pi/4); x = sigma L z with
              L = chol(Sigma), z ~ CN(0, I), sigma in {1, 3} per 64x64
              block; a fraction of pixels zeroed at one random date.
quadrant_regions / voronoi_regions   region maps of the benchmark scenes.

numpy generates on the host; the torch variant (same model, its own
random stream) generates large scenes on the GPU.
"""

import math

import numpy as np


def ds_stack(n, coherence, rows=30, cols=24, seed=0):
    """Complex stack (n, rows, cols) of the reference's fixture model."""
    rng = np.random.default_rng(seed)

    def gaussian(*shape):
        return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)

    y, x = np.mgrid[0:rows, 0:cols]
    phase = np.stack([0.7 * k + 0.05 * k * (x + 0.5 * y) for k in range(n)])
    common = gaussian(rows, cols)
    stack = np.sqrt(coherence) * common * np.exp(1j * phase) + np.sqrt(
        1 - coherence) * gaussian(n, rows, cols)
    return np.round(stack * 300.0)


def coherence_model(n, gamma0, tau, gamma_inf, spacing=12.0):
    """Gamma (n, n) with unit diagonal."""
    t = spacing * np.arange(n)
    g = (gamma0 - gamma_inf) * np.exp(-np.abs(t[:, None] - t[None, :]) / tau) + gamma_inf
    np.fill_diagonal(g, 1.0)
    return g


def quadrant_regions(rows, cols, gamma0=(0.9, 0.6, 0.3, 0.1), tau=48.0):
    """Four quadrants: region map and per-region (gamma0, tau)."""
    reg = np.zeros((rows, cols), np.int32)
    reg[:rows // 2, cols // 2:] = 1
    reg[rows // 2:, :cols // 2] = 2
    reg[rows // 2:, cols // 2:] = 3
    return reg, [(g, tau) for g in gamma0]


def voronoi_regions(rows, cols, cells, rng, gamma0_range, tau_range):
    """Voronoi cells of uniform random centres."""
    cy = rng.uniform(0, rows, cells)
    cx = rng.uniform(0, cols, cells)
    params = [(rng.uniform(*gamma0_range), rng.uniform(*tau_range))
              for _ in range(cells)]
    reg = np.empty((rows, cols), np.int32)
    yy = np.arange(rows)[:, None]
    for r0 in range(0, rows, 256):
        y = yy[r0:r0 + 256]
        xx = np.arange(cols)[None, :]
        d = (y[None] - cy[:, None, None]) ** 2 + (xx[None] - cx[:, None, None]) ** 2
        reg[r0:r0 + 256] = np.argmin(d, axis=0)
    return reg, params


def _region_factors(n, params, gamma_inf, rng):
    facs = []
    for g0, tau in params:
        v = rng.uniform(-np.pi / 4, np.pi / 4)
        phi = v * np.arange(n)
        sig = coherence_model(n, g0, tau, gamma_inf) * np.exp(
            1j * (phi[:, None] - phi[None, :]))
        facs.append(np.linalg.cholesky(sig))
    return facs


def scene_stack(n, rows, cols, regions, params, gamma_inf=0.05, seed=0,
                sigma_block=64, zero_frac=0.005):
    """complex64 stack (n, rows, cols) of the decorrelation model (numpy)."""
    rng = np.random.default_rng(seed)
    facs = _region_factors(n, params, gamma_inf, rng)
    z = (rng.standard_normal((n, rows, cols), np.float32) +
         1j * rng.standard_normal((n, rows, cols), np.float32)) / np.float32(math.sqrt(2))
    nby, nbx = -(-rows // sigma_block), -(-cols // sigma_block)
    sig = rng.choice(np.array([1.0, 3.0], np.float32), size=(nby, nbx))
    sigma = np.repeat(np.repeat(sig, sigma_block, 0), sigma_block, 1)[:rows, :cols]
    out = np.empty((n, rows, cols), np.complex64)
    for r, lo in enumerate(facs):
        m = regions == r
        out[:, m] = (lo.astype(np.complex64) @ z[:, m]) * sigma[m]
    nz = int(round(zero_frac * rows * cols))
    if nz:
        pix = rng.choice(rows * cols, nz, replace=False)
        date = rng.integers(0, n, nz)
        out.reshape(n, -1)[date, pix] = 0
    return out


def scene_stack_torch(n, rows, cols, regions, params, gamma_inf=0.05, seed=0,
                      sigma_block=64, zero_frac=0.005, device="cuda"):
    """Same model on the GPU: complex64 tensor (n, rows, cols).

    The region factors come from numpy with the same seed as scene_stack;
    the samples come from torch's generator on the device.
    """
    import torch

    rng = np.random.default_rng(seed)
    facs = _region_factors(n, params, gamma_inf, rng)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    z = torch.complex(
        torch.randn((n, rows * cols), generator=gen, device=device),
        torch.randn((n, rows * cols), generator=gen, device=device)) / math.sqrt(2)
    nby, nbx = -(-rows // sigma_block), -(-cols // sigma_block)
    sig = rng.choice(np.array([1.0, 3.0], np.float32), size=(nby, nbx))
    sigma = np.repeat(np.repeat(sig, sigma_block, 0), sigma_block, 1)[:rows, :cols]
    reg = torch.from_numpy(regions.reshape(-1)).to(device)
    sg = torch.from_numpy(np.ascontiguousarray(sigma.reshape(-1))).to(device)
    out = torch.empty((n, rows * cols), dtype=torch.complex64, device=device)
    for r, lo in enumerate(facs):
        idx = torch.nonzero(reg == r).squeeze(1)
        if idx.numel() == 0:
            continue
        lt = torch.from_numpy(lo.astype(np.complex64)).to(device)
        out[:, idx] = (lt @ z[:, idx]) * sg[idx]
    nz = int(round(zero_frac * rows * cols))
    if nz:
        pix = torch.from_numpy(rng.choice(rows * cols, nz, replace=False)).to(device)
        date = torch.from_numpy(rng.integers(0, n, nz)).to(device)
        out[date, pix] = 0
    return out.reshape(n, rows, cols)


def config_small(seed=0):
    """N=10 dates, 256 x 256, four quadrants g0 in {0.9, 0.6, 0.3, 0.1}."""
    reg, params = quadrant_regions(256, 256)
    return scene_stack(10, 256, 256, reg, params, 0.05, seed)


def config_voronoi(n=30, rows=2048, cols=2048, seed=0, gamma0=(0.05, 0.95),
                   tau=(24.0, 120.0), cells=16):
    """Region map and parameters of the Voronoi scenes (16 cells)."""
    rng = np.random.default_rng(seed + 1000)
    return voronoi_regions(rows, cols, cells, rng, gamma0, tau)


# ---- acquisition geometry (after tests/geometry_model.py:15-24) ----------

LIGHT = 299792458.0
ELLIPSOID_A = 6378137.0


def circular_orbits(t0s, baselines, radius=ELLIPSOID_A + 700e3, speed=7500.0,
                    latitude=25.0, sv_dt=10.0, half_span=60.0):
    """State vectors of circular polar orbits in an Earth-fixed frame over
    longitude 0, one per date: orbit k crosses `latitude` (degrees) at t0s[k]
    heading north, offset across track (+y) by baselines[k]; state vectors
    every sv_dt seconds within t0 +- half_span. Returns [(t, pos, vel)]."""
    omega = speed / radius
    out = []
    for t0, b in zip(t0s, baselines):
        t = t0 + np.arange(-half_span, half_span + 0.5 * sv_dt, sv_dt)
        th = math.radians(latitude) + omega * (t - t0)
        pos = np.stack([radius * np.cos(th), np.full_like(th, b), radius * np.sin(th)], 1)
        vel = np.stack([-radius * omega * np.sin(th), np.zeros_like(th),
                        radius * omega * np.cos(th)], 1)
        out.append((t, pos, vel))
    return out


def geometry_case(n=5, rows=300, cols=500, ref=None, seed=0, days=12.0):
    """A C-band stack geometry: dates every `days`, baselines ~ N(0, 100 m)
    (0 at the reference date), near range 850 km, the reference date's
    raster timing. Returns (orbits, scene dict, ref)."""
    rng = np.random.default_rng(seed)
    ref = n // 2 if ref is None else ref
    t0s = 1.0e5 + 86400.0 * days * np.arange(n)
    base = rng.normal(0.0, 100.0, n)
    base[ref] = 0.0
    scene = dict(t_first=t0s[ref], dt=2.055556e-3, tau0=2.0 * 850e3 / LIGHT,
                 fs=64.34523812571427e6, **{"lambda": LIGHT / 5.405e9},
                 line0=0.0, sample0=0.0, rows=rows, cols=cols)
    return circular_orbits(t0s, base), scene, ref


# ---- the documents orbit ingestion reads -------------------------------
# Layouts of a date's description.json (the r.in.s1slc sections the
# geometry reads) and of a Sentinel-1 POEORB/RESORB EOF file, after
# tests/geometry_model.py:145-220.


def iso_utc(sec):
    """ISO 8601 text with microseconds of seconds since 1970 UTC."""
    import datetime as _dt

    t = _dt.datetime(1970, 1, 1) + _dt.timedelta(seconds=float(sec))
    return t.isoformat(timespec="microseconds")


def s1_metadata(orbit, scene, first_line, mission="S1A"):
    """description.json text of one date: swath (mission, start/stop, radar
    constants, state vectors) and raster_geometry sections. orbit is
    (t, pos, vel); first_line the date's first line time (s since 1970)."""
    import json as _json

    t, pos, vel = orbit
    start = first_line
    stop = first_line + (scene["rows"] - 1) * scene["dt"]
    svs = [{"time": iso_utc(t[i]), "frame": "Earth Fixed",
            "position": [float(x) for x in pos[i]],
            "velocity": [float(x) for x in vel[i]]} for i in range(len(t))]
    doc = {
        "swath": {
            "mission": mission,
            "start_time": iso_utc(start),
            "stop_time": iso_utc(stop),
            "wavelength": scene["lambda"],
            "range_sampling_rate": scene["fs"],
            "terrain_height": 0.0,
            "orbit_state_vectors": svs,
        },
        "raster_geometry": {
            "first_line_time": iso_utc(start),
            "last_line_time": iso_utc(stop),
            "azimuth_time_interval": scene["dt"],
            "slant_range_time_first_sample": scene["tau0"],
        },
    }
    return _json.dumps(doc, indent=2)


def eof_text(orbit):
    """EOF document (Earth Explorer file) with one <OSV> per state vector,
    coordinates printed with 6 decimals as ESA does."""
    t, pos, vel = orbit
    recs = []
    for i in range(len(t)):
        u = iso_utc(t[i])
        recs.append(
            "      <OSV>\n"
            f"        <TAI>TAI={u}</TAI>\n"
            f"        <UTC>UTC={u}</UTC>\n"
            f"        <UT1>UT1={u}</UT1>\n"
            "        <Absolute_Orbit>+46742</Absolute_Orbit>\n"
            + "".join(f'        <{tag} unit="{unit}">{val:.6f}</{tag}>\n'
                      for tag, unit, val in zip(
                          ("X", "Y", "Z", "VX", "VY", "VZ"),
                          ("m",) * 3 + ("m/s",) * 3,
                          list(pos[i]) + list(vel[i])))
            + "        <Quality>NOMINAL</Quality>\n"
            "      </OSV>\n")
    return ('<?xml version="1.0" ?>\n<Earth_Explorer_File>\n  <Data_Block type="xml">\n'
            f'    <List_of_OSVs count="{len(recs)}">\n' + "".join(recs)
            + "    </List_of_OSVs>\n  </Data_Block>\n</Earth_Explorer_File>\n")


def eof_name(v0, v1, kind="POEORB", mission="S1A", produced="20990101T000000"):
    """<mission>_OPER_AUX_<kind>_OPOD_<production>_V<v0>_<v1>.EOF"""
    import datetime as _dt

    def tag(sec):
        t = _dt.datetime(1970, 1, 1) + _dt.timedelta(seconds=float(sec))
        return t.strftime("%Y%m%dT%H%M%S")

    return f"{mission}_OPER_AUX_{kind}_OPOD_{produced}_V{tag(v0)}_{tag(v1)}.EOF"
