"""ctypes binding of the C oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline leg use
it, as the checker or the timed CPU baseline.
"""

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OSettings(C.Structure):
    """struct settings of local_proto.h:33-49."""

    _fields_ = [
        ("ndates", C.c_int),
        ("win_az", C.c_int),
        ("win_rg", C.c_int),
        ("test", C.c_int),
        ("alpha", C.c_double),
        ("emi", C.c_int),
        ("bias", C.c_int),
        ("min_looks", C.c_int),
        ("ks_max", C.c_int),
        ("ad_norm", C.c_double),
        ("ad_sigma", C.c_double),
        ("ad_crit", C.c_double),
        ("tlog_crit", C.c_double),
        ("npairs", C.c_int),
        ("pairs_all", C.c_int),
        ("phase_reference", C.c_char_p),
        ("orbit", C.c_char_p),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
        L.oracle_shp_constants.argtypes = [C.POINTER(OSettings)]
        L.oracle_ad_sigma.argtypes = [C.c_int, C.c_int]
        L.oracle_ad_sigma.restype = C.c_double
        L.oracle_normal_quantile.argtypes = [C.c_double]
        L.oracle_normal_quantile.restype = C.c_double
        L.oracle_ad_critical.argtypes = [C.c_double]
        L.oracle_ad_critical.restype = C.c_double
        L.oracle_kernel_float.argtypes = [C.c_double]
        L.oracle_kernel_float.restype = C.c_float
        L.oracle_logf.argtypes = [C.c_float]
        L.oracle_logf.restype = C.c_float
        L.oracle_run.argtypes = [C.POINTER(OSettings), C.c_int, f32p, f32p, u8p,
                                 C.c_int, C.c_int, C.c_int, f32p, i32p,
                                 C.c_void_p, C.c_void_p, C.c_int]
        L.oracle_run.restype = C.c_long
        L.oracle_ks_accept.argtypes = [C.c_int, C.c_int, f32p, f32p]
        L.oracle_ad_statistic.argtypes = [C.c_int, C.c_float, C.c_float, f32p, f32p]
        L.oracle_ad_statistic.restype = C.c_float
        L.oracle_tlog_accept.argtypes = [C.c_int, C.c_float, f32p, f32p]
        L.oracle_prepare_strip.argtypes = [C.c_int, C.c_int, C.c_int, f32p, f32p,
                                           C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                           C.c_int, f32p, f32p, u8p]
        L.oracle_orbit_state.argtypes = [C.c_int, f64p, f64p, f64p, C.c_double,
                                         f64p, f64p]
        L.oracle_geometry_init.argtypes = [C.c_int, C.c_int, i32p, i32p, f64p, f64p,
                                           f64p] + [C.c_double] * 7 + [
                                               C.c_int, C.c_int, C.c_double,
                                               C.c_double, C.c_void_p,
                                               C.POINTER(C.c_int), C.POINTER(C.c_int),
                                               C.POINTER(C.c_int), f64p]
        L.oracle_geometry_init.restype = C.c_int
        L.oracle_geometry_phases.argtypes = [C.c_int, C.c_int, C.c_int, f64p, f64p,
                                             C.c_int, C.c_int, C.c_double, f64p]
        _lib = L
    return _lib


SHP_TESTS = {"ks": 0, "ad": 1, "tlog": 2}


def settings(n, window=(21, 7), shp_test="ks", alpha=0.05, min_shp=20,
             estimator="evd", bias=False, pairs=None):
    """Settings as main.c:553-608 derives them from the module options.

    pairs: None (no pair maps), "consecutive" or "all".
    """
    s = OSettings()
    s.ndates = n
    s.win_az, s.win_rg = window
    s.test = SHP_TESTS[shp_test]
    s.alpha = alpha
    s.emi = 1 if estimator == "emi" else 0
    s.bias = 1 if bias else 0
    s.min_looks = max(min_shp, n)
    s.pairs_all = 1 if pairs == "all" else 0
    s.npairs = 0 if pairs is None else (n * (n - 1) // 2 if pairs == "all" else n - 1)
    lib().oracle_shp_constants(C.byref(s))
    return s


def settings_dict(s):
    return {f: getattr(s, f) for f, _t in OSettings._fields_}


def run(s, padded_cols, slc, amp, valid, padded_rows, rows, cols, with_mask=True,
        nthreads=0):
    """FP32 kernel restatement over one strip. Returns (coh, count, pairs,
    mask, ad_flips)."""
    slc = np.ascontiguousarray(slc, np.float32).reshape(-1)
    amp = np.ascontiguousarray(amp, np.float32).reshape(-1)
    valid = np.ascontiguousarray(valid, np.uint8).reshape(-1)
    coh = np.empty(rows * cols, np.float32)
    count = np.empty(rows * cols, np.int32)
    pairs = np.empty(s.npairs * rows * cols, np.float32) if s.npairs else None
    words = (s.win_az * s.win_rg + 31) // 32
    mask = np.zeros((rows * cols, words), np.uint32) if with_mask else None
    flips = lib().oracle_run(
        C.byref(s), padded_cols, slc, amp, valid, padded_rows, rows, cols, coh,
        count, pairs.ctypes.data if pairs is not None else None,
        mask.ctypes.data if mask is not None else None, nthreads)
    if pairs is not None:
        pairs = pairs.reshape(s.npairs, rows * cols)
    return coh, count, pairs, mask, flips


def prepare_strip(i, q, row0, padded_rows, halo_cols, phase=None, elev_null=None):
    """main.c:52-127 on planar float32 (n, nrows, ncols) I/Q with NaN nulls."""
    i = np.ascontiguousarray(i, np.float32)
    q = np.ascontiguousarray(q, np.float32)
    n, nrows, ncols = i.shape
    pc = ncols + 2 * halo_cols
    slc = np.empty(padded_rows * pc * n * 2, np.float32)
    amp = np.empty(padded_rows * pc * n, np.float32)
    valid = np.empty(padded_rows * pc, np.uint8)
    ph = None if phase is None else np.ascontiguousarray(phase, np.float64)
    en = None if elev_null is None else np.ascontiguousarray(elev_null, np.uint8)
    lib().oracle_prepare_strip(n, nrows, ncols, i, q,
                               None if ph is None else ph.ctypes.data,
                               None if en is None else en.ctypes.data, row0,
                               padded_rows, halo_cols, slc, amp, valid)
    return slc, amp, valid


def pixel_phases(grid, heights, nrows, ncols, elevation=None, h0=0.0):
    """geometry_phases (geometry.c:256-291) of every region cell as planes
    (n, nrows, ncols) float64: the FP64 restatement at the cell's height
    (elevation[row, col]; NaN heights give 0) or at h0."""
    grid = np.ascontiguousarray(grid, np.float64)
    gr, gcn, nh, n = grid.shape
    hs = np.zeros(3)
    hs[:nh] = heights
    out = np.zeros((n, nrows, ncols))
    buf = np.zeros(n)
    flat = grid.reshape(-1)
    for r in range(nrows):
        for c in range(ncols):
            h = h0 if elevation is None else float(elevation[r, c])
            if np.isnan(h):
                continue
            lib().oracle_geometry_phases(n, gcn, nh, hs, flat, r, c, h, buf)
            out[:, r, c] = buf
    return out


def geometry_grid(orbits, scene, ref, hmin, hmax):
    """geometry.c:161-254 in FP64 on the CPU: the phase grid (grid_rows,
    grid_cols, nheights, n) and heights; orbits and scene as
    tcoh_amd.api.geometry_grid takes them."""
    n = len(orbits)
    nsv = np.array([len(o[0]) for o in orbits], np.int32)
    off = np.concatenate([[0], np.cumsum(nsv)[:-1]]).astype(np.int32)
    t = np.ascontiguousarray(np.concatenate([o[0] for o in orbits]), np.float64)
    pos = np.ascontiguousarray(np.concatenate([o[1] for o in orbits]), np.float64).reshape(-1)
    vel = np.ascontiguousarray(np.concatenate([o[2] for o in orbits]), np.float64).reshape(-1)
    gr, gc, nh = C.c_int(), C.c_int(), C.c_int()
    hs = np.zeros(3)
    args = [n, ref, nsv, off, t, pos, vel, scene["t_first"], scene["dt"], scene["tau0"],
            scene["fs"], scene["lambda"], scene["line0"], scene["sample0"],
            scene["rows"], scene["cols"], hmin, hmax]
    lib().oracle_geometry_init(*args, None, C.byref(gr), C.byref(gc), C.byref(nh), hs)
    phase = np.zeros(gr.value * gc.value * nh.value * n)
    status = lib().oracle_geometry_init(*args, phase.ctypes.data, C.byref(gr),
                                        C.byref(gc), C.byref(nh), hs)
    return phase.reshape(gr.value, gc.value, nh.value, n), hs[:nh.value], status
