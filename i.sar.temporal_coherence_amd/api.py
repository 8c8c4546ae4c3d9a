"""Python binding of libtcoh_hip.so, mirroring the module's interface.

The reference is a GRASS module (main.c) calling its OpenCL backend through
ocl_open / ocl_run / ocl_close (local_proto.h:104-112). This binding calls
the same entry points of the HIP backend:

  Settings / make_settings  struct settings and its derivation from the
                            module options (main.c:553-608, stats.c:61-95)
  Backend                   one ocl_open handle; run() = ocl_run on host
                            arrays, run_device() = tcoh_run_device on
                            device pointers (e.g. torch tensors)
  Batch                     several handles' strips (stacks of any depth)
                            as one unit of work (tcoh_batch_run)
  temporal_coherence        the module's strip loop (main.c:718-825) over a
                            planar complex stack, with read_strip's
                            preparation (main.c:52-127): returns the
                            coherence, SHP-count and pair-coherence maps

Errors raise RuntimeError (the library is put in its returning error mode)
instead of ending the process as G_fatal_error does. There is no CPU path:
without the built library or a HIP device every call fails.
"""

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# TCOH_LIB selects another build of the library (e.g. a debug build).
LIB_PATH = os.environ.get("TCOH_LIB") or os.path.join(HERE, "libtcoh_hip.so")

SHP_TESTS = {"ks": 0, "ad": 1, "tlog": 2}
PRECISIONS = {"plain": 0, "f16": 1, "bf16": 2}


class Settings(C.Structure):
    """struct settings, local_proto.h:33-49."""

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


class Utc(C.Structure):
    """struct utc, local_proto.h:15-18."""

    _fields_ = [
        ("year", C.c_int),
        ("month", C.c_int),
        ("day", C.c_int),
        ("hour", C.c_int),
        ("minute", C.c_int),
        ("second", C.c_double),
    ]


class Orbit(C.Structure):
    """struct orbit, local_proto.h:52-57."""

    _fields_ = [
        ("n", C.c_int),
        ("t", C.POINTER(C.c_double)),
        ("pos", C.POINTER(C.c_double)),
        ("vel", C.POINTER(C.c_double)),
        ("source", C.c_char * 32),
    ]


class Geometry(C.Structure):
    """struct geometry, local_proto.h:61-68."""

    _fields_ = [
        ("ndates", C.c_int),
        ("grid_rows", C.c_int),
        ("grid_cols", C.c_int),
        ("step", C.c_int),
        ("nheights", C.c_int),
        ("heights", C.c_double * 3),
        ("phase", C.POINTER(C.c_double)),
        ("h0", C.c_double),
    ]


class Scene(C.Structure):
    """struct tcoh_scene (include/tcoh_hip.h)."""

    _fields_ = [
        ("t_first", C.c_double),
        ("dt", C.c_double),
        ("tau0", C.c_double),
        ("fs", C.c_double),
        ("lambda_", C.c_double),
        ("line0", C.c_double),
        ("sample0", C.c_double),
        ("rows", C.c_int),
        ("cols", C.c_int),
    ]


_lib = None


def lib():
    """The loaded libtcoh_hip.so (RuntimeError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libtcoh_hip.so is not built: run make -C i.sar.temporal_coherence_amd "
            "(or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.shp_constants.argtypes = [C.POINTER(Settings)]
    L.ocl_list_devices.argtypes = []
    L.ocl_open.argtypes = [C.c_int, C.c_int, C.POINTER(Settings), C.c_int,
                           C.POINTER(C.c_size_t)]
    L.ocl_open.restype = vp
    L.ocl_run.argtypes = [vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, vp, vp, vp]
    L.ocl_close.argtypes = [vp]
    L.tcoh_run_device.argtypes = [vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, vp,
                                  vp, vp, vp]
    L.tcoh_run_device_phases.argtypes = [vp, vp, vp, vp, C.c_int, C.c_int, C.c_int,
                                         vp, vp, vp, vp, vp]
    L.tcoh_run_phases.argtypes = [vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, vp, vp,
                                  vp, vp]
    L.tcoh_set_precision.argtypes = [vp, C.c_int]
    L.tcoh_set_graph.argtypes = [vp, C.c_int]
    L.tcoh_set_timing.argtypes = [vp, C.c_int]
    L.tcoh_stage_times.argtypes = [vp, C.POINTER(C.c_double)]
    L.tcoh_stage_launches.argtypes = [vp, C.POINTER(C.c_int)]
    L.tcoh_set_tiles.argtypes = [vp, vp]
    L.tcoh_get_tiles.argtypes = [vp, vp]
    L.tcoh_batch_create.argtypes = [C.c_int]
    L.tcoh_batch_create.restype = vp
    L.tcoh_batch_set_graph.argtypes = [vp, C.c_int]
    L.tcoh_batch_run.argtypes = [vp, vp, C.c_int, vp]
    L.tcoh_batch_plan.argtypes = [vp, vp, vp]
    L.tcoh_batch_destroy.argtypes = [vp]
    L.tcoh_strip_cost.argtypes = [C.POINTER(Settings), C.c_int, C.c_int]
    L.tcoh_strip_cost.restype = C.c_double
    L.tcoh_batch_schedule.argtypes = [C.c_int, vp, C.c_int, vp, vp]
    L.tcoh_set_error_mode.argtypes = [C.c_int]
    L.tcoh_last_error.restype = C.c_char_p
    L.tcoh_prepare_strip.argtypes = [C.c_int, C.c_int, C.c_int, vp, vp, vp, vp,
                                     C.c_int, C.c_int, C.c_int, vp, vp, vp]
    L.tcoh_prepare_strip_device.argtypes = [vp, C.c_int, C.c_int, vp, vp,
                                            C.POINTER(Geometry), vp, C.c_int,
                                            C.c_int, C.c_int, vp, vp, vp, vp]
    L.tcoh_run_region.argtypes = [vp, C.c_int, C.c_int, vp, vp, C.POINTER(Geometry),
                                  vp, vp, vp, vp, C.c_int]
    L.orbit_state.argtypes = [C.POINTER(Orbit), C.c_double, vp, vp]
    L.tcoh_geometry_init.argtypes = [C.POINTER(Geometry), C.POINTER(Scene),
                                     C.c_int, C.c_int, C.POINTER(Orbit),
                                     C.c_double, C.c_double, C.c_double]
    L.geometry_phases.argtypes = [C.POINTER(Geometry), C.c_int, C.c_int,
                                  C.c_double, vp]
    L.geometry_free.argtypes = [C.POINTER(Geometry)]
    cp = C.c_char_p
    L.tcoh_utc_seconds.argtypes = [C.POINTER(Utc)]
    L.tcoh_utc_seconds.restype = C.c_double
    L.tcoh_utc_from_seconds.argtypes = [C.c_double, C.POINTER(Utc)]
    L.tcoh_parse_utc.argtypes = [cp, C.POINTER(C.c_double)]
    L.tcoh_json_get.argtypes = [cp, cp]
    L.tcoh_json_get.restype = vp
    L.tcoh_json_get_number.argtypes = [cp, cp]
    L.tcoh_json_get_number.restype = C.c_double
    L.tcoh_json_array_length.argtypes = [cp, cp]
    L.tcoh_free.argtypes = [vp]
    L.tcoh_acquisition_span.argtypes = [cp, cp, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]
    L.tcoh_orbit_from_annotation.argtypes = [cp, cp, C.POINTER(Orbit)]
    L.tcoh_orbit_from_eof.argtypes = [cp, C.c_double, C.c_double, cp,
                                      C.POINTER(Orbit)]
    L.tcoh_orbit_pick_file.argtypes = [cp, cp, C.c_double, C.c_double, cp,
                                       C.c_size_t]
    L.tcoh_orbit_from_file.argtypes = [cp, cp, cp, cp, C.c_int, C.POINTER(Orbit)]
    L.tcoh_orbit_best.argtypes = [cp, cp, cp, C.POINTER(Orbit)]
    L.tcoh_orbit_free.argtypes = [C.POINTER(Orbit)]
    L.tcoh_scene_from_metadata.argtypes = [C.POINTER(cp), C.POINTER(cp), C.c_int,
                                           C.c_int, C.c_double, C.c_double,
                                           C.c_int, C.c_int, C.POINTER(Scene)]
    L.tcoh_pair_names.argtypes = [cp, C.c_int, C.c_int, C.POINTER(Utc),
                                  C.POINTER(C.c_int), vp, C.c_size_t,
                                  C.POINTER(C.c_int)]
    L.tcoh_raw_map.argtypes = [cp, C.c_size_t, C.POINTER(vp)]
    L.tcoh_raw_unmap.argtypes = [vp, C.c_size_t]
    L.tcoh_raw_create.argtypes = [cp, cp, cp, C.POINTER(cp), C.c_int, C.c_int,
                                  C.c_int]
    L.tcoh_raw_create.restype = vp
    L.tcoh_raw_put_strip.argtypes = [vp, C.c_int, vp, vp, vp]
    L.tcoh_raw_close.argtypes = [vp, C.POINTER(C.c_long), C.POINTER(C.c_double)]
    # The HIP runtime the library links (found through the library's own
    # dependencies), for DeviceBuffer.
    L.hipMalloc.argtypes = [C.POINTER(vp), C.c_size_t]
    L.hipFree.argtypes = [vp]
    L.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
    L.hipMemset.argtypes = [vp, C.c_int, C.c_size_t]
    L.hipDeviceSynchronize.argtypes = []
    L.hipStreamCreateWithFlags.argtypes = [C.POINTER(vp), C.c_uint]
    L.hipStreamSynchronize.argtypes = [vp]
    L.hipStreamDestroy.argtypes = [vp]
    L.tcoh_set_error_mode(1)
    _lib = L
    return L


def _raise_if_error():
    msg = lib().tcoh_last_error()
    if msg:
        raise RuntimeError(msg.decode())


def _ptr(a):
    return None if a is None else a.ctypes.data


def make_settings(n, window=(21, 7), shp_test="ks", alpha=0.05, min_shp=20,
                  estimator="evd", bias=False, pairs=None):
    """Settings from the module options (main.c:553-608).

    window: (azimuth, range) odd sizes; pairs: None, "consecutive" or "all".
    """
    wa, wr = window
    if wa < 1 or wr < 1 or wa % 2 == 0 or wr % 2 == 0:
        raise ValueError("Window sizes must be odd, got %d,%d" % (wa, wr))
    if not 0 < alpha < 1:
        raise ValueError("Option alpha must be between 0 and 1")
    s = Settings()
    s.ndates = n
    s.win_az, s.win_rg = wa, wr
    s.test = SHP_TESTS[shp_test]
    s.alpha = alpha
    s.emi = 1 if estimator == "emi" else 0
    s.bias = 1 if bias else 0
    s.min_looks = max(min_shp, n)
    if wa * wr < s.min_looks:
        raise ValueError("The %dx%d window holds fewer pixels than the %d SHPs "
                         "needed" % (wa, wr, s.min_looks))
    s.pairs_all = 1 if pairs == "all" else 0
    s.npairs = 0 if pairs is None else (n * (n - 1) // 2 if pairs == "all" else n - 1)
    if shp_test == "ad" and not 0.001 <= alpha <= 0.25:
        raise ValueError("Anderson-Darling test: alpha must be between 0.001 and 0.25")
    lib().shp_constants(C.byref(s))
    return s


def pair_list(n, mode):
    """Date pairs in output order (tcoh.cl:19-23; main.c:610-615)."""
    if mode == "all":
        return [(i, j) for i in range(n - 1) for j in range(i + 1, n)]
    return [(i, i + 1) for i in range(n - 1)]


class Backend:
    """One ocl_open handle (ocl.c:191-254)."""

    def __init__(self, settings, padded_cols, platform=-1, device=-1,
                 precision=None):
        self.settings = settings
        self.padded_cols = padded_cols
        ma = C.c_size_t(0)
        self.h = lib().ocl_open(platform, device, C.byref(settings), padded_cols,
                                C.byref(ma))
        _raise_if_error()
        if not self.h:
            raise RuntimeError("ocl_open failed")
        self.max_alloc = ma.value
        if precision is not None:
            self.set_precision(precision)

    def set_precision(self, precision):
        lib().tcoh_set_precision(self.h, PRECISIONS.get(precision, precision))
        _raise_if_error()

    def set_timing(self, on):
        lib().tcoh_set_timing(self.h, 1 if on else 0)

    def set_tiles(self, tiles):
        """Launch shapes (shp_stripe, link_stripe, band_px), or None for the
        tuned table's (tcoh_set_tiles)."""
        t = None if tiles is None else (C.c_int * 3)(*tiles)
        lib().tcoh_set_tiles(self.h, t)
        _raise_if_error()

    def get_tiles(self):
        t = (C.c_int * 3)()
        lib().tcoh_get_tiles(self.h, t)
        return tuple(t)

    def set_graph(self, on):
        """Graph mode: strips replayed as one captured HIP graph."""
        lib().tcoh_set_graph(self.h, 1 if on else 0)
        _raise_if_error()

    def stage_times(self):
        """Last run's (prep, shp, covariance, eigen) device ms and launch
        counts (timing must be on)."""
        ms = (C.c_double * 4)()
        nl = (C.c_int * 4)()
        lib().tcoh_stage_times(self.h, ms)
        lib().tcoh_stage_launches(self.h, nl)
        return list(ms), list(nl)

    def run(self, slc, amp, valid, padded_rows, rows, cols, phases=False):
        """ocl_run on host arrays: (coh, count, pairs (npairs, rows*cols)),
        and with phases (tcoh_run_phases) also the linked phases (N,
        rows*cols) in radians."""
        s = self.settings
        slc = np.ascontiguousarray(slc, np.float32)
        amp = np.ascontiguousarray(amp, np.float32)
        valid = np.ascontiguousarray(valid, np.uint8)
        need = padded_rows * self.padded_cols
        if valid.size < need or amp.size < need * s.ndates or \
                slc.size < need * s.ndates * 2:
            raise ValueError("strip buffers smaller than padded_rows x padded_cols")
        coh = np.empty(rows * cols, np.float32)
        count = np.empty(rows * cols, np.int32)
        pairs = np.empty((s.npairs, rows * cols), np.float32) if s.npairs else None
        if phases:
            ph = np.empty((s.ndates, rows * cols), np.float32)
            lib().tcoh_run_phases(self.h, _ptr(slc), _ptr(amp), _ptr(valid), padded_rows,
                                  rows, cols, _ptr(coh), _ptr(count), _ptr(pairs),
                                  _ptr(ph))
            _raise_if_error()
            return coh, count, pairs, ph
        lib().ocl_run(self.h, _ptr(slc), _ptr(amp), _ptr(valid), padded_rows, rows,
                      cols, _ptr(coh), _ptr(count), _ptr(pairs))
        _raise_if_error()
        return coh, count, pairs

    def run_device(self, slc_ptr, amp_ptr, valid_ptr, padded_rows, rows, cols,
                   coh_ptr, count_ptr, pairs_ptr=None, stream=None, phases_ptr=None):
        """tcoh_run_device (tcoh_run_device_phases with phases_ptr) on device
        pointers (ints), asynchronous."""
        if phases_ptr:
            lib().tcoh_run_device_phases(self.h, slc_ptr, amp_ptr, valid_ptr,
                                         padded_rows, rows, cols, coh_ptr, count_ptr,
                                         pairs_ptr, phases_ptr, stream)
        else:
            lib().tcoh_run_device(self.h, slc_ptr, amp_ptr, valid_ptr, padded_rows,
                                  rows, cols, coh_ptr, count_ptr, pairs_ptr, stream)
        _raise_if_error()

    def prepare_strip_device(self, i_ptr, q_ptr, nrows, ncols, row0, padded_rows,
                             halo_cols, slc_ptr, amp_ptr, valid_ptr, geometry=None,
                             heights_ptr=None, stream=None):
        """tcoh_prepare_strip_device (read_strip, main.c:52-127, on the
        device) on device pointers (ints); geometry: a Geometry from
        make_geometry or None. Asynchronous."""
        g = None if geometry is None else C.byref(geometry)
        lib().tcoh_prepare_strip_device(self.h, nrows, ncols, i_ptr, q_ptr, g,
                                        heights_ptr, row0, padded_rows, halo_cols,
                                        slc_ptr, amp_ptr, valid_ptr, stream)
        _raise_if_error()

    def run_region(self, i, q, geometry=None, heights=None, strip_rows=0):
        """tcoh_run_region on host planes i, q (n, nrows, ncols) float32:
        (coh, count, pairs (npairs, nrows*ncols) or None), the region
        streamed through the device strip by strip."""
        s = self.settings
        i = np.ascontiguousarray(i, np.float32)
        q = np.ascontiguousarray(q, np.float32)
        n, nrows, ncols = i.shape
        if n != s.ndates or q.shape != i.shape:
            raise ValueError("date planes must be (%d, rows, cols)" % s.ndates)
        h = None if heights is None else np.ascontiguousarray(heights, np.float64)
        coh = np.empty(nrows * ncols, np.float32)
        count = np.empty(nrows * ncols, np.int32)
        pairs = np.empty((s.npairs, nrows * ncols), np.float32) if s.npairs else None
        g = None if geometry is None else C.byref(geometry)
        lib().tcoh_run_region(self.h, nrows, ncols, _ptr(i), _ptr(q), g, _ptr(h),
                              _ptr(coh), _ptr(count), _ptr(pairs), strip_rows)
        _raise_if_error()
        return coh, count, pairs

    def close(self):
        if self.h:
            lib().ocl_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchItem(C.Structure):
    _fields_ = [("o", C.c_void_p), ("slc", C.c_void_p), ("amp", C.c_void_p),
                ("valid", C.c_void_p), ("padded_rows", C.c_int), ("rows", C.c_int),
                ("cols", C.c_int), ("coh", C.c_void_p), ("count", C.c_void_p),
                ("pairs", C.c_void_p)]


def strip_cost(settings, rows, cols):
    """tcoh_strip_cost: the schedule's estimate of a strip's work."""
    return lib().tcoh_strip_cost(C.byref(settings), rows, cols)


def batch_schedule(costs, streams):
    """tcoh_batch_schedule: (order, lane) of the longest-first list
    schedule of the costs over `streams` lanes."""
    n = len(costs)
    c = (C.c_double * max(1, n))(*costs)
    order, lane = (C.c_int * max(1, n))(), (C.c_int * max(1, n))()
    lib().tcoh_batch_schedule(n, c, streams, order, lane)
    return list(order)[:n], list(lane)[:n]


class Batch:
    """A tcoh_batch: strips of several handles run as one unit of work.

    run() takes items (backend, slc_ptr, amp_ptr, valid_ptr, padded_rows,
    rows, cols, coh_ptr, count_ptr, pairs_ptr) with device pointers (ints)
    and returns without waiting."""

    def __init__(self, streams=0, graph=False):
        self.h = lib().tcoh_batch_create(streams)
        _raise_if_error()
        if not self.h:
            raise RuntimeError("tcoh_batch_create failed")
        self.n = 0
        if graph:
            self.set_graph(True)

    def set_graph(self, on):
        lib().tcoh_batch_set_graph(self.h, 1 if on else 0)
        _raise_if_error()

    def run(self, items, stream=None):
        arr = (BatchItem * max(1, len(items)))()
        for k, (be, slc, amp, valid, prow, rows, cols, coh, count, pairs) in enumerate(items):
            arr[k] = BatchItem(be.h if hasattr(be, "h") else be, slc, amp, valid, prow, rows,
                               cols, coh, count, pairs)
        lib().tcoh_batch_run(self.h, arr, len(items), stream)
        _raise_if_error()
        self.n = len(items)

    def plan(self):
        """The last run's (order, lane)."""
        order, lane = (C.c_int * max(1, self.n))(), (C.c_int * max(1, self.n))()
        lib().tcoh_batch_plan(self.h, order, lane)
        return list(order)[:self.n], list(lane)[:self.n]

    def close(self):
        if self.h:
            lib().tcoh_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def prepare_strip(i, q, row0, padded_rows, halo_cols, phase=None, elev_null=None):
    """read_strip (main.c:52-127) on planar float32 (n, nrows, ncols) I/Q."""
    i = np.ascontiguousarray(i, np.float32)
    q = np.ascontiguousarray(q, np.float32)
    n, nrows, ncols = i.shape
    pc = ncols + 2 * halo_cols
    slc = np.empty(padded_rows * pc * n * 2, np.float32)
    amp = np.empty(padded_rows * pc * n, np.float32)
    valid = np.empty(padded_rows * pc, np.uint8)
    ph = None if phase is None else np.ascontiguousarray(phase, np.float64)
    en = None if elev_null is None else np.ascontiguousarray(elev_null, np.uint8)
    lib().tcoh_prepare_strip(n, nrows, ncols, _ptr(i), _ptr(q), _ptr(ph), _ptr(en),
                             row0, padded_rows, halo_cols, _ptr(slc), _ptr(amp),
                             _ptr(valid))
    return slc, amp, valid


def strip_rows_for(settings, ncols, memory_mb, max_alloc, nrows):
    """Strip height of main.c:724-750."""
    n, wa, wr = settings.ndates, settings.win_az, settings.win_rg
    padded_cols = ncols + wr - 1
    per_row = padded_cols * (12 * n + 1) + ncols * 4 * (2 + settings.npairs)
    rows = int(memory_mb * 1024 * 1024 // per_row) - (wa - 1)
    if max_alloc > 0:
        dev_rows = int(max_alloc // (padded_cols * n * 8)) - (wa - 1)
        rows = min(rows, dev_rows)
        if settings.npairs:
            rows = min(rows, int(max_alloc // (ncols * settings.npairs * 4)))
    if rows < 1:
        raise RuntimeError("Not enough memory for one row of %d columns and %d "
                           "dates; increase memory or reduce the region" % (ncols, n))
    return min(rows, nrows)


def temporal_coherence(stack, window=(21, 7), shp_test="ks", alpha=0.05,
                       min_shp=20, estimator="evd", bias=False, pairs=None,
                       phase=None, elev_null=None, memory=300, precision=None,
                       platform=-1, device=-1):
    """The module's estimate over a complex stack (N, rows, cols).

    NaN samples are nulls. phase (N, rows, cols) is the geometric phase
    removed from every date (read_strip's rotation). Returns coh (rows,
    cols) float32 with NaN where not estimated, count (rows, cols) int32
    with -1 for invalid centres, and pairs (npairs, rows, cols) or None.
    """
    stack = np.asarray(stack)
    n, nrows, ncols = stack.shape
    s = make_settings(n, window, shp_test, alpha, min_shp, estimator, bias, pairs)
    ha, hr = s.win_az // 2, s.win_rg // 2
    i = np.ascontiguousarray(stack.real, np.float32)
    q = np.ascontiguousarray(stack.imag, np.float32)
    coh = np.empty((nrows, ncols), np.float32)
    count = np.empty((nrows, ncols), np.int32)
    pr = np.empty((s.npairs, nrows, ncols), np.float32) if s.npairs else None
    with Backend(s, ncols + 2 * hr, platform, device, precision) as be:
        strip = strip_rows_for(s, ncols, memory, be.max_alloc, nrows)
        for row0 in range(0, nrows, strip):
            rows = min(strip, nrows - row0)
            prow = rows + 2 * ha
            slc, amp, valid = prepare_strip(i, q, row0 - ha, prow, hr, phase,
                                            elev_null)
            c, k, p = be.run(slc, amp, valid, prow, rows, ncols)
            coh[row0:row0 + rows] = c.reshape(rows, ncols)
            count[row0:row0 + rows] = k.reshape(rows, ncols)
            if pr is not None:
                pr[:, row0:row0 + rows] = p.reshape(s.npairs, rows, ncols)
    return coh, count, pr


def geometry_grid(orbits, scene, ref, hmin, hmax, h0=0.0):
    """tcoh_geometry_init (geometry.c:161-254 on the GPU): the flat-earth /
    topographic phase grid of a scene.

    orbits: one (t (m,), pos (m, 3), vel (m, 3)) per date; scene: dict
    with t_first, dt, tau0, fs, lambda, line0, sample0, rows, cols.
    Returns (phase (grid_rows, grid_cols, nheights, n), heights).
    """
    keep = []
    ob = (Orbit * len(orbits))()
    for k, (t, pos, vel) in enumerate(orbits):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (t, pos, vel)]
        keep.append(arrs)
        ob[k].n = len(arrs[0])
        ob[k].t = arrs[0].ctypes.data_as(C.POINTER(C.c_double))
        ob[k].pos = arrs[1].ctypes.data_as(C.POINTER(C.c_double))
        ob[k].vel = arrs[2].ctypes.data_as(C.POINTER(C.c_double))
    sc = Scene(scene["t_first"], scene["dt"], scene["tau0"], scene["fs"],
               scene["lambda"], scene["line0"], scene["sample0"], scene["rows"],
               scene["cols"])
    g = Geometry()
    L = lib()
    L.tcoh_geometry_init(C.byref(g), C.byref(sc), len(orbits), ref, ob, hmin, hmax, h0)
    _raise_if_error()
    n, nh = len(orbits), g.nheights
    shape = (g.grid_rows, g.grid_cols, nh, n)
    phase = np.ctypeslib.as_array(g.phase, shape=(int(np.prod(shape)),)).copy()
    heights = np.array(g.heights[:nh])
    L.geometry_free(C.byref(g))
    return phase.reshape(shape), heights


class DeviceBuffer:
    """Device memory from the HIP runtime that libtcoh_hip.so itself uses,
    for the device entry points (tcoh_run_device,
    tcoh_prepare_strip_device) without a framework; ptr is the device
    address."""

    H2D, D2H = 1, 2

    def __init__(self, nbytes, fill=None):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        if lib().hipMalloc(C.byref(p), max(self.nbytes, 1)) != 0:
            raise RuntimeError("hipMalloc of %d bytes failed" % self.nbytes)
        self.ptr = p.value
        if fill is not None:
            lib().hipMemset(self.ptr, fill, max(self.nbytes, 1))

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if lib().hipMemcpy(b.ptr, a.ctypes.data, a.nbytes, cls.H2D) != 0:
            raise RuntimeError("hipMemcpy to the device failed")
        return b

    def to_array(self, dtype, shape=None):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        lib().hipDeviceSynchronize()
        if lib().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, self.D2H) != 0:
            raise RuntimeError("hipMemcpy from the device failed")
        return out if shape is None else out.reshape(shape)

    def free(self):
        if self.ptr:
            lib().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceStream:
    """A non-blocking hipStream_t of the library's runtime (one that graph
    capture accepts); ptr is the stream handle."""

    def __init__(self):
        p = C.c_void_p()
        if lib().hipStreamCreateWithFlags(C.byref(p), 1) != 0:
            raise RuntimeError("hipStreamCreateWithFlags failed")
        self.ptr = p.value

    def synchronize(self):
        if lib().hipStreamSynchronize(self.ptr) != 0:
            raise RuntimeError("hipStreamSynchronize failed")

    def destroy(self):
        if self.ptr:
            lib().hipStreamDestroy(self.ptr)
            self.ptr = None


def make_geometry(phase, heights, h0=0.0, step=16):
    """A struct geometry over a phase grid (grid_rows, grid_cols, nheights,
    n) float64 and its heights; the grid array must outlive the struct (it
    is kept as the struct's attribute _keep)."""
    phase = np.ascontiguousarray(phase, np.float64)
    gr, gcn, nh, n = phase.shape
    g = Geometry()
    g.ndates, g.grid_rows, g.grid_cols, g.step, g.nheights = n, gr, gcn, step, nh
    for k in range(nh):
        g.heights[k] = float(heights[k])
    g.phase = phase.ctypes.data_as(C.POINTER(C.c_double))
    g.h0 = h0
    g._keep = phase
    return g


def list_devices():
    """ocl_list_devices (ocl.c:105-136), printed to stdout."""
    lib().ocl_list_devices()
    _raise_if_error()


# ---- orbit ingestion (orbit.c, jsonget.c, epoch.c; host only) ----------


def _b(text):
    return None if text is None else (text.encode() if isinstance(text, str) else text)


def utc_seconds(year, month, day, hour=0, minute=0, second=0.0):
    """tcoh_utc_seconds (epoch.c:28-32): seconds since 1970 UTC."""
    return lib().tcoh_utc_seconds(C.byref(Utc(year, month, day, hour, minute, second)))


def utc_from_seconds(sec):
    """tcoh_utc_from_seconds (epoch.c:34-52): (year, month, day, hour,
    minute, second)."""
    u = Utc()
    lib().tcoh_utc_from_seconds(sec, C.byref(u))
    return (u.year, u.month, u.day, u.hour, u.minute, u.second)


def parse_utc(text):
    """tcoh_parse_utc (orbit.c:47-56): seconds, or None if unreadable."""
    x = C.c_double()
    return x.value if lib().tcoh_parse_utc(_b(text), C.byref(x)) else None


def json_get(text, path):
    """tcoh_json_get (jsonget.c:171-190): str or None."""
    L = lib()
    p = L.tcoh_json_get(_b(text), _b(path))
    if not p:
        return None
    try:
        return C.string_at(p).decode()
    finally:
        L.tcoh_free(p)


def json_get_number(text, path):
    return lib().tcoh_json_get_number(_b(text), _b(path))


def json_array_length(text, path):
    return lib().tcoh_json_array_length(_b(text), _b(path))


def acquisition_span(meta, name="date"):
    """tcoh_acquisition_span (orbit.c:59-71): (start, stop) seconds."""
    a, b = C.c_double(), C.c_double()
    rc = lib().tcoh_acquisition_span(_b(meta), _b(name), C.byref(a), C.byref(b))
    _raise_if_error()
    assert rc == 0
    return a.value, b.value


def _take_orbit(o):
    """Copy a library-filled struct orbit to (t, pos, vel, source) arrays and
    release it."""
    n = o.n
    if n == 0:
        t, pos, vel = np.zeros(0), np.zeros((0, 3)), np.zeros((0, 3))
    else:
        t = np.ctypeslib.as_array(o.t, shape=(n,)).copy()
        pos = np.ctypeslib.as_array(o.pos, shape=(3 * n,)).copy().reshape(n, 3)
        vel = np.ctypeslib.as_array(o.vel, shape=(3 * n,)).copy().reshape(n, 3)
    source = o.source.decode()
    lib().tcoh_orbit_free(C.byref(o))
    return t, pos, vel, source


def orbit_from_annotation(meta, name="date"):
    """tcoh_orbit_from_annotation (orbit.c:81-111)."""
    o = Orbit()
    rc = lib().tcoh_orbit_from_annotation(_b(meta), _b(name), C.byref(o))
    _raise_if_error()
    assert rc == 0
    return _take_orbit(o)


def orbit_from_eof(text, t0, t1, source="POEORB"):
    """tcoh_orbit_from_eof (orbit.c:175-203)."""
    o = Orbit()
    rc = lib().tcoh_orbit_from_eof(_b(text), t0, t1, _b(source), C.byref(o))
    _raise_if_error()
    assert rc == 0
    return _take_orbit(o)


def orbit_pick_file(directory, prefix, t0, t1):
    """tcoh_orbit_pick_file (orbit.c:114-151): file name or None."""
    buf = C.create_string_buffer(1024)
    found = lib().tcoh_orbit_pick_file(_b(str(directory)), _b(prefix), t0, t1, buf,
                                       len(buf))
    return buf.value.decode() if found else None


def orbit_from_file(meta, kind, cache_dir, required=False, name="date"):
    """tcoh_orbit_from_file (orbit.c:205-301, local files): the orbit, or
    None when no file covers the acquisition and required is false."""
    o = Orbit()
    rc = lib().tcoh_orbit_from_file(_b(meta), _b(name), _b(kind), _b(str(cache_dir)),
                                    int(required), C.byref(o))
    _raise_if_error()
    return _take_orbit(o) if rc == 1 else None


def orbit_best(meta, cache_dir=None, name="date"):
    """tcoh_orbit_best (orbit.c:304-318)."""
    o = Orbit()
    rc = lib().tcoh_orbit_best(_b(meta), _b(name),
                               None if cache_dir is None else _b(str(cache_dir)),
                               C.byref(o))
    _raise_if_error()
    assert rc == 0
    return _take_orbit(o)


def scene_from_metadata(metas, ref, line0, sample0, rows, cols, names=None):
    """tcoh_scene_from_metadata (geometry.c:165-191): the scene dict that
    geometry_grid takes."""
    n = len(metas)
    names = names or [f"date{k + 1}" for k in range(n)]
    m = (C.c_char_p * n)(*[_b(x) for x in metas])
    nm = (C.c_char_p * n)(*[_b(x) for x in names])
    sc = Scene()
    rc = lib().tcoh_scene_from_metadata(m, nm, n, ref, line0, sample0, rows, cols,
                                        C.byref(sc))
    _raise_if_error()
    assert rc == 0
    return dict(t_first=sc.t_first, dt=sc.dt, tau0=sc.tau0, fs=sc.fs,
                **{"lambda": sc.lambda_}, line0=sc.line0, sample0=sc.sample0,
                rows=sc.rows, cols=sc.cols)


# ---- the module's data either side of the estimate (main.c; host) --------


def pair_names(prefix, n, pairs_all=False, starts=None, has_time=None):
    """tcoh_pair_names (main.c:602-645): the pair maps' names in ocl_run's
    pair order, and whether they carry dates. starts: one datetime (or
    (year, month, day, hour, minute, second)) per date, or None."""
    npairs = n * (n - 1) // 2 if pairs_all else n - 1
    stride = len(prefix) + 40
    buf = C.create_string_buffer(max(npairs, 1) * stride)
    st = None
    if starts is not None:
        st = (Utc * n)()
        for k, t in enumerate(starts):
            f = (t.year, t.month, t.day, t.hour, t.minute,
                 t.second + t.microsecond * 1e-6) if hasattr(t, "year") else tuple(t)
            st[k] = Utc(*f)
    ht = None if has_time is None else (C.c_int * n)(*[int(bool(x)) for x in has_time])
    dated = C.c_int()
    got = lib().tcoh_pair_names(_b(prefix), n, int(pairs_all), st, ht, buf, stride,
                                C.byref(dated))
    _raise_if_error()
    assert got == npairs
    names = [buf.raw[p * stride:(p + 1) * stride].split(b"\0", 1)[0].decode()
             for p in range(npairs)]
    return names, bool(dated.value)


class RawMaps:
    """tcoh_raw_create / put_strip / close (main.c:796-824, 921-922): the
    output maps <dir>/<name>.raw + .hdr, strips appended in row order;
    close() returns (estimated pixels, mean gamma)."""

    def __init__(self, directory, output, rows, cols, shp_count=None, pair_maps=()):
        self.npairs = len(pair_maps)
        names = (C.c_char_p * max(self.npairs, 1))(*[_b(x) for x in pair_maps])
        self.h = lib().tcoh_raw_create(_b(str(directory)), _b(output), _b(shp_count),
                                       names, self.npairs, rows, cols)
        _raise_if_error()
        if not self.h:
            raise RuntimeError("tcoh_raw_create failed")
        self.cols = cols

    def put_strip(self, coh, count, pairs=None):
        coh = np.ascontiguousarray(coh, np.float32).reshape(-1)
        count = np.ascontiguousarray(count, np.int32).reshape(-1)
        rows = coh.size // self.cols
        pr = None if not self.npairs else np.ascontiguousarray(pairs, np.float32)
        rc = lib().tcoh_raw_put_strip(self.h, rows, _ptr(coh), _ptr(count), _ptr(pr))
        _raise_if_error()
        assert rc == 0

    def close(self):
        est, mean = C.c_long(), C.c_double()
        h, self.h = self.h, None
        lib().tcoh_raw_close(h, C.byref(est), C.byref(mean))
        _raise_if_error()
        return est.value, mean.value


def run_raw_region(backend, path_i, path_q, rows, cols, directory, output,
                   shp_count=None, pairs_prefix=None, starts=None,
                   strip_rows=0):
    """The module's loop without GRASS: the I and Q date planes
    [N][rows][cols] float32 of two raw files, mapped in place
    (tcoh_raw_map) and streamed through the GPU (tcoh_run_region); the maps
    written with the module's nulls and names. Returns (estimated pixels,
    mean gamma, pair map names)."""
    s = backend.settings
    n = s.ndates
    nbytes = n * rows * cols * 4
    L = lib()
    addr = [C.c_void_p(), C.c_void_p()]
    try:
        for k, path in enumerate((path_i, path_q)):
            L.tcoh_raw_map(_b(str(path)), nbytes, C.byref(addr[k]))
            _raise_if_error()
        names = []
        if s.npairs:
            if pairs_prefix is None:
                raise ValueError("pairs_prefix is needed with pair maps")
            names, _ = pair_names(pairs_prefix, n, bool(s.pairs_all), starts)
        coh = np.empty(rows * cols, np.float32)
        count = np.empty(rows * cols, np.int32)
        pairs = np.empty((s.npairs, rows * cols), np.float32) if s.npairs else None
        L.tcoh_run_region(backend.h, rows, cols, addr[0], addr[1], None, None,
                          _ptr(coh), _ptr(count), _ptr(pairs), strip_rows)
        _raise_if_error()
    finally:
        for a in addr:
            L.tcoh_raw_unmap(a, nbytes)
    maps = RawMaps(directory, output, rows, cols, shp_count, names)
    maps.put_strip(coh, count, pairs)
    est, mean = maps.close()
    return est, mean, names
