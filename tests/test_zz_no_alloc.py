"""The allocation-free entry point (tcoh_run_device_noalloc) against
tcoh_run_device on the same device buffers, through the C ABI: the same
maps bit for bit, on two calls in a row, the second with a smaller strip
(the fixed scratch array keeps the first call's contents), for the split
FP16 form with bias correction and pairs, the split BF16 form at 30 dates
with a 21x7 window, and for EMI (queues in the
scratch); and a strip beyond the scratch is refused with an error."""

import ctypes as C

import numpy as np
import pytest

import synth
from conftest import strip_of

pytestmark = pytest.mark.gpu


def noalloc(tcoh):
    L = tcoh.api.lib()
    f = L.tcoh_run_device_noalloc
    f.argtypes = [C.c_void_p] * 4 + [C.c_int] * 3 + [C.c_void_p] * 4
    return f


CASES = {
    "f16_evd_bias": dict(n=12, precision="f16", window=(9, 5), min_shp=12, bias=True,
                         pairs="consecutive", estimator="evd", shp_test="ks"),
    "bf16_evd_n30": dict(n=30, precision="bf16", window=(21, 7), min_shp=30, bias=True,
                         pairs="consecutive", estimator="evd", shp_test="ks"),
    "plain_emi_ad": dict(n=10, precision="plain", window=(9, 5), min_shp=10, bias=True,
                         pairs="all", estimator="emi", shp_test="ad"),
}


@pytest.mark.parametrize("name", list(CASES))
def test_noalloc_equals_allocating_entry(tcoh, name):
    c = CASES[name]
    D = tcoh.api.DeviceBuffer
    run_noalloc = noalloc(tcoh)
    n = c["n"]
    s = tcoh.make_settings(n, window=c["window"], shp_test=c["shp_test"],
                           min_shp=c["min_shp"], estimator=c["estimator"], bias=c["bias"],
                           pairs=c["pairs"])
    # a second, smaller strip of the same width: fewer rows
    for rows in (40, 17):
        stack = synth.ds_stack(n, 0.6, rows=rows, cols=36, seed=rows)
        slc_h, amp_h, valid_h, prow, pcols = strip_of(stack, c["window"])
        npx = rows * 36
        slc, amp, valid = D.from_array(slc_h), D.from_array(amp_h), D.from_array(valid_h)
        outs = []
        with tcoh.Backend(s, pcols, platform=0, device=0, precision=c["precision"]) as be:
            for entry in ("alloc", "noalloc"):
                coh, count = D(npx * 4, fill=0xFF), D(npx * 4, fill=0xFF)
                pairs = D(max(1, s.npairs) * npx * 4, fill=0xFF)
                args = (slc.ptr, amp.ptr, valid.ptr, prow, rows, 36, coh.ptr, count.ptr,
                        pairs.ptr)
                if entry == "alloc":
                    be.run_device(*args)
                else:
                    run_noalloc(be.h, *args, None)
                    tcoh.api._raise_if_error()
                outs.append((coh.to_array(np.float32), count.to_array(np.int32),
                             pairs.to_array(np.float32)))
        (c1, k1, p1), (c2, k2, p2) = outs
        assert np.array_equal(k1, k2)
        assert np.array_equal(c1.view(np.uint32), c2.view(np.uint32))
        assert np.array_equal(p1.view(np.uint32), p2.view(np.uint32))
        assert np.isfinite(c1).sum() > 100


def test_noalloc_refuses_oversize_strips(tcoh):
    run_noalloc = noalloc(tcoh)
    api = tcoh.api
    s = tcoh.make_settings(30, window=(21, 7), min_shp=30, bias=True, pairs="consecutive")
    rows, cols = 1024, 1024
    pcols, prow = cols + 6, rows + 20
    p = api.DeviceBuffer(64)
    api.lib().tcoh_set_error_mode(1)
    with tcoh.Backend(s, pcols, platform=0, device=0, precision="f16") as be:
        run_noalloc(be.h, p.ptr, p.ptr, p.ptr, prow, rows, cols, p.ptr, p.ptr, p.ptr, None)
        assert "allocation-free" in api.lib().tcoh_last_error().decode()
