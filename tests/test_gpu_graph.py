"""Graph mode (tcoh_set_graph): a strip captured into a HIP graph and
replayed gives the eager run's maps bit for bit, replay after replay, with
other eager work and synchronisations in between, and recaptures when the
call changes."""

import numpy as np
import pytest

import synth
from conftest import strip_of

pytestmark = pytest.mark.gpu

CASES = {
    "evd_f16": dict(n=12, g=0.7, seed=31, precision="f16",
                    opts=dict(window=(9, 5), min_shp=12, bias=True, pairs="consecutive")),
    "evd_bf16_tlog": dict(n=10, g=0.6, seed=32, precision="bf16",
                          opts=dict(window=(11, 5), shp_test="tlog", min_shp=10)),
    "evd_plain_deep": dict(n=40, g=0.8, seed=33, precision="plain",
                           opts=dict(window=(9, 7), min_shp=40, bias=True)),
    "emi_ad": dict(n=30, g=None, seed=0, precision="f16",
                   opts=dict(shp_test="ad", min_shp=20, estimator="emi", bias=True,
                             pairs="all")),
}


def case_stack(tcoh, c):
    if c["g"] is None:
        reg, params = synth.config_voronoi(n=30, rows=40, cols=40, seed=0,
                                                gamma0=(0.6, 0.99))
        return synth.scene_stack(30, 40, 40, reg, params, seed=0)
    return synth.ds_stack(c["n"], c["g"], seed=c["seed"])


@pytest.mark.parametrize("name", list(CASES))
def test_graph_replays_equal_eager(tcoh, name):
    c = CASES[name]
    D = tcoh.api.DeviceBuffer
    stack = case_stack(tcoh, c)
    n, rows, cols = stack.shape
    o = c["opts"]
    s = tcoh.make_settings(n, window=o.get("window", (21, 7)),
                           shp_test=o.get("shp_test", "ks"), min_shp=o["min_shp"],
                           estimator=o.get("estimator", "evd"), bias=o.get("bias", False),
                           pairs=o.get("pairs", "consecutive"))
    slc_h, amp_h, valid_h, prow, pcols = strip_of(stack, (s.win_az, s.win_rg))
    npx = rows * cols
    with tcoh.Backend(s, pcols, platform=0, device=0, precision=c["precision"]) as be, \
            tcoh.Backend(s, pcols, platform=0, device=0, precision="plain") as other:
        want = be.run(slc_h, amp_h, valid_h, prow, rows, cols)
        slc, amp, valid = D.from_array(slc_h), D.from_array(amp_h), D.from_array(valid_h)
        coh, count = D(npx * 4), D(npx * 4)
        pairs = D(max(1, s.npairs) * npx * 4)
        phases = D(n * npx * 4)
        be.set_graph(True)
        for rep in range(4):
            for b in (coh, count, pairs, phases):
                tcoh.api.lib().hipMemset(b.ptr, 0xFF, b.nbytes)
            be.run_device(slc.ptr, amp.ptr, valid.ptr, prow, rows, cols, coh.ptr,
                          count.ptr, pairs.ptr if s.npairs else None,
                          phases_ptr=phases.ptr)
            got_coh = coh.to_array(np.float32)
            got_count = count.to_array(np.int32)
            assert np.array_equal(got_count, want[1]), rep
            assert np.array_equal(got_coh.view(np.uint32), want[0].view(np.uint32)), rep
            if s.npairs:
                got_pairs = pairs.to_array(np.float32, (s.npairs, npx))
                assert np.array_equal(got_pairs.view(np.uint32), want[2].view(np.uint32))
            ph = phases.to_array(np.float32, (n, npx))
            assert np.array_equal(np.isnan(ph), np.broadcast_to(np.isnan(want[0]), ph.shape))
            # eager work of another handle between the replays
            other.run(slc_h, amp_h, valid_h, prow, rows, cols)
            tcoh.api.lib().hipDeviceSynchronize()
        # a different call (fewer output rows) captures anew, then the first
        # call's graph is captured again
        half = rows // 2
        be.run_device(slc.ptr, amp.ptr, valid.ptr, half + s.win_az - 1, half, cols,
                      coh.ptr, count.ptr, pairs.ptr if s.npairs else None)
        part = coh.to_array(np.float32)[:half * cols]
        assert np.array_equal(part.view(np.uint32), want[0][:half * cols].view(np.uint32))
        be.run_device(slc.ptr, amp.ptr, valid.ptr, prow, rows, cols, coh.ptr,
                      count.ptr, pairs.ptr if s.npairs else None)
        assert np.array_equal(coh.to_array(np.float32).view(np.uint32),
                              want[0].view(np.uint32))
