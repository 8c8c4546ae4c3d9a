"""Batches of stacks (tcoh_batch_run) on a GPU, through the C ABI.

Six stacks of depth N in {8, 12, 20, 30, 45, 60} (the batched workload's
depths; KS for three, tlog for the other three, EVD, bias correction,
21x7 windows, per-stack seeds and grid shapes), run as one batch in the
split BF16 form (the library's default). Every stack's maps must equal its own tcoh_run_device
maps bit for bit, and those are held to the oracle as the single-stack
parity tests hold them (SHP counts bit-exact against the FP32 restatement
of the reference kernel, gamma and pair coherences within 1e-4 of the
FP64 oracle). In graph mode the captured batch is replayed three times,
with eager work between the replays, and every replay must give the same
maps.
"""

import numpy as np
import pytest

import synth
from conftest import strip_of
from test_gpu_parity import check_against_oracles, settings_for

pytestmark = pytest.mark.gpu

STACKS = [
    dict(n=8, rows=40, cols=48, g=0.7, seed=81, opts=dict(shp_test="ks", min_shp=8)),
    dict(n=12, rows=36, cols=44, g=0.6, seed=82, opts=dict(shp_test="tlog", min_shp=12)),
    dict(n=20, rows=34, cols=40, g=0.8, seed=83, opts=dict(shp_test="ks", min_shp=20)),
    dict(n=30, rows=30, cols=36, g=0.7, seed=84, opts=dict(shp_test="tlog", min_shp=30)),
    dict(n=45, rows=26, cols=32, g=0.4, seed=85, opts=dict(shp_test="ks", min_shp=30)),
    dict(n=60, rows=24, cols=30, g=0.3, seed=86, opts=dict(shp_test="tlog", min_shp=30)),
]
for _c in STACKS:
    _c["opts"].update(window=(21, 7), estimator="evd", bias=True, pairs="consecutive")


def device_stack(tcoh, c):
    D = tcoh.api.DeviceBuffer
    stack = synth.ds_stack(c["n"], c["g"], rows=c["rows"], cols=c["cols"], seed=c["seed"])
    s = settings_for(tcoh, c["n"], c["opts"])
    slc, amp, valid, prow, pcols = strip_of(stack, (s.win_az, s.win_rg))
    npx = c["rows"] * c["cols"]
    bufs = dict(slc=D.from_array(slc), amp=D.from_array(amp), valid=D.from_array(valid),
                coh=D(npx * 4, fill=0xFF), count=D(npx * 4, fill=0xFF),
                pairs=D(max(1, s.npairs) * npx * 4, fill=0xFF))
    return stack, s, prow, pcols, bufs


def item(be, c, prow, b):
    return (be, b["slc"].ptr, b["amp"].ptr, b["valid"].ptr, prow, c["rows"], c["cols"],
            b["coh"].ptr, b["count"].ptr, b["pairs"].ptr)


def maps(b, npairs, npx):
    return (b["coh"].to_array(np.float32), b["count"].to_array(np.int32),
            b["pairs"].to_array(np.float32)[:max(1, npairs) * npx])


def same(m1, m2):
    return (np.array_equal(m1[1], m2[1]) and
            np.array_equal(m1[0].view(np.uint32), m2[0].view(np.uint32)) and
            np.array_equal(m1[2].view(np.uint32), m2[2].view(np.uint32)))


@pytest.fixture(scope="module")
def batch_setup(tcoh):
    stacks = [device_stack(tcoh, c) for c in STACKS]
    bes = [tcoh.Backend(s, pcols, platform=0, device=0, precision="bf16")
           for _, s, _, pcols, _ in stacks]
    yield stacks, bes
    for be in bes:
        be.close()


def test_batch_equals_each_stack_and_the_oracle(tcoh, cref, fp64, batch_setup):
    stacks, bes = batch_setup
    items = [item(be, c, prow, b) for be, c, (_, _, prow, _, b) in zip(bes, STACKS, stacks)]
    with tcoh.Batch() as batch:
        batch.run(items)
        got = [maps(b, s.npairs, c["rows"] * c["cols"])
               for c, (_, s, _, _, b) in zip(STACKS, stacks)]
        order, lane = batch.plan()
    # the schedule: longest first by tcoh_strip_cost, on the 4 lanes
    costs = [tcoh.api.strip_cost(s, c["rows"], c["cols"])
             for c, (_, s, _, _, _) in zip(STACKS, stacks)]
    assert order == sorted(range(len(STACKS)), key=lambda i: -costs[i])
    assert sorted(set(lane)) == [0, 1, 2, 3]
    for c, (stack, s, prow, pcols, b), be, m in zip(STACKS, stacks, bes, got):
        npx = c["rows"] * c["cols"]
        # alone, through tcoh_run_device on the same buffers
        be.run_device(*item(be, c, prow, b)[1:])
        assert same(maps(b, s.npairs, npx), m), "N=%d: batch differs from its own run" % c["n"]
        # ocl_run's maps are those of tcoh_run_device; they meet the oracle
        coh, count, pairs = check_against_oracles(tcoh, cref, fp64, stack, c["opts"], "bf16")
        assert np.array_equal(count, m[1])
        assert np.array_equal(coh.view(np.uint32), m[0].view(np.uint32))
        assert np.array_equal(pairs.reshape(-1).view(np.uint32),
                              m[2][:s.npairs * npx].view(np.uint32))
        assert np.isfinite(m[0]).sum() > npx // 4, "N=%d: too few estimated pixels" % c["n"]


def test_batch_graph_replays(tcoh, batch_setup):
    stacks, bes = batch_setup
    items = [item(be, c, prow, b) for be, c, (_, _, prow, _, b) in zip(bes, STACKS, stacks)]
    with tcoh.Batch() as eager:
        eager.run(items)
        want = [maps(b, s.npairs, c["rows"] * c["cols"])
                for c, (_, s, _, _, b) in zip(STACKS, stacks)]
    # a stream of the library's runtime that can be captured
    stream = tcoh.api.DeviceStream()
    try:
        with tcoh.Batch(graph=True) as batch:
            for rep in range(4):  # eager + capture, then three replays
                for _, _, _, _, b in stacks:
                    tcoh.api.lib().hipMemset(b["coh"].ptr, 0xFF, b["coh"].nbytes)
                tcoh.api.lib().hipDeviceSynchronize()
                batch.run(items, stream=stream.ptr)
                stream.synchronize()
                for k, (c, (_, s, _, _, b)) in enumerate(zip(STACKS, stacks)):
                    got = maps(b, s.npairs, c["rows"] * c["cols"])
                    assert same(got, want[k]), "replay %d, N=%d" % (rep, c["n"])
                # eager work between the replays
                bes[0].run_device(*items[0][1:])
                tcoh.api.lib().hipDeviceSynchronize()
    finally:
        stream.destroy()
