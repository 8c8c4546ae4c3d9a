"""Host logic of batches and launch shapes (no GPU): the cost estimate,
the longest-first list schedule against brute force and Graham's bound,
and the tuned table compiled into the library."""

import itertools
import os
import re

import numpy as np
import pytest

from conftest import PKG_DIR


def test_strip_cost_closed_form(tcoh):
    for n, win, rows, cols in [(8, (21, 7), 512, 512), (60, (21, 7), 24, 30),
                               (30, (9, 5), 2048, 3072)]:
        s = tcoh.make_settings(n, window=win, min_shp=n)
        w = win[0] * win[1]
        want = float(rows * cols) * (w * n * n + n ** 3)
        assert tcoh.api.strip_cost(s, rows, cols) == pytest.approx(want, rel=1e-15)


def lpt(costs, m):
    """Graham's LPT rule (1969): jobs by decreasing time, each to the
    machine free first (lowest index on ties)."""
    order = sorted(range(len(costs)), key=lambda i: (-costs[i], i))
    load = [0.0] * m
    lane = [0] * len(costs)
    for i in order:
        k = min(range(m), key=lambda j: (load[j], j))
        lane[i] = k
        load[k] += costs[i]
    return order, lane


def makespan(costs, lane, m):
    load = np.zeros(m)
    for c, k in zip(costs, lane):
        load[k] += c
    return load.max()


def test_schedule_is_lpt_and_within_grahams_bound(tcoh):
    rng = np.random.default_rng(7)
    for trial in range(60):
        m = int(rng.integers(1, 4))
        n = int(rng.integers(1, 8))
        # ties on purpose: integer costs from a small range
        costs = [float(c) for c in rng.integers(1, 6, size=n)]
        order, lane = tcoh.api.batch_schedule(costs, m)
        assert (order, lane) == lpt(costs, m)
        best = min(makespan(costs, a, m) for a in itertools.product(range(m), repeat=n))
        assert makespan(costs, lane, m) <= (4.0 / 3.0 - 1.0 / (3 * m)) * best + 1e-9


def test_schedule_of_the_batched_workload(tcoh):
    """The six stacks of the batched benchmark on 4 lanes: the two deepest
    (largest) first, each on its own lane."""
    shapes = {8: (2048, 3072), 12: (2048, 2048), 20: (1536, 2048), 30: (1024, 1024),
              45: (768, 768), 60: (512, 512)}
    ns = sorted(shapes)
    costs = [tcoh.api.strip_cost(tcoh.make_settings(n, min_shp=20), *shapes[n]) for n in ns]
    order, lane = tcoh.api.batch_schedule(costs, 4)
    assert order == sorted(range(6), key=lambda i: -costs[i])
    assert len(set(lane[i] for i in order[:4])) == 4


def test_tile_table_entries():
    text = open(os.path.join(PKG_DIR, "csrc", "tiles.h")).read()
    body = text[text.index("TILE_TABLE[] = {"):]
    rows = [tuple(int(v) for v in r.split(","))
            for r in re.findall(r"\{\s*(-?\d+(?:\s*,\s*-?\d+){5})\s*\}", body)]
    assert rows and rows[-1][2] == 0, "the table ends with its marker"
    keys = set()
    for emi, mixed, n, shp, link, band in rows[:-1]:
        assert emi in (0, 1) and mixed in (0, 1) and not (emi and mixed)
        assert 3 <= n <= 64 and shp >= 0 and link >= 0 and band > 0
        assert (emi, mixed, n) not in keys
        keys.add((emi, mixed, n))
