"""Batched domains (rsd_batch_*): several independent rain domains pulsed by one
persistent launch per round must give each domain exactly the state it gets
when run alone (bit-exact streams) and the same IQ up to reassociation of the
gate sums, with and without HIP-graph replay, through the autotuner's shape
sweep, and with a debris domain (launched on its own stream) in the batch."""
import numpy as np
import pytest

import harness as h
import simradar_amd as sa
from simradar_amd import workloads as wl

pytestmark = pytest.mark.gpu
IQ_TOL = 1e-4


def _domains(with_debris=False):
    cfgs = wl.c4(n_domains=4, seed=9, scale=1.0 / 32)
    if with_debris:
        cfgs[1] = h.Config(n_rain=3000, debris=[(sa.OBJ_LEAF, 256)], scan="P:3.0,-12:12:0.01",
                           les="suctvort:50.0,-30.0,77")
    return cfgs


def _run_alone(cfg, rounds):
    r = wl.make(cfg)
    iq = []
    for _ in range(rounds):
        r.make_pulse()
        r.download_pulse_only()
        iq.append(r.pulse_f64())
        r.advance_time()
        r.advance_beam()
    r.download()
    st = {s: r.stream(s) for s in ("pos", "vel", "ori", "tum", "rnd")}
    r.free()
    return np.array(iq), st


def _run_batched(cfgs, rounds, graph=True, shape=None):
    rs = [wl.make(c) for c in cfgs]
    b = sa.Batch(rs)
    b.set_graph(graph)
    if shape:
        b.set_launch(*shape)
    iq = [[] for _ in rs]
    for _ in range(rounds):
        for r in rs:
            r.make_pulse()
        b.launch()
        for k, r in enumerate(rs):
            r.download_pulse_only()
            iq[k].append(r.pulse_f64())
            r.advance_time()
            r.advance_beam()
    shape_used = b.launch_shape()
    out = []
    for k, r in enumerate(rs):
        r.download()
        out.append((np.array(iq[k]), {s: r.stream(s) for s in ("pos", "vel", "ori", "tum", "rnd")}))
    b.free()
    for r in rs:
        r.free()
    return out, shape_used


def _compare(alone, batched):
    for (iq_a, st_a), (iq_b, st_b) in zip(alone, batched):
        for s in st_a:
            assert np.array_equal(st_a[s].view(np.uint32), st_b[s].view(np.uint32)), s
        for k in range(len(iq_a)):
            assert h.rel_err(iq_b[k], iq_a[k]) < IQ_TOL


@pytest.mark.parametrize("graph", [True, False])
def test_batch_equals_domains_run_alone(gpu_ready, graph):
    cfgs = _domains()
    rounds = 6
    alone = [_run_alone(c, rounds) for c in cfgs]
    batched, _ = _run_batched(cfgs, rounds, graph=graph, shape=(1024, 4))
    _compare(alone, batched)


def test_batch_autotuner_sweep_keeps_results(gpu_ready):
    """30 rounds walk every autotuner shape (9 shapes x 3 launches) and settle."""
    cfgs = _domains()
    rounds = 30
    alone = [_run_alone(c, rounds) for c in cfgs]
    batched, shape = _run_batched(cfgs, rounds, graph=True)
    _compare(alone, batched)
    chunk, bpc, tuning = shape
    assert not tuning and chunk in (1024, 2048, 4096) and bpc in (2, 4, 8)


def test_batch_with_a_debris_domain(gpu_ready):
    cfgs = _domains(with_debris=True)
    rounds = 5
    alone = [_run_alone(c, rounds) for c in cfgs]
    batched, _ = _run_batched(cfgs, rounds, graph=True, shape=(2048, 2))
    _compare(alone, batched)
