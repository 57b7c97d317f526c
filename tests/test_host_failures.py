"""Failure paths and sizing rules of the RS_* host layer, on the CPU with
host-only handles (no device):

* the status machine of rs.c: setters refuse once the domain is populated
  (rs.c:1469-1547, 2363) and the online calls refuse before it (rs.c:4134,
  5183, 5409); a second RS_populate is refused (the reference exits the
  process, rs.c:4484-4487; a library keeps the populated state instead);
* RS_revise_population's rounding of the suggested population (rs.c:2134-2211):
  below 50,000 scatterers the count is kept; above, it is rounded down to a
  multiple of 2 x compute units x workers x 1024 work items, and a count that
  rounds to 0 takes one multiple (the reference compares the rounded count with
  body_per_cell x 9 / 10); beyond 20,000,000 the population stays undefined and
  RS_populate refuses (the reference exits, rs.c:2204-2207);
* RS_make_pulse_params (rs.c:1284-1385): the local-memory loop halves the
  work items of pass 1 while the gate count is even and fails for an odd one
  (the reference exits; here the pass-2 method reads 0xFFFFFFFF), and each of
  the three pass-2 methods is chosen where the reference chooses it.  The
  expected launch shapes are worked by hand from those lines."""
import numpy as np
import pytest

import harness as h
import simradar_amd as sa

POPULATION_DEFINED = 1 << 1
DOMAIN_POPULATED = 1 << 4
# host-only handles model one MI355X: 256 compute units, one worker
MULTIPLE = 256 * 1 * 1024 * 2
UNIVERSAL, IN_RANGE, IN_LOCAL = 0, 1, 2


def _configured(n_rain=4096, density=None):
    r = sa.Radar(verb=0, host_only=True)
    if density is not None:
        r.set_density(float(density))
    h.configure(r, h.Config(n_rain=n_rain, scan="P:3.0,-1:1:0.01"))
    return r


def _snapshot(r):
    p = r.params()
    fields = {name: getattr(p, name) for name, _ in p._fields_}
    inf = r.info()
    return fields, r.sim_desc().copy(), [int(c) for c in inf.counts], int(inf.num_scats)


def test_setters_after_populate_keep_the_configuration():
    r = _configured()
    r.populate()
    assert int(r.info().status) & DOMAIN_POPULATED
    before = _snapshot(r)
    pos = r.stream("pos").copy()
    r.set_density(7.0)
    r.set_antenna_params(3.0, 30.0)
    r.set_tx_params(1.0e-6, 1.0e3)
    box = r.suggest_scan_domain()
    box[4] *= 2.0
    r.set_scan_box(box)
    nd, d = h.mp_dsd(50.0)
    r.set_dsd(nd, d, 1)
    r.set_population(123456)
    after = _snapshot(r)
    assert before[0] == after[0]
    assert np.array_equal(before[1], after[1])
    assert before[2:] == after[2:]
    assert np.array_equal(r.stream("pos"), pos)
    r.free()


def test_online_calls_before_populate_are_refused():
    r = _configured()
    inf0 = r.info()
    assert not (int(inf0.status) & DOMAIN_POPULATED)
    sd0 = r.sim_desc().copy()
    r.advance_time()        # (RS_advance_beam has no such guard, rs.c:5397-5401)
    r.make_pulse()
    inf1 = r.info()
    assert inf1.sim_tic == inf0.sim_tic and inf1.status == inf0.status
    assert np.array_equal(r.sim_desc(), sd0)
    r.free()


def test_second_populate_keeps_the_population():
    r = _configured(n_rain=2048)
    r.populate()
    n = int(r.info().num_scats)
    pos = r.stream("pos").copy()
    r.set_random_seed(99)
    r.populate()
    assert int(r.info().num_scats) == n
    assert np.array_equal(r.stream("pos"), pos)
    r.free()


def _suggested_count(r):
    """body_per_cell x the box volume in resolution cells at the first gate,
    evaluated in binary32 in the order of rs.c:2158-2166."""
    p = r.params()
    f = np.float32
    sd = r.sim_desc()
    rr = f(p.range_start)
    bw = f(p.antenna_bw_rad)
    svol = f(f(bw * rr) * f(bw * rr)) * f(f(p.c) * f(p.tau) * f(0.5))
    nvol = f(f(f(sd[12]) * f(sd[13])) * f(sd[14])) / svol
    return int(f(p.body_per_cell) * nvol), float(nvol)


@pytest.mark.parametrize("target", [30_000, 200_000, 1_300_000])
def test_revise_population_rounds_to_the_device_multiple(target):
    probe = _configured(n_rain=0)
    _, nvol = _suggested_count(probe)
    probe.free()
    r = _configured(n_rain=0, density=target / nvol)
    n, _ = _suggested_count(r)
    assert abs(n - target) < 0.01 * target
    inf = r.info()
    assert int(inf.status) & POPULATION_DEFINED
    got = int(inf.counts[0])
    if n <= 50_000:
        assert got == n
    elif n < MULTIPLE:
        assert got == MULTIPLE
    else:
        assert got == n // MULTIPLE * MULTIPLE
    r.free()


def test_population_beyond_twenty_million_is_refused():
    probe = _configured(n_rain=0)
    _, nvol = _suggested_count(probe)
    probe.free()
    r = _configured(n_rain=0, density=25.0e6 / nvol)
    assert not (int(r.info().status) & POPULATION_DEFINED)
    r.populate()
    assert not (int(r.info().status) & DOMAIN_POPULATED)
    # an explicit population below the cap defines it again
    r.set_population(4096)
    r.set_scan_box(r.suggest_scan_domain())
    assert int(r.info().status) & POPULATION_DEFINED
    r.populate()
    assert int(r.info().status) & DOMAIN_POPULATED
    assert int(r.info().counts[0]) == 4096
    r.free()


def _params(count, multiple, max_groups, max_local, R):
    lib = sa.load()
    return lib.RS_make_pulse_params(count, multiple, max_groups, max_local, 1000.0, 15.0, R)


def test_make_pulse_params_halves_work_items_until_local_memory_fits():
    # 1,000,000 scatterers, 40 gates, 64 KiB: 256 work items need 160 KiB,
    # 128 need 80 KiB, 64 need 40 KiB; the groups stay capped at 1024
    p = _params(1_000_000, 256, 1024, 65536, 40)
    assert (p.local[0], p.group_counts[0], p.global_[0], p.local_mem_size[0]) == (64, 1024, 65536, 40960)
    assert p.entry_counts[0] == 1_000_000 and p.entry_counts[1] == 1024 * 40
    # pass 2: 40,960 / 80 = 512 work items; 64 % 40 != 0 and 512 > 256: in range
    assert p.cl_pass_2_method == IN_RANGE
    assert (p.group_counts[1], p.global_[1], p.local[1], p.local_mem_size[1]) == (1, 40, 1, 16)


def test_make_pulse_params_fails_for_an_odd_gate_count_that_does_not_fit():
    p = _params(1_000_000, 256, 1024, 65536, 41)
    assert p.cl_pass_2_method == 0xFFFFFFFF
    assert p.local[0] == 256 and p.local_mem_size[0] == 41 * 256 * 16


def test_make_pulse_params_chooses_each_pass_two_method():
    # 1,000 scatterers, 8 gates: 256 work items, 2 groups, 32 KiB; pass 2 has
    # 2 x 8 / 16 = 1 work item and 256 % 8 == 0: universal
    p = _params(1000, 256, 1024, 65536, 8)
    assert (p.local[0], p.group_counts[0], p.local_mem_size[0]) == (256, 2, 32768)
    assert p.cl_pass_2_method == UNIVERSAL
    assert (p.global_[1], p.local[1], p.local_mem_size[1]) == (1, 1, 16)
    # 100,000 scatterers, 24 gates: 196 groups of 256; pass 2 has 196 x 24 / 48
    # = 98 work items, 256 % 24 != 0, 196 >= 48 and 98 <= 256: in local memory
    p = _params(100_000, 256, 1024, 1 << 20, 24)
    assert (p.local[0], p.group_counts[0]) == (256, 196)
    assert p.cl_pass_2_method == IN_LOCAL
    assert (p.global_[1], p.local[1], p.local_mem_size[1]) == (98, 98, 98 * 16)
