"""The reference-style C caller (tests/c/rs_caller.c: the driver loop of
simradar.c:1011-1074 over the RS_* API, compiled against include/rs.h and
linked with the engine library) run on the device: it exits 0 and its mean H
power over the pulses equals the same run made through the Python binding."""
import os
import re
import subprocess

import numpy as np
import pytest

import simradar_amd as sa

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _api_power(pulses):
    r = sa.Radar(verb=0)
    L = r.lib
    r.set_concept(sa.CONCEPT_DRAGGED | sa.CONCEPT_BOUNDED | sa.CONCEPT_UNIFORM_DSD)
    r.set_antenna_params(1.0, 44.5)
    r.set_tx_params(0.2e-6, 50.0e3)
    r.set_prt(1.0e-3)
    r.set_lambda(0.1)
    r.set_sampling_spacing(15.0, 1.0, 1.0)
    L.RS_set_dsd_to_mp(r.H)
    r.set_vel_data_to_config("suctvort")
    r.set_scan_box(r.suggest_scan_domain())
    r.set_random_seed(19760520)
    r.set_population(65536)
    r.populate()
    r.check()
    power = 0.0
    R = int(r.params().range_count)
    for k in range(pulses):
        L.RS_set_beam_pos(r.H, np.float32(-1.0) + np.float32(0.01) * np.float32(k), 3.0)
        r.make_pulse()
        r.download_pulse_only()
        p = r.pulse().astype(np.float64)
        for g in range(R):
            power += p[g, 0] * p[g, 0] + p[g, 1] * p[g, 1]
        r.advance_time()
    r.check()
    n = int(r.info().num_scats)
    r.free()
    return n, R, power / (pulses * R)


def test_reference_style_caller_on_the_device(gpu_ready, tmp_path):
    exe = str(tmp_path / "rs_caller")
    subprocess.run(["gcc", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c", "rs_caller.c"), "-L", os.path.join(ROOT, "simradar_amd"),
                    "-lsimradar_amd", "-Wl,-rpath," + os.path.join(ROOT, "simradar_amd"), "-lm", "-o", exe],
                   check=True)
    p = subprocess.run([exe, "long"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, (p.returncode, p.stdout, p.stderr)
    m = re.search(r"(\d+) scatterers, (\d+) gates, (\d+) pulses, mean H power (\S+)", p.stdout)
    assert m, p.stdout
    n, R, pulses, power = int(m.group(1)), int(m.group(2)), int(m.group(3)), float(m.group(4))
    assert pulses == 10 and n == 65536 and R > 0 and np.isfinite(power) and power > 0.0
    n2, R2, power2 = _api_power(pulses)
    assert (n2, R2) == (n, R)
    assert power == pytest.approx(power2, rel=1e-5)      # printed with 7 significant digits
