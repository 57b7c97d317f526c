"""Shared set-up of the test workloads: the engine run (simradar_amd) and the
matching oracle tables and stepping."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import simradar_amd as sa  # noqa: E402
from simradar_amd.workloads import Config, configure, mp_dsd, SEED  # noqa: E402,F401
from oracle import oracle as orc  # noqa: E402


from oracle.run import OracleRun, adm_desc, rcs_desc, state_from_radar  # noqa: E402,F401


def rel_err(a, b):
    """max |a - b| over gates, relative to the RMS gate magnitude of b."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = np.sqrt(np.mean(np.sum(b * b, axis=-1)))
    if scale == 0.0:
        return float(np.max(np.abs(a - b)))
    return float(np.max(np.abs(a - b)) / scale)
