"""C3 precision study on the GPU (tests/precision_study.py): split hi/lo
positions with a binary64 range-to-phase agree with the full-binary64 oracle
within the IQ tolerance and the moment bounds, while the all-binary32 variant
(the reference's arithmetic) does not at 10 km range."""
import pytest

import precision_study as ps

pytestmark = pytest.mark.gpu


def test_c3_geometry_hilo_positions_match_the_binary64_oracle(gpu_ready):
    rows, live, R = ps.run(32768, 12)
    print()
    print(ps.table(rows))
    assert R == 202 and live >= 20
    hl = rows["hilo"]
    assert hl["iq"] < 1e-4, rows
    assert hl["z"] < 0.1 and hl["v"] < 0.05 and hl["zdr"] < 0.1, rows
    assert rows["fp32"]["iq"] > 10 * hl["iq"], rows
