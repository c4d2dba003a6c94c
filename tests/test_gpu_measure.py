"""The gather-rate probe behind the benchmark's gather fraction: both
patterns (random lane pairs, the NCC waves' tap-load layout) for both cell
sizes give a finite positive rate."""
import math

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cell_bytes", [16, 8])
@pytest.mark.parametrize("pattern", [0, 1])
def test_gather_peak_patterns(cell_bytes, pattern):
    from gmvs_amd import _native as N

    gbps = N.gather_peak(cell_bytes, pattern=pattern, rounds=50)
    assert math.isfinite(gbps) and gbps > 0.0
