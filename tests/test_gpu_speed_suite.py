"""The reference-shaped speed sweeps run and report one record per
measurement (smallest sizes only)."""
import json

import pytest

from cudaexactknn_amd import speed_suite

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("suite", ["construction", "speedup", "dimx"])
def test_speed_suite_quick(cuda, capsys, suite):
    speed_suite.main([suite, "--quick", "--reps", "1"])
    recs = [json.loads(line) for line in capsys.readouterr().out.splitlines() if line.startswith("{")]
    assert recs and all(r["suite"] == suite for r in recs)
    key = "build_ms" if suite == "construction" else "query_ms"
    assert all(r[key] > 0 for r in recs)
