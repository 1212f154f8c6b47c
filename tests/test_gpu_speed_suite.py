"""The reference-shaped speed sweeps run, report one record per measurement
(smallest sizes only), and return, at every sweep point, the oracle's rows
on a sample of 256 queries."""
import json

import numpy as np
import pytest

from cudaexactknn_amd import speed_suite

pytestmark = pytest.mark.gpu


def _checker(oracle, seen):
    def check(rec, pts, qs, rows):
        idx = np.random.Generator(np.random.Philox(len(seen))).choice(len(qs), min(256, len(qs)), replace=False)
        gi, gd = rows[0].cpu().numpy()[idx], rows[1].cpu().numpy()[idx]
        heap = rec["structure"] == "BallTreeGpu"   # the sweeps' ball trees run heap tracking
        wi, wd = oracle.bf_knn(pts, qs[idx], rec["k"], max_d2=rec["max_distance2"], heap=heap)
        assert np.array_equal(gi, wi), rec
        assert np.array_equal(gd.view(np.uint32), wd.view(np.uint32)), rec
        seen.append(rec)
    return check


@pytest.mark.parametrize("suite", ["construction", "speedup", "dimx", "dim3"])
def test_speed_suite_quick(oracle, cuda, capsys, suite):
    seen = []
    on_result = None if suite == "construction" else _checker(oracle, seen)
    speed_suite.SUITES[suite](True, 1, on_result)
    recs = [json.loads(line) for line in capsys.readouterr().out.splitlines() if line.startswith("{")]
    assert recs and all(r["suite"] == suite for r in recs)
    key = "build_ms" if suite == "construction" else "query_ms"
    assert all(r[key] > 0 for r in recs)
    if suite != "construction":
        assert len(seen) == len(recs)
    if suite == "dimx":
        assert all(r["max_distance2"] == float(np.float32(0.01 * r["dims"])) for r in recs)
