"""The HBM bandwidth probes the benchmark reports beside its roofline
figures: every kind runs and returns a positive duration, and an unknown
kind is refused."""
import math

import pytest

pytestmark = pytest.mark.gpu


def test_stream_probes_run(mlr):
    nb = 64 << 20
    for kind in range(5):
        ms = mlr.probe_stream(kind, nb, 3)
        gbs = (2 if kind in (0, 2) else 1) * nb / (ms * 1e-3) / 1e9
        print("kind %d: %.3f ms, %.0f GB/s" % (kind, ms, gbs))
        assert math.isfinite(ms) and ms > 0.0
    with pytest.raises(mlr.MallardError):
        mlr.probe_stream(5, nb, 1)
