"""The synthetic workloads have the shapes of the configurations they stand
for (no GPU needed)."""
import numpy as np

from painty_amd import workload as W
from tests.test_gpu_fullsize import served_stream


def test_served_stream_shape():
    reqs, follow = served_stream()
    assert follow >= 0.3 * len(reqs)
    sides = [q.side for q in reqs]
    assert min(sides) >= 512 and max(sides) <= 4096 and max(sides) >= 3000
    assert all(10 <= len(q.stamps) <= 1000 for q in reqs)
    assert all(16 <= q.radius <= 64 for q in reqs)
    assert all(q.stamps["flags"][0] & W.DIP for q in reqs)
    # stamps scaled to their canvas (strokes may leave it, as on the reference's canvas)
    for q in reqs:
        inside = (q.stamps["cx"] >= 0) & (q.stamps["cx"] < q.side)
        assert inside.mean() > 0.5


def test_served_requests_deterministic():
    a = W.served_requests(12, seed=3)
    b = W.served_requests(12, seed=3)
    for x, y in zip(a, b):
        assert x.canvas_id == y.canvas_id and x.side == y.side and x.radius == y.radius
        assert np.array_equal(x.stamps, y.stamps)
