"""Seeded serving request stream (host side, no GPU)."""
import numpy as np

from sors_amd import mesh


def test_requests_are_deterministic_and_in_range():
    a = [mesh.serving_request(i, seed=3) for i in range(24)]
    b = [mesh.serving_request(i, seed=3) for i in range(24)]
    kinds = set()
    for x, y in zip(a, b):
        assert x.name == y.name and x.steps == y.steps
        assert np.array_equal(x.X, y.X) and np.array_equal(x.conn, y.conn)
        assert x.youngs[0] == y.youngs[0] and x.poisson[0] == y.poisson[0]
        # node counts follow the log-uniform target up to the grid rounding
        assert 1_000 <= x.n_nodes <= 260_000
        assert 1e5 <= x.youngs[0] <= 1e7 and 0.30 <= x.poisson[0] <= 0.48
        assert 10 <= x.steps <= 50
        assert x.bcmask.reshape(-1, 3).any(1).sum() > 0  # clamped face
        kinds.add(x.name.split("_")[1])
    assert kinds == {"beam", "cube"}
    assert mesh.serving_request(0, seed=4).youngs[0] != a[0].youngs[0]


def test_cube_request_volume_and_orientation():
    sc = mesh.serving_request(1, seed=0, n_lo=300, n_hi=400)
    X = sc.X
    for e in sc.conn[:200]:
        a, b, c, d = X[e]
        assert np.dot(np.cross(b - a, c - a), d - a) != 0.0


def test_poisson_arrivals():
    t = mesh.serving_arrivals(4000, rate=200.0, seed=1)
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 1 / 200.0) < 0.1 / 200.0
    assert abs(gaps.std() - 1 / 200.0) < 0.1 / 200.0  # exponential: std = mean
