"""The serving stream generator: deterministic from its seed, sizes inside
the stated range on the mesh ladder, both case kinds, Re in [50, 1000],
arrivals increasing with the stated mean rate."""
from __future__ import annotations

import numpy as np

import serve_cases as sc


def test_stream_is_deterministic_and_in_range():
    a = sc.stream()
    b = sc.stream()
    assert a == b
    assert len(a) == 256
    cells = np.array([q.cells for q in a])
    assert cells.min() >= 1e4 * 0.9 and cells.max() <= 1e6 * 1.01
    assert {q.kind for q in a} == {"cavity", "channel"}
    for q in a:
        sides = sc.CAVITY_SIDES if q.kind == "cavity" else sc.CHANNEL_SIDES
        assert q.side in sides
        assert q.cells == (q.side ** 3 if q.kind == "cavity" else 4 * q.side ** 3)
        assert 50.0 <= q.re <= 1000.0
    t = np.array([q.arrival for q in a])
    assert np.all(np.diff(t) > 0)
    # Mean inter-arrival 1 / rate within a few standard errors.
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 0.25) < 4 * 0.25 / np.sqrt(len(gaps))


def test_stream_sizes_are_log_uniform():
    q = sc.stream(count=4000, seed=3)
    logs = np.log10([r.cells for r in q])
    # Rounded to the ladder, each decade holds about half of the requests.
    lower = np.mean(logs < 5.0)
    assert 0.4 < lower < 0.6


def test_serve_config_carries_the_precision():
    assert sc.serve_config(100.0, precision=1).pre.precision == 1
    assert sc.serve_config(100.0).mom.precision == 0
    assert abs(sc.serve_config(400.0).mu - 1.0 / 400.0) < 1e-15
