"""Synthetic workload generators (host): seeded, ragged ES population ranges."""
from __future__ import annotations

import numpy as np

from microcosmos_amd import workloads as W


def test_es_population_ranges_and_determinism():
    a = W.es_population(24, seed=2)
    b = W.es_population(24, seed=2)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.position, y.position)
        np.testing.assert_array_equal(x.act_omega, y.act_omega)
    for w in a:
        assert w.grid_shape[0] == w.grid_shape[1] and w.grid_shape[0] in (128, 192, 256)
        # line filaments are trees: components = nodes - edges
        n_fil = w.num_nodes - w.pairs.shape[0]
        assert 1 <= n_fil <= 16
        assert 0.0 <= w.act_amp.min() and w.act_amp.max() <= 0.5
        per = 2 * np.pi / w.act_omega
        assert per.min() >= 100.0 - 1e-3 and per.max() <= 400.0 + 1e-3
        assert 16 * n_fil <= w.num_nodes <= 64 * n_fil
