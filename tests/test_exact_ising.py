"""The exact ferromagnet energies the statistics tests compare with: Kaufman's partition function
(tests/exact_ising.py) against brute-force enumeration of small periodic lattices, including odd
and two-row ones, across the critical coupling; and the committed 16x16 table."""
import json
import os

import numpy as np
import pytest

import oracle as O
from tests.exact_ising import brute_force_mean_energy, mean_energy

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("m,n", [(4, 4), (3, 5), (2, 6), (4, 5)])
def test_kaufman_equals_enumeration(m, n):
    for beta in [0.05, 0.2, 0.4, np.log(1 + np.sqrt(2)) / 2, 0.5, 0.8, 1.5, 3.0]:
        assert abs(mean_energy(beta, m, n) - brute_force_mean_energy(beta, m, n)) < 1e-6 * m * n


def test_kaufman_table_for_the_16x16_ladder():
    t = json.load(open(os.path.join(HERE, "golden", "kaufman_L16.json")))
    assert np.array_equal(np.array(t["betas"], np.float32), O.beta_geometric(0.2, 5.0, 32))
    got = [mean_energy(b, 16, 16) for b in t["betas"]]
    assert np.allclose(got, t["mean_energy"], rtol=0, atol=1e-9)
    # limits: high temperature -> -2N tanh(beta) to leading order; ground state -2N
    assert abs(t["mean_energy"][-1] + 512) < 1e-6
    assert abs(mean_energy(1e-3, 16, 16) + 512 * np.tanh(1e-3)) < 1e-5  # next term is O(N beta^3)
