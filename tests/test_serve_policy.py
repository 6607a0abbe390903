"""Host logic of the serving loop's size policy (no GPU): which running
requests count as large, and the per-step cost it compares."""
import types

import numpy as np

from nanogw_amd.batch import large_requests, step_cost


def test_large_requests_against_the_median_of_the_others():
    assert large_requests({}, 50.0) == set()
    assert large_requests({7: 1e12}, 50.0) == set()               # alone: never held
    assert large_requests({1: 1.0, 2: 1000.0}, 50.0) == {2}        # one small, one large
    assert large_requests({1: 1.0, 2: 40.0}, 50.0) == set()        # within the factor
    costs = {i: (1000.0 if i < 5 else 1.0) for i in range(10)}     # half large, half small
    assert large_requests(costs, 50.0) == {0, 1, 2, 3, 4}
    costs = {i: float(i + 1) for i in range(20)}                    # a smooth spread: none
    assert large_requests(costs, 50.0) == set()


def test_step_cost_counts_columns_times_operator_flop():
    def mol(nmu, nv, nb):
        return types.SimpleNamespace(psi_intp=np.zeros((nmu, nb)), nv=nv, e1=np.zeros(nb))

    # the c3 molecule: 8 states x 1600 bands x (4 Nmu nv nc + 2 Nmu^2)
    c3 = step_cost(mol(5000, 100, 1600), 8)
    assert c3 == 8 * 1600 * (4.0 * 5000 * 100 * 1500 + 2.0 * 5000 ** 2)
    # a c4-distribution molecule (nv = 30, nc = 300, Nmu = 360, 330 bands, 10 states)
    small = step_cost(mol(360, 30, 330), 10)
    assert small == 10 * 330 * (4.0 * 360 * 30 * 300 + 2.0 * 360 ** 2)
    assert large_requests({0: c3, 1: small, 2: small}, 50.0) == {0}
    # the largest c4 molecule (nv = 60) against a median one (nv = 38): not held
    big4 = step_cost(mol(720, 60, 660), 16)
    med4 = step_cost(mol(456, 38, 418), 16)
    assert large_requests({0: big4, 1: med4, 2: med4}, 50.0) == set()
