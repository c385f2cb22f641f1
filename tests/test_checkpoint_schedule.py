"""The binomial checkpoint schedule of the autograd trajectory (host code of
openpfc_amd/csrc/torch_pfc.cpp; no GPU needed).

The schedule's recurrence T(n, s) (forward steps of the reversal of n steps
with the start held and s free slots) is pinned by the closed form of
Griewank and Walther (Algorithm 799 "revolve", ACM TOMS 26(1), 2000): with
S = s + 1 checkpoints counting the start and t the integer with
beta(S, t-1) < n <= beta(S, t), beta(S, t) = C(S + t, S), the minimum is
t*n - beta(S + 1, t - 1).  T(n, 0) = n(n-1)/2 is the textbook no-snapshot
cost.  The uniform schedule's counts are checked by hand.
"""
from math import comb

import pytest


def _closed(n, s):
    S = s + 1
    t = 0
    while comb(S + t, S) < n:
        t += 1
    return 0 if t == 0 else t * n - comb(S + t, S + 1)


@pytest.fixture(scope="module")
def A():
    from openpfc_amd import autograd as A
    A.load_extension()
    return A


@pytest.mark.parametrize("snaps", [0, 1, 2, 3, 5, 8])
def test_binomial_cost_matches_revolve_closed_form(A, snaps):
    costs = A.binomial_costs(240, snaps)
    for n in range(1, 241):
        assert costs[n - 1] == _closed(n, snaps), (n, snaps, costs[n - 1], _closed(n, snaps))
    if snaps == 0:
        assert costs == [n * (n - 1) // 2 for n in range(1, 241)]


def test_schedule_costs_of_the_adjoint_workload(A):
    # 200 steps with a checkpoint every 20: 10 segments of 19 recomputed steps
    # each, holding 9 checkpoints besides psi0 plus one segment of 20 states
    assert A.schedule_cost(200, 20, 0) == (190, 29)
    # the binomial schedule with fewer fields recomputes fewer steps
    rec, held = A.schedule_cost(200, 20, 19)
    assert held == 19 and rec < 190
    # more snapshots never cost more
    prev = None
    for s in range(1, 40):
        r, _ = A.schedule_cost(200, 20, s)
        assert prev is None or r <= prev
        prev = r
    # with a slot for every step nothing is recomputed
    assert A.schedule_cost(50, 20, 49)[0] == 0


def test_schedule_rejects_bad_arguments(A):
    with pytest.raises(Exception):
        A.schedule_cost(0, 20, 0)
    with pytest.raises(Exception):
        A.schedule_cost(10, 20, -1)
