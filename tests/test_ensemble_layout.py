"""Host-side layout of ensembles (pfc_hip_ensemble_layout; no device calls):
runs with equal extents, spacing and dt share a group, groups are ordered by
their first run and a group's fields are consecutive in the shared buffer."""
import math

import pytest

from openpfc_amd import PfcHipError, Run
from openpfc_amd.ensemble import layout

DX = math.pi / 4


def test_groups_are_consecutive_in_first_run_order():
    sizes = [32, 48, 32, 64, 48]
    runs = [Run.cubic(n, DX, -0.25, 0.5) for n in sizes]
    offs, total = layout(runs)
    a, b, c = 32 ** 3, 48 ** 3, 64 ** 3
    assert offs == [0, 2 * a, a, 2 * a + 2 * b, 2 * a + b]
    assert total == 2 * a + 2 * b + c


def test_eps_and_mean_do_not_split_groups_but_dt_and_spacing_do():
    runs = [Run.cubic(16, DX, -0.25, 0.5, -0.2), Run.cubic(16, DX, -0.30, 0.5, -0.25),
            Run.cubic(16, DX, -0.25, 0.25), Run.cubic(16, 0.7, -0.25, 0.5),
            Run((16, 16, 8), (DX, DX, DX), -0.25, 0.5)]
    offs, total = layout(runs)
    n = 16 ** 3
    # group 0: runs 0 and 1; then run 2 (dt), run 3 (dx), run 4 (extents)
    assert offs == [0, n, 2 * n, 3 * n, 4 * n]
    assert total == 4 * n + 16 * 16 * 8


@pytest.mark.parametrize("bad", [
    [],
    [Run((16, 16, 16), (DX, DX, DX), float("nan"), 0.5)],
    [Run((16, 16, 16), (0.0, DX, DX), -0.25, 0.5)],
    [Run((1, 16, 16), (DX, DX, DX), -0.25, 0.5)],
])
def test_invalid_runs_are_rejected(bad):
    with pytest.raises(PfcHipError) as e:
        layout(bad)
    assert e.value.code == 1
