"""Pins the oracle's Tungsten mean-field PFC step against the reference's own
golden norms (apps/tungsten/tests/test_tungsten.cpp:180-354): 32^3 grid,
spacing 1.1107207345395915, origin at the centre, dt = 1, constant -0.4 then a
one-mode seed (rho -0.047, amplitude 0.215936, radius 18, untouched outside),
sum(psi^2) after each of 10 steps within 0.1 of the recorded values.

The derived parameters restate tungsten_params.hpp:339-411 and the operator
scalars make_operator_params (tungsten_spectral.hpp:68-73).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

# tungsten_single_seed.json values used by the golden test
PARAMS = dict(n0=-0.10, alpha=0.50, n_sol=-0.047, n_vap=-0.464, T0=156000.0, T=3300.0,
              Bx=0.8582, alpha_farTol=0.001, alpha_highOrd=4, lambda_=0.22, stabP=0.2,
              shift_s=0.1898, shift_u=0.3341, p2=1.0, p3=-0.5, p4=0.333333333,
              q20=-0.0037, q21=1.0, q30=-12.4567, q31=20.0, q40=45.0)

GOLDEN = [11965.0889218507, 11259.9705028338, 11050.6245088282, 10903.2783913748,
          10782.9029639299, 10678.0834708269, 10583.6114287826, 10496.5733807390,
          10415.1258360383, 10338.0182669812, 10264.3672697202]


def derived(p):
    s, u = p["shift_s"], p["shift_u"]
    tau = p["T"] / p["T0"]
    p2b = p["p2"] + 2 * s * p["p3"] + 3 * s ** 2 * p["p4"]
    p3b = u * (p["p3"] + 3 * s * p["p4"])
    p4b = u ** 2 * p["p4"]
    q20b = p["q20"] + 2.0 * s * p["q30"] + 3.0 * s ** 2 * p["q40"]
    q21b = p["q21"] + 2.0 * s * p["q31"]
    q30b = u * (p["q30"] + 3.0 * s * p["q40"])
    q31b = u * p["q31"]
    q40b = u ** 2 * p["q40"]
    q2b = q21b * tau + q20b
    q3b = q31b * tau + q30b
    q4b = q40b
    op = [2.0 * p["alpha"] * p["alpha"], 2.0 * p["lambda_"] * p["lambda_"], p["alpha_farTol"],
          float(p["alpha_highOrd"]), p["Bx"], p["T"], p["T0"], p["stabP"], p2b, q2b]
    nl = [p3b, p4b, q3b, q4b, p["stabP"]]
    return op, nl


def initial_field(n, dx, rho=-0.047, amp=0.215936, radius=18.0):
    a = np.full((n, n, n), -0.4)
    x0 = -0.5 * dx * n
    O.ic_single_seed_masked(a, (x0,) * 3, (dx,) * 3, rho, amp, radius, 1)
    return a


def test_tungsten_golden_norms():
    n, dx, dt = 32, 1.1107207345395915, 1.0
    op, nl = derived(PARAMS)
    fMF, expL, nW = O.tungsten_operators((n, n, n), (dx, dx, dx), op, dt)
    psi = initial_field(n, dx)
    norms = [float((psi * psi).sum())]
    for _ in range(10):
        psi, _mf = O.tungsten_step(psi, fMF, expL, nW, nl, 1)
        norms.append(float((psi * psi).sum()))
    for i, (g, r) in enumerate(zip(norms, GOLDEN)):
        assert abs(g - r) <= 0.1, (i, g, r)


def test_tungsten_operator_provenance():
    """n_weight = k_lap * phi1(L) with L = k_lap*opCk equals the legacy
    expm1(L dt)/opCk form (tungsten_spectral.hpp:126-147; the reference checks
    the two against each other in test_tungsten.cpp:585-609)."""
    n, dx, dt = 16, 1.1107207345395915, 1.0
    op, _ = derived(PARAMS)
    fMF, expL, nW = O.tungsten_operators((n, n, n), (dx, dx, dx), op, dt)
    f = 2 * math.pi / (dx * n)
    for (k, j, i) in [(0, 0, 1), (1, 2, 3), (5, 0, 8), (15, 15, 4)]:
        kx, ky, kz = [O.k_component(v, n, f) for v in (i, j, k)]
        klap = -(kx * kx + ky * ky + kz * kz)
        fm, opCk = O.tungsten_physics_for_mode(klap, op)
        assert fm == fMF[k, j, i]
        arg = klap * opCk * dt
        legacy = math.expm1(arg) / opCk
        assert abs(nW[k, j, i] - legacy) <= 1e-12 * abs(legacy)
        assert abs(expL[k, j, i] - math.exp(arg)) <= 1e-15 * max(1.0, abs(math.exp(arg)))
    # k = 0: exp(0) = 1 and n_weight = 0 (mass conservation)
    assert expL[0, 0, 0] == 1.0 and nW[0, 0, 0] == 0.0


def test_tungsten_conserves_mean():
    n, dx = 32, 1.1107207345395915
    op, nl = derived(PARAMS)
    fMF, expL, nW = O.tungsten_operators((n, n, n), (dx, dx, dx), op, 1.0)
    psi = initial_field(n, dx)
    out, _ = O.tungsten_step(psi, fMF, expL, nW, nl, 5)
    assert abs(out.mean() - psi.mean()) <= 1e-12


def test_host_tungsten_params_match_golden_derivation():
    """openpfc_amd.TungstenParams (the host mirror of tungsten_params.hpp) gives
    the parameter vectors that reproduce the golden norms above."""
    from openpfc_amd.tungsten import TungstenParams
    prm = TungstenParams.from_dict({**{k: v for k, v in PARAMS.items() if k != "lambda_"},
                                    "lambda": PARAMS["lambda_"]})
    op, nl = derived(PARAMS)
    assert prm.operator_params() == op
    assert prm.nonlinear_params() == nl
    assert prm.abi_params((1.0, 2.0, 3.0)) == [1.0, 2.0, 3.0] + op + nl[:4]
    # defaults are the tungsten_single_seed.json values
    assert TungstenParams() == TungstenParams.from_dict(
        {k: v for k, v in PARAMS.items() if k != "lambda_"} | {"lambda": 0.22})
    with pytest.raises(KeyError):
        TungstenParams.from_dict({"no_such": 1.0})


@pytest.mark.parametrize("n", [(16, 8, 12), (8, 6, 9)])
def test_tungsten_operators_same_bits_on_mirrored_kz_planes(n):
    """As for the PFC operators: the mean-field filter, exp(L dt) and the ETD
    weight depend on kz only through kz*kz, so the planes k and nz - k agree
    bit for bit (the premise of the tabulated z pass)."""
    dx = 1.1107207345395915
    op, _ = derived(PARAMS)
    fMF, expL, nW = O.tungsten_operators(n, (dx, 1.2 * dx, 0.9 * dx), op, 1.0)
    mirror = (-np.arange(n[2])) % n[2]
    for a in (fMF, expL, nW):
        assert np.array_equal(a, a[mirror])
        assert not np.array_equal(a[1], a[0])
