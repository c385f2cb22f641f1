"""The oracle's adjoint of the plain PFC step, pinned by central finite
differences of the oracle's own forward step (no reference code exists for
the adjoint; the forward step is 10_complete_pfc_model.cpp:121-144).

Loss J(psi0, eps) = sum (psi_T - target)^2 over a few steps at 16x8x8.
The adjoint chain gives dJ/dpsi0 (checked along random directions) and
dJ/deps; both must match the central differences to 1e-6 relative.  The
symmetry of A = F^-1 opL F (<A u, v> = <u, A v>) is checked on its own.
"""
import math

import numpy as np

from oracle import oracle as O

N = (16, 8, 8)
DX = math.pi / 4
DT = 0.5


def forward(psi0, eps, steps):
    opL, opN = O.pfc_operators(N, (DX,) * 3, eps, DT)
    traj = [psi0.copy()]
    psi = psi0.copy()
    for _ in range(steps):
        psi = O.pfc_step(psi, opL, opN, 1)
        traj.append(psi.copy())
    return traj


def loss(psi0, eps, target, steps):
    psiT = forward(psi0, eps, steps)[-1]
    return float(((psiT - target) ** 2).sum())


def adjoint(psi0, eps, target, steps):
    traj = forward(psi0, eps, steps)
    opL, opN = O.pfc_operators(N, (DX,) * 3, eps, DT)
    dL, dN = O.pfc_operator_derivs(N, (DX,) * 3, eps, DT)
    lam = 2.0 * (traj[-1] - target)
    dr = 0.0
    for n in range(steps - 1, -1, -1):
        lam, d = O.pfc_adjoint_step(traj[n], lam, opL, opN, dL, dN)
        dr += d
    return lam, dr


def fields(seed):
    rng = np.random.default_rng(seed)
    shape = (N[2], N[1], N[0])
    psi0 = -0.2 + 0.1 * rng.uniform(-1, 1, shape)
    target = -0.2 + 0.1 * rng.uniform(-1, 1, shape)
    return psi0, target


def test_adjoint_gradient_psi0_matches_central_differences():
    psi0, target = fields(1)
    eps, steps = -0.25, 4
    g, _ = adjoint(psi0, eps, target, steps)
    rng = np.random.default_rng(2)
    for _ in range(3):
        v = rng.standard_normal(psi0.shape)
        h = 1e-6
        fd = (loss(psi0 + h * v, eps, target, steps) -
              loss(psi0 - h * v, eps, target, steps)) / (2 * h)
        ad = float((g * v).sum())
        assert abs(fd - ad) <= 1e-6 * abs(ad), (fd, ad)


def test_adjoint_gradient_eps_matches_central_differences():
    psi0, target = fields(3)
    eps, steps = -0.3, 4
    _, dr = adjoint(psi0, eps, target, steps)
    h = 1e-6
    fd = (loss(psi0, eps + h, target, steps) - loss(psi0, eps - h, target, steps)) / (2 * h)
    assert abs(fd - dr) <= 1e-6 * abs(dr), (fd, dr)


def test_linear_part_is_symmetric():
    """<A u, v> = <u, A v> for A = F^-1 diag(opL) F (zero field: the adjoint
    step at psi = 0 is exactly A)."""
    opL, opN = O.pfc_operators(N, (DX,) * 3, -0.25, DT)
    rng = np.random.default_rng(4)
    shape = (N[2], N[1], N[0])
    u, v = rng.standard_normal(shape), rng.standard_normal(shape)
    zero = np.zeros(shape)
    Au, _ = O.pfc_adjoint_step(zero, u, opL, opN)
    Av, _ = O.pfc_adjoint_step(zero, v, opL, opN)
    assert abs(float((Au * v).sum()) - float((u * Av).sum())) <= 1e-12 * abs(float((Au * v).sum()))
