"""Tungsten mean-field PFC parameters (host side).

``TungstenParams`` mirrors the reference's parameter struct
(apps/tungsten/include/tungsten/common/tungsten_params.hpp:34-412): the base
parameters with their defaults (the values of tungsten_single_seed.json) and
the derived vapor-model coefficients, computed on access in the reference's
formulas and order (tungsten_params.hpp:339-411).  ``abi_params`` packs what
``pfc_hip_pfc_set_operators`` takes for PFC_HIP_MODEL_TUNGSTEN: the spacing,
the OperatorParams scalars of make_operator_params (tungsten_spectral.hpp:
68-73) and the nonlinear coefficients p3_bar, p4_bar, q3_bar, q4_bar.
"""
from __future__ import annotations

from dataclasses import dataclass, fields


@dataclass
class TungstenParams:
    n0: float = -0.10
    n_sol: float = -0.047
    n_vap: float = -0.464
    T: float = 3300.0
    T0: float = 156000.0
    Bx: float = 0.8582
    alpha: float = 0.50
    alpha_farTol: float = 0.001
    alpha_highOrd: int = 4
    lambda_: float = 0.22
    stabP: float = 0.2
    shift_u: float = 0.3341
    shift_s: float = 0.1898
    p2: float = 1.0
    p3: float = -0.5
    p4: float = 0.333333333
    q20: float = -0.0037
    q21: float = 1.0
    q30: float = -12.4567
    q31: float = 20.0
    q40: float = 45.0

    # -- derived (tungsten_params.hpp:339-411) ---------------------------------
    @property
    def tau(self) -> float:
        return self.T / self.T0

    @property
    def p2_bar(self) -> float:
        return self.p2 + 2 * self.shift_s * self.p3 + 3 * self.shift_s ** 2 * self.p4

    @property
    def p3_bar(self) -> float:
        return self.shift_u * (self.p3 + 3 * self.shift_s * self.p4)

    @property
    def p4_bar(self) -> float:
        return self.shift_u ** 2 * self.p4

    @property
    def q20_bar(self) -> float:
        return self.q20 + 2.0 * self.shift_s * self.q30 + 3.0 * self.shift_s ** 2 * self.q40

    @property
    def q21_bar(self) -> float:
        return self.q21 + 2.0 * self.shift_s * self.q31

    @property
    def q30_bar(self) -> float:
        return self.shift_u * (self.q30 + 3.0 * self.shift_s * self.q40)

    @property
    def q31_bar(self) -> float:
        return self.shift_u * self.q31

    @property
    def q40_bar(self) -> float:
        return self.shift_u ** 2 * self.q40

    @property
    def q2_bar(self) -> float:
        return self.q21_bar * self.tau + self.q20_bar

    @property
    def q3_bar(self) -> float:
        return self.q31_bar * self.tau + self.q30_bar

    @property
    def q4_bar(self) -> float:
        return self.q40_bar

    # -- packing ---------------------------------------------------------------
    def operator_params(self) -> list:
        """OperatorParams of make_operator_params (tungsten_spectral.hpp:68-73):
        alpha2, lambda2, alpha_farTol, alpha_highOrd, Bx, T, T0, stabP,
        p2_bar, q2_bar."""
        return [2.0 * self.alpha * self.alpha, 2.0 * self.lambda_ * self.lambda_,
                self.alpha_farTol, float(int(self.alpha_highOrd)), self.Bx, self.T,
                self.T0, self.stabP, self.p2_bar, self.q2_bar]

    def nonlinear_params(self) -> list:
        """p3_bar, p4_bar, q3_bar, q4_bar, stabP (cpu/tungsten_model.hpp:326-345)."""
        return [self.p3_bar, self.p4_bar, self.q3_bar, self.q4_bar, self.stabP]

    def abi_params(self, spacing) -> list:
        """The 17 parameters of PFC_HIP_MODEL_TUNGSTEN."""
        dx, dy, dz = (float(s) for s in spacing)
        return [dx, dy, dz] + self.operator_params() + self.nonlinear_params()[:4]

    @classmethod
    def from_dict(cls, d: dict) -> "TungstenParams":
        """Base parameters from a mapping with the reference's JSON keys
        ("lambda" for lambda_); unknown keys raise KeyError."""
        names = {f.name for f in fields(cls)}
        kw = {}
        for k, v in d.items():
            key = "lambda_" if k == "lambda" else k
            if key not in names:
                raise KeyError(f"TungstenParams: unknown parameter '{k}'")
            kw[key] = int(v) if key == "alpha_highOrd" else float(v)
        return cls(**kw)
