"""Operating point and march options, with the reference's field names and defaults.

``FlowConditions`` and ``SolverOptions`` mirror ``src/config.py:43-419`` so
the device backend can be constructed from either package's objects; only
attributes are read (duck typing).  Validation follows the reference's
messages for the fields the device path consumes.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, replace
from typing import Any

FLUXES = ("roe", "hllc", "ausmup")
SCHEMES = ("euler", "ssprk3")
LIMITERS = ("none", "positivity", "superbee")


@dataclass(frozen=True)
class FlowConditions:
    gamma: float = 1.4
    Rgas: float = 0.4
    total_temperature: float = 1.0
    total_pressure: float = 1.0
    back_pressure_ratio: float = 0.15
    inflow_angle: float = 0.0
    entropy_fix: float = 0.05
    flux: str = "roe"
    hllc_low_mach: float = 0.0
    backflow: bool = False
    allow_shock_in_nozzle: bool = False

    def __post_init__(self) -> None:
        if self.gamma <= 1.0:
            raise ValueError(f"gamma must exceed 1, got {self.gamma}")
        if self.Rgas <= 0.0 or self.total_pressure <= 0.0 or self.total_temperature <= 0.0:
            raise ValueError("Rgas, total_pressure and total_temperature must be positive")
        if not 0.0 < self.back_pressure_ratio <= 1.0:
            raise ValueError(
                f"back_pressure_ratio must lie in (0, 1], got {self.back_pressure_ratio}")
        if self.entropy_fix < 0.0:
            raise ValueError("entropy_fix must be non-negative")
        if self.flux not in FLUXES:
            raise ValueError(f"flux must be one of {sorted(FLUXES)}, got {self.flux!r}")
        if self.hllc_low_mach < 0.0:
            raise ValueError("hllc_low_mach must be non-negative")

    @property
    def back_pressure(self) -> float:
        return self.back_pressure_ratio * self.total_pressure

    @property
    def stagnation_sound_speed(self) -> float:
        return math.sqrt(self.gamma * self.Rgas * self.total_temperature)

    @property
    def stagnation_density(self) -> float:
        return self.gamma * self.total_pressure / self.stagnation_sound_speed**2

    def replace(self, **kw: Any) -> "FlowConditions":
        return replace(self, **kw)


@dataclass(frozen=True)
class SolverOptions:
    cfl: float | None = None
    tolerance: float = 1e-6
    max_iterations: int = 200_000
    scheme: str | None = None
    limiter: str = "positivity"
    tvb_constant: float = 50.0
    positivity_cfl: float = 0.0
    initial_condition: str = "quasi1d"
    p_continuation: bool = False
    check_interval: int = 50
    print_interval: int = 500
    divergence_factor: float = 1e4
    stall_window: int = 40
    stall_ratio: float = 0.98

    def __post_init__(self) -> None:
        if self.cfl is not None and self.cfl <= 0.0:
            raise ValueError(f"cfl must be positive, got {self.cfl}")
        if self.tolerance <= 0.0:
            raise ValueError(f"tolerance must be positive, got {self.tolerance}")
        if self.max_iterations < 1:
            raise ValueError("max_iterations must be at least 1")
        if self.check_interval < 1:
            raise ValueError("check_interval must be at least 1")
        if self.stall_window < 2:
            raise ValueError("stall_window must be at least 2")
        if not 0.0 < self.stall_ratio < 1.0:
            raise ValueError(f"stall_ratio must lie in (0, 1), got {self.stall_ratio}")
        if self.scheme is not None and self.scheme not in SCHEMES:
            raise ValueError(f"unknown scheme {self.scheme!r}; choose from {SCHEMES}")
        if not 0.0 <= self.positivity_cfl <= 1.0:
            raise ValueError(f"positivity_cfl must lie in [0, 1], got {self.positivity_cfl}")
        if self.tvb_constant < 0.0:
            raise ValueError(f"tvb_constant must be non-negative, got {self.tvb_constant}")
        if self.limiter not in LIMITERS:
            raise ValueError(
                f"limiter must be 'none', 'positivity' or 'superbee', got {self.limiter!r}")

    def replace(self, **kw: Any) -> "SolverOptions":
        return replace(self, **kw)


# ----------------------------------------------------------------- time-step constants
# Measured stability limits and the 70 % margin (src/timestepping/step.py:34-37,52-55,80).
STABILITY_LIMIT = {"ssprk3": (1.437, 2.337, 2.240), "euler": (1.169, 1.337, 0.775)}
STABILITY_FALLBACK = {"ssprk3": ("grow", 1.32), "euler": ("flat", 0.70)}
CFL_MARGIN = 0.7
SSP_FROM_ORDER = 1
FLOOR = 1e-10


def recommended_scheme(order: int) -> str:
    return "euler" if order < SSP_FROM_ORDER else "ssprk3"


def stability_limit(order: int, scheme: str = "ssprk3") -> float:
    if scheme not in STABILITY_LIMIT:
        raise ValueError(f"unknown scheme {scheme!r}; choose from {tuple(STABILITY_LIMIT)}")
    table = STABILITY_LIMIT[scheme]
    if order < len(table):
        return table[order]
    rule, k = STABILITY_FALLBACK[scheme]
    return k if rule == "flat" else k * (2 * order + 1) / (order + 1)


def recommended_cfl(order: int, scheme: str = "ssprk3") -> float:
    return CFL_MARGIN * stability_limit(order, scheme)


def step_coefficient(order: int, cfl: float | None, scheme: str = "ssprk3") -> float:
    """``c`` in ``dt_e = c * 2 A_e / sum_f s_f l_f`` (``step.py:150-168``)."""
    if cfl is None:
        cfl = recommended_cfl(order, scheme)
    return cfl / (2 * order + 1)


def residual_scale(flow: FlowConditions, length: float) -> float:
    return float(flow.stagnation_density * flow.stagnation_sound_speed / max(length, 1e-30))


def component_weights(flow) -> tuple[float, float, float, float]:
    """``(1, 1/a_t, 1/a_t, 1/a_t^2)`` on the state axis (``src/backends/base.py:42-78``)."""
    a_t = flow.stagnation_sound_speed
    return (1.0, 1.0 / a_t, 1.0 / a_t, 1.0 / a_t**2)


def reference_floors(flow, fraction: float = 1e-8) -> tuple[float, float]:
    """Absolute density and pressure floors (``src/spatial/limiter.py:98-100``)."""
    return fraction * flow.stagnation_density, fraction * flow.total_pressure
