"""Boundary observables of a converged state: thrust, mass flow, exit Mach.

Restates ``src/post/postprocess.py:56-75,155-171,193-226``: boundary traces
on the owner side at the face quadrature points, integrated with
``w * |dX/dsigma|``, times two for the full planar channel.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .mesh import BoundaryTag


@dataclass(frozen=True)
class BoundaryTrace:
    state: np.ndarray  # (n, nqf, 4)
    normal: np.ndarray  # (n, nqf, 2)
    weight: np.ndarray  # (n, nqf)


def boundary_trace(U, ops, tag) -> BoundaryTrace:
    topo = ops.topology
    sl = topo.tag_slice(tag)
    ed = topo.edges
    phi = ops.ref.phi_face[0][ed.bedge_face[sl]]
    state = np.einsum("kiq,kis->kqs", phi, np.asarray(U)[ed.bedge_elem[sl]])
    g = ops.n_interior + np.arange(sl.start, sl.stop)
    ds = ops.edge_jac[g] * ops.ref.w_face[None, :]
    return BoundaryTrace(state, ops.edge_normal[g], ds)


def _prim(S, gamma):
    rho = np.maximum(S[..., 0], 1e-10)
    vx = S[..., 1] / rho
    vy = S[..., 2] / rho
    p = np.maximum((gamma - 1.0) * (S[..., 3] - 0.5 * rho * (vx * vx + vy * vy)), 1e-10)
    return rho, vx, vy, p


def axial_momentum_flux(tr: BoundaryTrace, gamma: float) -> float:
    rho, vx, vy, p = _prim(tr.state, gamma)
    vn = vx * tr.normal[..., 0] + vy * tr.normal[..., 1]
    return float(((rho * vx * vn + p * tr.normal[..., 0]) * tr.weight).sum())


def mass_flux(tr: BoundaryTrace, gamma: float) -> float:
    rho, vx, vy, _ = _prim(tr.state, gamma)
    vn = vx * tr.normal[..., 0] + vy * tr.normal[..., 1]
    return float((rho * vn * tr.weight).sum())


@dataclass(frozen=True)
class Performance:
    """The reference's integral metrics of a solution (field names of
    ``src/post/postprocess.py:89-150``); fields that need the geometry (the
    throat and the quasi-1D ideal thrust) are NaN without it."""

    thrust: float
    thrust_wall: float
    mass_flow_in: float
    mass_flow_out: float
    exit_mach_area_averaged: float
    exit_pressure_ratio: float
    thrust_imbalance: float = float("nan")
    thrust_coefficient: float = float("nan")
    mass_imbalance: float = float("nan")
    entropy_error: float = float("nan")
    exit_temperature_ratio: float = float("nan")
    specific_thrust: float = float("nan")
    specific_thrust_ratio: float = float("nan")
    discharge_coefficient: float = float("nan")
    ideal_thrust: float = float("nan")
    thrust_efficiency: float = float("nan")


def entropy_integral(U, ops, flow) -> float:
    """Integral over the domain of ``(s / s_t - 1)^2``, ``s = p / rho^gamma``
    (``entropy_error``, ``src/post/postprocess.py:174-190``, before the
    division by the area and the square root)."""
    g = flow.gamma
    s_t = flow.total_pressure ** (1.0 - g) * (flow.Rgas * flow.total_temperature) ** g
    uq = np.einsum("iq,eis->eqs", ops.ref.phi_vol, np.asarray(U))
    rho, _, _, p = _prim(uq, g)
    ratio = (p / rho**g) / s_t
    return float(((ratio - 1.0) ** 2 * np.asarray(ops.weighted_det)).sum())


def ideal_thrust(q1d, flow) -> float:
    """Quasi-1D axial force of the same geometry, momentum form, full nozzle
    (``_ideal_thrust``, ``src/post/postprocess.py:258-263``)."""
    rho, vx, p, a = q1d.density, q1d.velocity, q1d.pressure, q1d.area
    return float((rho[-1] * vx[-1] ** 2 + p[-1]) * a[-1] - (rho[0] * vx[0] ** 2 + p[0]) * a[0])


def assemble(flow, *, thrust, thrust_wall, mdot_in, mdot_out, mach_int, p_int, t_int, area_e,
             entropy_int, area, geom=None, quasi1d=None) -> Performance:
    """The reference's performance fields (``postprocess.py:193-255``) from
    the integrals, shared by the host and the device paths."""
    denom = abs(thrust) if abs(thrust) > 0.0 else 1.0
    nan = float("nan")
    a_t = cf = cd = ideal = nan
    if geom is not None:
        a_t = 2.0 * geom.throat_height()
        cf = thrust / (flow.total_pressure * a_t)
        cd = mdot_in / (flow.stagnation_density * flow.stagnation_sound_speed * a_t)
        if quasi1d is None:
            from .initial import solve_quasi1d

            quasi1d = solve_quasi1d(geom, flow)
        ideal = ideal_thrust(quasi1d, flow)
    area_e = area_e if area_e != 0.0 else nan
    return Performance(
        thrust=thrust, thrust_wall=thrust_wall, mass_flow_in=mdot_in, mass_flow_out=mdot_out,
        exit_mach_area_averaged=mach_int / area_e,
        exit_pressure_ratio=p_int / area_e / flow.total_pressure,
        thrust_imbalance=abs(thrust - thrust_wall) / denom,
        thrust_coefficient=cf,
        mass_imbalance=abs(mdot_out - mdot_in) / (abs(mdot_in) if mdot_in else 1.0),
        entropy_error=float(np.sqrt(entropy_int / area)),
        exit_temperature_ratio=t_int / area_e / flow.total_temperature,
        specific_thrust=thrust / mdot_in if mdot_in else nan,
        specific_thrust_ratio=(thrust / mdot_in / flow.stagnation_sound_speed
                               if mdot_in else nan),
        discharge_coefficient=cd,
        ideal_thrust=ideal,
        thrust_efficiency=thrust / ideal if ideal == ideal and ideal else nan,
    )


def performance(U, ops, flow, geom=None, quasi1d=None) -> Performance:
    """All performance fields of ``U`` on the host (``postprocess.py:193-255``)."""
    g = flow.gamma
    inlet = boundary_trace(U, ops, BoundaryTag.INFLOW)
    outlet = boundary_trace(U, ops, BoundaryTag.OUTFLOW)
    wall = boundary_trace(U, ops, BoundaryTag.WALL)
    thrust = 2.0 * (axial_momentum_flux(outlet, g) + axial_momentum_flux(inlet, g))
    _, _, _, pw = _prim(wall.state, g)
    thrust_wall = 2.0 * float(-(pw * wall.normal[..., 0] * wall.weight).sum())
    rho, vx, vy, p = _prim(outlet.state, g)
    mach = np.sqrt((vx * vx + vy * vy) * rho / (g * p))
    t_e = p / (flow.Rgas * rho)
    return assemble(
        flow, thrust=thrust, thrust_wall=thrust_wall, mdot_in=2.0 * -mass_flux(inlet, g),
        mdot_out=2.0 * mass_flux(outlet, g), mach_int=float((mach * outlet.weight).sum()),
        p_int=float((p * outlet.weight).sum()), t_int=float((t_e * outlet.weight).sum()),
        area_e=float(outlet.weight.sum()), entropy_int=entropy_integral(U, ops, flow),
        area=float(np.asarray(ops.elem_area).sum()), geom=geom, quasi1d=quasi1d)
