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
    thrust: float
    thrust_wall: float
    mass_flow_in: float
    mass_flow_out: float
    exit_mach_area_averaged: float
    exit_pressure_ratio: float


def performance(U, ops, flow) -> Performance:
    g = flow.gamma
    inlet = boundary_trace(U, ops, BoundaryTag.INFLOW)
    outlet = boundary_trace(U, ops, BoundaryTag.OUTFLOW)
    wall = boundary_trace(U, ops, BoundaryTag.WALL)
    thrust = 2.0 * (axial_momentum_flux(outlet, g) + axial_momentum_flux(inlet, g))
    _, _, _, pw = _prim(wall.state, g)
    thrust_wall = 2.0 * float(-(pw * wall.normal[..., 0] * wall.weight).sum())
    rho, vx, vy, p = _prim(outlet.state, g)
    area = float(outlet.weight.sum())
    mach = np.sqrt((vx * vx + vy * vy) * rho / (g * p))
    return Performance(
        thrust=thrust,
        thrust_wall=thrust_wall,
        mass_flow_in=2.0 * -mass_flux(inlet, g),
        mass_flow_out=2.0 * mass_flux(outlet, g),
        exit_mach_area_averaged=float((mach * outlet.weight).sum() / area),
        exit_pressure_ratio=float((p * outlet.weight).sum() / area / flow.total_pressure),
    )
