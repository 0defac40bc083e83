"""Exact discrete-adjoint design gradients on the device.

Restates the reference's ``DifferentiableCase`` (``src/sensitivity.py:134-460``)
around the hand-written residual VJP:

1. forward solve on a frozen logical grid (``case_from_grids``,
   ``src/api.py:213-244``), warm-started by the caller;
2. the objective ``J(U, X)`` and its partials on the host (boundary integrals
   of the owner trace, ``src/post/postprocess.py:56-75,155-171``);
3. ``(dR/dU)^T lambda = dJ/dU`` by the RK4 pseudo-time march of
   ``HipBackend.adjoint_solve``;
4. ``grad = dJ/da - lambda^T dR/da``: the device returns the cotangents of
   the metrics, edge normals/Jacobians and back pressure; they are chained to
   node coordinates here (reverse of ``build_operators``,
   ``src/spatial/operators.py:244-305``) and to design variables through the
   exact tangent of the closed-form node map ``nozzle_node_coords``
   (``mesh.node_coords_tangent``, a complex-step derivative: no truncation
   error, no cancellation).
"""

from __future__ import annotations

from collections.abc import Mapping, Sequence
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from .config import FlowConditions, SolverOptions
from .mesh import (BoundaryTag, NozzleGeometry, build_nozzle_mesh, node_coords_tangent,
                   nozzle_node_coords)
from .operators import build_operators

DESIGN_PARAMETERS = ("area_ratio", "throat_x", "throat_half_height", "inlet_half_height",
                     "theta_initial", "theta_exit", "bezier_w1", "bezier_w2", "length")
# the reference's objectives (src/sensitivity.py:90) and their ids in nz_objective
OBJECTIVES = ("thrust", "thrust_coefficient", "exit_mach", "mass_flow", "exit_pressure")
OBJECTIVE_IDS = {"thrust": 0, "thrust_coefficient": 1, "mass_flow": 2, "exit_mach": 3,
                 "exit_pressure": 4}


# ------------------------------------------------------------------ geometry chain
def metric_vjp(ops, gmet: np.ndarray) -> np.ndarray:
    """Node-coordinate cotangent from per-point metric cotangents.

    ``gmet[e, q] = d/d(xi_x, xi_y, eta_x, eta_y, w det)``; the forward map is
    ``J = [[x_xi, x_eta], [y_xi, y_eta]] = X^T dgeom``, ``det = x_xi y_eta -
    x_eta y_xi``, ``xi_x = y_eta/det``, ``xi_y = -x_eta/det``,
    ``eta_x = -y_xi/det``, ``eta_y = x_xi/det``, ``w det``.
    """
    rd = ops.ref
    topo = ops.topology
    X = np.asarray(ops.node_coords)
    Xe = X[topo.elem_nodes]  # (E, NQ, 2)
    x_xi = Xe[:, :, 0] @ rd.dgeom_dxi_vol
    x_eta = Xe[:, :, 0] @ rd.dgeom_deta_vol
    y_xi = Xe[:, :, 1] @ rd.dgeom_dxi_vol
    y_eta = Xe[:, :, 1] @ rd.dgeom_deta_vol
    det = x_xi * y_eta - x_eta * y_xi
    xix, xiy, etx, ety = y_eta / det, -x_eta / det, -y_xi / det, x_xi / det
    g = np.asarray(gmet)
    gxix, gxiy, getx, gety, gwd = (g[..., k] for k in range(5))
    gdet = rd.w_vol[None, :] * gwd - (gxix * xix + gxiy * xiy + getx * etx + gety * ety) / det
    g_yeta = gxix / det + gdet * x_xi
    g_xeta = -gxiy / det - gdet * y_xi
    g_yxi = -getx / det - gdet * x_eta
    g_xxi = gety / det + gdet * y_eta
    gXe = np.empty_like(Xe)
    gXe[:, :, 0] = g_xxi @ rd.dgeom_dxi_vol.T + g_xeta @ rd.dgeom_deta_vol.T
    gXe[:, :, 1] = g_yxi @ rd.dgeom_dxi_vol.T + g_yeta @ rd.dgeom_deta_vol.T
    gX = np.zeros_like(X)
    np.add.at(gX, topo.elem_nodes.reshape(-1), gXe.reshape(-1, 2))
    return gX


def edge_vjp(ops, gedge: np.ndarray) -> np.ndarray:
    """Node-coordinate cotangent from ``d/d(n_x, n_y, |dX/dsigma|)`` per edge point.

    Forward: ``t = sum_n dgeom_dsigma[f, n, q] X_n`` on the owner,
    ``jac = |t|``, ``n = (t_y, -t_x) / jac`` (``operators.py:293-302``).
    """
    rd = ops.ref
    topo = ops.topology
    ed = topo.edges
    X = np.asarray(ops.node_coords)
    own_e = np.concatenate([ed.iedge_elem[:, 0], ed.bedge_elem])
    own_f = np.concatenate([ed.iedge_face[:, 0], ed.bedge_face])
    nodes = topo.elem_nodes[own_e]  # (F, NQ)
    dg = rd.dgeom_dsigma_face[own_f]  # (F, NQ, nqf)
    Xo = X[nodes]
    tx = np.einsum("enq,en->eq", dg, Xo[:, :, 0])
    ty = np.einsum("enq,en->eq", dg, Xo[:, :, 1])
    jac = np.sqrt(tx * tx + ty * ty)
    j3 = jac**3
    g = np.asarray(gedge)
    gnx, gny, gj = g[..., 0], g[..., 1], g[..., 2]
    gtx = gnx * (-ty * tx / j3) + gny * (-1.0 / jac + tx * tx / j3) + gj * tx / jac
    gty = gnx * (1.0 / jac - ty * ty / j3) + gny * (tx * ty / j3) + gj * ty / jac
    gXo = np.stack([np.einsum("enq,eq->en", dg, gtx), np.einsum("enq,eq->en", dg, gty)], axis=-1)
    gX = np.zeros_like(X)
    np.add.at(gX, nodes.reshape(-1), gXo.reshape(-1, 2))
    return gX


# ------------------------------------------------------------------ autograd
_DEVICE_RESIDUAL = None


def _device_residual():
    """The ``torch.autograd.Function`` behind :func:`residual`, built on first use."""
    global _DEVICE_RESIDUAL
    if _DEVICE_RESIDUAL is not None:
        return _DEVICE_RESIDUAL
    import torch
    from torch.autograd.function import once_differentiable

    class DeviceResidual(torch.autograd.Function):
        """``R(U, X)`` by ``nz_residual``; backward by ``nz_residual_vjp``,
        which re-evaluates the flux Jacobians per face point instead of storing
        them.  The node-coordinate cotangent chains the kernel's metric and
        edge cotangents back to the nodes (``metric_vjp``, ``edge_vjp``)."""

        @staticmethod
        def forward(ctx, U, X, bk):
            R, _ = bk.residual(U)
            ctx.bk = bk
            ctx.x_device = None if X is None else X.device
            ctx.save_for_backward(U)
            return R

        @staticmethod
        @once_differentiable
        def backward(ctx, gR):
            (U,) = ctx.saved_tensors
            bk = ctx.bk
            gR = gR.to(dtype=torch.float64).contiguous()
            if not ctx.needs_input_grad[1]:
                return bk.residual_vjp(U, gR), None, None
            dU, gmet, gedge, _ = bk.residual_vjp(U, gR, geometry=True)
            gX = metric_vjp(bk.ops, gmet.cpu().numpy()) + edge_vjp(bk.ops, gedge.cpu().numpy())
            gX = torch.from_numpy(gX).to(device=ctx.x_device)
            return (dU if ctx.needs_input_grad[0] else None), gX, None

    _DEVICE_RESIDUAL = DeviceResidual
    return DeviceResidual


def residual(backend, U, X=None):
    """The unfolded DG residual ``R(U)`` (``(E, N, 4)``, the reference's
    ``DifferentiableCase.residual``, ``src/sensitivity.py:230-234``) as a
    differentiable torch operation on the device.

    ``backend`` is a ``HipBackend``; ``U`` an ``(E, N, 4)`` float64 tensor on
    its device.  ``X`` (optional, ``(n_nodes, 2)`` float64, any device) is the
    node-coordinate tensor the backend's operators were built from: the
    forward pass uses the backend's metrics, so ``X`` only carries the
    gradient ``d(<g, R>)/dX`` back to the caller; coordinates that differ from
    the backend's raise ``ValueError``.  Backward gives ``(dR/dU)^T g`` and,
    when ``X`` requires it, ``(dR/dX)^T g``; double backward is not supported.
    """
    import torch

    if not isinstance(U, torch.Tensor) or U.dtype != torch.float64 or U.device != backend.device:
        raise ValueError(f"U must be a float64 tensor on {backend.device}")
    fn = _device_residual()
    if X is None:
        return fn.apply(U, None, backend)
    Xref = np.asarray(backend.ops.node_coords)
    if tuple(X.shape) != Xref.shape or not np.array_equal(X.detach().cpu().numpy(), Xref):
        raise ValueError("X differs from the node coordinates the backend was built from")
    return fn.apply(U, X, backend)


# ------------------------------------------------------------------ objectives
def _trace(U, ops, tag):
    topo = ops.topology
    ed = topo.edges
    sl = topo.tag_slice(tag)
    phi = ops.ref.phi_face[0][ed.bedge_face[sl]]  # (n, N, nqf)
    S = np.einsum("kiq,kis->kqs", phi, np.asarray(U)[ed.bedge_elem[sl]])
    g = ops.n_interior + np.arange(sl.start, sl.stop)
    return S, phi, ed.bedge_elem[sl], g


def objective_and_partials(U, ops, flow, name: str = "thrust", geom=None):
    """``J``, ``dJ/dU`` ``(E, N, 4)`` and ``dJ/d(edge geometry)`` ``(F, nqf, 3)``
    on the host (``objective_value``, ``src/sensitivity.py:96-130``):
    ``thrust = 2 sum_{out, in} int (rho v_x v_n + p n_x) ds``, the thrust
    coefficient ``thrust / (p_t 2 y_throat)`` (needs ``geom``),
    ``mass_flow = -2 int_in rho v_n ds``, and the outflow averages
    ``int M ds / int ds`` (exit Mach) and ``int p ds / int ds / p_t`` (exit
    pressure).  Floors contribute no derivative when active.  The device
    computes the same in ``HipBackend.objective``.
    """
    if name not in OBJECTIVES:
        raise ValueError(f"unknown objective {name!r}; choose from {OBJECTIVES}")
    g = flow.gamma
    w = ops.ref.w_face
    J = 0.0
    dU = np.zeros(np.asarray(U).shape)
    dG = np.zeros((ops.n_edges_total, ops.ref.n_qface, 3))
    if name in ("thrust", "thrust_coefficient"):
        tags, fac = (BoundaryTag.OUTFLOW, BoundaryTag.INFLOW), 2.0
        if name == "thrust_coefficient":
            if geom is None:
                raise ValueError("the thrust coefficient needs the geometry (throat height)")
            fac /= flow.total_pressure * 2.0 * geom.throat_height()
    elif name == "mass_flow":
        tags, fac = (BoundaryTag.INFLOW,), -2.0
    else:
        tags, fac = (BoundaryTag.OUTFLOW,), 1.0
    averaged = name in ("exit_mach", "exit_pressure")
    num = area = 0.0
    keep = []
    for tag in tags:
        S, phi, belem, gidx = _trace(U, ops, tag)
        if S.size == 0:
            continue
        nx = ops.edge_normal[gidx][..., 0]
        ny = ops.edge_normal[gidx][..., 1]
        jac = ops.edge_jac[gidx]
        ds = jac * w[None, :]
        rho_raw = S[..., 0]
        rho = np.maximum(rho_raw, 1e-10)
        drho = (rho_raw > 1e-10).astype(float)
        m1, m2, E = S[..., 1], S[..., 2], S[..., 3]
        ke = 0.5 * (m1 * m1 + m2 * m2) / rho
        p_raw = (g - 1.0) * (E - ke)
        p = np.maximum(p_raw, 1e-10)
        dp_on = (p_raw > 1e-10).astype(float)
        vn_num = m1 * nx + m2 * ny  # rho v_n
        # d p / d S, with the density floor
        dp = np.stack([(g - 1.0) * ke / rho * drho, -(g - 1.0) * m1 / rho, -(g - 1.0) * m2 / rho,
                       np.full_like(rho, g - 1.0)], axis=-1) * dp_on[..., None]
        if averaged:
            if name == "exit_mach":
                q2 = m1 * m1 + m2 * m2
                Q = np.sqrt(((m1 / rho) ** 2 + (m2 / rho) ** 2) * rho / (g * p))
                safe = np.where(q2 > 0.0, q2, 1.0)
                h = 0.5 * Q
                dQ = np.stack([h * (-drho / rho - dp[..., 0] / p), h * (2 * m1 / safe - dp[..., 1] / p),
                               h * (2 * m2 / safe - dp[..., 2] / p), h * (-dp[..., 3] / p)], axis=-1)
                dQ = np.where((q2 > 0.0)[..., None], dQ, 0.0)
            else:
                Q = p / flow.total_pressure
                dQ = dp / flow.total_pressure
            num += float((Q * ds).sum())
            area += float(ds.sum())
            keep.append((phi, belem, gidx, Q, dQ, ds))
            continue
        if name in ("thrust", "thrust_coefficient"):
            I = m1 * vn_num / rho + p * nx
            dI = np.stack([-m1 * vn_num / rho**2 * drho, (vn_num + m1 * nx) / rho, m1 * ny / rho,
                           np.zeros_like(rho)], axis=-1) + dp * nx[..., None]
            dG[gidx, :, 0] += fac * (m1 * m1 / rho + p) * ds
            dG[gidx, :, 1] += fac * (m1 * m2 / rho) * ds
        else:
            I = vn_num
            dI = np.stack([np.zeros_like(rho), nx, ny, np.zeros_like(rho)], axis=-1)
            dG[gidx, :, 0] += fac * m1 * ds
            dG[gidx, :, 1] += fac * m2 * ds
        J += fac * float((I * ds).sum())
        dG[gidx, :, 2] += fac * I * w[None, :]
        gS = fac * dI * ds[..., None]  # (n, nqf, 4)
        np.add.at(dU, belem, np.einsum("kiq,kqs->kis", phi, gS))
    if averaged:
        J = num / area
        for phi, belem, gidx, Q, dQ, ds in keep:
            dG[gidx, :, 2] += (Q - J) * w[None, :] / area
            np.add.at(dU, belem, np.einsum("kiq,kqs->kis", phi, dQ * (ds / area)[..., None]))
    return J, dU, dG


class _Throat:
    """The throat height of a design, as the objectives read it."""

    def __init__(self, y):
        self.y = float(y)

    def throat_height(self) -> float:
        return self.y


# ------------------------------------------------------------------ the case
@dataclass
class DifferentiableCase:
    """A nozzle whose objective can be differentiated w.r.t. design variables.

    The logical grid (axial and transverse node fractions) is frozen at the
    base design, as in the reference, so that a perturbed design moves the
    nodes through ``nozzle_node_coords`` only.
    """

    geometry: NozzleGeometry = field(default_factory=lambda: NozzleGeometry(contour="smooth"))
    flow: FlowConditions = field(default_factory=FlowConditions)
    order: int = 1
    refine: int = 0
    kind: str = "tri"
    geometry_order: int = 1
    options: SolverOptions = field(default_factory=lambda: SolverOptions(tolerance=1e-11,
                                                                         print_interval=0))
    nx: int = 14
    nr: int = 5

    def __post_init__(self) -> None:
        topo, coords, grids = build_nozzle_mesh(self.geometry, kind=self.kind, nx=self.nx,
                                                nr=self.nr, refine=self.refine,
                                                geometry_order=self.geometry_order)
        self.topology = topo
        self.x_frac, self.r_frac = grids["x_frac"], grids["r_frac"]

    def default_params(self) -> dict[str, float]:
        d = dict(self.geometry.params())
        d["back_pressure"] = float(self.flow.back_pressure)
        return d

    def _merged(self, params: Mapping[str, Any] | None) -> dict[str, float]:
        d = self.default_params()
        for k, v in (params or {}).items():
            if k not in d:
                raise ValueError(f"{k!r} is not a design parameter")
            d[k] = float(v)
        return d

    def coords(self, params: Mapping[str, Any] | None = None) -> np.ndarray:
        return nozzle_node_coords(self.x_frac, self.r_frac, self._merged(params),
                                  self.geometry.contour)

    def flow_at(self, params=None) -> FlowConditions:
        pb = self._merged(params)["back_pressure"]
        return self.flow.replace(back_pressure_ratio=pb / self.flow.total_pressure)

    def operators(self, params=None):
        return build_operators(self.topology, self.coords(params), self.order, dense=False)

    def solve_forward(self, params=None, *, U0, backend=None):
        from .march import solve_steady

        ops = self.operators(params)
        return solve_steady(ops, self.flow_at(params), self.options, U0=U0, backend=backend or "hip")

    def value_and_gradient(self, objective: str = "thrust", params=None, *,
                           names: Sequence[str] | None = None, U0, adjoint_tolerance: float = 1e-8,
                           verbose: bool = False,
                           adjoint_method: str = "rk4"):
        """``(J, {name: dJ/dname}, forward_result, adjoint_info)``.

        ``adjoint_method``: ``"rk4"``, the reference's pseudo-time march, or
        ``"gmres"``, restarted GMRES on the same system (see
        ``HipBackend.adjoint_solve``).
        """
        from .backend import HipBackend

        base = self._merged(params)
        names = tuple(names) if names else ("area_ratio",)
        fwd = self.solve_forward(params, U0=U0)
        if not fwd.converged:
            raise RuntimeError(
                f"the forward solve did not converge, so no gradient can be taken: {fwd.message}")
        ops = fwd.operators
        flow = self.flow_at(params)
        bk = HipBackend(ops, flow, self.options)
        # objective and partials on the device (nz_objective); the geometry
        # cotangent comes back for the host chain to node coordinates
        J, J_U, J_edge = bk.objective(fwd.U, objective, geom=_Throat(base["throat_half_height"]))
        J_edge = J_edge.cpu().numpy()
        lam, rel, its, ok = bk.adjoint_solve(fwd.U, J_U, tolerance=adjoint_tolerance,
                                             verbose=verbose, method=adjoint_method)
        if not ok:
            raise RuntimeError(f"the adjoint march stalled at relative residual {rel:.2e} after "
                               f"{its} iterations")
        _, gmet, gedge, gpb = bk.residual_vjp(fwd.U, lam, geometry=True)
        gX = edge_vjp(ops, J_edge) - metric_vjp(ops, gmet) - edge_vjp(ops, gedge)
        grad = {}
        for k in names:
            if k == "back_pressure":
                grad[k] = -gpb
                continue
            dX = node_coords_tangent(self.x_frac, self.r_frac, base, self.geometry.contour, k)
            grad[k] = float((gX * dX).sum())
            if k == "throat_half_height" and objective == "thrust_coefficient":
                grad[k] -= J / base[k]  # the coefficient's scale 1 / (p_t 2 y_throat)
        info = {"adjoint_iterations": its, "adjoint_relative_residual": rel}
        return J, grad, fwd, info
