"""Per-mesh DG operators (host side, once per mesh and order).

``build_operators`` returns an object with the field names of the reference's
``Operators`` (``src/spatial/operators.py:138-206``), so that the device
backend accepts either the reference's operator set or this one.  The device
path itself only reads the topology, the node coordinates, the reference
tables and the edge metrics from it, and derives its compact per-element
metrics on its own.

``dense=True`` also materialises the reference's element-batched arrays
(``grad_x/y`` and the ``M^-1``-folded ``grad_x_m/grad_y_m/face_basis_m``,
``operators.py:270-318``).  They are about eleven times the size of the state
at order 3 and only the CPU oracle reads them, so large meshes skip them.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import numpy as np

from .mesh import MeshTopology
from .reference_element import ReferenceData, reference_data


@dataclass(frozen=True)
class Operators:
    ref: ReferenceData
    topology: MeshTopology
    node_coords: Any
    det_jac: Any  # (E, nqv)
    weighted_det: Any  # (E, nqv)
    elem_area: Any  # (E,)
    mean_weights: Any  # (E, N)
    xy_vol: Any  # (E, nqv, 2)
    edge_normal: Any  # (F, nqf, 2)
    edge_jac: Any  # (F, nqf)
    edge_length: Any  # (F,)
    xy_face: Any  # (F, nqf, 2)
    face_edge: np.ndarray  # (E, nface)
    face_sign: Any  # (E, nface)
    grad_x: Any = None
    grad_y: Any = None
    inv_mass: Any = None
    grad_x_m: Any = None
    grad_y_m: Any = None
    face_basis: Any = None
    face_basis_m: Any = None

    @property
    def n_elem(self) -> int:
        return int(self.topology.n_elem)

    @property
    def n_interior(self) -> int:
        return int(self.topology.edges.n_interior)

    @property
    def n_boundary(self) -> int:
        return int(self.topology.edges.n_boundary)

    @property
    def n_edges_total(self) -> int:
        return self.n_interior + self.n_boundary


def element_metrics(rd: ReferenceData, Xe: np.ndarray):
    """Jacobian determinant and inverse Jacobian at the volume points.

    ``Xe`` is ``(E, NQ, 2)``.  Returns ``det, dxi_dx, dxi_dy, deta_dx, deta_dy``,
    each ``(E, nqv)`` (``operators.py:245-268``).
    """
    dx_dxi = Xe[:, :, 0] @ rd.dgeom_dxi_vol
    dx_deta = Xe[:, :, 0] @ rd.dgeom_deta_vol
    dy_dxi = Xe[:, :, 1] @ rd.dgeom_dxi_vol
    dy_deta = Xe[:, :, 1] @ rd.dgeom_deta_vol
    det = dx_dxi * dy_deta - dx_deta * dy_dxi
    inv = 1.0 / det
    return det, dy_deta * inv, -dx_deta * inv, -dy_dxi * inv, dx_dxi * inv


def edge_metrics(rd: ReferenceData, Xe: np.ndarray, owner_elem, owner_face):
    """Unit owner-outward normal, ``|dX/dsigma|`` and length per edge."""
    Xo = Xe[owner_elem]
    dg = rd.dgeom_dsigma_face[owner_face]  # (F, NQ, nqf)
    tx = np.einsum("enq,en->eq", dg, Xo[:, :, 0])
    ty = np.einsum("enq,en->eq", dg, Xo[:, :, 1])
    jac = np.sqrt(tx * tx + ty * ty)
    normal = np.stack([ty, -tx], axis=-1) / jac[..., None]
    length = (jac * rd.w_face[None, :]).sum(axis=1)
    return normal, jac, length, Xo


def build_operators(topology: MeshTopology, node_coords, order: int, *, dense: bool = True,
                    check_jacobian: bool = True) -> Operators:
    rd = reference_data(topology.kind, order, topology.geometry_order)
    X = np.asarray(node_coords, dtype=np.float64)
    Xe = X[topology.elem_nodes]
    det, xix, xiy, etx, ety = element_metrics(rd, Xe)
    if check_jacobian and np.any(det <= 0.0):
        worst = int(np.argmin(det.min(axis=1)))
        raise ValueError(
            f"{int(np.sum(det <= 0.0))} quadrature point(s) have a non-positive Jacobian "
            f"determinant (worst element {worst}); the mesh is inverted or tangled"
        )
    wdet = rd.w_vol[None, :] * det
    area = wdet.sum(axis=1)
    mean_w = np.einsum("iq,eq->ei", rd.phi_vol, wdet) / area[:, None]
    xy_vol = np.einsum("nq,end->eqd", rd.geom_vol, Xe)

    ed = topology.edges
    own_e = np.concatenate([ed.iedge_elem[:, 0], ed.bedge_elem])
    own_f = np.concatenate([ed.iedge_face[:, 0], ed.bedge_face])
    normal, jac, length, Xo = edge_metrics(rd, Xe, own_e, own_f)
    xy_face = np.einsum("enq,end->eqd", rd.geom_face[own_f], Xo)
    face_sign = np.where(ed.face_side == 0, 1.0, -1.0)

    extra: dict[str, Any] = {}
    if dense:
        gx = wdet[:, None, :] * (rd.dphi_dxi_vol[None] * xix[:, None, :]
                                 + rd.dphi_deta_vol[None] * etx[:, None, :])
        gy = wdet[:, None, :] * (rd.dphi_dxi_vol[None] * xiy[:, None, :]
                                 + rd.dphi_deta_vol[None] * ety[:, None, :])
        mass = np.einsum("iq,jq,eq->eij", rd.phi_vol, rd.phi_vol, wdet)
        minv = np.linalg.inv(mass)
        nface = topology.n_faces
        fb = rd.phi_face[ed.face_side, np.arange(nface)[None, :]]
        extra = dict(
            grad_x=gx, grad_y=gy, inv_mass=minv,
            grad_x_m=np.einsum("eij,ejq->eiq", minv, gx),
            grad_y_m=np.einsum("eij,ejq->eiq", minv, gy),
            face_basis=fb,
            face_basis_m=np.einsum("eij,efjq->efiq", minv, fb),
        )
    return Operators(
        ref=rd, topology=topology, node_coords=X, det_jac=det, weighted_det=wdet,
        elem_area=area, mean_weights=mean_w, xy_vol=xy_vol, edge_normal=normal,
        edge_jac=jac, edge_length=length, xy_face=xy_face, face_edge=ed.face_edge,
        face_sign=face_sign, **extra,
    )
