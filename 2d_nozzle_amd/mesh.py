"""Mesh connectivity, face maps and node coordinates (host side, once per mesh).

The integer maps built here are what the device kernels gather through, and
they must equal the reference's bit for bit, so the ordering rules are kept:

* faces are matched on their sorted corner-node pair, visited in
  ``(lo, hi)`` lexicographic order (``src/geometry/mesh.py:119-127``);
* of the two elements on an interior face, the lower element index is the
  owner ("left", side 0) (``mesh.py:142-151``);
* interior edges are numbered in that sorted-key order; boundary edges follow,
  first in key order, then stably regrouped by tag INFLOW < OUTFLOW < WALL <
  AXIS (``mesh.py:153-176``);
* ``face_neighbour`` points back at the element itself on a boundary
  (``mesh.py:192-196``);
* structured nozzle nodes are numbered ``j * n_ix + i``; cells ``(ci, cj)``
  run with ``cj`` fastest, lower-left triangles first, then upper-right
  (``mesh.py:393-412``); boundary tags come from face midpoints on the
  straight logical grid with tolerance ``0.25 / max(n_ix, n_iy)``
  (``mesh.py:340-358,414-421``).

The matching here sorts a single packed 64-bit key ``lo * n_nodes + hi``
(one stable argsort instead of a two-key lexsort); both give the same order
because node ids are below ``n_nodes``.
"""

from __future__ import annotations

import math
from collections.abc import Callable, Mapping
from dataclasses import dataclass, field
from enum import IntEnum

import numpy as np

from . import reference_element as el


class BoundaryTag(IntEnum):
    INFLOW = 1
    OUTFLOW = 2
    WALL = 3
    AXIS = 4


@dataclass(frozen=True)
class EdgeConnectivity:
    iedge_elem: np.ndarray  # (Fi, 2) owner, neighbour
    iedge_face: np.ndarray  # (Fi, 2)
    bedge_elem: np.ndarray  # (Fb,)
    bedge_face: np.ndarray
    bedge_tag: np.ndarray
    face_edge: np.ndarray  # (E, nface)
    face_side: np.ndarray  # (E, nface) 0 owner, 1 neighbour
    face_neighbour: np.ndarray  # (E, nface)
    tag_slices: dict[int, slice] = field(default_factory=dict)

    @property
    def n_interior(self) -> int:
        return int(self.iedge_elem.shape[0])

    @property
    def n_boundary(self) -> int:
        return int(self.bedge_elem.shape[0])


@dataclass(frozen=True)
class MeshTopology:
    kind: str
    geometry_order: int
    elem_nodes: np.ndarray
    n_nodes: int
    edges: EdgeConnectivity
    logical: dict[str, int] = field(default_factory=dict)

    @property
    def n_elem(self) -> int:
        return int(self.elem_nodes.shape[0])

    @property
    def n_faces(self) -> int:
        return el.n_edges(self.kind)

    def tag_slice(self, tag) -> slice:
        return self.edges.tag_slices.get(int(tag), slice(0, 0))


def build_edges(
    kind: str,
    elem_nodes: np.ndarray,
    geometry_order: int,
    tagger: Callable[[np.ndarray], np.ndarray],
    node_coords: np.ndarray,
) -> EdgeConnectivity:
    """Edge lists and the element-face gather map of a conforming mesh."""
    nface = el.n_edges(kind)
    elem_nodes = np.asarray(elem_nodes, dtype=np.int64)
    nelem = elem_nodes.shape[0]
    corners = elem_nodes[:, np.asarray(el.vertex_indices(kind, geometry_order))]
    ends = np.array([el.edge_vertex_local_indices(kind, f) for f in range(nface)])
    a = corners[:, ends[:, 0]].reshape(-1)  # slot = e * nface + f
    b = corners[:, ends[:, 1]].reshape(-1)
    lo = np.minimum(a, b)
    hi = np.maximum(a, b)
    n_nodes = int(max(int(elem_nodes.max()) + 1, 1))
    key = lo * n_nodes + hi
    order = np.argsort(key, kind="stable")
    sk = key[order]

    dup = sk[1:] == sk[:-1]
    if np.any(dup[1:] & dup[:-1]):
        raise ValueError("non-conforming mesh: a face is shared by more than two elements")
    first = np.flatnonzero(dup)
    paired = np.zeros(sk.size, dtype=bool)
    paired[first] = True
    paired[first + 1] = True

    s0, s1 = order[first], order[first + 1]
    e0, e1 = s0 // nface, s1 // nface
    own = np.where(e0 <= e1, s0, s1)
    nbr = np.where(e0 <= e1, s1, s0)
    iedge_elem = np.column_stack([own // nface, nbr // nface]).astype(np.int64)
    iedge_face = np.column_stack([own % nface, nbr % nface]).astype(np.int64)

    bslot = order[~paired]
    mid = 0.5 * (node_coords[a[bslot]] + node_coords[b[bslot]])
    btag = np.asarray(tagger(np.asarray(mid, dtype=float)), dtype=np.int64)
    if btag.shape != bslot.shape:
        raise ValueError("tagger must return one tag per boundary face")
    regroup = np.argsort(btag, kind="stable")
    bslot, btag = bslot[regroup], btag[regroup]
    bedge_elem = (bslot // nface).astype(np.int64)
    bedge_face = (bslot % nface).astype(np.int64)

    tag_slices: dict[int, slice] = {}
    if btag.size:
        starts = np.flatnonzero(np.r_[True, btag[1:] != btag[:-1]])
        stops = np.r_[starts[1:], btag.size]
        tag_slices = {int(btag[s]): slice(int(s), int(t)) for s, t in zip(starts, stops)}

    ni = iedge_elem.shape[0]
    face_edge = np.full(nelem * nface, -1, dtype=np.int64)
    face_side = np.zeros(nelem * nface, dtype=np.int64)
    face_edge[own] = np.arange(ni)
    face_edge[nbr] = np.arange(ni)
    face_side[nbr] = 1
    face_edge[bslot] = ni + np.arange(bslot.size)
    if np.any(face_edge < 0):
        raise ValueError("internal error: some element face was left unassigned")
    face_nb = np.repeat(np.arange(nelem, dtype=np.int64), nface)
    face_nb[own] = nbr // nface
    face_nb[nbr] = own // nface
    return EdgeConnectivity(
        iedge_elem=iedge_elem,
        iedge_face=iedge_face,
        bedge_elem=bedge_elem,
        bedge_face=bedge_face,
        bedge_tag=btag,
        face_edge=face_edge.reshape(nelem, nface),
        face_side=face_side.reshape(nelem, nface),
        face_neighbour=face_nb.reshape(nelem, nface),
        tag_slices=tag_slices,
    )


def nozzle_boundary_tagger(length: float, tol: float) -> Callable[[np.ndarray], np.ndarray]:
    def tagger(mid: np.ndarray) -> np.ndarray:
        x, y = mid[:, 0], mid[:, 1]
        t = np.full(x.shape, int(BoundaryTag.WALL), dtype=np.int64)
        t[np.abs(y) < tol] = int(BoundaryTag.AXIS)
        t[x < tol] = int(BoundaryTag.INFLOW)
        t[x > length - tol] = int(BoundaryTag.OUTFLOW)
        return t

    return tagger


def _cell_offsets(kind: str, q: int) -> list[np.ndarray]:
    """Per-element node offsets ``(n_nodes_elem, 2)`` inside one structured cell."""
    lat = np.asarray(el._lattice(kind, q), dtype=np.int64)  # (NQ, 2) in units of 1/q
    if kind == "quad":
        return [lat]
    ix, iy = lat[:, 0], lat[:, 1]
    # lower-left triangle (0,0),(q,0),(0,q): identity on the lattice;
    # upper-right triangle (q,0),(q,q),(0,q): (ix, iy) -> (q - iy, ix + iy)
    return [lat, np.column_stack([q - iy, ix + iy])]


def structured_nozzle_topology(
    kind: str = "tri", *, nx: int = 14, nr: int = 5, refine: int = 0, geometry_order: int = 1
) -> tuple[MeshTopology, np.ndarray, np.ndarray]:
    """Topology of the ``nx*2^refine`` by ``nr*2^refine`` channel grid."""
    if kind not in el.ELEMENT_KINDS:
        raise ValueError(f"unknown element kind {kind!r}; use one of {el.ELEMENT_KINDS}")
    if refine < 0 or geometry_order < 1 or nx < 1 or nr < 1:
        raise ValueError("invalid structured-mesh parameters")
    q = geometry_order
    ncx, ncy = nx * 2**refine, nr * 2**refine
    n_ix, n_iy = q * ncx + 1, q * ncy + 1
    ci, cj = np.meshgrid(np.arange(ncx), np.arange(ncy), indexing="ij")
    bi, bj = (ci * q).reshape(-1), (cj * q).reshape(-1)
    elem_nodes = np.concatenate(
        [(bj[:, None] + off[None, :, 1]) * n_ix + (bi[:, None] + off[None, :, 0])
         for off in _cell_offsets(kind, q)],
        axis=0,
    ).astype(np.int64)
    x_frac = np.linspace(0.0, 1.0, n_ix)
    r_frac = np.linspace(0.0, 1.0, n_iy)
    probe = np.column_stack(
        [np.tile(x_frac, n_iy), np.repeat(r_frac, n_ix)]
    )  # node j*n_ix+i at (x_i, r_j)
    tol = 0.25 / max(n_ix, n_iy)
    conn = build_edges(kind, elem_nodes, q, nozzle_boundary_tagger(1.0, tol), probe)
    topo = MeshTopology(
        kind=kind,
        geometry_order=q,
        elem_nodes=elem_nodes,
        n_nodes=n_ix * n_iy,
        edges=conn,
        logical={"nx_cells": ncx, "nr_cells": ncy, "nx_nodes": n_ix, "nr_nodes": n_iy,
                 "refine": refine, "nx_base": nx, "nr_base": nr},
    )
    return topo, x_frac, r_frac


# --------------------------------------------------------------------------- axial spacing
def _throat_clustered(n_int: int, x_t: float, strength: float, width: float) -> np.ndarray:
    """Nodes with spacing ``1 + (strength-1)(1 - exp(-((x-x_t)/w)^2))`` (``mesh.py:263-282``)."""
    if strength < 1.0:
        raise ValueError(f"cluster strength must be >= 1, got {strength}")
    fine = np.linspace(0.0, 1.0, 20001)
    dens = 1.0 / (1.0 + (strength - 1.0) * (1.0 - np.exp(-(((fine - x_t) / width) ** 2))))
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (dens[1:] + dens[:-1]) * np.diff(fine))])
    cdf /= cdf[-1]
    x = np.interp(np.linspace(0.0, 1.0, n_int + 1), cdf, fine)
    x[0], x[-1] = 0.0, 1.0
    return x


def _inlet_clustered(n_int: int) -> np.ndarray:
    a, b = 0.7, 1.2
    v = np.logspace(a, b, n_int + 1)
    return (v - 10.0**a) / (10.0**b - 10.0**a)


def _subdivide(base: np.ndarray, n_sub: int) -> np.ndarray:
    if n_sub == 1:
        return base.copy()
    frac = np.arange(n_sub) / n_sub
    inner = base[:-1, None] * (1.0 - frac[None, :]) + base[1:, None] * frac[None, :]
    return np.concatenate([inner.reshape(-1), base[-1:]])


def nozzle_x_distribution(n_nodes_x, *, nx_base, n_sub, spacing, x_throat,
                          cluster_strength, cluster_width) -> np.ndarray:
    if spacing == "uniform":
        base = np.linspace(0.0, 1.0, nx_base + 1)
    elif spacing == "inlet":
        base = _inlet_clustered(nx_base)
    elif spacing == "throat":
        base = _throat_clustered(nx_base, x_throat, cluster_strength, cluster_width)
    else:
        raise ValueError(f"unknown x_spacing {spacing!r}; use 'throat', 'inlet' or 'uniform'")
    x = _subdivide(base, n_sub)
    if x.size != n_nodes_x:
        raise AssertionError(f"expected {n_nodes_x} axial nodes, built {x.size}")
    return x


# --------------------------------------------------------------------------- contours
CONTOURS = ("bell", "smooth", "conical", "moc", "bezier", "bezier9", "analytic")
BEZIER9_DEGREE = 9
ANALYTIC_THROAT_X = 0.13880048
ANALYTIC_THROAT_Y = 0.13989434
ANALYTIC_INLET_Y = 0.15
ANALYTIC_EXIT_Y = 0.35


def analytic_contour(x):
    t = math.pi * np.log1p((math.exp(2.0) - 1.0) * x)
    half = np.cos(0.5 * t)
    return 0.01 * (4.0 - half) * (np.cos(t) + 5.0 - half)


def bezier9_default_heights(y_th: float, y_ex: float) -> np.ndarray:
    """Interior control heights of the degree-9 Bernstein diverging wall that
    reproduces the ``smooth`` cubic ``y_th + (y_ex - y_th)(3 s^2 - 2 s^3)``
    exactly: the cubic's Bernstein coefficients ``(y_th, y_th, y_ex, y_ex)``
    raised to degree 9."""
    b = np.array([y_th, y_th, y_ex, y_ex])
    n, m = 3, BEZIER9_DEGREE
    c = np.zeros(m + 1)
    for k in range(m + 1):
        for j in range(max(0, k - (m - n)), min(n, k) + 1):
            c[k] += math.comb(n, j) * math.comb(m - n, k - j) / math.comb(m, k) * b[j]
    return c[1:m]


def _bernstein9(s, ctrl):
    """``sum_k B_k^9(s) ctrl[k]`` for ten control values."""
    m = BEZIER9_DEGREE
    out = np.zeros_like(s)
    for k in range(m + 1):
        out = out + math.comb(m, k) * s**k * (1.0 - s) ** (m - k) * ctrl[k]
    return out


def _clip01(v):
    """``numpy.clip(v, 0, 1)`` with the bounds decided on the real part, so a
    complex-step tangent passes through the branch the value takes."""
    r = np.real(v)
    return np.where(r < 0.0, 0.0, np.where(r > 1.0, 1.0, v))


def _tan(v):
    return np.tan(v) if isinstance(v, complex) else math.tan(v)


def _hermite(s, y0, y1, m0, m1):
    s2 = s * s
    s3 = s2 * s
    return ((2.0 * s3 - 3.0 * s2 + 1.0) * y0 + (s3 - 2.0 * s2 + s) * m0
            + (-2.0 * s3 + 3.0 * s2) * y1 + (s3 - s2) * m1)


@dataclass(frozen=True)
class NozzleGeometry:
    """Design variables of the planar contour (``src/geometry/contour.py:111-247``)."""

    area_ratio: float = ANALYTIC_EXIT_Y / ANALYTIC_THROAT_Y
    throat_x: float = ANALYTIC_THROAT_X
    throat_half_height: float = ANALYTIC_THROAT_Y
    inlet_half_height: float = ANALYTIC_INLET_Y
    length: float = 1.0
    contour: str = "bell"
    theta_initial_deg: float | None = None
    theta_exit_deg: float = 0.0
    bezier_w1: float = 0.55
    bezier_w2: float = 0.90
    # bezier9: the eight interior control heights of the diverging wall
    # (None: the exact degree elevation of the smooth contour)
    bezier9_heights: tuple | None = None

    def __post_init__(self) -> None:
        if self.contour not in CONTOURS:
            raise ValueError(f"unknown contour {self.contour!r}; choose from {CONTOURS}")
        if self.contour != "analytic":
            if not 0.0 < self.throat_x < 1.0:
                raise ValueError(f"throat_x must lie in (0, 1), got {self.throat_x}")
            if self.area_ratio <= 1.0:
                raise ValueError(f"area_ratio must exceed 1, got {self.area_ratio}")
            if self.inlet_half_height <= self.throat_half_height:
                raise ValueError("inlet_half_height must exceed throat_half_height")

    def _theta_initial_rad(self) -> float:
        if self.contour == "smooth":
            return 0.0
        if self.theta_initial_deg is not None:
            return math.radians(float(self.theta_initial_deg))
        if self.contour == "moc":
            return math.radians(30.0)
        y_th = self.throat_half_height
        span = max(self.length * (1.0 - self.throat_x), 1e-12)
        return math.atan(1.6 * (y_th * self.area_ratio - y_th) / span)

    def params(self) -> dict[str, float]:
        return {
            "area_ratio": float(self.area_ratio),
            "throat_x": float(self.throat_x),
            "throat_half_height": float(self.throat_half_height),
            "inlet_half_height": float(self.inlet_half_height),
            "theta_initial": float(self._theta_initial_rad()),
            "theta_exit": math.radians(float(self.theta_exit_deg)),
            "bezier_w1": float(self.bezier_w1),
            "bezier_w2": float(self.bezier_w2),
            "length": float(self.length),
        } | self._bezier9_params()

    def _bezier9_params(self) -> dict[str, float]:
        if self.contour != "bezier9":
            return {}
        y_th = float(self.throat_half_height)
        h = (bezier9_default_heights(y_th, y_th * float(self.area_ratio))
             if self.bezier9_heights is None else np.asarray(self.bezier9_heights, dtype=float))
        if h.shape != (BEZIER9_DEGREE - 1,):
            raise ValueError(f"bezier9 needs {BEZIER9_DEGREE - 1} interior control heights")
        return {f"bz{i + 1}": float(v) for i, v in enumerate(h)}

    def wall(self, x):
        return wall_half_height(x, self.params(), self.contour)

    def throat_location(self) -> float:
        if self.contour == "analytic":
            return ANALYTIC_THROAT_X * self.length
        return float(self.throat_x) * self.length

    def throat_height(self) -> float:
        if self.contour == "analytic":
            return ANALYTIC_THROAT_Y
        return float(self.throat_half_height)


def wall_half_height(x, params: Mapping[str, float], contour: str = "bell"):
    """``y_wall(x)`` for a contour family (``contour.py:264-311``).

    Complex ``x`` or parameters are carried through (branches follow the
    real part), which gives exact complex-step tangents of the wall."""
    if contour not in CONTOURS:
        raise ValueError(f"unknown contour {contour!r}; choose from {CONTOURS}")
    x = np.asarray(x)
    if not np.iscomplexobj(x):
        x = x.astype(float)
    length = params.get("length", 1.0)
    xi = x / length
    if contour == "analytic":
        return analytic_contour(xi)
    y_th = params["throat_half_height"]
    y_in = params["inlet_half_height"]
    y_ex = y_th * params["area_ratio"]
    x_th = params["throat_x"]
    y_conv = _hermite(_clip01(xi / x_th), y_in, y_th, 0.0, 0.0)
    span = 1.0 - x_th
    s = _clip01((xi - x_th) / span)
    dx_ds = span * length
    if contour == "conical":
        y_div = y_th + (y_ex - y_th) * s
    elif contour in ("bell", "smooth", "moc"):
        y_div = _hermite(s, y_th, y_ex, _tan(params["theta_initial"]) * dx_ds,
                         _tan(params["theta_exit"]) * dx_ds)
    elif contour == "bezier9":
        ctrl = [y_th] + [params[f"bz{i}"] for i in range(1, BEZIER9_DEGREE)] + [y_ex]
        y_div = _bernstein9(s, ctrl)
    else:  # bezier
        c1 = y_th + params["bezier_w1"] * (y_ex - y_th)
        c2 = y_th + params["bezier_w2"] * (y_ex - y_th)
        u = 1.0 - s
        y_div = u**3 * y_th + 3.0 * u**2 * s * c1 + 3.0 * u * s**2 * c2 + s**3 * y_ex
    return np.where(np.real(xi) <= np.real(x_th), y_conv, y_div)


def nozzle_node_coords(x_frac, r_frac, params, contour) -> np.ndarray:
    """Nodes ``(x_i, r_j * y_wall(x_i))`` numbered ``j * n_ix + i`` (complex
    parameters carried through, see :func:`node_coords_tangent`)."""
    x = np.asarray(x_frac, dtype=float) * params.get("length", 1.0)
    yw = wall_half_height(x, params, contour)
    r = np.asarray(r_frac, dtype=float)
    xx = np.broadcast_to(x[None, :], (r.size, x.size))
    yy = r[:, None] * yw[None, :]
    return np.stack([xx.reshape(-1), yy.reshape(-1)], axis=-1)


def node_coords_tangent(x_frac, r_frac, params, contour, name: str) -> np.ndarray:
    """``d nozzle_node_coords / d params[name]`` at frozen node fractions,
    exact to rounding: the complex-step derivative ``Im f(p + i h) / h``
    (h = 1e-30) of the closed-form node map, which has no subtractive
    cancellation.  Clipped wall parameters differentiate through the branch
    each node takes, as the reference's automatic differentiation does."""
    h = 1e-30
    p = dict(params)
    p[name] = complex(float(p[name]), h)
    return np.imag(nozzle_node_coords(x_frac, r_frac, p, contour)) / h


def build_nozzle_mesh(geom: NozzleGeometry, *, kind="tri", nx=14, nr=5, refine=0,
                      geometry_order=1, x_spacing="throat", cluster_strength=3.0,
                      cluster_width=0.12):
    topo, _, r_frac = structured_nozzle_topology(kind, nx=nx, nr=nr, refine=refine,
                                                 geometry_order=geometry_order)
    x_frac = nozzle_x_distribution(
        topo.logical["nx_nodes"], nx_base=nx, n_sub=geometry_order * 2**refine,
        spacing=x_spacing, x_throat=geom.throat_location() / geom.length,
        cluster_strength=cluster_strength, cluster_width=cluster_width)
    coords = nozzle_node_coords(x_frac, r_frac, geom.params(), geom.contour)
    return topo, coords, {"x_frac": x_frac, "r_frac": r_frac}


def synthetic_channel(ncx: int, ncy: int, *, kind: str = "tri", geometry_order: int = 1):
    """Structured channel with wall ``y = 1 + 1.5 (x - 0.3)^2`` on ``x in [0, 1]``.

    Same topology and tags as the nozzle mesh (``nx=ncx, nr=ncy, refine=0``);
    node ``(i, j)`` sits at ``x = i/n``, ``y = (j/m) * y_wall(x)`` on the
    ``q``-refined node grid.
    """
    topo, x_frac, r_frac = structured_nozzle_topology(kind, nx=ncx, nr=ncy, refine=0,
                                                      geometry_order=geometry_order)
    x = x_frac
    yw = 1.0 + 1.5 * (x - 0.3) ** 2
    X = np.broadcast_to(x[None, :], (r_frac.size, x.size)).reshape(-1)
    Y = (r_frac[:, None] * yw[None, :]).reshape(-1)
    return topo, np.stack([X, Y], axis=-1)
