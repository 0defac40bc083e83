"""Reference elements and quadrature rules (host side, built once per order).

Tables here are baked into the device kernels as small constant blocks:
the Lagrange basis at the volume and face quadrature points, its reference
gradients, and the quadrature weights.

Conventions are those of the reference solver, because the nodal coefficient
layout of the state ``U (E, N, 4)`` and the face numbering depend on them:

* triangle vertices ``(0,0), (1,0), (0,1)``; local face ``f`` is opposite
  vertex ``f`` and runs from vertex ``(f+1)%3`` to ``(f+2)%3``
  (``src/reference/elements.py:8-16,170-199``);
* quadrilateral vertices ``(0,0), (1,0), (1,1), (0,1)``; face ``f`` runs from
  vertex ``f`` to ``(f+1)%4``;
* Lagrange nodes on the equispaced lattice, row by row in ``eta``
  (``elements.py:57-75``); ``p = 0`` has one node at the centroid;
* quadrature degrees ``2p+1+2(Q-1)`` (volume) and ``2p+Q`` (edge), at least
  ``2p+1`` for quadrilaterals (``elements.py:224-241``); Gauss-Legendre on
  ``[0, 1]``; triangles use the symmetric 1- and 3-point rules up to degree 2
  and a Duffy-collapsed Gauss-Jacobi(1,0) x Gauss-Legendre rule above
  (``src/reference/quadrature.py:47-144``).

The Gauss-Jacobi nodes are computed here by the Golub-Welsch eigenvalue
method on the Jacobi three-term recurrence (no SciPy).
"""

from __future__ import annotations

from dataclasses import dataclass
from functools import lru_cache

import numpy as np

ELEMENT_KINDS = ("tri", "quad")


def n_edges(kind: str) -> int:
    if kind == "tri":
        return 3
    if kind == "quad":
        return 4
    raise ValueError(f"unknown element kind {kind!r}; use one of {ELEMENT_KINDS}")


def n_basis(kind: str, order: int) -> int:
    if order < 0:
        raise ValueError(f"order must be non-negative, got {order}")
    if kind == "tri":
        return (order + 1) * (order + 2) // 2
    if kind == "quad":
        return (order + 1) ** 2
    raise ValueError(f"unknown element kind {kind!r}; use one of {ELEMENT_KINDS}")


def _lattice(kind: str, order: int) -> list[tuple[int, int]]:
    """Integer lattice indices ``(ix, iy)`` of the Lagrange nodes, in node order."""
    if kind == "tri":
        return [(ix, iy) for iy in range(order + 1) for ix in range(order + 1 - iy)]
    if kind == "quad":
        return [(ix, iy) for iy in range(order + 1) for ix in range(order + 1)]
    raise ValueError(f"unknown element kind {kind!r}; use one of {ELEMENT_KINDS}")


@lru_cache(maxsize=64)
def reference_nodes(kind: str, order: int) -> np.ndarray:
    """Lagrange node coordinates ``(n_basis, 2)`` (``elements.py:57-75``)."""
    n_edges(kind)
    if order == 0:
        c = (1.0 / 3.0, 1.0 / 3.0) if kind == "tri" else (0.5, 0.5)
        return np.asarray([c], dtype=float)
    h = 1.0 / order
    return np.asarray([(ix * h, iy * h) for ix, iy in _lattice(kind, order)], dtype=float)


def vertex_indices(kind: str, order: int) -> tuple[int, ...]:
    """Node indices of the element corners, counter-clockwise (``elements.py:79-92``)."""
    if order == 0:
        raise ValueError("an order-0 node list has no vertices; geometry order must be >= 1")
    if kind == "tri":
        return (0, order, n_basis("tri", order) - 1)
    if kind == "quad":
        return (0, order, (order + 1) * order + order, (order + 1) * order)
    raise ValueError(f"unknown element kind {kind!r}; use one of {ELEMENT_KINDS}")


def _exponents(kind: str, order: int) -> list[tuple[int, int]]:
    # the monomial space has the same lattice shape as the nodes
    return _lattice(kind, order)


@lru_cache(maxsize=64)
def _monomial_to_lagrange(kind: str, order: int) -> np.ndarray:
    """``C`` with ``phi_j = sum_k C[k, j] xi^r_k eta^s_k`` (Vandermonde inverse)."""
    nodes = reference_nodes(kind, order)
    ex = _exponents(kind, order)
    V = np.array([[x**r * y**s for (r, s) in ex] for (x, y) in nodes], dtype=float)
    return np.linalg.solve(V, np.eye(len(ex)))


def shape_functions(kind: str, order: int, points) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Basis values and reference gradients at ``points``, each ``(n_basis, npts)``."""
    pts = np.atleast_2d(np.asarray(points, dtype=float))
    if pts.shape[1] != 2:
        raise ValueError(f"points must have shape (npts, 2), got {pts.shape}")
    npts = pts.shape[0]
    nbf = n_basis(kind, order)
    if order == 0:
        return np.ones((1, npts)), np.zeros((1, npts)), np.zeros((1, npts))
    C = _monomial_to_lagrange(kind, order)
    xi, eta = pts[:, 0], pts[:, 1]
    phi = np.zeros((nbf, npts))
    dxi = np.zeros((nbf, npts))
    deta = np.zeros((nbf, npts))
    for k, (r, s) in enumerate(_exponents(kind, order)):
        ck = C[k, :]
        phi += ck[:, None] * (xi**r * eta**s)[None, :]
        if r > 0:
            dxi += (ck * r)[:, None] * (xi ** (r - 1) * eta**s)[None, :]
        if s > 0:
            deta += (ck * s)[:, None] * (xi**r * eta ** (s - 1))[None, :]
    return phi, dxi, deta


_TRI_EDGES = {0: ((1.0, 0.0), (-1.0, 1.0)), 1: ((0.0, 1.0), (0.0, -1.0)), 2: ((0.0, 0.0), (1.0, 0.0))}
_QUAD_EDGES = {
    0: ((0.0, 0.0), (1.0, 0.0)),
    1: ((1.0, 0.0), (0.0, 1.0)),
    2: ((1.0, 1.0), (-1.0, 0.0)),
    3: ((0.0, 1.0), (0.0, -1.0)),
}


def _edge_line(kind: str, edge: int):
    table = _TRI_EDGES if kind == "tri" else _QUAD_EDGES if kind == "quad" else None
    if table is None:
        raise ValueError(f"unknown element kind {kind!r}; use one of {ELEMENT_KINDS}")
    if edge not in table:
        raise ValueError(f"{kind} element has no local edge {edge}")
    return table[edge]


def edge_reference_coords(kind: str, edge: int, sigma) -> np.ndarray:
    """Reference coordinates of edge parameter ``sigma`` (start vertex at 0)."""
    start, direction = _edge_line(kind, edge)
    s = np.asarray(sigma, dtype=float).ravel()

    def walk(c0: float, d: float) -> np.ndarray:
        # direction components are exactly -1, 0 or +1
        if d == 0.0:
            return np.full_like(s, c0)
        return c0 + s if d > 0.0 else c0 - s

    return np.column_stack([walk(start[0], direction[0]), walk(start[1], direction[1])])


def edge_dreference_dsigma(kind: str, edge: int) -> tuple[float, float]:
    return _edge_line(kind, edge)[1]


def edge_vertex_local_indices(kind: str, edge: int) -> tuple[int, int]:
    if kind == "tri":
        return ((edge + 1) % 3, (edge + 2) % 3)
    if kind == "quad":
        return (edge, (edge + 1) % 4)
    raise ValueError(f"unknown element kind {kind!r}; use one of {ELEMENT_KINDS}")


def quadrature_degree(kind: str, order: int, geometry_order: int) -> tuple[int, int]:
    vol = 2 * order + 1 + 2 * (geometry_order - 1)
    edge = 2 * order + geometry_order
    if kind == "quad":
        vol = max(vol, 2 * order + 1)
        edge = max(edge, 2 * order + 1)
    return vol, edge


# --------------------------------------------------------------------------- quadrature
@dataclass(frozen=True)
class QuadratureRule:
    points: np.ndarray
    weights: np.ndarray

    @property
    def n_points(self) -> int:
        return int(self.weights.shape[0])


def _n_gauss(degree: int) -> int:
    if degree < 0:
        raise ValueError(f"degree must be non-negative, got {degree}")
    return max(1, (degree + 2) // 2)


@lru_cache(maxsize=64)
def gauss_legendre_1d(degree: int) -> QuadratureRule:
    """Gauss-Legendre on ``[0, 1]``, exact to ``degree``."""
    x, w = np.polynomial.legendre.leggauss(_n_gauss(degree))
    return QuadratureRule(points=(0.5 * (x + 1.0)).reshape(-1, 1), weights=0.5 * w)


def gauss_jacobi_10(n: int) -> tuple[np.ndarray, np.ndarray]:
    """Gauss-Jacobi nodes/weights on ``[-1, 1]`` for the weight ``(1 - t)``.

    Golub-Welsch: the nodes are the eigenvalues of the symmetric Jacobi matrix
    of the monic Jacobi(alpha=1, beta=0) recurrence, the weights are
    ``mu0 * v0^2`` with ``mu0 = int (1-t) dt = 2``.  Newton polishing on
    ``P_n^{(1,0)}`` brings the nodes to round-off.
    """
    a, b = 1.0, 0.0
    k = np.arange(n, dtype=float)
    ab = 2.0 * k + a + b
    diag = (b * b - a * a) / np.where(ab * (ab + 2.0) == 0.0, 1.0, ab * (ab + 2.0))
    if n >= 1 and ab[0] * (ab[0] + 2.0) == 0.0:  # pragma: no cover - not hit for a=1
        diag[0] = (b - a) / (a + b + 2.0)
    kk = np.arange(1, n, dtype=float)
    abk = 2.0 * kk + a + b
    off = np.sqrt(
        4.0 * kk * (kk + a) * (kk + b) * (kk + a + b) / (abk * abk * (abk + 1.0) * (abk - 1.0))
    )
    J = np.diag(diag) + np.diag(off, 1) + np.diag(off, -1)
    t, V = np.linalg.eigh(J)
    w = 2.0 * V[0, :] ** 2

    def jac(nn, x):
        # P_n^{(1,0)} and its derivative by the standard recurrence
        p0 = np.ones_like(x)
        if nn == 0:
            return p0, np.zeros_like(x)
        p1 = 0.5 * ((a - b) + (a + b + 2.0) * x)
        for m in range(2, nn + 1):
            c1 = 2.0 * m * (m + a + b) * (2.0 * m + a + b - 2.0)
            c2 = (2.0 * m + a + b - 1.0) * (a * a - b * b)
            c3 = (2.0 * m + a + b - 2.0) * (2.0 * m + a + b - 1.0) * (2.0 * m + a + b)
            c4 = 2.0 * (m + a - 1.0) * (m + b - 1.0) * (2.0 * m + a + b)
            p0, p1 = p1, ((c2 + c3 * x) * p1 - c4 * p0) / c1
        # derivative: d/dx P_n^{(a,b)} = (n+a+b+1)/2 P_{n-1}^{(a+1,b+1)}; use the
        # identity (1-x^2) P' = n[(a-b) - (2n+a+b)x] P_n/(2n+a+b) + 2(n+a)(n+b)P_{n-1}/(2n+a+b)
        nn_ = float(nn)
        d = (nn_ * ((a - b) - (2 * nn_ + a + b) * x) * p1 + 2.0 * (nn_ + a) * (nn_ + b) * p0) / (
            (2 * nn_ + a + b) * (1.0 - x * x)
        )
        return p1, d

    for _ in range(3):
        p, dp = jac(n, t)
        t = t - p / dp
    return t, w


@lru_cache(maxsize=64)
def triangle_rule(degree: int) -> QuadratureRule:
    """Positive-weight rule on the unit triangle (weights sum to 1/2)."""
    if degree <= 1:
        return QuadratureRule(np.array([[1.0 / 3.0, 1.0 / 3.0]]), np.array([0.5]))
    if degree == 2:
        return QuadratureRule(
            np.array([[1.0 / 6.0, 1.0 / 6.0], [2.0 / 3.0, 1.0 / 6.0], [1.0 / 6.0, 2.0 / 3.0]]),
            np.array([1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0]),
        )
    n = _n_gauss(degree)
    t, wgj = gauss_jacobi_10(n)
    x = 0.5 * (t + 1.0)
    wx = wgj / 4.0
    leg = gauss_legendre_1d(degree)
    y, wy = leg.points[:, 0], leg.weights
    X, Y = np.meshgrid(x, y, indexing="ij")
    W = np.outer(wx, wy)
    return QuadratureRule(np.column_stack([X.ravel(), ((1.0 - X) * Y).ravel()]), W.ravel().copy())


@lru_cache(maxsize=64)
def square_rule(degree: int) -> QuadratureRule:
    leg = gauss_legendre_1d(degree)
    x, w = leg.points[:, 0], leg.weights
    X, Y = np.meshgrid(x, x, indexing="ij")
    return QuadratureRule(np.column_stack([X.ravel(), Y.ravel()]), np.outer(w, w).ravel().copy())


def volume_rule(kind: str, degree: int) -> QuadratureRule:
    if kind == "tri":
        return triangle_rule(degree)
    if kind == "quad":
        return square_rule(degree)
    raise ValueError(f"unknown element kind {kind!r}; use 'tri' or 'quad'")


# --------------------------------------------------------------------------- tables
@dataclass(frozen=True)
class ReferenceData:
    """Mesh-independent tables for one ``(kind, p, Q)`` (``operators.py:46-135``)."""

    kind: str
    order: int
    geometry_order: int
    n_basis: int
    n_basis_geom: int
    points_vol: np.ndarray
    w_vol: np.ndarray
    phi_vol: np.ndarray  # (N, nqv)
    dphi_dxi_vol: np.ndarray
    dphi_deta_vol: np.ndarray
    dgeom_dxi_vol: np.ndarray  # (NQ, nqv)
    dgeom_deta_vol: np.ndarray
    geom_vol: np.ndarray
    w_face: np.ndarray
    phi_face: np.ndarray  # (2, nface, N, nqf): 0 walks sigma, 1 walks 1 - sigma
    geom_face: np.ndarray  # (nface, NQ, nqf)
    dgeom_dsigma_face: np.ndarray

    @property
    def n_qvol(self) -> int:
        return int(self.w_vol.size)

    @property
    def n_qface(self) -> int:
        return int(self.w_face.size)

    @property
    def n_faces(self) -> int:
        return n_edges(self.kind)


@lru_cache(maxsize=64)
def reference_data(kind: str, order: int, geometry_order: int) -> ReferenceData:
    dv, df = quadrature_degree(kind, order, geometry_order)
    vol = volume_rule(kind, dv)
    face = gauss_legendre_1d(df)
    sigma = face.points[:, 0]
    nface = n_edges(kind)
    phi_v, dxi_v, deta_v = shape_functions(kind, order, vol.points)
    g_v, gdxi_v, gdeta_v = shape_functions(kind, geometry_order, vol.points)
    nbf, nbq, nqf = n_basis(kind, order), n_basis(kind, geometry_order), sigma.size
    phi_face = np.zeros((2, nface, nbf, nqf))
    geom_face = np.zeros((nface, nbq, nqf))
    dgeom = np.zeros((nface, nbq, nqf))
    for f in range(nface):
        dxs, dys = edge_dreference_dsigma(kind, f)
        phi_face[0, f] = shape_functions(kind, order, edge_reference_coords(kind, f, sigma))[0]
        phi_face[1, f] = shape_functions(kind, order, edge_reference_coords(kind, f, 1.0 - sigma))[0]
        g, gx, gy = shape_functions(kind, geometry_order, edge_reference_coords(kind, f, sigma))
        geom_face[f] = g
        dgeom[f] = gx * dxs + gy * dys
    return ReferenceData(
        kind=kind,
        order=order,
        geometry_order=geometry_order,
        n_basis=nbf,
        n_basis_geom=nbq,
        points_vol=vol.points,
        w_vol=vol.weights,
        phi_vol=phi_v,
        dphi_dxi_vol=dxi_v,
        dphi_deta_vol=deta_v,
        dgeom_dxi_vol=gdxi_v,
        dgeom_deta_vol=gdeta_v,
        geom_vol=g_v,
        w_face=face.weights,
        phi_face=phi_face,
        geom_face=geom_face,
        dgeom_dsigma_face=dgeom,
    )
