"""Synthetic workloads: the converging-diverging channel and its seeded state.

Channel: the reference's structured triangle topology with ``ncx x ncy``
cells (``nx=ncx, nr=ncy, refine=0, Q=1``) and wall ``y = 1 + 1.5 (x-0.3)^2``
on ``x in [0, 1]``.

State at the Lagrange nodes of order ``p`` (nodal coefficients):
``s_k = 1/2 + 1/2 sin(2 pi (a_k x + b_k eta) + phi_k)``, ``eta = y / y_wall(x)``,
with ``a_k, b_k`` in {1, 2, 3} and ``phi_k`` in ``[0, 2 pi)`` drawn from
``numpy.random.default_rng(seed)``; then density and pressure ``0.5 + 1.5 s``,
Mach ``3 s_3`` and flow angle ``(pi/4)(2 s_4 - 1)``.
"""

from __future__ import annotations

import numpy as np

from . import reference_element as el
from .mesh import synthetic_channel
from .operators import build_operators


def channel_state(topo, coords: np.ndarray, order: int, *, gamma: float = 1.4,
                  seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    a = rng.integers(1, 4, size=4)
    b = rng.integers(1, 4, size=4)
    ph = rng.random(4) * 2.0 * np.pi
    nodes = el.reference_nodes("tri", order)
    g, _, _ = el.shape_functions("tri", 1, nodes)
    Xe = coords[topo.elem_nodes]
    E = Xe.shape[0]
    out = np.empty((E, nodes.shape[0], 4))
    step = 1 << 18
    for lo in range(0, E, step):
        hi = min(E, lo + step)
        xy = np.einsum("nq,end->eqd", g, Xe[lo:hi])
        x = xy[..., 0]
        eta = xy[..., 1] / (1.0 + 1.5 * (x - 0.3) ** 2)
        s = [0.5 + 0.5 * np.sin(2 * np.pi * (a[k] * x + b[k] * eta) + ph[k]) for k in range(4)]
        rho = 0.5 + 1.5 * s[0]
        p = 0.5 + 1.5 * s[1]
        M = 3.0 * s[2]
        th = (np.pi / 4) * (2 * s[3] - 1)
        c = np.sqrt(gamma * p / rho)
        vx = M * c * np.cos(th)
        vy = M * c * np.sin(th)
        out[lo:hi] = np.stack([rho, rho * vx, rho * vy, p / (gamma - 1) + 0.5 * rho * (vx * vx + vy * vy)],
                              axis=-1)
    return out


def channel_case(ncx: int, ncy: int, order: int, *, dense: bool = False, seed: int = 0):
    """``(ops, U)`` of the synthetic channel at order ``p``."""
    topo, coords = synthetic_channel(ncx, ncy)
    ops = build_operators(topo, coords, order, dense=dense)
    return ops, channel_state(topo, coords, order, seed=seed)


def residual_bytes(ops) -> dict:
    """Algorithmic HBM bytes of one residual and of each fused stage.

    ``B_res = 64 N E + 8 g_e E + 8 g_f F + 4 nface E + 12 F`` with
    ``g_e = 5, g_f = 3`` for affine triangles and ``5 nqv, 3 nqf`` otherwise;
    ``B_stage(k) = B_res + 32 N E [k >= 2] + 8 E``.
    """
    rd = ops.ref
    E, N, F = ops.n_elem, rd.n_basis, ops.n_edges_total
    affine = ops.topology.kind == "tri" and ops.topology.geometry_order == 1
    ge = 5 if affine else 5 * rd.n_qvol
    gf = 3 if ops.topology.geometry_order == 1 else 3 * rd.n_qface
    nface = ops.topology.n_faces
    b_res = 64 * N * E + 8 * ge * E + 8 * gf * F + 4 * nface * E + 12 * F
    st = [b_res + 8 * E, b_res + 32 * N * E + 8 * E, b_res + 32 * N * E + 8 * E]
    return {"residual": b_res, "stage": st, "step": sum(st)}


def bezier9_geometry(seed: int = 0, spread: float = 0.05):
    """The degree-9 Bezier wall with its eight interior control heights
    perturbed as ``c_i (1 + spread u_i)``, ``u_i ~ U[-1, 1]`` from
    ``default_rng(seed)``; the unperturbed heights are the exact degree
    elevation of the smooth contour."""
    from .mesh import NozzleGeometry, bezier9_default_heights

    base = NozzleGeometry(contour="bezier9")
    c = bezier9_default_heights(base.throat_half_height,
                                base.throat_half_height * base.area_ratio)
    u = np.random.default_rng(seed).uniform(-1.0, 1.0, c.size)
    return NozzleGeometry(contour="bezier9", bezier9_heights=tuple(c * (1.0 + spread * u)))
