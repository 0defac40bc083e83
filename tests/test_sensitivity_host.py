"""Host-side pieces of the adjoint gradient, against finite differences (CPU).

The reverse maps (metric and edge-geometry cotangents to node coordinates,
objective partials) are checked by dot-product tests: for random directions
``<vjp(g), dX> = <g, (f(X + h dX) - f(X - h dX)) / 2h>``.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import build_case, golden, nz_module


def _metrics(ops, X):
    op = nz_module("operators")
    rd = ops.ref
    det, xix, xiy, etx, ety = op.element_metrics(rd, X[ops.topology.elem_nodes])
    return np.stack([xix, xiy, etx, ety, rd.w_vol[None, :] * det], axis=-1)


def _edges(ops, X):
    op = nz_module("operators")
    ed = ops.topology.edges
    own_e = np.concatenate([ed.iedge_elem[:, 0], ed.bedge_elem])
    own_f = np.concatenate([ed.iedge_face[:, 0], ed.bedge_face])
    n, jac, _, _ = op.edge_metrics(ops.ref, X[ops.topology.elem_nodes], own_e, own_f)
    return np.concatenate([n, jac[..., None]], axis=-1)


@pytest.mark.parametrize("name", ["cfg1", "quad_q2_p2", "tri_p2_q2_backflow"])
def test_geometry_reverse_maps(name):
    sens = nz_module("sensitivity")
    ops, _ = build_case(name, dense=False)
    X = np.asarray(ops.node_coords)
    rng = np.random.default_rng(1)
    dX = rng.standard_normal(X.shape)
    h = 1e-6
    gm = rng.standard_normal(_metrics(ops, X).shape)
    fd = ((_metrics(ops, X + h * dX) - _metrics(ops, X - h * dX)) / (2 * h) * gm).sum()
    an = (sens.metric_vjp(ops, gm) * dX).sum()
    assert abs(an - fd) <= 1e-6 * abs(fd)
    ge = rng.standard_normal(_edges(ops, X).shape)
    fd = ((_edges(ops, X + h * dX) - _edges(ops, X - h * dX)) / (2 * h) * ge).sum()
    an = (sens.edge_vjp(ops, ge) * dX).sum()
    assert abs(an - fd) <= 1e-6 * abs(fd)


GOLDEN_OBJECTIVE = {"thrust": "perf_thrust", "thrust_coefficient": "perf_thrust_coefficient",
                    "mass_flow": "perf_mass_flow_in", "exit_mach": "perf_exit_mach_area_averaged",
                    "exit_pressure": "perf_exit_pressure_ratio"}


@pytest.mark.parametrize("objective", list(GOLDEN_OBJECTIVE))
def test_objective_partials(objective):
    """Each of the reference's objectives (src/sensitivity.py:96-130): its
    value on the reference's converged state against the reference's own
    performance figure, and its state and geometry partials against central
    differences."""
    sens = nz_module("sensitivity")
    op, mesh = nz_module("operators"), nz_module("mesh")
    g = golden("cfg1_converged")
    ops, flow = build_case("cfg1", dense=False)
    geom = mesh.NozzleGeometry(contour="smooth", area_ratio=2.5)
    U = g["U"]
    J, dU, dG = sens.objective_and_partials(U, ops, flow, objective, geom=geom)
    assert abs(J / float(g[GOLDEN_OBJECTIVE[objective]]) - 1) <= 1e-13
    rng = np.random.default_rng(3)
    v = rng.standard_normal(U.shape) * np.abs(U).mean(axis=(0, 1))
    h = 1e-6
    Jp = sens.objective_and_partials(U + h * v, ops, flow, objective, geom=geom)[0]
    Jm = sens.objective_and_partials(U - h * v, ops, flow, objective, geom=geom)[0]
    assert abs((Jp - Jm) / (2 * h) - (dU * v).sum()) <= 1e-7 * abs((dU * v).sum())
    # edge geometry: through node coordinates
    X = np.asarray(ops.node_coords)
    dX = rng.standard_normal(X.shape) * 1e-2
    topo = ops.topology

    def J_of(Xn):
        o = op.build_operators(topo, Xn, 1, dense=False)
        return sens.objective_and_partials(U, o, flow, objective, geom=geom)[0]

    fd = (J_of(X + h * dX) - J_of(X - h * dX)) / (2 * h)
    an = (sens.edge_vjp(ops, dG) * dX).sum()
    assert abs(fd - an) <= 1e-6 * max(abs(fd), 1e-12)


def test_autograd_residual_rejects_host_inputs():
    """``sensitivity.residual`` refuses a state that is not a float64 tensor
    on the backend's device, before any device call (host logic only)."""
    import torch

    sens = nz_module("sensitivity")

    class _Stub:  # only the attributes the input checks read
        device = torch.device("cuda", 0)

    U = np.zeros((4, 3, 4))
    with pytest.raises(ValueError, match="float64 tensor"):
        sens.residual(_Stub(), U)
    with pytest.raises(ValueError, match="float64 tensor"):
        sens.residual(_Stub(), torch.zeros((4, 3, 4), dtype=torch.float64))
    with pytest.raises(ValueError, match="float64 tensor"):
        sens.residual(_Stub(), torch.zeros((4, 3, 4), dtype=torch.float32))


# ------------------------------------------------------------ node-map tangent
def _grids(contour="bell", **gk):
    mesh = nz_module("mesh")
    geom = mesh.NozzleGeometry(contour=contour, **gk)
    _, coords, grids = mesh.build_nozzle_mesh(geom, refine=1)
    return mesh, geom, coords, grids["x_frac"], grids["r_frac"]


def test_node_map_real_path_unchanged_by_complex_support():
    """The complex-capable wall gives the same doubles as the mesh builder."""
    for contour in ("bell", "smooth", "conical", "moc", "bezier", "bezier9", "analytic"):
        mesh, geom, coords, xf, rf = _grids(contour)
        X = mesh.nozzle_node_coords(xf, rf, geom.params(), contour)
        assert X.dtype == np.float64 and np.array_equal(X, coords), contour


@pytest.mark.parametrize("contour", ["conical", "smooth", "bell"])
def test_node_map_tangent_closed_form_area_ratio(contour):
    """d y / d area_ratio = r y_th h(s): the diverging wall is linear in its
    exit height y_th * area_ratio, with h(s) = s (conical) or the Hermite
    basis -2 s^3 + 3 s^2 (smooth, bell); the converging wall does not move."""
    mesh, geom, _, xf, rf = _grids(contour, area_ratio=2.5)
    p = geom.params()
    d = mesh.node_coords_tangent(xf, rf, p, contour, "area_ratio")
    x_th, y_th = p["throat_x"], p["throat_half_height"]
    s = np.clip((xf - x_th) / (1.0 - x_th), 0.0, 1.0)
    h = s if contour == "conical" else -2.0 * s**3 + 3.0 * s**2
    h = np.where(xf <= x_th, 0.0, h)
    want = (rf[:, None] * (y_th * h)[None, :]).reshape(-1)
    assert np.all(d[:, 0] == 0.0)
    np.testing.assert_allclose(d[:, 1], want, rtol=1e-14, atol=1e-16)


@pytest.mark.parametrize("contour,name", [("bell", "throat_x"), ("bell", "theta_initial"),
                                          ("bezier", "bezier_w1"), ("bezier9", "bz3"),
                                          ("smooth", "throat_half_height"),
                                          ("analytic", "length"), ("moc", "inlet_half_height")])
def test_node_map_tangent_against_five_point_differences(contour, name):
    """Against the fourth-order central difference at h = 1e-4 (its error
    falls as h^4 towards the tangent: 7e-5, 1e-8, 2e-12 at h = 1e-2, 1e-3,
    1e-4 for the throat position) away from the clip kinks."""
    mesh, geom, _, xf, rf = _grids(contour)
    p = geom.params()
    d = mesh.node_coords_tangent(xf, rf, p, contour, name)
    h = 1e-4 * max(1.0, abs(p[name]))

    def at(v):
        return mesh.nozzle_node_coords(xf, rf, {**p, name: v}, contour)

    fd = (-at(p[name] + 2 * h) + 8 * at(p[name] + h) - 8 * at(p[name] - h)
          + at(p[name] - 2 * h)) / (12 * h)
    # nodes whose branch (converging / diverging wall) flips within 2h are
    # not differentiable there: compare the rest
    x = np.broadcast_to(xf[None, :] * p.get("length", 1.0), (rf.size, xf.size)).reshape(-1)
    near = np.abs(x - p["throat_x"] * p.get("length", 1.0)) < 3 * h if contour != "analytic" \
        else np.zeros_like(x, dtype=bool)
    scale = np.abs(fd[~near]).max()
    assert scale > 0
    assert np.abs(d[~near] - fd[~near]).max() <= 1e-8 * scale
