"""Host-side data formats: reference tables, mesh connectivity, operators.

Mesh maps must equal the reference's bit for bit (they are what the kernels
gather through); tables and operators agree to round-off.  Besides the golden
vectors, a brute-force face matcher and the reference's own invariants
(``tests/test_mesh_operators.py:29-161``, ``tests/test_quadrature.py``) pin
the ordering rules and the geometry independently.
"""

from __future__ import annotations

import itertools

import numpy as np
import pytest

from conftest import CASES, build_case, golden, nz_module, rel_err

MAPS = ("elem_nodes", "iedge_elem", "iedge_face", "bedge_elem", "bedge_face", "bedge_tag",
        "face_edge", "face_side", "face_neighbour")


def _maps(topo):
    ed = topo.edges
    out = {k: getattr(ed, k) for k in MAPS if hasattr(ed, k)}
    out["elem_nodes"] = topo.elem_nodes
    return out


@pytest.mark.parametrize("name", ["cfg1", "quad_q2_p2", "tri_p0", "tri_p2_q2_backflow",
                                  "synth_p3", "synth_p1"])
def test_maps_bit_exact(name):
    g = golden(name)
    ops, _ = build_case(name, dense=False)
    for k, v in _maps(ops.topology).items():
        assert v.dtype == np.int64
        assert np.array_equal(v, g[k]), k
    ts = np.array([[t, s.start, s.stop] for t, s in sorted(ops.topology.edges.tag_slices.items())])
    assert np.array_equal(ts, g["tag_slices"])
    assert rel_err(ops.node_coords, g["node_coords"]) <= 1e-15


def test_more_meshes_bit_exact():
    g = golden("meshes")
    mesh = nz_module("mesh")
    for contour in ("smooth", "bell", "conical", "bezier"):
        geom = mesh.NozzleGeometry(contour=contour, area_ratio=2.5)
        for kind in ("tri", "quad"):
            for q in (1, 2):
                topo, coords, _ = mesh.build_nozzle_mesh(geom, kind=kind, geometry_order=q)
                key = f"{contour}_{kind}_q{q}"
                assert rel_err(coords, g[f"{key}_node_coords"]) <= 1e-15, key
                if contour == "smooth":
                    for k, v in _maps(topo).items():
                        assert np.array_equal(v, g[f"{key}_{k}"]), (key, k)


@pytest.mark.parametrize("kind,q", [("tri", 1), ("tri", 2), ("quad", 1), ("quad", 2)])
def test_face_matching_brute_force(kind, q):
    """O(F^2) matching on a tiny mesh fixes the pairing and the owner rule."""
    mesh = nz_module("mesh")
    el = nz_module("reference_element")
    topo, _, _ = mesh.structured_nozzle_topology(kind, nx=3, nr=2, geometry_order=q)
    ed = topo.edges
    corners = topo.elem_nodes[:, list(el.vertex_indices(kind, q))]
    nface = el.n_edges(kind)
    faces = []
    for e in range(topo.n_elem):
        for f in range(nface):
            a, b = el.edge_vertex_local_indices(kind, f)
            faces.append((e, f, frozenset((corners[e, a], corners[e, b]))))
    pairs = {}
    for (e1, f1, k1), (e2, f2, k2) in itertools.combinations(faces, 2):
        if k1 == k2:
            pairs[(e1, f1)] = (e2, f2)
            pairs[(e2, f2)] = (e1, f1)
    got = {}
    for k in range(ed.n_interior):
        (le, re), (lf, rf) = ed.iedge_elem[k], ed.iedge_face[k]
        assert le < re  # lower index owns the face
        got[(le, lf)] = (re, rf)
        got[(re, rf)] = (le, lf)
    assert got == pairs
    nb = 2 * ed.n_interior + ed.n_boundary
    assert nb == topo.n_elem * nface  # face-count identity
    assert ed.n_boundary == len(faces) - len(pairs)
    # boundary edges grouped by tag, in increasing tag order
    assert np.all(np.diff(ed.bedge_tag) >= 0)
    # gather map is consistent with the edge lists
    for e in range(topo.n_elem):
        for f in range(nface):
            k = ed.face_edge[e, f]
            if k < ed.n_interior:
                side = ed.face_side[e, f]
                assert ed.iedge_elem[k, side] == e and ed.iedge_face[k, side] == f
            else:
                assert ed.bedge_elem[k - ed.n_interior] == e and ed.face_side[e, f] == 0


def test_element_counts():
    """140/560/2240 triangles and 70/280/1120 quads (reference tests)."""
    mesh = nz_module("mesh")
    for kind, counts in (("tri", (140, 560, 2240)), ("quad", (70, 280, 1120))):
        for r, n in enumerate(counts):
            topo, _, _ = mesh.structured_nozzle_topology(kind, refine=r)
            assert topo.n_elem == n


def test_tables_match_reference():
    g = golden("tables")
    el = nz_module("reference_element")
    for kind in ("tri", "quad"):
        for q in (1, 2):
            for p in range(4):
                rd = el.reference_data(kind, p, q)
                key = f"{kind}_p{p}_q{q}"
                for f in ("points_vol", "w_vol", "phi_vol", "dphi_dxi_vol", "dphi_deta_vol",
                          "w_face", "phi_face", "dgeom_dxi_vol", "dgeom_deta_vol",
                          "dgeom_dsigma_face"):
                    a, b = getattr(rd, f), g[f"{key}_{f}"]
                    assert a.shape == b.shape, (key, f)
                    assert np.abs(a - b).max() <= 1e-13, (key, f)


@pytest.mark.parametrize("deg", range(0, 12))
def test_quadrature_exactness(deg):
    """Monomials up to the degree are integrated exactly (closed forms)."""
    from math import factorial

    el = nz_module("reference_element")
    tri = el.triangle_rule(deg)
    sq = el.square_rule(deg)
    line = el.gauss_legendre_1d(deg)
    assert np.all(tri.weights > 0) and np.all(sq.weights > 0)
    for a in range(deg + 1):
        assert abs((line.points[:, 0] ** a * line.weights).sum() - 1.0 / (a + 1)) < 1e-14
        for b in range(deg + 1 - a):
            exact = factorial(a) * factorial(b) / factorial(a + b + 2)
            got = (tri.points[:, 0] ** a * tri.points[:, 1] ** b * tri.weights).sum()
            assert abs(got - exact) < 1e-14, (deg, a, b)
            exact_sq = 1.0 / ((a + 1) * (b + 1))
            got_sq = (sq.points[:, 0] ** a * sq.points[:, 1] ** b * sq.weights).sum()
            assert abs(got_sq - exact_sq) < 1e-14


@pytest.mark.parametrize("kind", ["tri", "quad"])
@pytest.mark.parametrize("p", [0, 1, 2, 3])
def test_lagrange_property(kind, p):
    el = nz_module("reference_element")
    nodes = el.reference_nodes(kind, p)
    phi, dx, dy = el.shape_functions(kind, p, nodes)
    assert np.abs(phi - np.eye(len(nodes))).max() < 1e-12
    # partition of unity and its derivative at arbitrary points
    pts = np.array([[0.1, 0.2], [0.3, 0.05], [0.25, 0.5]])
    phi, dx, dy = el.shape_functions(kind, p, pts)
    assert np.abs(phi.sum(axis=0) - 1).max() < 1e-12
    assert np.abs(dx.sum(axis=0)).max() < 1e-11 and np.abs(dy.sum(axis=0)).max() < 1e-11


@pytest.mark.parametrize("name", list(CASES))
def test_operators_match_reference(name):
    g = golden(name)
    ops, _ = build_case(name)
    for k in ("det_jac", "elem_area", "mean_weights", "inv_mass", "edge_normal", "edge_jac",
              "edge_length", "xy_vol"):
        assert rel_err(getattr(ops, k), g[k]) <= 1e-13, k


@pytest.mark.parametrize("kind,q,p", [("tri", 1, 0), ("tri", 1, 2), ("tri", 2, 1), ("quad", 1, 1),
                                      ("quad", 2, 2)])
def test_operator_invariants(kind, q, p):
    """Divergence theorem, M M^-1 = I and freestream preservation (1e-12)."""
    mesh = nz_module("mesh")
    op = nz_module("operators")
    geom = mesh.NozzleGeometry(contour="smooth", area_ratio=2.5)
    topo, coords, _ = mesh.build_nozzle_mesh(geom, kind=kind, geometry_order=q)
    ops = op.build_operators(topo, coords, p)
    # sum_i grad_x[e,i,q] = 0 pointwise (partition of unity) and the
    # divergence theorem: sum_q grad . 1 = boundary integral of n
    E = ops.n_elem
    mass = np.einsum("iq,jq,eq->eij", ops.ref.phi_vol, ops.ref.phi_vol, ops.weighted_det)
    assert np.abs(np.einsum("eij,ejk->eik", mass, ops.inv_mass) - np.eye(ops.ref.n_basis)).max() < 1e-9
    w = ops.ref.w_face
    flux_n = np.zeros((E, 2))
    for f in range(topo.n_faces):
        k = ops.face_edge[:, f]
        sg = ops.face_sign[:, f]
        nds = ops.edge_normal[k] * (ops.edge_jac[k] * w[None, :])[..., None]
        flux_n += sg[:, None] * nds.sum(axis=1)
    assert np.abs(flux_n).max() < 1e-13  # closed elements: sum of n ds is zero
    # freestream: constant state has zero volume term when integrated against 1
    gx = ops.grad_x.sum(axis=1)
    assert np.abs(gx).max() < 1e-12


def test_bezier9_default_is_the_smooth_contour():
    """Degree elevation is exact: the unperturbed degree-9 wall is the smooth cubic."""
    mesh = nz_module("mesh")
    x = np.linspace(0.0, 1.0, 2001)
    g9 = mesh.NozzleGeometry(contour="bezier9")
    g3 = mesh.NozzleGeometry(contour="smooth")
    assert np.abs(g9.wall(x) - g3.wall(x)).max() < 1e-15
    # closed form of the cubic on the diverging section
    p = g9.params()
    s = np.clip((x - p["throat_x"]) / (1 - p["throat_x"]), 0, 1)
    y_th, y_ex = p["throat_half_height"], p["throat_half_height"] * p["area_ratio"]
    div = x > p["throat_x"]
    np.testing.assert_allclose(g9.wall(x)[div], (y_th + (y_ex - y_th) * (3 * s**2 - 2 * s**3))[div],
                               rtol=0, atol=1e-15)


def test_bezier9_perturbation_moves_only_the_diverging_wall():
    mesh, wl = nz_module("mesh"), nz_module("workloads")
    g = wl.bezier9_geometry(0)
    base = mesh.NozzleGeometry(contour="bezier9")
    x = np.linspace(0.0, 1.0, 2001)
    p = g.params()
    conv = x <= p["throat_x"]
    np.testing.assert_array_equal(g.wall(x)[conv], base.wall(x)[conv])
    assert g.wall(np.array([1.0]))[0] == pytest.approx(base.wall(np.array([1.0]))[0], abs=1e-15)
    h = np.array([p[f"bz{i}"] for i in range(1, 9)])
    h0 = np.array([base.params()[f"bz{i}"] for i in range(1, 9)])
    assert np.all(np.abs(h / h0 - 1.0) <= 0.05) and np.any(h != h0)
    with pytest.raises(ValueError, match="eight|8 interior"):
        mesh.NozzleGeometry(contour="bezier9", bezier9_heights=(0.2, 0.3)).params()
