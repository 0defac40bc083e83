"""Mesh connectivity and operators built on the device (``nz_build_edges``,
``nz_build_geometry``), against the reference's fixtures and the host build.

Bars: every map bit-identical to the reference's golden meshes and to the
host ``build_edges`` (nozzle grids of both kinds, geometry orders 1 and 2,
refinements up to 7 and the synthetic channel); metrics, areas, cell-average
weights, inverse mass matrices, edge normals, Jacobians, lengths and points
within 1e-13 relative of the golden and host operators (at refine 7, within
the rounding that O(1) coordinates leave in O(h) metric derivatives); a backend built on
device operators gives the rate of one built on host operators to 1e-13 and
the oracle's to 1e-12; a second design on a reused topology gets its own
geometry.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import CASES, build_case, golden, nz_module, rel_err

pytestmark = pytest.mark.gpu

MAPS = ("iedge_elem", "iedge_face", "bedge_elem", "bedge_face", "bedge_tag", "face_edge",
        "face_side", "face_neighbour")


def _device_nozzle(contour, kind, q, refine=0, order=1, area_ratio=2.5):
    mesh, dm = nz_module("mesh"), nz_module("device_mesh")
    geom = mesh.NozzleGeometry(contour=contour, area_ratio=area_ratio)
    ops, topo, _ = dm.build_nozzle_operators(geom, kind=kind, refine=refine, geometry_order=q,
                                             order=order)
    return geom, ops, topo


def _assert_maps_equal(ed, ref, key=""):
    for k in MAPS:
        a, b = np.asarray(getattr(ed, k)), np.asarray(ref[k] if isinstance(ref, dict)
                                                      else getattr(ref, k))
        assert a.dtype == np.int64 and a.shape == b.shape and np.array_equal(a, b), (key, k)
    ts_ref = ref["tag_slices"] if isinstance(ref, dict) else ref.tag_slices
    if isinstance(ref, dict):
        ts = np.array([[t, s.start, s.stop] for t, s in sorted(ed.tag_slices.items())])
        assert np.array_equal(ts, ts_ref), key
    else:
        assert ed.tag_slices == ts_ref, key


def test_device_edges_match_golden_meshes():
    g = golden("meshes")
    for kind in ("tri", "quad"):
        for q in (1, 2):
            key = f"smooth_{kind}_q{q}"
            _, ops, topo = _device_nozzle("smooth", kind, q)
            assert np.array_equal(topo.elem_nodes, g[f"{key}_elem_nodes"]), key
            assert rel_err(ops.node_coords, g[f"{key}_node_coords"]) <= 1e-15
            _assert_maps_equal(topo.edges, {k: g[f"{key}_{k}"] for k in (*MAPS, "tag_slices")}, key)


@pytest.mark.parametrize("name", ["cfg1", "quad_q2_p2", "tri_p2_q2_backflow", "synth_p3"])
def test_device_edges_match_golden_cases(name):
    g = golden(name)
    dm = nz_module("device_mesh")
    if name.startswith("synth"):
        ops = dm.build_channel_operators(32, 16, 3)
        topo = ops.topology
    else:
        mk, order, _ = CASES[name]
        _, ops, topo = _device_nozzle(mk["contour"], mk["kind"], mk["q"], mk["refine"], order)
    assert np.array_equal(topo.elem_nodes, g["elem_nodes"])
    _assert_maps_equal(topo.edges, {k: g[k] for k in (*MAPS, "tag_slices")}, name)


@pytest.mark.parametrize("kind,q,refine", [("tri", 1, 3), ("tri", 2, 2), ("quad", 1, 5),
                                           ("quad", 2, 4)])
def test_device_edges_match_host_build(kind, q, refine):
    mesh = nz_module("mesh")
    geom = mesh.NozzleGeometry(contour="bell", area_ratio=2.5)
    topo_h, coords_h, _ = mesh.build_nozzle_mesh(geom, kind=kind, refine=refine, geometry_order=q)
    _, ops, topo = _device_nozzle("bell", kind, q, refine)
    assert np.array_equal(topo.elem_nodes, topo_h.elem_nodes)
    assert np.array_equal(ops.node_coords, coords_h)
    _assert_maps_equal(topo.edges, topo_h.edges, f"{kind} q{q} refine {refine}")


@pytest.mark.parametrize("name", list(CASES))
def test_device_operators_match_golden(name):
    g = golden(name)
    mk, order, _ = CASES[name]
    _, ops, _ = _device_nozzle(mk["contour"], mk["kind"], mk["q"], mk["refine"], order)
    for k in ("det_jac", "elem_area", "mean_weights", "edge_normal", "edge_jac", "edge_length",
              "xy_vol"):
        err = rel_err(getattr(ops, k), g[k])
        print(name, k, err)
        assert err <= 1e-13, k
    if ops.inv_mass_packed is not None:
        N = ops.ref.n_basis
        iu = np.triu_indices(N)
        assert rel_err(ops.inv_mass_packed, np.asarray(g["inv_mass"])[:, iu[0], iu[1]]) <= 1e-12


@pytest.mark.parametrize("kind,q,order", [("tri", 1, 3), ("tri", 2, 2), ("quad", 1, 3),
                                          ("quad", 2, 3)])
def test_device_operators_match_host_build(kind, q, order):
    mesh, opm = nz_module("mesh"), nz_module("operators")
    geom = mesh.NozzleGeometry(contour="bell", area_ratio=2.5)
    topo_h, coords, _ = mesh.build_nozzle_mesh(geom, kind=kind, refine=2, geometry_order=q)
    host = opm.build_operators(topo_h, coords, order, dense=True)
    _, dev, _ = _device_nozzle("bell", kind, q, 2, order)
    for k in ("det_jac", "weighted_det", "elem_area", "mean_weights", "xy_vol", "edge_normal",
              "edge_jac", "edge_length", "xy_face"):
        err = rel_err(getattr(dev, k), getattr(host, k))
        assert err <= 1e-13, (k, err)
    if dev.inv_mass_packed is not None:
        N = dev.ref.n_basis
        iu = np.triu_indices(N)
        err = rel_err(dev.inv_mass_packed, host.inv_mass[:, iu[0], iu[1]])
        print(kind, q, order, "inverse mass", err)
        assert err <= 1e-12


@pytest.mark.parametrize("name,flux", [("cfg1", "roe"), ("quad_q2_p2", "hllc"),
                                       ("tri_p2_q2_backflow", "ausmup")])
def test_backend_on_device_operators(name, flux):
    """The rate through a backend built on device operators: equal to the one
    built on host operators to 1e-13 (the geometry is computed with another
    summation order), and to the oracle and the golden to 1e-12."""
    from oracle.oracle import CpuOracle

    cf, be = nz_module("config"), nz_module("backend")
    g = golden(name)
    U = g["U0"] if "U0" in g else g["U"]
    mk, order, fk = CASES[name]
    _, dops, _ = _device_nozzle(mk["contour"], mk["kind"], mk["q"], mk["refine"], order)
    hops, flow = build_case(name, dense=True)
    flow = flow.replace(flux=flux)
    bd = be.HipBackend(dops, flow, cf.SolverOptions())
    bh = be.HipBackend(hops, flow, cf.SolverOptions())
    Fd, wd = bd.rate(U)
    Fh, wh = bh.rate(U)
    Fo, wo = CpuOracle(hops, flow).rate(U)
    print(name, flux, "device vs host operators", rel_err(Fd, Fh), "vs oracle", rel_err(Fd, Fo))
    assert rel_err(Fd, Fh) <= 1e-13 and rel_err(wd, wh) <= 1e-13
    assert rel_err(Fd, Fo) <= 1e-12
    if name == "cfg1":  # a smooth march (limiter idle): the steps agree too
        Ud, rd_ = bd.run(U, 3, "ssprk3")
        Uh, rh = bh.run(U, 3, "ssprk3")
        assert rel_err(Ud, Uh) <= 1e-12 and abs(rd_ / rh - 1) <= 1e-11


def test_topology_reuse_across_designs():
    """A second design on the same grid reuses the device topology: its
    geometry matches a fresh host build of that design."""
    mesh, opm, dm = nz_module("mesh"), nz_module("operators"), nz_module("device_mesh")
    g1 = mesh.NozzleGeometry(contour="smooth", area_ratio=2.0)
    g2 = mesh.NozzleGeometry(contour="smooth", area_ratio=3.5)
    ops1, topo, _ = dm.build_nozzle_operators(g1, kind="quad", refine=2, order=2)
    ops2, topo2, _ = dm.build_nozzle_operators(g2, kind="quad", refine=2, order=2, topology=topo)
    assert topo2 is topo
    t_h, c_h, _ = mesh.build_nozzle_mesh(g2, kind="quad", refine=2)
    host = opm.build_operators(t_h, c_h, 2, dense=False)
    for k in ("det_jac", "elem_area", "mean_weights", "edge_normal", "edge_jac", "xy_vol"):
        assert rel_err(getattr(ops2, k), getattr(host, k)) <= 1e-13, k
    assert rel_err(ops1.elem_area, ops2.elem_area) > 1e-3  # a different design


def test_device_build_of_the_refine7_quad_mesh():
    """Quadrilaterals, Q2, refine 7 (1,146,880 elements): maps bit-identical
    to the host build, geometry within 1e-13, and the setup times."""
    import time

    import torch

    mesh, opm = nz_module("mesh"), nz_module("operators")
    geom = mesh.NozzleGeometry(contour="bell", area_ratio=2.5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, dev, topo = _device_nozzle("bell", "quad", 2, 7, 2)
    torch.cuda.synchronize()
    t_dev = time.perf_counter() - t0
    t0 = time.perf_counter()
    topo_h, coords, _ = mesh.build_nozzle_mesh(geom, kind="quad", refine=7, geometry_order=2)
    host = opm.build_operators(topo_h, coords, 2, dense=False)
    t_host = time.perf_counter() - t0
    print(f"1.15M-quad Q2 mesh + operators: device {t_dev:.2f} s, host {t_host:.2f} s", flush=True)
    assert dev.n_elem == 1146880
    _assert_maps_equal(topo.edges, topo_h.edges, "refine 7")
    # The metric derivatives sum O(1) node coordinates into O(h) results
    # (h ~ 5e-4 here), so any summation order carries about eps |X| / h of
    # relative rounding; the bar is 20 times that.
    h = float(np.sqrt(host.elem_area.min()))
    bar = 20 * np.finfo(float).eps * float(np.abs(coords).max()) / h
    for k in ("det_jac", "elem_area", "edge_normal", "edge_jac", "mean_weights"):
        err = rel_err(getattr(dev, k), getattr(host, k))
        print(f"  {k}: {err:.2e} (bar {bar:.1e})", flush=True)
        assert err <= bar, k
