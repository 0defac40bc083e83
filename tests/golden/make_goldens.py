"""Write the golden input/output vectors that pin the oracle and the HIP path.

Run once, in the development container only (the reference package is not on
the GPU machines):

    python tests/golden/make_goldens.py /root/reference

It imports the reference's own FP64 NumPy path (``src``, package version
1.0.0) and writes small ``.npz`` fixtures next to this script.  Nothing else in
this repository imports the reference; the fixtures are plain numeric arrays.

What the fixtures pin, and why an import is needed for each:

* mesh connectivity and face maps (``src/geometry/mesh.py:77-208``): the
  ordering conventions (lexsort key order, lower-index owner, stable tag sort)
  are not fixed by any invariant;
* reference tables and operators (``src/spatial/operators.py:83-342``);
* the residual ``R(U)`` and wave-speed sum for each flux
  (``src/spatial/assembly.py:153-157``): the entropy fix, HLLC branch logic
  and AUSM+-up splittings are only partly fixed by consistency/conservation;
* limiter outputs and activation flags (``src/spatial/limiter.py:103-243``);
* SSP-RK3 / Euler steps (``src/backends/base.py:161-201``);
* the converged smooth-nozzle solve at tolerance 1e-10 and its thrust, mass
  flow and exit Mach (``src/api.py:247``, ``src/post/postprocess.py:193``).
"""

from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _import_reference(root: str):
    sys.path.insert(0, root)
    import src  # noqa: F401  (the reference package)

    return src


def _mesh_arrays(topo, coords) -> dict:
    ed = topo.edges
    tags = np.array(
        [[t, s.start, s.stop] for t, s in sorted(ed.tag_slices.items())], dtype=np.int64
    ).reshape(-1, 3)
    return {
        "node_coords": np.asarray(coords, dtype=np.float64),
        "elem_nodes": topo.elem_nodes.astype(np.int64),
        "iedge_elem": ed.iedge_elem.astype(np.int64),
        "iedge_face": ed.iedge_face.astype(np.int64),
        "bedge_elem": ed.bedge_elem.astype(np.int64),
        "bedge_face": ed.bedge_face.astype(np.int64),
        "bedge_tag": ed.bedge_tag.astype(np.int64),
        "face_edge": ed.face_edge.astype(np.int64),
        "face_side": ed.face_side.astype(np.int64),
        "face_neighbour": ed.face_neighbour.astype(np.int64),
        "tag_slices": tags,
    }


def _ops_arrays(ops) -> dict:
    rd = ops.ref
    return {
        "ref_points_vol": rd.points_vol,
        "ref_w_vol": rd.w_vol,
        "ref_phi_vol": rd.phi_vol,
        "ref_dphi_dxi_vol": rd.dphi_dxi_vol,
        "ref_dphi_deta_vol": rd.dphi_deta_vol,
        "ref_w_face": rd.w_face,
        "ref_phi_face": rd.phi_face,
        "ref_dgeom_dsigma_face": rd.dgeom_dsigma_face,
        "det_jac": np.asarray(ops.det_jac),
        "elem_area": np.asarray(ops.elem_area),
        "mean_weights": np.asarray(ops.mean_weights),
        "inv_mass": np.asarray(ops.inv_mass),
        "edge_normal": np.asarray(ops.edge_normal),
        "edge_jac": np.asarray(ops.edge_jac),
        "edge_length": np.asarray(ops.edge_length),
        "xy_vol": np.asarray(ops.xy_vol),
    }


def _limiter_flags(U, ops, flow, lim):
    """Per-element activation flags of the Zhang-Shu limiter (uint8 bit field).

    bit 0: the cell average needed repair (``limiter.py:125-129``);
    bit 1: some probe violates the density or pressure bound after the repair
           (``limiter.py:222-224``).
    """
    rho_floor, p_floor = lim.reference_floors(flow)
    g = flow.gamma
    ubar = lim.cell_means(U, ops)
    rho_new = np.maximum(ubar[:, 0], rho_floor)
    e_min = p_floor / (g - 1.0) + 0.5 * (ubar[:, 1] ** 2 + ubar[:, 2] ** 2) / rho_new
    repaired = (ubar[:, 0] < rho_floor) | (ubar[:, 3] < e_min)
    Uf = lim.floor_cell_means(U, ops, flow)
    ubar = lim.cell_means(Uf, ops)
    probes = lim.probe_values(Uf, ops)
    p_bar = (g - 1.0) * (ubar[:, 3] - 0.5 * (ubar[:, 1] ** 2 + ubar[:, 2] ** 2) / ubar[:, 0])
    eps_rho = np.maximum(1e-8 * ubar[:, 0], 0.5 * rho_floor)
    eps_p = np.maximum(1e-8 * p_bar, 0.5 * p_floor)
    rmin = probes[..., 0].min(axis=1)
    pmin = lim._pressure_at(probes, g, np).min(axis=1)
    violates = (rmin < eps_rho) | (pmin < eps_p)
    flags = repaired.astype(np.uint8) | (violates.astype(np.uint8) << 1)
    margin = np.minimum((rmin - eps_rho) / np.abs(ubar[:, 0]), (pmin - eps_p) / np.abs(p_bar))
    return flags, margin


def _stress(U, ops, rng, n_violate=12, n_repair=3):
    """A state with negative density/pressure at some probe points, and a few
    elements whose cell-average density is non-physical."""
    Us = U.copy()
    E = Us.shape[0]
    pick = rng.choice(E, size=n_violate + n_repair, replace=False)
    nb = Us.shape[1]
    for k, e in enumerate(pick):
        if k < n_violate:
            # large zero-mean deviation: mean kept, probes go negative
            dev = rng.standard_normal((nb, 4))
            dev -= (ops.mean_weights[e][:, None] * dev).sum(axis=0)[None, :]
            scale = np.abs(Us[e]).max(axis=0) * (1.5 + rng.random())
            Us[e] = Us[e] + dev * scale[None, :]
        else:
            Us[e, :, 0] = -0.1 * np.abs(Us[e, :, 0])
            Us[e, :, 3] = 0.5 * Us[e, :, 3]
    return Us


def _flow_fluxes(flow):
    return [flow.replace(flux=f) for f in ("roe", "hllc", "ausmup")]


def case_cfg1(src, out: dict) -> tuple:
    from src import NozzleGeometry, build_case
    from src.backends import get_backend
    from src.config import FlowConditions, SolverOptions
    from src.spatial import assembly as asm
    from src.spatial import limiter as lim
    from src.timestepping import initialize as ini

    geom = NozzleGeometry(contour="smooth", area_ratio=2.5)
    flow = FlowConditions()
    case = build_case(geom, flow, order=1, refine=1)
    ops = case.operators
    out.update(_mesh_arrays(case.topology, case.node_coords))
    out.update(_ops_arrays(ops))
    out["x_frac"] = case.grids["x_frac"]
    out["r_frac"] = case.grids["r_frac"]
    U0 = np.ascontiguousarray(ini.initial_state(ops, flow, geom, "quasi1d"))
    out["U0"] = U0
    rng = np.random.default_rng(0)
    base = np.array([2.0, 0.6, 0.02, 5.0])
    Ur = np.ascontiguousarray(base * (1.0 + 0.03 * rng.standard_normal(U0.shape)))
    out["Ur"] = Ur
    for fl in _flow_fluxes(flow):
        for tagname, U in (("U0", U0), ("Ur", Ur)):
            R, wave = asm.residual(U, ops, fl)
            out[f"R_{fl.flux}_{tagname}"] = R
            out[f"wave_{fl.flux}_{tagname}"] = wave
    bk = get_backend("numpy", ops, flow, SolverOptions())
    R, wave = bk.residual(U0)
    out["rate_roe_U0"] = -bk.inverse_mass(R)
    out["dt_U0"] = bk.time_step(wave)[:, 0, 0]
    out["norm_rate_roe_U0"] = np.array(bk.norm(out["rate_roe_U0"]))
    # steps
    for scheme in ("ssprk3", "euler"):
        b = get_backend("numpy", ops, flow, SolverOptions(scheme=scheme))
        U1, res = b.run(U0.copy(), 1, scheme)
        out[f"{scheme}_1step_U"] = U1
        out[f"{scheme}_1step_res"] = np.array(res)
    b = get_backend("numpy", ops, flow, SolverOptions())
    U10, res10 = b.run(U0.copy(), 10, "ssprk3")
    out["ssprk3_10step_U"] = U10
    out["ssprk3_10step_res"] = np.array(res10)
    # limiter on a stressed state
    Us = _stress(U0, ops, np.random.default_rng(7))
    out["Us"] = Us
    flags, margin = _limiter_flags(Us, ops, flow, lim)
    out["Us_flags"] = flags
    out["Us_margin"] = margin
    out["Us_positivity"] = lim.positivity_limiter(Us, ops, flow)
    Ub = Ur.copy()
    Ub[::7, :, :] *= 1.0 + 0.2 * np.random.default_rng(3).standard_normal(Ub[::7].shape)
    out["Ub"] = Ub
    out["Ub_superbee"] = lim.superbee_limiter(Ub, ops, flow, tvb_constant=0.0)
    out["Ub_superbee_tvb50"] = lim.superbee_limiter(Ub, ops, flow, tvb_constant=50.0)
    out["Ub_apply_superbee"] = lim.apply_limiter(Ub, ops, flow, "superbee", tvb_constant=0.0)
    return case, U0


def case_cfg1_converged(src, out: dict) -> None:
    from src import solve_nozzle
    from src.post.postprocess import performance

    t0 = time.perf_counter()
    res = solve_nozzle(
        contour="smooth",
        area_ratio=2.5,
        order=1,
        refine=1,
        tolerance=1e-10,
        backend="numpy",
        verbose=False,
    )
    out["wall_s"] = np.array(time.perf_counter() - t0)
    out["converged"] = np.array(bool(res.converged))
    out["iterations"] = np.array(res.iterations)
    out["residual"] = np.array(res.residual)
    out["residual_initial"] = np.array(res.residual_initial)
    out["residual_scale"] = np.array(res.residual_scale)
    it, rs = res.history.as_arrays()
    out["hist_iterations"] = it
    out["hist_residual"] = rs
    out["U"] = res.U
    perf = performance(res)
    for k in (
        "thrust",
        "thrust_wall",
        "mass_flow_in",
        "mass_flow_out",
        "exit_mach_area_averaged",
        "exit_pressure_ratio",
        "exit_temperature_ratio",
        "thrust_coefficient",
        "discharge_coefficient",
    ):
        out[f"perf_{k}"] = np.array(getattr(perf, k))
    out["mean_repairs"] = np.array(res.mean_repairs)


def synthetic_channel(ncx: int, ncy: int, order: int, seed: int = 0):
    """Synthetic converging-diverging channel state generator (the 2^20-triangle benchmark form)."""
    from src.geometry.mesh import structured_nozzle_topology
    from src.reference import elements as el

    topo, _, _ = structured_nozzle_topology("tri", nx=ncx, nr=ncy, refine=0, geometry_order=1)
    n_ix, n_iy = ncx + 1, ncy + 1
    i = np.arange(n_ix)
    j = np.arange(n_iy)
    x = i / ncx
    yw = 1.0 + 1.5 * (x - 0.3) ** 2
    X = np.broadcast_to(x[None, :], (n_iy, n_ix)).reshape(-1)
    Y = ((j[:, None] / ncy) * yw[None, :]).reshape(-1)
    coords = np.stack([X, Y], axis=-1)
    return topo, coords


def synthetic_state(ops, coords, topo, order: int, gamma: float = 1.4, seed: int = 0):
    from src.reference import elements as el

    rng = np.random.default_rng(seed)
    a = rng.integers(1, 4, size=4)
    b = rng.integers(1, 4, size=4)
    ph = rng.random(4) * 2.0 * np.pi
    nodes = el.reference_nodes("tri", order)
    g, _, _ = el.shape_functions("tri", 1, nodes)
    Xe = coords[topo.elem_nodes]  # (E, 3, 2)
    xy = np.einsum("nq,end->eqd", g, Xe)  # physical Lagrange node positions
    x = xy[..., 0]
    yw = 1.0 + 1.5 * (x - 0.3) ** 2
    eta = xy[..., 1] / yw
    s = [0.5 + 0.5 * np.sin(2 * np.pi * (a[k] * x + b[k] * eta) + ph[k]) for k in range(4)]
    rho = 0.5 + 1.5 * s[0]
    p = 0.5 + 1.5 * s[1]
    M = 3.0 * s[2]
    th = (np.pi / 4) * (2 * s[3] - 1)
    c = np.sqrt(gamma * p / rho)
    vx = M * c * np.cos(th)
    vy = M * c * np.sin(th)
    E = p / (gamma - 1) + 0.5 * rho * (vx * vx + vy * vy)
    return np.ascontiguousarray(np.stack([rho, rho * vx, rho * vy, E], axis=-1))


def case_synth(src, out: dict, ncx=32, ncy=16, order=3) -> None:
    from src.config import FlowConditions
    from src.spatial import assembly as asm
    from src.spatial.operators import build_operators

    topo, coords = synthetic_channel(ncx, ncy, order)
    ops = build_operators(topo, coords, order)
    out.update(_mesh_arrays(topo, coords))
    out["det_jac"] = np.asarray(ops.det_jac)
    out["edge_normal"] = np.asarray(ops.edge_normal)
    out["edge_jac"] = np.asarray(ops.edge_jac)
    out["mean_weights"] = np.asarray(ops.mean_weights)
    U = synthetic_state(ops, coords, topo, order)
    out["U"] = U
    flow = FlowConditions()
    for fl in _flow_fluxes(flow):
        R, wave = asm.residual(U, ops, fl)
        out[f"R_{fl.flux}"] = R
        out[f"wave_{fl.flux}"] = wave
    out["rate_roe"] = -asm.apply_inverse_mass(out["R_roe"], ops)


def case_quad(src, out: dict) -> None:
    from src import NozzleGeometry, build_case
    from src.backends import get_backend
    from src.config import FlowConditions, SolverOptions
    from src.spatial import assembly as asm
    from src.spatial import limiter as lim
    from src.timestepping import initialize as ini

    geom = NozzleGeometry(area_ratio=2.5)  # bell contour, as cases/shocked.json
    flow = FlowConditions(back_pressure_ratio=0.7, allow_shock_in_nozzle=True, flux="hllc")
    case = build_case(geom, flow, element="quad", order=2, geometry_order=2, refine=0)
    ops = case.operators
    out.update(_mesh_arrays(case.topology, case.node_coords))
    out.update(_ops_arrays(ops))
    U0 = np.ascontiguousarray(ini.initial_state(ops, flow, geom, "quasi1d"))
    out["U0"] = U0
    for fl in _flow_fluxes(flow):
        R, wave = asm.residual(U0, ops, fl)
        out[f"R_{fl.flux}_U0"] = R
        out[f"wave_{fl.flux}_U0"] = wave
    bk = get_backend("numpy", ops, flow, SolverOptions())
    R, _ = bk.residual(U0)
    out["rate_hllc_U0"] = -bk.inverse_mass(R)
    U1, res = bk.run(U0.copy(), 1, "ssprk3")
    out["ssprk3_1step_U"] = U1
    out["ssprk3_1step_res"] = np.array(res)
    Us = _stress(U0, ops, np.random.default_rng(11), n_violate=8, n_repair=2)
    out["Us"] = Us
    flags, margin = _limiter_flags(Us, ops, flow, lim)
    out["Us_flags"] = flags
    out["Us_margin"] = margin
    out["Us_positivity"] = lim.positivity_limiter(Us, ops, flow)


def case_p0(src, out: dict) -> None:
    from src import NozzleGeometry, build_case
    from src.backends import get_backend
    from src.config import FlowConditions, SolverOptions
    from src.spatial import assembly as asm
    from src.timestepping import initialize as ini

    geom = NozzleGeometry(contour="smooth", area_ratio=2.5)
    flow = FlowConditions()
    case = build_case(geom, flow, order=0, refine=0)
    ops = case.operators
    out.update(_mesh_arrays(case.topology, case.node_coords))
    out.update(_ops_arrays(ops))
    U0 = np.ascontiguousarray(ini.initial_state(ops, flow, geom, "quasi1d"))
    out["U0"] = U0
    R, wave = asm.residual(U0, ops, flow)
    out["R_roe_U0"] = R
    out["wave_roe_U0"] = wave
    bk = get_backend("numpy", ops, flow, SolverOptions())
    U5, res = bk.run(U0.copy(), 5, bk.scheme)
    out["scheme"] = np.array(bk.scheme)
    out["5step_U"] = U5
    out["5step_res"] = np.array(res)


def case_p2_backflow(src, out: dict) -> None:
    from src import NozzleGeometry, build_case
    from src.config import FlowConditions
    from src.spatial import assembly as asm
    from src.timestepping import initialize as ini

    geom = NozzleGeometry(contour="smooth", area_ratio=2.5)
    flow = FlowConditions(back_pressure_ratio=0.9, backflow=True, hllc_low_mach=0.3)
    case = build_case(geom, flow, order=2, refine=0, geometry_order=2)
    ops = case.operators
    out.update(_mesh_arrays(case.topology, case.node_coords))
    out.update(_ops_arrays(ops))
    U0 = np.ascontiguousarray(ini.initial_state(ops, flow, geom, "uniform"))
    # push part of the exit into reverse flow and part to supersonic outflow
    xy = np.asarray(ops.xy_vol)
    rng = np.random.default_rng(5)
    Up = U0 * (1.0 + 0.05 * rng.standard_normal(U0.shape))
    near_exit = xy[..., 0].mean(axis=1) > 0.9
    Up[near_exit, :, 1] *= np.where(rng.random(near_exit.sum()) < 0.5, -0.3, 2.5)[:, None]
    out["U"] = Up
    for fl in _flow_fluxes(flow):
        R, wave = asm.residual(Up, ops, fl)
        out[f"R_{fl.flux}"] = R
        out[f"wave_{fl.flux}"] = wave


def tables(src, out: dict) -> None:
    from src.spatial.operators import reference_data

    for kind in ("tri", "quad"):
        for q in (1, 2):
            for p in (0, 1, 2, 3):
                rd = reference_data(kind, p, q)
                key = f"{kind}_p{p}_q{q}"
                out[f"{key}_points_vol"] = rd.points_vol
                out[f"{key}_w_vol"] = rd.w_vol
                out[f"{key}_phi_vol"] = rd.phi_vol
                out[f"{key}_dphi_dxi_vol"] = rd.dphi_dxi_vol
                out[f"{key}_dphi_deta_vol"] = rd.dphi_deta_vol
                out[f"{key}_w_face"] = rd.w_face
                out[f"{key}_phi_face"] = rd.phi_face
                out[f"{key}_dgeom_dxi_vol"] = rd.dgeom_dxi_vol
                out[f"{key}_dgeom_deta_vol"] = rd.dgeom_deta_vol
                out[f"{key}_dgeom_dsigma_face"] = rd.dgeom_dsigma_face


def meshes(src, out: dict) -> None:
    """Connectivity of several structured meshes (topology only)."""
    from src import NozzleGeometry
    from src.geometry.mesh import build_nozzle_mesh

    for contour in ("smooth", "bell", "conical", "bezier"):
        geom = NozzleGeometry(contour=contour, area_ratio=2.5)
        for kind in ("tri", "quad"):
            for q in (1, 2):
                topo, coords, grids = build_nozzle_mesh(
                    geom, kind=kind, refine=0, geometry_order=q
                )
                key = f"{contour}_{kind}_q{q}"
                out[f"{key}_node_coords"] = np.asarray(coords)
                if contour == "smooth":
                    for k, v in _mesh_arrays(topo, coords).items():
                        if k != "node_coords":
                            out[f"{key}_{k}"] = v


def case_adjoint(src, out: dict) -> None:
    """The reference's captured sensitivity case (examples/03_sensitivity.py):
    smooth contour (default area ratio), p=1, refine 0, back pressure 0.15,
    forward tolerance 1e-11, Roe.  Its adjoint gradients are only known from
    docs/validation/gradients.md (JAX is not available here)."""
    from src import NozzleGeometry, build_case, solve_nozzle
    from src.config import FlowConditions, SolverOptions
    from src.post.postprocess import performance
    from src.timestepping import initialize as ini

    geom = NozzleGeometry(contour="smooth")
    flow = FlowConditions(back_pressure_ratio=0.15)
    case = build_case(geom, flow, order=1, refine=0)
    ops = case.operators
    out.update(_mesh_arrays(case.topology, case.node_coords))
    out["x_frac"] = case.grids["x_frac"]
    out["r_frac"] = case.grids["r_frac"]
    U0 = np.ascontiguousarray(ini.initial_state(ops, flow, geom, "quasi1d"))
    out["U0"] = U0
    res = solve_nozzle(case=case, flow=flow, options=SolverOptions(tolerance=1e-11),
                       backend="numpy", verbose=False)
    out["U"] = res.U
    out["iterations"] = np.array(res.iterations)
    out["converged"] = np.array(bool(res.converged))
    out["thrust"] = np.array(performance(res).thrust)
    out["area_ratio"] = np.array(geom.area_ratio)


def main() -> None:
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = _import_reference(root)
    only = set(sys.argv[2:])
    jobs = {
        "tables": tables,
        "meshes": meshes,
        "cfg1": lambda s, o: case_cfg1(s, o),
        "synth_p3": lambda s, o: case_synth(s, o, 32, 16, 3),
        "synth_p1": lambda s, o: case_synth(s, o, 12, 6, 1),
        "quad_q2_p2": case_quad,
        "tri_p0": case_p0,
        "tri_p2_q2_backflow": case_p2_backflow,
        "cfg1_converged": case_cfg1_converged,
        "adjoint_case": case_adjoint,
    }
    meta = {"reference_version": getattr(src, "__version__", "?"), "numpy": np.__version__}
    for name, fn in jobs.items():
        if only and name not in only:
            continue
        t0 = time.perf_counter()
        out: dict = {}
        fn(src, out)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"{name}: {len(out)} arrays, {os.path.getsize(path)/1e3:.1f} kB, "
              f"{time.perf_counter()-t0:.1f} s", flush=True)
    with open(os.path.join(HERE, "meta.json"), "w") as fh:
        json.dump(meta, fh, indent=1)


if __name__ == "__main__":
    main()
