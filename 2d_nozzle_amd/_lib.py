"""ctypes binding of ``libnozzle_hip.so`` (the C ABI in ``include/nozzle_hip.h``).

The library is built in-tree by ``build.py`` (``hipcc --offload-arch=gfx950``)
into ``2d_nozzle_amd/_lib/``.  There is no fallback: if the library is
missing or fails to load, ``load()`` raises, and so does every entry point
of the device backend.
"""

from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libnozzle_hip.so")

NZ_OK = 0


class TableOffsets(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "phi_vol", "dxi", "deta", "axi", "aeta", "phi_face", "lift", "w_vol", "w_face",
        "mean_w", "minv_ref", "vol_t")]


class Problem(C.Structure):
    """Mirror of ``nz_problem``; field order and types must match the header."""

    _fields_ = [
        ("n_elem", C.c_int64), ("n_edge", C.c_int64),
        ("nface", C.c_int32), ("nbf", C.c_int32), ("nqv", C.c_int32), ("nqf", C.c_int32),
        ("affine", C.c_int32), ("straight", C.c_int32),
        ("geo", C.c_void_p), ("area", C.c_void_p), ("minv", C.c_void_p), ("mean_w", C.c_void_p),
        ("edge_geo", C.c_void_p), ("fe_nbr", C.c_void_p), ("fe_info", C.c_void_p),
        ("n_int", C.c_int64), ("bedge", C.c_void_p), ("bflux", C.c_void_p),
        ("tables", C.c_void_p),
        ("off", TableOffsets),
        ("gamma", C.c_double), ("efix", C.c_double), ("low_mach", C.c_double),
        ("flux_id", C.c_int32), ("limiter", C.c_int32),
        ("rho_t_bf", C.c_double), ("p_t_bf", C.c_double), ("band_bf", C.c_double),
        ("at2", C.c_double), ("at", C.c_double), ("rho_t", C.c_double),
        ("ca", C.c_double), ("sa", C.c_double), ("p_back", C.c_double),
        ("rho_floor", C.c_double), ("p_floor", C.c_double), ("fraction", C.c_double),
        ("lebesgue", C.c_double),
        ("bisect_steps", C.c_int32), ("bflux_chain", C.c_int32),
        ("wnorm", C.c_double * 4),
        ("tiles", C.c_void_p), ("n_tiles", C.c_int64),
        ("vjp_tiles", C.c_void_p), ("n_vjp_tiles", C.c_int64),
        ("vjp_kernel", C.c_int32), ("tiles_cover_mesh", C.c_int32),
        ("patches", C.c_void_p), ("n_patches", C.c_int64), ("bjac", C.c_void_p),
    ]


class MeshDesc(C.Structure):
    """Mirror of ``nz_mesh_desc``."""

    _fields_ = [
        ("n_elem", C.c_int64), ("n_nodes", C.c_int64),
        ("nface", C.c_int32), ("nodes_per_elem", C.c_int32),
        ("elem_nodes", C.c_void_p), ("face_ends", C.c_int32 * 8),
        ("tag_coords", C.c_void_p), ("tag_length", C.c_double), ("tag_tol", C.c_double),
    ]


class EdgeMaps(C.Structure):
    """Mirror of ``nz_edge_maps`` (device pointers)."""

    _fields_ = [(n, C.c_void_p) for n in (
        "iedge_elem", "iedge_face", "bedge_elem", "bedge_face", "bedge_tag", "face_edge",
        "face_side", "face_neighbour", "edge_elem", "edge_face", "fe_nbr", "fe_info", "bedge",
        "counts")]


class GeometryOffsets(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "dgeom_dxi_vol", "dgeom_deta_vol", "geom_vol", "w_vol", "phi_vol", "dgeom_dsigma_face",
        "geom_face", "w_face")]


class GeometryDesc(C.Structure):
    """Mirror of ``nz_geometry_desc``."""

    _fields_ = [
        ("n_elem", C.c_int64), ("n_edge", C.c_int64),
        ("nq", C.c_int32), ("nqv", C.c_int32), ("nqf", C.c_int32), ("nbf", C.c_int32),
        ("nface", C.c_int32), ("affine", C.c_int32), ("straight", C.c_int32), ("pad0", C.c_int32),
        ("elem_nodes", C.c_void_p), ("coords", C.c_void_p),
        ("edge_elem", C.c_void_p), ("edge_face", C.c_void_p),
        ("tables", C.c_void_p), ("off", GeometryOffsets),
    ]


class GeometryOut(C.Structure):
    """Mirror of ``nz_geometry_out`` (device pointers, the reference-layout ones nullable)."""

    _fields_ = [(n, C.c_void_p) for n in (
        "geo", "area", "minv", "mean_w", "edge_geo", "det", "metrics", "wdet", "xy_vol",
        "mean_w_full", "normal", "jac", "length", "xy_face", "bad_jacobian")]


class Quasi1DDesc(C.Structure):
    """Mirror of ``nz_quasi1d_desc``."""

    _fields_ = [
        ("contour", C.c_int32), ("n_points", C.c_int32),
        ("x", C.c_void_p), ("density", C.c_void_p), ("velocity", C.c_void_p),
        ("pressure", C.c_void_p),
        ("length", C.c_double), ("throat_x", C.c_double), ("y_throat", C.c_double),
        ("y_inlet", C.c_double), ("y_exit", C.c_double),
        ("slope_throat", C.c_double), ("slope_exit", C.c_double),
        ("bezier_c1", C.c_double), ("bezier_c2", C.c_double),
        ("ctrl9", C.c_double * 10), ("binom9", C.c_double * 10),
        ("pi", C.c_double), ("e2m1", C.c_double), ("gamma", C.c_double), ("h", C.c_double),
    ]


# name -> (restype, argtypes); every exported symbol of the header
P = C.c_void_p
SIGNATURES = {
    "nz_scratch_doubles": (C.c_int64, [C.POINTER(Problem)]),
    "nz_residual": (C.c_int, [C.POINTER(Problem), P, P, P, P]),
    "nz_rate": (C.c_int, [C.POINTER(Problem), P, P, P, P]),
    "nz_rate_noalloc": (C.c_int, [C.POINTER(Problem), P, P, P, P]),
    "nz_rate_noalloc_max_boundary_points": (C.c_int64, []),
    "nz_inverse_mass": (C.c_int, [C.POINTER(Problem), P, P, P]),
    "nz_project": (C.c_int, [C.POINTER(Problem), P, P, P]),
    "nz_boundary_integrals": (C.c_int, [C.POINTER(Problem), P, P, P]),
    "nz_stage": (C.c_int, [C.POINTER(Problem), P, P, P, P, C.c_double, C.c_double, C.c_double,
                           C.c_int, P, P, P, P]),
    "nz_norm_finish": (C.c_int, [C.POINTER(Problem), P, P, P]),
    "nz_norm_segments": (C.c_int, [C.POINTER(Problem), P, P, P, C.c_int64, P, P]),
    "nz_norm": (C.c_int, [C.POINTER(Problem), P, P, P, P]),
    "nz_limit": (C.c_int, [C.POINTER(Problem), P, P, P, P]),
    "nz_time_step": (C.c_int, [C.POINTER(Problem), P, C.c_double, P, P]),
    "nz_superbee_work_doubles": (C.c_int64, [C.POINTER(Problem)]),
    "nz_adjoint_work_doubles": (C.c_int64, [C.POINTER(Problem)]),
    "nz_adjoint_tile_elements": (C.c_int64, [C.POINTER(Problem)]),
    "nz_face_node": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int]),
    "nz_residual_vjp": (C.c_int, [C.POINTER(Problem), P, P, P, P, P, P, P, P]),
    "nz_adjoint_rk4_stage": (C.c_int, [C.POINTER(Problem), P, P, P, P, P, P, P, C.c_int, P, P]),
    "nz_adjoint_defect": (C.c_int, [C.POINTER(Problem), P, P, P, P, P, P, P]),
    "nz_boundary_jacobian": (C.c_int, [C.POINTER(Problem), P, P, P]),
    "nz_superbee": (C.c_int, [C.POINTER(Problem), P, C.c_double, C.c_double, P, P, P]),
    "nz_last_error": (C.c_char_p, []),
    "nz_build_edges_workspace_bytes": (C.c_int64, [C.c_int64, C.c_int32]),
    "nz_build_edges": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(EdgeMaps), P, C.c_int64, P]),
    "nz_build_geometry": (C.c_int, [C.POINTER(GeometryDesc), C.POINTER(GeometryOut), P]),
    "nz_mesh_last_error": (C.c_char_p, []),
    "nz_quasi1d_values": (C.c_int, [C.POINTER(Quasi1DDesc), P, C.c_int64, P, P]),
    "nz_initial_last_error": (C.c_char_p, []),
    "nz_observables_work_doubles": (C.c_int64, [C.POINTER(Problem)]),
    "nz_performance_integrals": (C.c_int, [C.POINTER(Problem), P, C.c_double, P, P, P]),
    "nz_objective": (C.c_int, [C.POINTER(Problem), P, C.c_int32, C.c_double, P, P, P, P, P]),
    "nz_observables_last_error": (C.c_char_p, []),
}

_LIB: C.CDLL | None = None


def load() -> C.CDLL:
    """Load the HIP library (once).  Raises ``ImportError`` if it is absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"the HIP backend library is not built ({LIB_PATH} missing); "
            "run `python -c 'import __graft_entry__ as g; g.build()'` or "
            "`python 2d_nozzle_amd/build.py`"
        )
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


class HipError(RuntimeError):
    pass


def check(rc: int, what: str) -> None:
    if rc != NZ_OK:
        msg = load().nz_last_error().decode(errors="replace")
        raise HipError(f"{what} failed ({rc}): {msg}")
