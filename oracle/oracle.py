"""ctypes front end of the CPU oracle.  TEST INFRASTRUCTURE ONLY.

Imported only by ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
CPU-baseline leg, as the checker (or the timed CPU reference) of the device
path; the product never calls it.

``CpuOracle`` mirrors the reference ``Backend`` primitives
(``src/backends/base.py:81-201``) and the fused SSP-RK3/Euler steps of the
reference's compiled CPU backend (``src/backends/numba_backend.py:283-337``),
on the reference's dense operator layout.  The arithmetic lives in
``nozzle_oracle.c``.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(HERE, "_build")
_LIBS: dict[str, C.CDLL] = {}

FLUX_IDS = {"roe": 0, "hllc": 1, "ausmup": 2}


class OrFlow(C.Structure):
    _fields_ = [
        ("gamma", C.c_double), ("efix", C.c_double), ("low_mach", C.c_double),
        ("flux_id", C.c_int),
        ("rho_t_bf", C.c_double), ("p_t_bf", C.c_double), ("band_bf", C.c_double),
        ("at2", C.c_double), ("at", C.c_double), ("rho_t", C.c_double),
        ("ca", C.c_double), ("sa", C.c_double), ("p_back", C.c_double),
    ]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load(fast: bool = False) -> C.CDLL:
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name in _LIBS:
        return _LIBS[name]
    path = os.path.join(_BUILD, name)
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    P = C.c_void_p
    i64 = C.c_longlong
    lib.or_edge_pass.argtypes = [i64, i64, C.c_int, C.c_int, C.c_int] + [P] * 10 + [C.POINTER(OrFlow), P, P]
    lib.or_element_pass.argtypes = [i64, C.c_int, C.c_int, C.c_int, C.c_int, P, P, P, P, P, P, P,
                                    C.c_int, P, P, P, P, C.c_double, P, C.c_int, P, P]
    lib.or_inverse_mass.argtypes = [i64, C.c_int, P, P, P]
    lib.or_rms_weighted.argtypes = [i64, C.c_int, P, P]
    lib.or_rms_weighted.restype = C.c_double
    lib.or_local_dt.argtypes = [i64, P, P, C.c_double, P]
    lib.or_stage.argtypes = [i64, C.c_int, P, P, C.c_double, P, P]
    lib.or_combine_ssp.argtypes = [i64, C.c_int, P, P, P, P, C.c_double, C.c_double, P]
    lib.or_positivity_limit.argtypes = [i64, C.c_int, C.c_int, C.c_int, C.c_int, P, P, P, P, P,
                                        C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                        C.c_double, P, P]
    lib.or_superbee.argtypes = [i64, C.c_int, C.c_int, C.c_int, P, P, P, P, P, P, C.c_double, P]
    lib.or_set_threads.argtypes = [C.c_int]
    lib.or_set_threads.restype = C.c_int
    _LIBS[name] = lib
    return lib


def set_threads(n: int, fast: bool = False) -> int:
    """OpenMP threads of the oracle's loops (0: leave as is); returns the
    previous maximum."""
    return int(load(fast).or_set_threads(int(n)))


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _i64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64))


def or_flow(flow) -> OrFlow:
    at2 = flow.gamma * flow.Rgas * flow.total_temperature
    at = float(np.sqrt(at2))
    return OrFlow(
        gamma=flow.gamma, efix=flow.entropy_fix, low_mach=getattr(flow, "hllc_low_mach", 0.0),
        flux_id=FLUX_IDS[flow.flux],
        rho_t_bf=float(flow.stagnation_density) if getattr(flow, "backflow", False) else 0.0,
        p_t_bf=flow.total_pressure, band_bf=0.05,
        at2=at2, at=at, rho_t=flow.gamma * flow.total_pressure / at2,
        ca=float(np.cos(flow.inflow_angle)), sa=float(np.sin(flow.inflow_angle)),
        p_back=flow.back_pressure_ratio * flow.total_pressure,
    )


def positivity_limit(U, ops, flow, *, screen: bool = True):
    """The Zhang-Shu limiter alone, from the limiter's own data (any operator
    layout): ``(U_limited, flags, (n_scaled, n_repaired))``; input untouched
    (``src/backends/_numba_kernels.py:927-1093``)."""
    lib = load(False)
    rd, ed = ops.ref, ops.topology.edges
    out = _f64(U).copy()
    E, N = out.shape[0], out.shape[1]
    flags = np.zeros(E, dtype=np.uint8)
    counts = np.zeros(2, dtype=np.int64)
    if rd.order == 0:
        return out, flags, (0, 0)
    phi_vol, phi_face = _f64(rd.phi_vol), _f64(rd.phi_face)
    face_side, mean_w = _i64(ed.face_side), _f64(ops.mean_weights)
    leb = float(max(np.abs(rd.phi_vol).sum(axis=0).max(),
                    np.abs(rd.phi_face).sum(axis=2).max())) if screen else 0.0
    lib.or_positivity_limit(E, N, rd.n_qvol, rd.n_qface, ops.topology.n_faces, _p(out),
                            _p(phi_vol), _p(phi_face), _p(face_side), _p(mean_w), flow.gamma,
                            1e-8, 1e-8 * flow.stagnation_density, 1e-8 * flow.total_pressure, 12,
                            leb, _p(flags), _p(counts))
    return out, flags, (int(counts[0]), int(counts[1]))


class CpuOracle:
    """Backend primitives on the CPU oracle (dense reference operator layout)."""

    def __init__(self, ops, flow, opts=None, *, fast: bool = False, screen: bool = True):
        self.lib = load(fast)
        self.ops, self.flow, self.opts = ops, flow, opts
        rd, ed = ops.ref, ops.topology.edges
        self.E, self.N = ops.n_elem, rd.n_basis
        self.nqv, self.nqf, self.nface = rd.n_qvol, rd.n_qface, ops.topology.n_faces
        self.ni, self.nb = ops.n_interior, ops.n_boundary
        self.fl = or_flow(flow)
        self.phi_vol, self.phi_face, self.w_face = _f64(rd.phi_vol), _f64(rd.phi_face), _f64(rd.w_face)
        self.iedge_elem, self.iedge_face = _i64(ed.iedge_elem), _i64(ed.iedge_face)
        self.bedge_elem, self.bedge_face, self.bedge_tag = (
            _i64(ed.bedge_elem), _i64(ed.bedge_face), _i64(ed.bedge_tag))
        self.face_edge, self.face_side = _i64(ed.face_edge), _i64(ed.face_side)
        self.face_neighbour = _i64(ed.face_neighbour)
        self.face_sign = _f64(ops.face_sign)
        self.edge_normal, self.edge_jac = _f64(ops.edge_normal), _f64(ops.edge_jac)
        self.edge_length, self.elem_area = _f64(ops.edge_length), _f64(ops.elem_area)
        self.mean_weights = _f64(ops.mean_weights)
        self.grad_x, self.grad_y = _f64(ops.grad_x), _f64(ops.grad_y)
        self.grad_x_m, self.grad_y_m = _f64(ops.grad_x_m), _f64(ops.grad_y_m)
        self.face_basis, self.face_basis_m = _f64(ops.face_basis), _f64(ops.face_basis_m)
        self.inv_mass = _f64(ops.inv_mass)
        a_t = flow.stagnation_sound_speed
        self.weights = np.array([1.0, 1.0 / a_t, 1.0 / a_t, 1.0 / a_t**2])
        self.rho_floor = 1e-8 * flow.stagnation_density
        self.p_floor = 1e-8 * flow.total_pressure
        self.lebesgue = float(max(np.abs(rd.phi_vol).sum(axis=0).max(),
                                  np.abs(rd.phi_face).sum(axis=2).max())) if screen else 0.0
        F = self.ni + self.nb
        self.fw = np.zeros((F, self.nqf, 4))
        self.smax = np.zeros(F)

    # -- primitives ------------------------------------------------------
    def edges(self, U: np.ndarray) -> None:
        self.lib.or_edge_pass(self.ni, self.nb, self.N, self.nqf, self.nface, _p(U),
                              _p(self.phi_face), _p(self.iedge_elem), _p(self.iedge_face),
                              _p(self.bedge_elem), _p(self.bedge_face), _p(self.bedge_tag),
                              _p(self.edge_normal), _p(self.edge_jac), _p(self.w_face),
                              C.byref(self.fl), _p(self.fw), _p(self.smax))

    def residual(self, U, *, apply_mass: bool = False):
        U = _f64(U)
        self.edges(U)
        out = np.empty_like(U)
        wave = np.empty(self.E)
        self.lib.or_element_pass(self.E, self.N, self.nqv, self.nqf, self.nface, _p(U),
                                 _p(self.fw), _p(self.smax), _p(self.phi_vol), _p(self.grad_x),
                                 _p(self.grad_y), _p(self.phi_face), 0, _p(self.face_edge),
                                 _p(self.face_side), _p(self.face_sign), _p(self.edge_length),
                                 self.flow.gamma, _p(self.inv_mass), int(apply_mass), _p(out),
                                 _p(wave))
        return out, wave

    def rate(self, U):
        """``-M^-1 R`` with the folded operators (``_numba_kernels.py:586-664``)."""
        U = _f64(U)
        self.edges(U)
        out = np.empty_like(U)
        wave = np.empty(self.E)
        self.lib.or_element_pass(self.E, self.N, self.nqv, self.nqf, self.nface, _p(U),
                                 _p(self.fw), _p(self.smax), _p(self.phi_vol), _p(self.grad_x_m),
                                 _p(self.grad_y_m), _p(self.face_basis_m), 1, _p(self.face_edge),
                                 _p(self.face_side), _p(self.face_sign), _p(self.edge_length),
                                 self.flow.gamma, _p(self.inv_mass), 0, _p(out), _p(wave))
        return out, wave

    def inverse_mass(self, R):
        R = _f64(R)
        out = np.empty_like(R)
        self.lib.or_inverse_mass(self.E, self.N, _p(self.inv_mass), _p(R), _p(out))
        return out

    def limit(self, U, kind: str = "positivity", tvb_constant: float = 50.0):
        """Returns ``(U_limited, flags, (n_scaled, n_repaired))``; input untouched."""
        out = _f64(U).copy()
        flags = np.zeros(self.E, dtype=np.uint8)
        counts = np.zeros(2, dtype=np.int64)
        if kind == "none" or self.ops.ref.order == 0:
            return out, flags, (0, 0)
        if kind == "superbee":
            scratch = np.empty((self.E, 4))
            self.lib.or_superbee(self.E, self.N, self.nqf, self.nface, _p(out),
                                 _p(self.mean_weights), _p(self.face_basis), _p(self.w_face),
                                 _p(self.face_neighbour), _p(self.elem_area), tvb_constant,
                                 _p(scratch))
        self.lib.or_positivity_limit(self.E, self.N, self.nqv, self.nqf, self.nface, _p(out),
                                     _p(self.phi_vol), _p(self.phi_face), _p(self.face_side),
                                     _p(self.mean_weights), self.flow.gamma, 1e-8,
                                     self.rho_floor, self.p_floor, 12, self.lebesgue,
                                     _p(flags), _p(counts))
        return out, flags, (int(counts[0]), int(counts[1]))

    def time_step(self, wave, factor: float):
        wave = _f64(wave)
        dt = np.empty(self.E)
        self.lib.or_local_dt(self.E, _p(wave), _p(self.elem_area), factor, _p(dt))
        return dt

    def norm(self, A) -> float:
        A = _f64(A)
        with np.errstate(over="ignore", invalid="ignore"):
            return float(self.lib.or_rms_weighted(self.E, self.N, _p(A), _p(self.weights)))

    # -- fused steps (numba_backend.py:305-337) --------------------------
    def step(self, U, scheme: str, factor: float, limiter: str = "positivity",
             tvb_constant: float = 50.0):
        U = _f64(U)
        F0, wave = self.rate(U)
        dt = self.time_step(wave, factor)
        res = self.norm(F0)
        U1 = np.empty_like(U)
        self.lib.or_stage(self.E, self.N, _p(U), _p(F0), 1.0, _p(dt), _p(U1))
        U1 = self.limit(U1, limiter, tvb_constant)[0]
        if scheme == "euler":
            return U1, res
        F1, _ = self.rate(U1)
        U2 = np.empty_like(U)
        self.lib.or_combine_ssp(self.E, self.N, _p(U), _p(U1), _p(F1), _p(dt), 0.75, 0.25, _p(U2))
        U2 = self.limit(U2, limiter, tvb_constant)[0]
        F2, _ = self.rate(U2)
        U3 = np.empty_like(U)
        self.lib.or_combine_ssp(self.E, self.N, _p(U), _p(U2), _p(F2), _p(dt), 1.0 / 3.0,
                                2.0 / 3.0, _p(U3))
        return self.limit(U3, limiter, tvb_constant)[0], res

    def run(self, U, n_steps: int, scheme: str, factor: float, limiter: str = "positivity",
            tvb_constant: float = 50.0):
        res = float("nan")
        for _ in range(n_steps):
            U, res = self.step(U, scheme, factor, limiter, tvb_constant)
        return U, res
