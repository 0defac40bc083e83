"""Initial states for the pseudo-time march (host side, once per solve).

The march starts from the quasi-one-dimensional nozzle solution for the
actual contour and back pressure, projected onto the DG basis
(``src/timestepping/initialize.py:41-111``), or from a uniform isentropic
state at one Mach number (``initialize.py:49-66``).

Quasi-1D theory (``src/physics/quasi1d.py``): the planar area is the channel
height ``A(x) = 2 y_wall(x)``; the isentropic area relation is inverted on the
requested branch by bracketed bisection; the regime follows from the three
critical back-pressure ratios; a normal shock in the diverging section is
located by bisection on its position (the exit pressure falls monotonically as
the shock moves downstream).  The sampled solution is interpolated linearly to
the volume quadrature points, a transverse velocity ``vy = vx (y / y_wall)
y_wall'`` is added, and the pointwise state is L2-projected,
``U_e = M_e^-1 int phi_i u0``.

The regime logic and the quasi-1D bisections are scalar NumPy work on the
host, once per design.  The per-point work has device forms:
``quasi1d_values_device`` evaluates the point values with ``nz_quasi1d_values``
and ``quasi1d_initial(..., backend=bk)`` projects with ``nz_project``; the
host functions here are their pinned counterparts.
"""

from __future__ import annotations

import enum
from dataclasses import dataclass

import numpy as np

from .config import FlowConditions


class Regime(str, enum.Enum):
    """Operating regime set by the back-pressure ratio (``quasi1d.py:42-58``)."""

    SUBSONIC = "subsonic"
    SHOCK_IN_NOZZLE = "shock-in-nozzle"
    OVEREXPANDED = "overexpanded"
    DESIGN = "design"
    UNDEREXPANDED = "underexpanded"


# ------------------------------------------------------------ isentropic flow
def area_over_throat(mach, gamma: float):
    """``A / A*`` as a function of the Mach number (``quasi1d.py:64-70``)."""
    m = np.maximum(np.asarray(mach, dtype=float), 1e-12)
    k = (gamma + 1.0) / (2.0 * (gamma - 1.0))
    return (1.0 / m) * (2.0 / (gamma + 1.0) * (1.0 + 0.5 * (gamma - 1.0) * m * m)) ** k


def pressure_ratio(mach, gamma: float):
    """``p / p_t`` (``quasi1d.py:73-76``)."""
    m = np.asarray(mach, dtype=float)
    return (1.0 + 0.5 * (gamma - 1.0) * m * m) ** (-gamma / (gamma - 1.0))


def temperature_ratio(mach, gamma: float):
    """``T / T_t`` (``quasi1d.py:84-86``)."""
    m = np.asarray(mach, dtype=float)
    return 1.0 / (1.0 + 0.5 * (gamma - 1.0) * m * m)


def mach_from_pressure_ratio(p_over_pt, gamma: float):
    """Invert ``p / p_t`` (``quasi1d.py:89-92``)."""
    r = np.clip(np.asarray(p_over_pt, dtype=float), 1e-300, 1.0)
    return np.sqrt(2.0 / (gamma - 1.0) * (r ** (-(gamma - 1.0) / gamma) - 1.0))


def mach_from_area(area_ratio, gamma: float, supersonic: bool = False, tol: float = 1e-13):
    """Mach number on one branch of the area relation (``quasi1d.py:95-136``).

    Bisection on a bracket: ``[1e-10, 1]`` for the subsonic branch; for the
    supersonic branch ``[1, 1.2]`` with the upper end grown by 1.5x until it
    spans the target.  Ratios a hair below 1 (a station at the throat) are
    clipped to 1.
    """
    ar = np.atleast_1d(np.asarray(area_ratio, dtype=float))
    if np.any(ar < 1.0 - 1e-9):
        raise ValueError(f"A/A* must be >= 1; got a minimum of {float(ar.min())}")
    ar = np.maximum(ar, 1.0)
    if supersonic:
        lo, hi = np.ones_like(ar), np.full_like(ar, 1.2)
        for _ in range(200):
            short = area_over_throat(hi, gamma) < ar
            if not short.any():
                break
            hi = np.where(short, hi * 1.5, hi)
    else:
        lo, hi = np.full_like(ar, 1e-10), np.ones_like(ar)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        below = area_over_throat(mid, gamma) < ar
        if supersonic:  # A/A* grows with M above 1
            lo, hi = np.where(below, mid, lo), np.where(below, hi, mid)
        else:  # and falls with M below 1
            lo, hi = np.where(below, lo, mid), np.where(below, mid, hi)
        if np.all(hi - lo < tol):
            break
    m = 0.5 * (lo + hi)
    return m if np.ndim(area_ratio) else float(m[0])


# -------------------------------------------------------------- normal shocks
def normal_shock_mach(m1, gamma: float):
    """Mach number behind a normal shock (``quasi1d.py:142-145``)."""
    m1 = np.asarray(m1, dtype=float)
    return np.sqrt((1.0 + 0.5 * (gamma - 1.0) * m1 * m1) / (gamma * m1 * m1 - 0.5 * (gamma - 1.0)))


def normal_shock_static_pressure_ratio(m1, gamma: float):
    """``p2 / p1`` across a normal shock (``quasi1d.py:148-151``)."""
    m1 = np.asarray(m1, dtype=float)
    return (2.0 * gamma * m1 * m1 - (gamma - 1.0)) / (gamma + 1.0)


def normal_shock_total_pressure_ratio(m1, gamma: float):
    """``p_t2 / p_t1`` across a normal shock (``quasi1d.py:154-160``)."""
    m1 = np.asarray(m1, dtype=float)
    g = gamma
    a = ((g + 1.0) * 0.5 * m1 * m1 / (1.0 + 0.5 * (g - 1.0) * m1 * m1)) ** (g / (g - 1.0))
    b = (2.0 * g / (g + 1.0) * m1 * m1 - (g - 1.0) / (g + 1.0)) ** (1.0 / (g - 1.0))
    return a / b


@dataclass(frozen=True)
class CriticalRatios:
    """The three critical back-pressure ratios (``quasi1d.py:166-188``)."""

    first: float   # throat just sonic, subsonic exit
    second: float  # normal shock at the exit plane
    third: float   # shock-free supersonic exit (design)
    exit_mach_subsonic: float
    exit_mach_design: float


def critical_ratios(area_ratio: float, gamma: float = 1.4) -> CriticalRatios:
    """``quasi1d.py:191-206``."""
    if area_ratio <= 1.0:
        raise ValueError(f"area_ratio must exceed 1, got {area_ratio}")
    m_sub = float(mach_from_area(area_ratio, gamma, supersonic=False))
    m_sup = float(mach_from_area(area_ratio, gamma, supersonic=True))
    third = float(pressure_ratio(m_sup, gamma))
    return CriticalRatios(
        first=float(pressure_ratio(m_sub, gamma)),
        second=third * float(normal_shock_static_pressure_ratio(m_sup, gamma)),
        third=third, exit_mach_subsonic=m_sub, exit_mach_design=m_sup)


def is_shock_free(area_ratio: float, back_pressure_ratio: float, gamma: float = 1.4) -> bool:
    """No normal shock inside the diverging section (``quasi1d.py:232-243``)."""
    c = critical_ratios(area_ratio, gamma)
    return back_pressure_ratio < c.second or back_pressure_ratio >= c.first


def operating_regime(area_ratio: float, back_pressure_ratio: float, gamma: float = 1.4):
    """``(regime, critical ratios)`` (``quasi1d.py:246-261``)."""
    c = critical_ratios(area_ratio, gamma)
    pb = back_pressure_ratio
    if pb >= c.first:
        reg = Regime.SUBSONIC
    elif pb > c.second:
        reg = Regime.SHOCK_IN_NOZZLE
    elif pb > c.third * (1.0 + 1e-9):
        reg = Regime.OVEREXPANDED
    elif pb < c.third * (1.0 - 1e-9):
        reg = Regime.UNDEREXPANDED
    else:
        reg = Regime.DESIGN
    return reg, c


@dataclass(frozen=True)
class Quasi1DSolution:
    """Quasi-1D state sampled along the axis (``quasi1d.py:267-314``)."""

    x: np.ndarray
    area: np.ndarray
    mach: np.ndarray
    pressure: np.ndarray
    density: np.ndarray
    temperature: np.ndarray
    velocity: np.ndarray
    regime: Regime
    critical: CriticalRatios
    shock_x: float | None
    shock_mach: float | None
    mass_flow: float
    exit_mach: float
    exit_pressure: float


def solve_quasi1d(geom, flow: FlowConditions, n_points: int = 601) -> Quasi1DSolution:
    """The quasi-1D solution on ``n_points`` equispaced stations (``quasi1d.py:317-401``).

    Throat and exit areas come from the analytic contour, not the samples.
    """
    g = flow.gamma
    x = np.linspace(0.0, geom.length, n_points)
    area = 2.0 * np.asarray(geom.wall(x), dtype=float)
    x_t = geom.throat_location()
    a_t = 2.0 * geom.throat_height()
    a_e = 2.0 * float(geom.wall(np.asarray([geom.length]))[0])
    regime, crit = operating_regime(a_e / a_t, flow.back_pressure_ratio, g)
    pt = flow.total_pressure
    shock_x = shock_mach = None
    pt_local = np.full_like(x, pt)
    if regime is Regime.SUBSONIC:
        m_e = float(mach_from_pressure_ratio(flow.back_pressure_ratio, g))
        a_star = a_e / float(area_over_throat(m_e, g))
        mach = mach_from_area(np.maximum(area / a_star, 1.0), g, supersonic=False)
    else:
        a_star = a_t
        ratio = np.maximum(area / a_star, 1.0)
        mach = np.where(x > x_t, mach_from_area(ratio, g, supersonic=True),
                        mach_from_area(ratio, g, supersonic=False))
        if regime is Regime.SHOCK_IN_NOZZLE:

            def upstream(xs):
                a1 = 2.0 * float(geom.wall(np.asarray([xs]))[0])
                m1 = float(mach_from_area(a1 / a_star, g, supersonic=True))
                return m1, float(normal_shock_total_pressure_ratio(m1, g))

            def exit_pressure(xs):
                _, ptr = upstream(xs)
                m_e = float(mach_from_area(a_e / (a_star / ptr), g, supersonic=False))
                return float(pressure_ratio(m_e, g)) * ptr

            lo, hi = x_t + 1e-9 * geom.length, geom.length
            for _ in range(200):
                mid = 0.5 * (lo + hi)
                if exit_pressure(mid) > flow.back_pressure_ratio:
                    lo = mid
                else:
                    hi = mid
                if hi - lo < 1e-12 * geom.length:
                    break
            shock_x = 0.5 * (lo + hi)
            shock_mach, ptr = upstream(shock_x)
            behind = x > shock_x
            mach = np.where(behind, mach_from_area(np.maximum(area / (a_star / ptr), 1.0), g,
                                                   supersonic=False), mach)
            pt_local = np.where(behind, pt * ptr, pt)
    p = pt_local * pressure_ratio(mach, g)
    T = flow.total_temperature * temperature_ratio(mach, g)
    rho = p / (flow.Rgas * T)
    vx = mach * np.sqrt(g * flow.Rgas * T)
    return Quasi1DSolution(
        x=x, area=area, mach=mach, pressure=p, density=rho, temperature=T, velocity=vx,
        regime=regime, critical=crit, shock_x=shock_x, shock_mach=shock_mach,
        mass_flow=float(rho[0] * vx[0] * area[0]), exit_mach=float(mach[-1]),
        exit_pressure=float(p[-1] / pt))


# ----------------------------------------------------------------- projection
def _inverse_mass(ops) -> np.ndarray:
    minv = getattr(ops, "inv_mass", None)
    if minv is not None:
        return np.asarray(minv)
    phi = np.asarray(ops.ref.phi_vol)
    mass = np.einsum("iq,jq,eq->eij", phi, phi, np.asarray(ops.weighted_det))
    return np.linalg.inv(mass)


def project(values, ops) -> np.ndarray:
    """L2 projection of point values ``(E, nqv, 4)`` onto the basis (``initialize.py:41-46``)."""
    rhs = np.einsum("iq,eq,eqs->eis", np.asarray(ops.ref.phi_vol), np.asarray(ops.weighted_det),
                    np.asarray(values))
    return np.einsum("eij,ejs->eis", _inverse_mass(ops), rhs)


def uniform_state(flow: FlowConditions, mach: float = 0.95) -> np.ndarray:
    """Isentropic conserved state at one Mach number (``initialize.py:49-58``)."""
    g = flow.gamma
    fac = 1.0 + 0.5 * (g - 1.0) * mach * mach
    p = flow.total_pressure * fac ** (-g / (g - 1.0))
    rho = flow.stagnation_density * fac ** (-1.0 / (g - 1.0))
    vx = mach * np.sqrt(g * p / rho)
    return np.array([rho, rho * vx, 0.0, p / (g - 1.0) + 0.5 * rho * vx * vx])


def uniform_initial(ops, flow: FlowConditions, mach: float = 0.95, *, backend=None) -> np.ndarray:
    """``initialize.py:61-65``; with ``backend`` the projection runs on the device."""
    values = np.broadcast_to(uniform_state(flow, mach), (ops.n_elem, ops.ref.n_qvol, 4))
    if backend is not None:
        return backend.project(values)
    return project(values, ops)


_CONTOUR_IDS = {"bell": 0, "smooth": 0, "moc": 0, "conical": 1, "bezier": 2, "bezier9": 3,
                "analytic": 4}


def quasi1d_values_device(ops, flow: FlowConditions, geom, backend, *, solution=None):
    """Conserved point values ``(E, nqv, 4)`` of the quasi-1D initial state at
    the volume points, formed on the device (``nz_quasi1d_values``); returns a
    device tensor.  The quasi-1D solution itself (601 stations) is solved on
    the host, as it is once per design and independent of the mesh."""
    import ctypes as C
    import math

    import torch

    from . import _lib

    sol = solution if solution is not None else solve_quasi1d(geom, flow)
    dev = backend.device
    f64 = dict(dtype=torch.float64, device=dev)
    fields = getattr(ops, "device_fields", None) or {}
    if fields.get("xy_vol") is not None:
        xy = fields["xy_vol"]
    else:
        xy = torch.from_numpy(np.ascontiguousarray(np.asarray(ops.xy_vol), dtype=np.float64)).to(dev)
    st = [torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
          for a in (sol.x, sol.density, sol.velocity, sol.pressure)]
    prm = geom.params()
    d = _lib.Quasi1DDesc()
    d.contour, d.n_points = _CONTOUR_IDS[geom.contour], len(sol.x)
    d.x, d.density, d.velocity, d.pressure = (C.c_void_p(t.data_ptr()) for t in st)
    d.length = prm.get("length", 1.0)
    if geom.contour != "analytic":
        y_th = prm["throat_half_height"]
        y_ex = y_th * prm["area_ratio"]
        d.throat_x, d.y_throat, d.y_inlet, d.y_exit = prm["throat_x"], y_th, prm["inlet_half_height"], y_ex
        dx_ds = (1.0 - prm["throat_x"]) * d.length
        d.slope_throat = math.tan(prm["theta_initial"]) * dx_ds
        d.slope_exit = math.tan(prm["theta_exit"]) * dx_ds
        d.bezier_c1 = y_th + prm["bezier_w1"] * (y_ex - y_th)
        d.bezier_c2 = y_th + prm["bezier_w2"] * (y_ex - y_th)
        if geom.contour == "bezier9":
            ctrl = [y_th] + [prm[f"bz{i}"] for i in range(1, 9)] + [y_ex]
            for k in range(10):
                d.ctrl9[k] = ctrl[k]
                d.binom9[k] = float(math.comb(9, k))
    d.pi, d.e2m1 = math.pi, math.exp(2.0) - 1.0
    d.gamma, d.h = flow.gamma, 1e-6 * geom.length
    E, nqv = ops.n_elem, ops.ref.n_qvol
    out = torch.empty((E, nqv, 4), **f64)
    lib = _lib.load()
    rc = lib.nz_quasi1d_values(C.byref(d), C.c_void_p(xy.data_ptr()), E * nqv,
                               C.c_void_p(out.data_ptr()), backend._stream())
    if rc != 0:
        raise _lib.HipError(f"nz_quasi1d_values failed ({rc}): "
                            f"{lib.nz_initial_last_error().decode(errors='replace')}")
    return out


def quasi1d_initial(ops, flow: FlowConditions, geom, *, solution=None, backend=None,
                    as_tensor: bool = False):
    """Projected quasi-1D state with a wall-slope transverse velocity (``initialize.py:68-102``).

    With ``backend`` (a ``HipBackend`` on these operators) the point values
    are formed on the device (``nz_quasi1d_values``) and projected there
    (``nz_project``); ``as_tensor`` then returns the device tensor instead of
    a NumPy copy.  Without a backend both steps run here.
    """
    if backend is not None:
        u = backend.project(quasi1d_values_device(ops, flow, geom, backend, solution=solution))
        return u if as_tensor else u.cpu().numpy()
    return project(quasi1d_values(ops, flow, geom, solution=solution), ops)


def quasi1d_values(ops, flow: FlowConditions, geom, *, solution=None) -> np.ndarray:
    """Conserved point values ``(E, nqv, 4)`` of the quasi-1D initial state at
    the volume points, on the host (``initialize.py:80-100``)."""
    sol = solution if solution is not None else solve_quasi1d(geom, flow)
    xy = np.asarray(ops.xy_vol)
    x, y = xy[..., 0], xy[..., 1]
    rho = np.interp(x, sol.x, sol.density)
    vx = np.interp(x, sol.x, sol.velocity)
    p = np.interp(x, sol.x, sol.pressure)
    y_wall = np.asarray(geom.wall(x))
    h = 1e-6 * geom.length
    # one-sided at the inlet, central elsewhere
    slope = (np.asarray(geom.wall(x + h)) - np.asarray(geom.wall(np.maximum(x - h, 0.0)))) / (
        h + np.minimum(x, h))
    vy = vx * np.clip(y / np.maximum(y_wall, 1e-30), 0.0, 1.0) * slope
    rhoE = p / (flow.gamma - 1.0) + 0.5 * rho * (vx * vx + vy * vy)
    return np.stack([rho, rho * vx, rho * vy, rhoE], axis=-1)


def initial_state(ops, flow: FlowConditions, geom, kind: str = "quasi1d") -> np.ndarray:
    """``initialize.py:105-118``."""
    if kind == "quasi1d":
        return quasi1d_initial(ops, flow, geom)
    if kind == "uniform":
        return uniform_initial(ops, flow)
    raise ValueError(f"unknown initial_condition {kind!r}; use 'quasi1d' or 'uniform'")


def change_order(U, ops_from, ops_to, *, backend=None) -> np.ndarray:
    """Re-project a solution onto another order on the same mesh (``initialize.py:121-147``).

    With ``backend`` (a ``HipBackend`` on ``ops_to``) the projection runs on the device.
    """
    from .reference_element import shape_functions

    if ops_from.n_elem != ops_to.n_elem:
        raise ValueError(
            f"change_order needs the same mesh: {ops_from.n_elem} vs {ops_to.n_elem} elements")
    phi, _, _ = shape_functions(ops_from.ref.kind, ops_from.ref.order, ops_to.ref.points_vol)
    values = np.einsum("iq,eis->eqs", phi, np.asarray(U))
    if backend is not None:
        return backend.project(values)
    return project(values, ops_to)
