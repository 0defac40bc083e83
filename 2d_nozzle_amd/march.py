"""The host pseudo-time march around a backend, and a steady solve.

``march`` keeps the reference's control logic (``src/timestepping/march.py:125-230``):
one step alone for the initial residual, then chunks of ``check_interval``
steps handed to ``backend.run``, with the non-finite, tolerance, divergence
and stall checks in the same order and with the same messages.  Between
chunks the state stays on the device when the backend returns device arrays.
"""

from __future__ import annotations

import time
from collections.abc import Callable
from dataclasses import dataclass, field

import numpy as np

from .config import SolverOptions, residual_scale


@dataclass
class SolveHistory:
    iterations: list[int] = field(default_factory=list)
    residual: list[float] = field(default_factory=list)
    wall_time: float = 0.0

    def as_arrays(self):
        return np.asarray(self.iterations), np.asarray(self.residual)


@dataclass
class SolveResult:
    U: np.ndarray
    operators: object
    flow: object
    converged: bool
    iterations: int
    residual: float
    residual_initial: float
    history: SolveHistory
    backend: str
    residual_scale: float = 1.0
    message: str = ""
    min_density: float = float("nan")
    min_pressure: float = float("nan")
    limiter_activity: float = float("nan")
    mean_repairs: int = 0
    geometry: object = None

    @property
    def residual_scaled(self) -> float:
        return self.residual / self.residual_scale


def march(U, backend, opts: SolverOptions, scale: float, *, label: str = "",
          progress: Callable[[int, float], None] | None = None):
    """Advance to steady state; returns ``(U, history, converged, iters, res, res0, message)``."""
    history = SolveHistory()
    chunk = min(max(1, opts.check_interval), opts.max_iterations)
    next_print = opts.print_interval
    window: list[float] = []
    U, res0 = backend.run(U, 1, backend.scheme)
    res = res0
    done = 1
    history.iterations.append(done)
    history.residual.append(float(res0))
    converged = False
    message = ""
    t0 = time.perf_counter()
    if not np.isfinite(res0):
        history.wall_time = time.perf_counter() - t0
        return U, history, False, done, float(res0), float(res0), "initial residual is not finite"
    if res0 <= opts.tolerance * scale:
        history.wall_time = time.perf_counter() - t0
        return U, history, True, done, float(res0), float(res0), ""
    while done < opts.max_iterations:
        n = min(chunk, opts.max_iterations - done)
        U, res = backend.run(U, n, backend.scheme)
        done += n
        history.iterations.append(done)
        history.residual.append(float(res))
        if progress is not None:
            progress(done, float(res))
        elif opts.print_interval > 0 and done >= next_print:
            print(f"  {label}{done:>7d}   {res / scale:.6e}   ({res / res0:.3e} of initial)")
            next_print = done + opts.print_interval
        if not np.isfinite(res):
            message = (f"residual became non-finite after {done} iterations; reduce cfl "
                       "(try 0.5) or enable limiter='superbee'")
            break
        if res <= opts.tolerance * scale:
            converged = True
            break
        if res > opts.divergence_factor * res0:
            message = (f"residual grew by {res / res0:.1e}x after {done} iterations, which is "
                       f"past divergence_factor={opts.divergence_factor:g}; reduce cfl or "
                       "start from initial_condition='quasi1d'")
            break
        window.append(res)
        if len(window) > opts.stall_window:
            window.pop(0)
        if len(window) == opts.stall_window:
            best, worst = min(window), max(window)
            if worst > 0.0 and best / worst > opts.stall_ratio:
                active = backend.limiter_activity()
                span = opts.stall_window * chunk
                message = (f"the residual has stalled at {res / scale:.3e} (scaled; target "
                           f"{opts.tolerance:g}) -- it changed by less than "
                           f"{100 * (1 - opts.stall_ratio):.0f}% over the last {span} iterations")
                if np.isfinite(active) and active > 0.0:
                    message += (f", with the limiter active on ~{active:.1f} element(s) per call. "
                                "That is a limit cycle, not slow convergence")
                else:
                    message += ". Try a smaller cfl, scheme='ssprk3', or a finer mesh"
                break
    if not converged and not message:
        message = (f"hit max_iterations={opts.max_iterations} at scaled residual "
                   f"{res / scale:.2e} (target {opts.tolerance:g}); raise max_iterations or "
                   "relax tolerance")
    history.wall_time = time.perf_counter() - t0
    return U, history, converged, done, float(res), float(res0), message


def probe_minima(U: np.ndarray, ops, gamma: float) -> tuple[float, float]:
    """Minimum density and pressure over all probe points (``limiter.py:432-445``)."""
    rd = ops.ref
    ed = ops.topology.edges
    vol = np.einsum("iq,eis->eqs", rd.phi_vol, U)
    nface = ops.topology.n_faces
    fb = rd.phi_face[ed.face_side, np.arange(nface)[None, :]]  # (E, nface, N, nqf)
    face = np.einsum("efiq,eis->efqs", fb, U).reshape(U.shape[0], -1, 4)
    pr = np.concatenate([vol, face], axis=1)
    rho = pr[..., 0]
    p = (gamma - 1.0) * (pr[..., 3] - 0.5 * (pr[..., 1] ** 2 + pr[..., 2] ** 2)
                         / np.where(rho != 0.0, rho, 1.0))
    return float(rho.min()), float(p.min())


def solve_steady(ops, flow, opts: SolverOptions | None = None, *, U0, backend: str = "hip",
                 length: float = 1.0, progress=None, geometry=None) -> SolveResult:
    """Steady solve on a prepared operator set from the initial state ``U0``
    (``march.py:233-299``)."""
    opts = opts or SolverOptions()
    from . import get_backend

    bk = get_backend(backend, ops, flow, opts)
    U = np.asarray(U0)
    if U.shape != (ops.n_elem, ops.ref.n_basis, 4):
        raise ValueError(f"U0 has shape {U.shape}, expected {(ops.n_elem, ops.ref.n_basis, 4)}")
    U = bk.asarray(U)
    scale = residual_scale(flow, length)
    U, hist, conv, iters, res, res0, msg = march(U, bk, opts, scale, progress=progress)
    U_np = np.asarray(bk.to_numpy(U))
    rmin, pmin = probe_minima(U_np, ops, flow.gamma)
    if conv and (rmin <= 0.0 or pmin <= 0.0):
        conv = False
        msg = (f"residual converged but the solution is not physical (min rho {rmin:.3e}, "
               f"min p {pmin:.3e}); enable a limiter or refine the mesh")
    return SolveResult(
        U=U_np, operators=ops, flow=flow, converged=conv, iterations=iters, residual=res,
        residual_initial=res0, history=hist, backend=backend, residual_scale=scale, message=msg,
        min_density=rmin, min_pressure=pmin, limiter_activity=bk.limiter_activity(),
        mean_repairs=int(bk.n_mean_repaired), geometry=geometry,
    )
